"""Shared helpers of the GPU tests: run the HIP library through the C ABI on
device arrays and bring results back to the host."""
from __future__ import annotations

import numpy as np


def to_dev(a, dtype):
    import torch
    return torch.tensor(np.ascontiguousarray(a).reshape(-1), dtype=dtype, device="cuda")


def gpu_compute(dmgx, nep, box, types, pos, **kw):
    """Runs nep.compute_domain (defaults: every atom a centre); returns
    (pe[N], force[3,N], virial[9,N]) as numpy."""
    import torch
    N = len(types)
    t = to_dev(types, torch.int32)
    x = to_dev(pos, torch.float64)
    pe = torch.zeros(N, dtype=torch.float64, device="cuda")
    f = torch.zeros(3 * N, dtype=torch.float64, device="cuda")
    v = torch.zeros(9 * N, dtype=torch.float64, device="cuda")
    nep.compute_domain(box, N, t, x, pe, f, v, **kw)
    torch.cuda.synchronize()
    return pe.cpu().numpy(), f.cpu().numpy().reshape(3, N), v.cpu().numpy().reshape(9, N)


def errors(gpu, ref, n=None):
    """max |dF|, |dE_total|/N, max |dE_atom|, max |dW| over the first n atoms."""
    pe, f, v = gpu
    sl = slice(0, n)
    N = pe[sl].shape[0]
    return dict(force=float(np.abs(f[:, sl] - ref["force"][:, sl]).max()),
                energy_per_atom=float(abs(pe[sl].sum() - ref["pe"][sl].sum()) / N),
                energy_atom_max=float(np.abs(pe[sl] - ref["pe"][sl]).max()),
                virial=float(np.abs(v[:, sl] - ref["virial"][:, sl]).max()))
