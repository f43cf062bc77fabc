"""Pins of the FP64 oracle that come from the mathematics rather than from
reference outputs: forces are minus the gradient of the energy, descriptors
are rotation invariant, momentum is conserved, the per-atom virial sums to
the box virial, the angular constants equal their closed forms (Legendre
addition theorem, Wigner-3j couplings), and the ZBL pair term is the
textbook universal screened Coulomb potential."""
import os

import numpy as np
import pytest

from conftest import BASELINE, POTENTIALS
from oracle import OracleNEP, constant_tables

C_NEP = os.path.join(BASELINE, "inputs", "potentials", "nep_C.txt")
W_NEP = os.path.join(BASELINE, "inputs", "potentials", "nep_water.txt")
Z_NEP = os.path.join(BASELINE, "inputs", "potentials", "nep_BaTiO3_zbl.txt")
TC_NEP = os.path.join(POTENTIALS, "nep4_typecut_ch.txt")  # per-type cutoffs C 4.5/3.5, H 3.6/2.9


def cluster(n, L, seed, dmin=1.1):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, L, 3)
        if all(np.linalg.norm(((p - q) + L / 2) % L - L / 2) > dmin for q in pts):
            pts.append(p)
    return np.ascontiguousarray(np.array(pts).T)


@pytest.mark.parametrize("path,ntypes", [(C_NEP, 1), (W_NEP, 2), (Z_NEP, 3), (TC_NEP, 2)])
def test_force_is_minus_energy_gradient(path, ntypes):
    nep = OracleNEP(path)
    L = 15.0
    pos = cluster(18, L, 3, dmin=1.3)
    t = (np.arange(pos.shape[1]) % ntypes).astype(np.int32)
    h = np.diag([L, L, L])
    o = nep.compute(h, [1, 1, 1], t, pos)
    eps = 1e-5
    for i in (0, 5, 11):
        for c in range(3):
            p1 = pos.copy(); p1[c, i] += eps
            p2 = pos.copy(); p2[c, i] -= eps
            e1 = nep.compute(h, [1, 1, 1], t, p1)["pe"].sum()
            e2 = nep.compute(h, [1, 1, 1], t, p2)["pe"].sum()
            fd = -(e1 - e2) / (2 * eps)
            assert abs(fd - o["force"][c, i]) < 2e-6 * max(1.0, abs(fd)), (i, c, fd, o["force"][c, i])


def test_virial_is_strain_derivative():
    """sum of the per-atom virial = -dE/d(strain) (homogeneous deformation)."""
    nep = OracleNEP(C_NEP)
    L = 16.0
    pos = cluster(24, L, 7, dmin=1.3)
    t = np.zeros(pos.shape[1], np.int32)
    h = np.diag([L, L, L])
    W = nep.compute(h, [1, 1, 1], t, pos)["virial"].sum(1)
    eps = 1e-6
    # xx and xy components: W_ab = -dE/de_ab with x -> (I + e) x applied to box and positions
    for (a, b, k) in ((0, 0, 0), (0, 1, 3), (2, 1, 8)):
        E = []
        for s in (+1, -1):
            F = np.eye(3)
            F[a, b] += s * eps
            E.append(nep.compute(F @ h, [1, 1, 1], t, F @ pos)["pe"].sum())
        dEde = (E[0] - E[1]) / (2 * eps)
        assert abs(-dEde - W[k]) < 1e-5 * max(1.0, abs(dEde)), (a, b, -dEde, W[k])


def test_momentum_conservation_and_translation():
    nep = OracleNEP(W_NEP)
    L = 14.0
    pos = cluster(40, L, 9)
    t = (np.arange(pos.shape[1]) % 2).astype(np.int32)
    h = np.diag([L, L, L])
    o = nep.compute(h, [1, 1, 1], t, pos)
    assert np.abs(o["force"].sum(1)).max() < 1e-10
    o2 = nep.compute(h, [1, 1, 1], t, (pos + np.array([[0.3], [-1.7], [2.2]])))
    assert np.abs(o2["pe"] - o["pe"]).max() < 1e-10


def test_descriptor_rotation_and_permutation_invariance():
    nep = OracleNEP(C_NEP)
    pos = cluster(14, 10.0, 4, dmin=1.3) - 5.0
    t = np.zeros(pos.shape[1], np.int32)
    h = np.diag([80.0, 80.0, 80.0])  # isolated cluster in a large box
    q0 = nep.compute(h, [0, 0, 0], t, pos + 40, want_descriptor=True)["q"]
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    R = R @ np.array([[1, 0, 0], [0, np.cos(1.1), -np.sin(1.1)], [0, np.sin(1.1), np.cos(1.1)]])
    q1 = nep.compute(h, [0, 0, 0], t, R @ pos + 40, want_descriptor=True)["q"]
    assert np.abs(q1 - q0).max() < 1e-10 * max(1.0, np.abs(q0).max())
    perm = np.random.default_rng(1).permutation(pos.shape[1])
    q2 = nep.compute(h, [0, 0, 0], t, pos[:, perm] + 40, want_descriptor=True)["q"]
    assert np.abs(q2 - q0[perm]).max() < 1e-10 * max(1.0, np.abs(q0).max())


def test_angular_constants_closed_forms():
    sp = pytest.importorskip("sympy")
    from sympy.physics.wigner import wigner_3j
    c3, c4, c5, zc = constant_tables()
    z = sp.symbols("z")
    for L in range(1, 9):
        for m in range(L + 1):
            p = sp.Poly(sp.diff(sp.legendre(L, z), z, m), z)
            coeffs = [int(sp.Rational(c) / sp.gcd_list([sp.Rational(x) for x in p.all_coeffs() if x != 0]))
                      for c in p.all_coeffs()[::-1]]
            got = [int(v) for v in zc[L * (L + 1) // 2 + m][:len(coeffs)]]
            assert got == coeffs, (L, m)
    # q_L through the Legendre addition theorem: sum_jk g_j g_k P_L(u_j.u_k) (2L+1)/(4 pi)
    rng = np.random.default_rng(0)
    u = rng.normal(size=(6, 3))
    u /= np.linalg.norm(u, axis=1)[:, None]
    g = rng.normal(size=6)
    for L in range(1, 9):
        direct = sum(g[a] * g[b] * np.polynomial.legendre.legval(u[a] @ u[b], [0] * L + [1])
                     for a in range(6) for b in range(6)) * (2 * L + 1) / (4 * np.pi)
        s = np.zeros(2 * L + 1)
        for a in range(6):
            x, y, zz = u[a]
            for m in range(L + 1):
                Z = sum(zc[L * (L + 1) // 2 + m][k] * zz ** k for k in range(L - m + 1))
                P = (x + 1j * y) ** m
                if m == 0:
                    s[0] += g[a] * Z
                else:
                    s[2 * m - 1] += g[a] * Z * P.real
                    s[2 * m] += g[a] * Z * P.imag
        qL = sum(c3[L * L - 1 + k] * s[k] ** 2 for k in range(2 * L + 1))
        assert abs(qL - direct) < 1e-10 * max(1.0, abs(direct)), L
    # q_222 and q_1111 weights from the Wigner-3j coupling of complex moments
    rt14 = np.sqrt(14.0) / (448 * np.pi ** 1.5)
    assert np.allclose(c4, np.array([-5, -90, 45, 270, -540]) * rt14, rtol=1e-14)
    s1 = rng.normal(size=3)
    A1 = {0: np.sqrt(3 / (4 * np.pi)) * s1[0],
          1: -np.sqrt(3 / (8 * np.pi)) * (s1[1] + 1j * s1[2])}
    A1[-1] = -np.conj(A1[1])
    q = 0.0
    for Lc in (0, 1, 2):
        for m1 in (-1, 0, 1):
            for m2 in (-1, 0, 1):
                for m3 in (-1, 0, 1):
                    m4 = -m1 - m2 - m3
                    M = m1 + m2
                    if abs(m4) > 1 or abs(M) > Lc:
                        continue
                    w = float(wigner_3j(1, 1, Lc, m1, m2, -M)) * float(wigner_3j(Lc, 1, 1, M, m3, m4))
                    q += (-1) ** (Lc - M) * w * A1[m1] * A1[m2] * A1[m3] * A1[m4]
    assert abs(q.real - c5 * (s1 ** 2).sum() ** 2) < 1e-12 and abs(q.imag) < 1e-12


def test_zbl_pair_is_universal_screened_coulomb():
    """Two atoms well inside rc_inner: E = K Zi Zj phi(r/a)/r with the
    universal ZBL screening function, and F = -dE/dr."""
    nep = OracleNEP(Z_NEP)
    h = np.diag([30.0, 30.0, 30.0])
    r = 0.5
    pos = np.array([[10.0, 10.0 + r], [10.0, 10.0], [10.0, 10.0]])
    t = np.array([0, 1], np.int32)  # Ba, Ti
    o = nep.compute(h, [1, 1, 1], t, pos)
    # separate the ZBL term: the NEP part at the same geometry from a
    # no-ZBL copy of the model
    txt = open(Z_NEP).read().splitlines()
    import tempfile
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as fh:
        fh.write("\n".join([txt[0].replace("nep4_zbl", "nep4")] + txt[2:]) + "\n")
        plain = fh.name
    o0 = OracleNEP(plain).compute(h, [1, 1, 1], t, pos)
    os.unlink(plain)
    Zi, Zj = 56, 22
    a = 0.46850 / (Zi ** 0.23 + Zj ** 0.23)
    x = r / a
    phi = (0.18175 * np.exp(-3.1998 * x) + 0.50986 * np.exp(-0.94229 * x) +
           0.28022 * np.exp(-0.4029 * x) + 0.02817 * np.exp(-0.20162 * x))
    E = 14.399645 * Zi * Zj * phi / r
    ez = (o["pe"] - o0["pe"]).sum()
    assert abs(ez - E) < 2e-4 * E  # 0.46850 = 1/2.134563 to 5 digits
    fz = o["force"][0, 1] - o0["force"][0, 1]
    dr = 1e-6
    x1, x2 = (r + dr) / a, (r - dr) / a
    ph = lambda x: (0.18175 * np.exp(-3.1998 * x) + 0.50986 * np.exp(-0.94229 * x) +
                    0.28022 * np.exp(-0.4029 * x) + 0.02817 * np.exp(-0.20162 * x))
    dE = 14.399645 * Zi * Zj * (ph(x1) / (r + dr) - ph(x2) / (r - dr)) / (2 * dr)
    assert abs(fz + dE) < 2e-4 * abs(dE)


def test_seeded_models_parse_and_are_finite():
    for name in ("nep_carbon.txt", "nep_cu.txt", "nep_carbon_zbl.txt"):
        nep = OracleNEP(os.path.join(POTENTIALS, name))
        assert nep.version == 4


def test_seeded_models_are_scaled_for_their_configuration(workloads):
    """Each configuration's model has q_scaler = 1/range of its descriptor
    over a sample of that configuration, so on an independent sample of the
    same configuration every scaled component spans a range of order one
    (the carbon model serves its lattice and the serving stream's lattices
    with 30% vacancies; fit on the perfect lattice alone, the vacancy
    lattices' scaled components spanned up to ~48).  The amorphous-carbon
    model fit on diamond instead saw 4-/5-body invariants up to ~2,400x their
    range on the melt and dE/dq up to 6e6; FP32 rounding of the angular
    moments, amplified by those gradients, then moved forces by ~1e-3 eV/A."""
    cases = [("nep_carbon.txt", workloads.carbon_diamond_4096(seed=3)),
             ("nep_cu.txt", workloads.copper_fcc(n=6, seed=3)),
             ("nep_carbon.txt", workloads.vacancy_lattice("carbon", 5, np.random.default_rng(3))),
             ("nep_carbon_zbl.txt", workloads.amorphous_carbon(3000, seed=3))]
    for name, s in cases:
        o = OracleNEP(os.path.join(POTENTIALS, name)).compute(s["box"].h, s["box"].pbc, s["types"], s["pos"],
                                                              want_descriptor=True)
        spread = o["q"].max(0) - o["q"].min(0)
        assert spread.max() < 3.0, (name, spread)


def test_per_type_cutoffs_use_the_pair_mean():
    """`cutoff rcR_C rcA_C rcR_H rcA_H ...`: a C-H pair interacts within the
    mean of the two types' radii (nep.cu find_neighbor_list_large_box: pair
    cutoffs are type-pair averages).  An isolated C-H pair between the mean
    radial cutoff (4.05 A) and the larger one (4.5 A) leaves both atoms'
    energies exactly those of isolated atoms; just inside the mean it does not."""
    nep = OracleNEP(TC_NEP)
    L = 40.0
    h = np.diag([L, L, L])
    t = np.array([0, 1], np.int32)
    far = nep.compute(h, [1, 1, 1], t, np.array([[5.0, 25.0], [5.0, 5.0], [5.0, 5.0]]))["pe"]
    for d, same in ((4.2, True), (4.45, True), (3.95, False)):
        pe = nep.compute(h, [1, 1, 1], t, np.array([[5.0, 5.0 + d], [5.0, 5.0], [5.0, 5.0]]))["pe"]
        assert np.array_equal(pe, far) == same, (d, pe, far)


def embed_in_lmax4(path, out_path):
    """Writes the 3-body order 4 model with a 4-body block equal to the
    model at `path` (order 2 or 3, with or without the 4-body block): every
    descriptor component keeps its ANN weights and scaler in the order-4
    layout, added components (L above the model's order, a missing 4-body
    block) get zero input weights, and the 4-/5-body blocks sit behind L = 4
    in that order."""
    lines = open(path).read().split("\n")
    head, vals = lines[:6], np.array([float(x) for x in lines[6:] if x.strip()])
    T = int(head[0].split()[1])
    nR, nA = [int(x) + 1 for x in head[2].split()[1:3]]
    kR, kA = [int(x) + 1 for x in head[3].split()[1:3]]
    L, f4, f5 = [int(x) for x in head[4].split()[1:4]]
    H = int(head[5].split()[1])
    q4, q5 = int(f4 != 0), int(f5 != 0)
    D = nR + nA * (L + q4 + q5)
    D4 = nR + nA * (4 + 1 + q5)
    blocks = list(range(L)) + ([4] if q4 else []) + ([5] if q5 else [])
    col = list(range(nR)) + [nR + b4 * nA + n for b4 in blocks for n in range(nA)]
    out, p = [], 0
    for _ in range(T):
        w0 = vals[p:p + H * D].reshape(H, D)
        p += H * D
        w = np.zeros((H, D4))
        w[:, col] = w0
        out += [w.reshape(-1), vals[p:p + 2 * H]]
        p += 2 * H
    out.append(vals[p:p + 1])
    p += 1
    nd = T * T * (nR * kR + nA * kA)
    out.append(vals[p:p + nd])
    p += nd
    qs = np.ones(D4)
    qs[col] = vals[p:p + D]
    out += [qs, vals[p + D:]]
    head[4] = f"l_max 4 2 {1 if q5 else 0}"
    with open(out_path, "w") as fh:
        fh.write("\n".join(head) + "\n" + "\n".join(f"{v:.9e}" for v in np.concatenate(out)) + "\n")


@pytest.mark.parametrize("name", ["nep4_lmax2_ch.txt", "nep4_lmax3_ch.txt"])
def test_low_3body_orders_equal_their_order4_embedding(name, tmp_path):
    """3-body orders 2 and 3 (nep_utilities.cuh:1335-1420: components for
    L <= l_max, then the 4-/5-body invariants that are switched on), and a
    5-body block without the 4-body one: the reference's CPU evaluator fixes
    the order at 4 and takes a single extra block for the 4-body one
    (nep_cpu.cpp:883,1016-1022), so these are pinned by equivalence: the same
    model written at order 4 with both blocks, zero weights on the added
    components (checked against that evaluator itself where it is built),
    has the same energies, forces and virials."""
    from test_oracle_golden import coverage_sample
    path = os.path.join(POTENTIALS, name)
    emb = str(tmp_path / "emb.txt")
    embed_in_lmax4(path, emb)
    h, pos, t = coverage_sample()
    a = OracleNEP(path).compute(h, [1, 1, 1], t, pos)
    b = OracleNEP(emb).compute(h, [1, 1, 1], t, pos)
    for k in ("pe", "force", "virial"):
        scale = max(1.0, np.abs(b[k]).max())
        assert np.abs(a[k] - b[k]).max() <= 1e-12 * scale, (k, np.abs(a[k] - b[k]).max())
    ref = pytest.importorskip("ref")
    if ref.available():
        r = ref.RefNEP(emb).compute(h, [1, 1, 1], t, pos)
        for k in ("pe", "force", "virial"):
            scale = max(1.0, np.abs(r[k]).max())
            assert np.abs(b[k] - r[k]).max() <= 1e-12 * scale, (k, np.abs(b[k] - r[k]).max())


L6_NEP = os.path.join(POTENTIALS, "nep4_lmax6_ch.txt")


def split_high_orders(path, zeroed_path, l4_path):
    """From a model of 3-body order L > 4 (with its 4-/5-body blocks) writes
    (a) the same model with zero ANN input weights on the components of
    orders 5..L, and (b) the order-4 model without those components (every
    other weight and scaler unchanged).  Components are ordered q_1..q_L per
    radial function, then the 4-body and 5-body blocks (nep.cu:297-306,
    nep_utilities.cuh:1583-1730)."""
    lines = open(path).read().split("\n")
    head, vals = lines[:6], np.array([float(x) for x in lines[6:] if x.strip()])
    T = int(head[0].split()[1])
    nR, nA = [int(x) + 1 for x in head[2].split()[1:3]]
    kR, kA = [int(x) + 1 for x in head[3].split()[1:3]]
    L, f4, f5 = [int(x) for x in head[4].split()[1:4]]
    H = int(head[5].split()[1])
    q4, q5 = int(f4 != 0), int(f5 != 0)
    D = nR + nA * (L + q4 + q5)
    high = [nR + b * nA + n for b in range(4, L) for n in range(nA)]
    keep = [d for d in range(D) if d not in high]
    za, zb, p = [], [], 0
    for _ in range(T):
        w0 = vals[p:p + H * D].reshape(H, D).copy()
        p += H * D
        w0[:, high] = 0.0
        za += [w0.reshape(-1), vals[p:p + 2 * H]]
        zb += [w0[:, keep].reshape(-1), vals[p:p + 2 * H]]
        p += 2 * H
    nd = T * T * (nR * kR + nA * kA)
    rest = vals[p:p + 1 + nd]
    qs = vals[p + 1 + nd:p + 1 + nd + D]
    za += [rest, qs, vals[p + 1 + nd + D:]]
    zb += [rest, qs[keep], vals[p + 1 + nd + D:]]
    for out, arrs, lmax in ((zeroed_path, za, head[4]), (l4_path, zb, f"l_max 4 {f4} {f5}")):
        h = list(head)
        h[4] = lmax
        with open(out, "w") as fh:
            fh.write("\n".join(h) + "\n" + "\n".join(f"{v:.9e}" for v in np.concatenate(arrs)) + "\n")


def test_order6_model_layout_forces_and_invariance(tmp_path):
    """3-body order 6 (no reference CPU evaluator goes past 4): the
    components of orders 5 and 6 sit between q_4 and the 4-body block (with
    zero weights on them the model equals the order-4 model without them),
    forces are minus the energy gradient, and the descriptor is rotation
    invariant.  The order-5..8 angular constants are pinned by the Legendre
    addition theorem (test_angular_constants_closed_forms)."""
    from test_oracle_golden import coverage_sample
    za, zb = str(tmp_path / "zeroed.txt"), str(tmp_path / "l4.txt")
    split_high_orders(L6_NEP, za, zb)
    h, pos, t = coverage_sample()
    a = OracleNEP(za).compute(h, [1, 1, 1], t, pos)
    b = OracleNEP(zb).compute(h, [1, 1, 1], t, pos)
    for k in ("pe", "force", "virial"):
        scale = max(1.0, np.abs(b[k]).max())
        assert np.abs(a[k] - b[k]).max() <= 1e-12 * scale, k
    full = OracleNEP(L6_NEP).compute(h, [1, 1, 1], t, pos)
    assert np.abs(full["pe"] - b["pe"]).max() > 1e-6  # the order-5/6 components do contribute
    nep = OracleNEP(L6_NEP)
    Lb = 15.0
    p0 = cluster(18, Lb, 5, dmin=1.3)
    t0 = (np.arange(p0.shape[1]) % 2).astype(np.int32)
    hb = np.diag([Lb, Lb, Lb])
    o = nep.compute(hb, [1, 1, 1], t0, p0)
    eps = 1e-5
    for i in (0, 7, 13):
        for c in range(3):
            p1 = p0.copy(); p1[c, i] += eps
            p2 = p0.copy(); p2[c, i] -= eps
            fd = -(nep.compute(hb, [1, 1, 1], t0, p1)["pe"].sum() - nep.compute(hb, [1, 1, 1], t0, p2)["pe"].sum()) / (2 * eps)
            assert abs(fd - o["force"][c, i]) < 2e-6 * max(1.0, abs(fd)), (i, c, fd, o["force"][c, i])
    pc = cluster(14, 10.0, 4, dmin=1.3) - 5.0
    tc = (np.arange(pc.shape[1]) % 2).astype(np.int32)
    hc = np.diag([80.0, 80.0, 80.0])
    q0 = nep.compute(hc, [0, 0, 0], tc, pc + 40, want_descriptor=True)["q"]
    R = np.array([[0.6, -0.8, 0.0], [0.8, 0.6, 0.0], [0.0, 0.0, 1.0]]) @ \
        np.array([[1.0, 0.0, 0.0], [0.0, np.cos(1.1), -np.sin(1.1)], [0.0, np.sin(1.1), np.cos(1.1)]])
    q1 = nep.compute(hc, [0, 0, 0], tc, R @ pc + 40, want_descriptor=True)["q"]
    assert np.abs(q1 - q0).max() < 1e-10 * max(1.0, np.abs(q0).max())
