// device_common.hpp -- device helpers shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace dmgx {

// Workgroup -> atom tile mapping that keeps consecutive tiles on one XCD:
// the dispatcher deals workgroups round-robin over the 8 XCDs, so workgroup
// b runs on XCD b%8; tile = (b%8)*per_xcd + b/8 gives each XCD one contiguous
// (spatially coherent) range of atoms and its own L2 working set.
__device__ __forceinline__ int xcd_tile(int nblocks)
{
  const int per = (nblocks + 7) >> 3;
  const int b = blockIdx.x;
  return (b & 7) * per + (b >> 3);
}

__host__ inline int xcd_grid(int nblocks) { return ((nblocks + 7) >> 3) << 3; }

// FP64 minimum image (box.cuh:77-123), no FMA contraction so that the
// predicate is bit-identical to the FP64 oracle.
__device__ __forceinline__ void mic(const DevBox& b, double& x, double& y, double& z)
{
#pragma clang fp contract(off)
  if (b.ortho) {
    if (b.pbc[0]) {
      const double L = b.h[0];
      if (x < -L * 0.5) x += L;
      else if (x > L * 0.5) x -= L;
    }
    if (b.pbc[1]) {
      const double L = b.h[4];
      if (y < -L * 0.5) y += L;
      else if (y > L * 0.5) y -= L;
    }
    if (b.pbc[2]) {
      const double L = b.h[8];
      if (z < -L * 0.5) z += L;
      else if (z > L * 0.5) z -= L;
    }
  } else {
    double sx = b.hi[0] * x + b.hi[1] * y + b.hi[2] * z;
    double sy = b.hi[3] * x + b.hi[4] * y + b.hi[5] * z;
    double sz = b.hi[6] * x + b.hi[7] * y + b.hi[8] * z;
    if (b.pbc[0]) sx -= rint(sx);
    if (b.pbc[1]) sy -= rint(sy);
    if (b.pbc[2]) sz -= rint(sz);
    x = b.h[0] * sx + b.h[1] * sy + b.h[2] * sz;
    y = b.h[3] * sx + b.h[4] * sy + b.h[5] * sz;
    z = b.h[6] * sx + b.h[7] * sy + b.h[8] * sz;
  }
}

__device__ __forceinline__ double dist2(double x, double y, double z)
{
#pragma clang fp contract(off)
  return (x * x + y * y) + z * z;
}

// cell coordinates of a position; returns false when a non-periodic
// coordinate lies outside the box
__device__ __forceinline__ bool cell_coords(const DevBox& b, double x, double y, double z, int* c)
{
  const double s[3] = {b.hi[0] * x + b.hi[1] * y + b.hi[2] * z, b.hi[3] * x + b.hi[4] * y + b.hi[5] * z,
                       b.hi[6] * x + b.hi[7] * y + b.hi[8] * z};
  bool ok = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double u = s[d];
    if (b.pbc[d]) {
      u -= floor(u);
    } else if (u < 0.0 || u >= 1.0) {
      ok = false;
      u = fmin(fmax(u, 0.0), 0.999999999);
    }
    int k = (int)(u * b.nc[d]);
    c[d] = k < 0 ? 0 : (k >= b.nc[d] ? b.nc[d] - 1 : k);
  }
  return ok;
}

// type index clamped into [0, T) (invalid caller types are flagged by the filter)
__device__ __forceinline__ int safe_type(int t, int T) { return (t < 0 || t >= T) ? 0 : t; }

}  // namespace dmgx
