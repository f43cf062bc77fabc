// neighbor.hip -- Verlet-list build for gfx950 with a device-side rebuild
// decision.
//
// Every kernel of the rebuild is launched every step and returns at once
// unless the device flag flags[kFlagRebuild] is set (by the displacement
// check or by the host for a first call / box change).  The host therefore
// never has to read the decision back: a whole MD chunk is enqueued without a
// synchronisation.  Behaviour follows the reference's Neighbor class
// (neighbor.cu:90-134 cell list, :141-218 rows, :446-582 skin policy,
// neighbor.cuh:133-196 row order): rows hold every slot j != i with
// FP64 d^2 < (rc + skin)^2, sorted ascending by slot (or by global id), plus
// the position of the centre inside each neighbour's row.
#include "device_common.hpp"

namespace dmgx {

namespace {

__global__ void k_set_flag(int* __restrict__ flag, int value) { *flag = value; }

__global__ void k_zero_if(const int* __restrict__ go, int* __restrict__ buf, int n)
{
  if (!*go) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) buf[i] = 0;
}

__global__ void k_cell_count(const int* __restrict__ go, DevBox b, int N, const double* __restrict__ pos,
                             int* __restrict__ cell_of, int* __restrict__ count, int* __restrict__ flags)
{
  if (!*go) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int c[3];
  if (!cell_coords(b, pos[i], pos[N + i], pos[2 * N + i], c)) atomicOr(&flags[kFlagOutOfBox], 1);
  const int id = c[0] + b.nc[0] * (c[1] + b.nc[1] * c[2]);
  cell_of[i] = id;
  atomicAdd(&count[id], 1);
}

// three-phase exclusive scan, 1024 counts per workgroup
__global__ void k_scan_local(const int* __restrict__ go, int n, const int* __restrict__ in, int* __restrict__ out,
                             int* __restrict__ bsum)
{
  if (!*go) return;
  __shared__ int sh[256];
  const int base = blockIdx.x * 1024 + threadIdx.x * 4;
  int v[4], sum = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    sum += v[k];
  }
  sh[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int run = sh[threadIdx.x] - sum;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) bsum[blockIdx.x] = sh[255];
}

__global__ void k_scan_blocks(const int* __restrict__ go, int nb, int* __restrict__ bsum, int* __restrict__ total)
{
  if (!*go) return;
  __shared__ int sh[1024];
  const int chunk = (nb + 1023) / 1024;
  const int b0 = threadIdx.x * chunk;
  int sum = 0;
  for (int k = 0; k < chunk; ++k)
    if (b0 + k < nb) sum += bsum[b0 + k];
  sh[threadIdx.x] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  int run = sh[threadIdx.x] - sum;
  for (int k = 0; k < chunk; ++k)
    if (b0 + k < nb) {
      const int v = bsum[b0 + k];
      bsum[b0 + k] = run;
      run += v;
    }
  if (threadIdx.x == 1023) *total = sh[1023];
}

__global__ void k_scan_add(const int* __restrict__ go, int n, int* __restrict__ out, const int* __restrict__ bsum)
{
  if (!*go) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += bsum[i >> 10];
}

__global__ void k_cell_fill(const int* __restrict__ go, int N, const int* __restrict__ cell_of,
                            const int* __restrict__ start, int* __restrict__ fill, int* __restrict__ atoms)
{
  if (!*go) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int c = cell_of[i];
  atoms[start[c] + atomicAdd(&fill[c], 1)] = i;
}

// One thread per centre: scan the 5x5x5 cells around it (cell edge >= (rc+skin)/2)
// and keep every slot with FP64 d2 < (rc+skin)^2.  Rows past the capacity MV
// are counted but not written; the count is reported through the flags.
__global__ __launch_bounds__(kBlock) void k_verlet_build(const int* __restrict__ go, DevBox b, int N, int c1,
                                                         int c2, const double* __restrict__ pos, double rv2,
                                                         DevWork w)
{
  if (!*go) return;
  const int i = c1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c2) return;
  const double xi = pos[i], yi = pos[N + i], zi = pos[2 * N + i];
  int c[3];
  cell_coords(b, xi, yi, zi, c);
  int cnt = 0;
  for (int dz = -2; dz <= 2; ++dz) {
    int cz = c[2] + dz;
    if (b.pbc[2]) cz = (cz + b.nc[2]) % b.nc[2];
    else if (cz < 0 || cz >= b.nc[2]) continue;
    for (int dy = -2; dy <= 2; ++dy) {
      int cy = c[1] + dy;
      if (b.pbc[1]) cy = (cy + b.nc[1]) % b.nc[1];
      else if (cy < 0 || cy >= b.nc[1]) continue;
      for (int dx = -2; dx <= 2; ++dx) {
        int cx = c[0] + dx;
        if (b.pbc[0]) cx = (cx + b.nc[0]) % b.nc[0];
        else if (cx < 0 || cx >= b.nc[0]) continue;
        const int cell = cx + b.nc[0] * (cy + b.nc[1] * cz);
        const int e = w.cell_start[cell + 1];
        for (int k = w.cell_start[cell]; k < e; ++k) {
          const int j = w.cell_atoms[k];
          if (j == i) continue;
          double x = pos[j] - xi, y = pos[N + j] - yi, z = pos[2 * N + j] - zi;
          mic(b, x, y, z);
          if (dist2(x, y, z) < rv2) {
            if (cnt < w.MV) w.vl[(size_t)cnt * N + i] = j;
            ++cnt;
          }
        }
      }
    }
  }
  w.vn[i] = cnt;
  if (cnt > w.MV) atomicMax(&w.flags[kFlagVerletOverflow], cnt);
}

__device__ __forceinline__ bool key_less(const uint64_t* gid, int a, int b)
{
  if (gid) {
    const uint64_t ga = gid[a], gb = gid[b];
    if (ga != gb) return ga < gb;
  }
  return a < b;
}

// One wavefront per row: rank sort in LDS (keys unique: slot, or (gid, slot)).
__global__ __launch_bounds__(256) void k_verlet_sort(const int* __restrict__ go, int N, int c1, int c2,
                                                     const uint64_t* __restrict__ gid, DevWork w)
{
  if (!*go) return;
  extern __shared__ int sh[];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = c1 + blockIdx.x * 4 + wave;
  int* row = sh + wave * w.MV;
  const int n = i < c2 ? w.vn[i] : 0;
  const int nn = n < w.MV ? n : w.MV;
  for (int k = lane; k < nn; k += 64) row[k] = w.vl[(size_t)k * N + i];
  __syncthreads();
  for (int k = lane; k < nn; k += 64) {
    const int v = row[k];
    int rank = 0;
    for (int t = 0; t < nn; ++t) rank += key_less(gid, row[t], v) ? 1 : 0;
    w.vl[(size_t)rank * N + i] = v;
  }
}

// Position of the centre inside each neighbour's row (binary search on the
// row key); -1 when the neighbour has no row (outside the centre range).
__global__ __launch_bounds__(kBlock) void k_verlet_reverse(const int* __restrict__ go, int N, int c1, int c2,
                                                           const uint64_t* __restrict__ gid, DevWork w)
{
  if (!*go) return;
  const int i = c1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= c2) return;
  const int n = min(w.vn[i], w.MV);
  for (int p = 0; p < n; ++p) {
    const int j = w.vl[(size_t)p * N + i];
    int q = -1;
    if (j >= c1 && j < c2) {
      int lo = 0, hi = min(w.vn[j], w.MV) - 1;
      while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const int v = w.vl[(size_t)mid * N + j];
        if (v == i) {
          q = mid;
          break;
        }
        if (key_less(gid, v, i)) lo = mid + 1;
        else hi = mid - 1;
      }
    }
    w.vrev[(size_t)p * N + i] = q;
  }
}

__global__ void k_save_positions(const int* __restrict__ go, int N, const double* __restrict__ pos,
                                 double* __restrict__ x0)
{
  if (!*go) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * N) x0[i] = pos[i];
}

// closes a rebuild: clears the request and counts it
__global__ void k_rebuild_done(int* __restrict__ flags)
{
  if (flags[kFlagRebuild]) {
    flags[kFlagRebuild] = 0;
    flags[kFlagRebuildCount] += 1;
  }
}

// any atom displaced by more than skin/2 since the last rebuild requests one
// (neighbor.cu:446-490); the reduction is an OR over the wavefront
__global__ void k_check_moved(DevBox b, int N, const double* __restrict__ pos, const double* __restrict__ x0,
                              double lim2, int* __restrict__ flags)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool moved = false;
  if (i < N) {
    double x = pos[i] - x0[i], y = pos[N + i] - x0[N + i], z = pos[2 * N + i] - x0[2 * N + i];
    mic(b, x, y, z);
    moved = !(dist2(x, y, z) <= lim2);  // NaN displacements request a rebuild too
  }
  if (__any(moved) && (threadIdx.x & 63) == 0) atomicOr(&flags[kFlagRebuild], 1);
}

inline int nblk(int n) { return (n + kBlock - 1) / kBlock; }

}  // namespace

hipError_t launch_request_rebuild(DevWork& w, hipStream_t s)
{
  k_set_flag<<<1, 1, 0, s>>>(w.flags + kFlagRebuild, 1);
  return hipGetLastError();
}

hipError_t launch_check_moved(const DevBox& b, int N, const double* pos, DevWork& w, hipStream_t s)
{
  const double half = 0.5 * kSkin;
  if (N > 0) k_check_moved<<<nblk(N), kBlock, 0, s>>>(b, N, pos, w.x0, half * half, w.flags);
  return hipGetLastError();
}

hipError_t launch_rebuild(const DevBox& b, int N, int c1, int c2, const double* pos, double rv2,
                          const uint64_t* gid, DevWork& w, hipStream_t s)
{
  const int* go = w.flags + kFlagRebuild;
  const int ncell = b.nc[0] * b.nc[1] * b.nc[2];
  const int nb = (ncell + 1023) / 1024;
  k_zero_if<<<(ncell + 255) / 256, 256, 0, s>>>(go, w.cell_count, ncell);
  if (N > 0) k_cell_count<<<nblk(N), kBlock, 0, s>>>(go, b, N, pos, w.cell_of, w.cell_count, w.flags);
  // block sums use cell_fill as scratch (nb <= ncell), cleared again before the fill
  k_scan_local<<<nb, 256, 0, s>>>(go, ncell, w.cell_count, w.cell_start, w.cell_fill);
  k_scan_blocks<<<1, 1024, 0, s>>>(go, nb, w.cell_fill, w.cell_start + ncell);
  k_scan_add<<<(ncell + 255) / 256, 256, 0, s>>>(go, ncell, w.cell_start, w.cell_fill);
  k_zero_if<<<(ncell + 255) / 256, 256, 0, s>>>(go, w.cell_fill, ncell);
  if (N > 0) k_cell_fill<<<nblk(N), kBlock, 0, s>>>(go, N, w.cell_of, w.cell_start, w.cell_fill, w.cell_atoms);
  if (c2 > c1) {
    k_verlet_build<<<nblk(c2 - c1), kBlock, 0, s>>>(go, b, N, c1, c2, pos, rv2, w);
    k_verlet_sort<<<(c2 - c1 + 3) / 4, 256, sizeof(int) * 4 * (size_t)w.MV, s>>>(go, N, c1, c2, gid, w);
    k_verlet_reverse<<<nblk(c2 - c1), kBlock, 0, s>>>(go, N, c1, c2, gid, w);
  }
  if (N > 0) k_save_positions<<<nblk(3 * N), kBlock, 0, s>>>(go, N, pos, w.x0);
  k_rebuild_done<<<1, 1, 0, s>>>(w.flags);
  return hipGetLastError();
}

}  // namespace dmgx
