// neighbor.hip -- Verlet-list build for gfx950 with a device-side rebuild
// decision.
//
// Every kernel of the rebuild is launched every step and returns at once
// unless the device flag flags[kFlagRebuild] is set (by the displacement
// check in k_pack or by the host for a first call / box change).  The host
// therefore never has to read the decision back: a whole MD chunk is enqueued
// without a synchronisation.  Behaviour follows the reference's Neighbor
// class (neighbor.cu:90-134 cell list, :141-218 rows, :446-582 skin policy,
// neighbor.cuh:133-196 row order): rows hold every slot j != i with FP64
// d^2 < (rc + skin)^2, sorted ascending by slot (or by global id), plus the
// position of the centre inside each neighbour's row.
//
// Layout: rows are row-major (vl[i*MV + p]) so that a group of lanes that
// works on one centre reads and writes contiguous memory.  Build, sort and
// reverse search run with kRowLanes lanes per centre (4 centres per wave).
#include <algorithm>
#include "device_common.hpp"

namespace dmgx {

namespace {

constexpr int kRowLanes = 16;  // lanes cooperating on one Verlet row
constexpr int kScanOneBlock = 262144;  // larger cell grids are scanned by many workgroups

__global__ void k_set_flag(int* __restrict__ flag, int value) { *flag = value; }

__global__ void k_clear_moved(const int* __restrict__ go, ImageMap m)
{
  if (!*go) return;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < m.n_req) m.req_moved[r] = 0;
}

__global__ void k_clear_ints(int* __restrict__ p, long long n)
{
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
    p[k] = 0;
}

// Per-step packing of positions and types into one 32-byte record per atom
// (x, y, z, type bits) so that every neighbour gather is a single 32-byte
// access; also validates the types and, unless a rebuild is forced, tests
// whether any atom moved more than skin/2 since the last rebuild
// (neighbor.cu:446-490).
__global__ __launch_bounds__(kBlock) void k_pack(DevBox b, int N, int T, const double* __restrict__ pos,
                                                 const int* __restrict__ type, const double* __restrict__ x0,
                                                 double lim2, int forced, double4* __restrict__ xq,
                                                 int* __restrict__ flags, ImageMap m)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (forced && i == 0) flags[kFlagRebuild] = kRebuildAsked | kRebuildAll;
  bool moved = false, bad = false;
  if (i < N) {
    const double x = pos[i], y = pos[N + i], z = pos[2 * N + i];
    const int t = type[i];
    bad = t < 0 || t >= T;
    xq[i] = make_double4(x, y, z, __longlong_as_double((long long)t));
    if (!forced) {
      double dx = x - x0[i], dy = y - x0[N + i], dz = z - x0[2 * N + i];
      mic(b, dx, dy, dz);
      moved = !(dist2(dx, dy, dz) <= lim2);  // NaN displacements request a rebuild too
      if (moved && m.slot_req) m.req_moved[m.slot_req[i < m.n_owned ? i : m.owner[i]]] = 1;
    }
  }
  const bool lead = (threadIdx.x & 63) == 0;
  if (__any(moved) && lead) atomicOr(&flags[kFlagRebuild], kRebuildAsked);
  if (__any(bad) && lead) atomicOr(&flags[kFlagBadType], 1);
}

// cell of every atom, its rank inside the cell, and the positions the next
// displacement checks compare against (for the slots in [a1, b1) and
// [a2, b2) whose rows this rebuild makes)
// With compacted lists (a rebuild of some requests of a batch) only the
// listed slots enter the cell list: every row it makes holds slots of its
// own request only.
__global__ void k_cell_count(const int* __restrict__ go, DevBox b, int N, const double4* __restrict__ xq,
                             int* __restrict__ cell_of, int* __restrict__ rank, int* __restrict__ count,
                             double* __restrict__ x0, int a1, int b1, int a2, int b2, int* __restrict__ flags,
                             ImageMap m)
{
  const int g = *go;
  if (!g) return;
  const bool listed = rebuild_listed(m);
  const int n = listed ? m.req_off[2 * m.n_req + 1] : N;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int i = listed ? m.lists[m.n_owned + k] : k;
    const double4 p = xq[i];
    int c[3];
    if (!cell_coords(b, p.x, p.y, p.z, c)) atomicOr(&flags[kFlagOutOfBox], 1);
    const int id = c[0] + b.nc[0] * (c[1] + b.nc[1] * c[2]);
    cell_of[i] = id;
    rank[i] = atomicAdd(&count[id], 1);
    if (((i >= a1 && i < b1) || (i >= a2 && i < b2)) && (listed || slot_rebuilt(m, g, i))) {
      x0[i] = p.x;
      x0[N + i] = p.y;
      x0[2 * N + i] = p.z;
    }
  }
}

// Exclusive scan of the cell counts by one workgroup, which also clears the
// counts for the next rebuild; start[ncell] = total.  Each pass covers 4
// sub-tiles of 4096 cells, 4 consecutive cells per thread in each (one
// 16-byte access per thread, contiguous across the wavefront), scanned
// together; the next pass's counts are loaded before this pass is scanned.
__global__ __launch_bounds__(1024) void k_scan_cells(const int* __restrict__ go, int ncell,
                                                     int* __restrict__ count, int* __restrict__ start)
{
  if (!*go) return;
  constexpr int kSub = 4096, kChunks = 4, kTile = kSub * kChunks;
  __shared__ int wsum[kChunks][16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  auto load = [&](int base, int4* v) {
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
      const int k = base + c * kSub + threadIdx.x * 4;
      if (k + 3 < ncell) {
        v[c] = *reinterpret_cast<const int4*>(count + k);
      } else {
        v[c].x = k < ncell ? count[k] : 0;
        v[c].y = k + 1 < ncell ? count[k + 1] : 0;
        v[c].z = k + 2 < ncell ? count[k + 2] : 0;
        v[c].w = 0;
      }
    }
  };
  int4 v[kChunks], nv[kChunks];
  load(0, v);
  int carry = 0;
  for (int base = 0; base < ncell; base += kTile) {
    if (base + kTile < ncell) load(base + kTile, nv);
    int s[kChunks], x[kChunks];
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
      s[c] = v[c].x + v[c].y + v[c].z + v[c].w;
      x[c] = s[c];
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
      for (int c = 0; c < kChunks; ++c) {
        const int y = __shfl_up(x[c], off, 64);
        if (lane >= off) x[c] += y;
      }
    }
    if (lane == 63) {
#pragma unroll
      for (int c = 0; c < kChunks; ++c) wsum[c][wave] = x[c];
    }
    __syncthreads();
    int run_base = carry;
#pragma unroll
    for (int c = 0; c < kChunks; ++c) {
      int before = 0, total = 0;
      for (int q = 0; q < 16; ++q) {
        const int t = wsum[c][q];
        if (q < wave) before += t;
        total += t;
      }
      const int k = base + c * kSub + threadIdx.x * 4;
      int4 o;
      o.x = run_base + before + x[c] - s[c];
      o.y = o.x + v[c].x;
      o.z = o.y + v[c].y;
      o.w = o.z + v[c].z;
      if (k + 3 < ncell) {
        *reinterpret_cast<int4*>(start + k) = o;
        *reinterpret_cast<int4*>(count + k) = make_int4(0, 0, 0, 0);
      } else {
        if (k < ncell) start[k] = o.x, count[k] = 0;
        if (k + 1 < ncell) start[k + 1] = o.y, count[k + 1] = 0;
        if (k + 2 < ncell) start[k + 2] = o.z, count[k + 2] = 0;
      }
      run_base += total;
    }
    carry = run_base;
#pragma unroll
    for (int c = 0; c < kChunks; ++c) v[c] = nv[c];
    __syncthreads();
  }
  if (threadIdx.x == 0) start[ncell] = carry;
}

// Multi-block form of the same scan for large grids: each workgroup scans
// one tile of 4096 cells (and clears its counts) and stores the tile total;
// k_scan_cells then scans the tile totals into offsets and k_scan_offsets
// adds each tile's offset.
__global__ __launch_bounds__(1024) void k_scan_tiles(const int* __restrict__ go, int ncell, int* __restrict__ count,
                                                     int* __restrict__ start, int* __restrict__ tile_sum)
{
  if (!*go) return;
  __shared__ int wsum[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int k0 = blockIdx.x * 4096 + threadIdx.x * 4;
  int v[4], s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = k0 + k < ncell ? count[k0 + k] : 0;
    s += v[k];
  }
  int x = s;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  int before = 0, total = 0;
  for (int q = 0; q < 16; ++q) {
    if (q < wave) before += wsum[q];
    total += wsum[q];
  }
  int run = before + x - s;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k0 + k < ncell) {
      start[k0 + k] = run;
      count[k0 + k] = 0;
    }
    run += v[k];
  }
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

__global__ void k_scan_offsets(const int* __restrict__ go, int ncell, const int* __restrict__ tile_off,
                               int* __restrict__ start)
{
  if (!*go) return;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < ncell) start[k] += tile_off[k / 4096];
  if (k == 0) start[ncell] = tile_off[(ncell + 4095) / 4096];
}

__global__ void k_cell_fill(const int* __restrict__ go, int N, const int* __restrict__ cell_of,
                            const int* __restrict__ rank, const int* __restrict__ start,
                            const double4* __restrict__ xq, int* __restrict__ atoms, double4* __restrict__ xs,
                            ImageMap m)
{
  if (!*go) return;
  const bool listed = rebuild_listed(m);
  const int n = listed ? m.req_off[2 * m.n_req + 1] : N;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const int i = listed ? m.lists[m.n_owned + k] : k;
    const int at = start[cell_of[i]] + rank[i];
    atoms[at] = i;
    xs[at] = xq[i];
  }
}

// Offsets of each request's centres and slots in the compacted lists of the
// requests this rebuild remakes (exclusive scans over the requests, one
// workgroup), then the totals.
__global__ __launch_bounds__(1024) void k_list_scan(const int* __restrict__ go, ImageMap m)
{
  const int g = *go;
  if (!g) return;
  __shared__ int wsum[2][16];
  __shared__ int carry[2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (threadIdx.x == 0) carry[0] = carry[1] = 0;
  __syncthreads();
  for (int base = 0; base < m.n_req; base += blockDim.x) {
    const int r = base + threadIdx.x;
    int c = 0, s = 0;
    if (r < m.n_req && ((g & kRebuildAll) || m.req_moved[r])) {
      const int4 q = m.req_range[r];
      c = q.y;
      s = q.y + q.w;
    }
    int ic = c, is = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int vc = __shfl_up(ic, d), vs = __shfl_up(is, d);
      if (lane >= d) {
        ic += vc;
        is += vs;
      }
    }
    if (lane == 63) {
      wsum[0][wave] = ic;
      wsum[1][wave] = is;
    }
    __syncthreads();
    int pc = carry[0], ps = carry[1];
    for (int v = 0; v < wave; ++v) {
      pc += wsum[0][v];
      ps += wsum[1][v];
    }
    if (r < m.n_req) {
      m.req_off[2 * r] = pc + ic - c;
      m.req_off[2 * r + 1] = ps + is - s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int v = 0; v < nw; ++v) {
        carry[0] += wsum[0][v];
        carry[1] += wsum[1][v];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    m.req_off[2 * m.n_req] = carry[0];
    m.req_off[2 * m.n_req + 1] = carry[1];
  }
}

// The compacted lists: one workgroup per request
__global__ void k_list_fill(const int* __restrict__ go, ImageMap m)
{
  const int g = *go;
  if (!g) return;
  const int r = blockIdx.x;
  if (!((g & kRebuildAll) || m.req_moved[r])) return;
  const int4 q = m.req_range[r];
  int* cl = m.lists + m.req_off[2 * r];
  int* sl = m.lists + m.n_owned + m.req_off[2 * r + 1];
  for (int k = threadIdx.x; k < q.y; k += blockDim.x) {
    cl[k] = q.x + k;
    sl[k] = q.x + k;
  }
  for (int k = threadIdx.x; k < q.w; k += blockDim.x) sl[q.y + k] = q.z + k;
}

// Row key order: slot, or (global id, slot).
__device__ __forceinline__ bool key_less(const uint64_t* gid, int a, int b)
{
  if (gid) {
    const uint64_t ga = gid[a], gb = gid[b];
    if (ga != gb) return ga < gb;
  }
  return a < b;
}

// Candidate ranges of one centre: the 5x5 rows of cells around it (cell edge
// >= (rc+skin)/2, so +-2 cells cover rc+skin), each a contiguous range of
// cell_atoms, split in two where a periodic row wraps: at most 50 ranges.
constexpr int kRanges = 64;
constexpr int kRangesPerLane = kRanges / kRowLanes;
constexpr size_t kRangeLds = 2 * kRanges * sizeof(int);  // starts + prefix, per centre

// LDS bytes of one centre's row buffers: 64-bit keys (rows keyed by global
// id only), the accepted slots and the sorted row; a multiple of 8
__host__ __device__ inline size_t build_row_lds(int mv, bool keyed)
{
  return (keyed ? 16 : 8) * (size_t)mv;
}

// One kRowLanes group per centre.  The group first reads the bounds of all
// its ranges at once and prefix-sums their lengths, then streams the
// candidates as one list: candidate k is slot rs[q] + k - rp[q] of
// cell_atoms / xs (the cell-ordered copy of the positions), so a candidate
// costs two independent contiguous loads; each step tests two candidates
// per lane whose loads were issued a step earlier.  Accepted slots are compacted into LDS
// with a ballot, rank-sorted there on (key, slot) and written out as one
// contiguous row.  Rows past the capacity MV are counted but not stored;
// the count is reported through the flags.
template <bool LISTED>
__global__ __launch_bounds__(256) void k_verlet_build(const int* __restrict__ go, DevBox b, int c1, int c2,
                                                      int groups, const double4* __restrict__ xq, double rv2,
                                                      const uint64_t* __restrict__ gid, DevWork w)
{
  const int go_word = *go;
  if (!go_word) return;
  extern __shared__ uint64_t shk[];
  const int g = threadIdx.x / kRowLanes, lg = threadIdx.x % kRowLanes;
  const int gw = g % (64 / kRowLanes);  // group inside the wavefront
  // per centre: [keys (with gid only) | row | sorted row], then the ranges
  const size_t per = build_row_lds(w.MV, gid != nullptr) / sizeof(uint64_t);
  uint64_t* key = shk + (size_t)g * per;
  int* row = reinterpret_cast<int*>(gid ? key + w.MV : key);
  int* sorted = row + w.MV;
  int* rs = reinterpret_cast<int*>(shk + (size_t)groups * per) + g * 2 * kRanges;
  int* rp = rs + kRanges;
  // centres c1 + k, or with compacted lists the listed centres (of the
  // requests this rebuild remakes) inside [c1, c2): a pass per group of
  // `groups` centres, the grid striding over the groups
  const int ncent = LISTED ? w.map.req_off[2 * w.map.n_req] : c2 - c1;
  for (int pass = 0;; ++pass) {
  const int base = (blockIdx.x + pass * gridDim.x) * groups;
  if (base >= ncent) break;
  const int kc = base + g;
  const int i = LISTED ? (kc < ncent ? w.map.lists[kc] : -1) : c1 + kc;
  const bool active = LISTED ? (i >= c1 && i < c2) : (i < c2 && slot_rebuilt(w.map, go_word, i));
  if (!__syncthreads_or(active)) continue;  // no centre of this pass is rebuilt
  double4 pi = make_double4(0, 0, 0, 0);
  int c[3] = {0, 0, 0};
  if (active) {
    pi = xq[i];
    cell_coords(b, pi.x, pi.y, pi.z, c);
  }
  // ranges q = 2 * (5 (dz+2) + dy+2) + part, lane lg holding q in
  // [lg*kRangesPerLane, (lg+1)*kRangesPerLane)
  int len[kRangesPerLane];
  int tot = 0;
#pragma unroll
  for (int u = 0; u < kRangesPerLane; ++u) {
    const int q = lg * kRangesPerLane + u;
    int s = 0, e = 0;
    if (active && q < 50) {
      const int r = q >> 1, part = q & 1;
      int cz = c[2] + r / 5 - 2, cy = c[1] + r % 5 - 2;
      bool ok = true;
      if (b.pbc[2]) cz = (cz + b.nc[2]) % b.nc[2];
      else ok = ok && cz >= 0 && cz < b.nc[2];
      if (b.pbc[1]) cy = (cy + b.nc[1]) % b.nc[1];
      else ok = ok && cy >= 0 && cy < b.nc[1];
      int lo0 = c[0] - 2, hi0 = c[0] + 2, lo1 = 0, hi1 = -1;
      if (b.pbc[0]) {
        if (lo0 < 0) {
          lo1 = lo0 + b.nc[0];
          hi1 = b.nc[0] - 1;
          lo0 = 0;
        } else if (hi0 >= b.nc[0]) {
          lo1 = 0;
          hi1 = hi0 - b.nc[0];
          hi0 = b.nc[0] - 1;
        }
      } else {
        lo0 = max(lo0, 0);
        hi0 = min(hi0, b.nc[0] - 1);
      }
      const int lo = part ? lo1 : lo0, hi = part ? hi1 : hi0;
      if (ok && hi >= lo) {
        const int base = b.nc[0] * (cy + b.nc[1] * cz);
        s = w.cell_start[base + lo];
        e = w.cell_start[base + hi + 1];
      }
    }
    rs[q] = s;
    len[u] = e - s;
    tot += e - s;
  }
  int incl = tot;
#pragma unroll
  for (int d = 1; d < kRowLanes; d <<= 1) {
    const int v = __shfl_up(incl, d, kRowLanes);
    if (lg >= d) incl += v;
  }
  int off = incl - tot;
#pragma unroll
  for (int u = 0; u < kRangesPerLane; ++u) {
    rp[lg * kRangesPerLane + u] = off;
    off += len[u];
  }
  const int total = __shfl(incl, kRowLanes - 1, kRowLanes);
  __syncthreads();
  // k < total lies in range q with rp[q] <= k < rp[q+1] (q < 50, and
  // rp[50] = total); a lane's k only grows, so q only moves forward
  int q = 0;
  auto slot_of = [&](int k) {
    while (rp[q + 1] <= k) ++q;
    return rs[q] + (k - rp[q]);
  };
  // each step tests candidates k and k + kRowLanes of every lane, whose
  // loads were issued in the step before
  constexpr int kStep = 2 * kRowLanes;
  int jn[2] = {0, 0};
  double4 pn[2] = {make_double4(0, 0, 0, 0), make_double4(0, 0, 0, 0)};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int k = lg + h * kRowLanes;
    if (k < total) {
      const int sl = slot_of(k);
      jn[h] = w.cell_atoms[sl];
      pn[h] = w.xs[sl];
    }
  }
  int found = 0;
  for (int k0 = 0; __any(k0 < total); k0 += kStep) {
    int jc[2];
    double4 pc[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      jc[h] = jn[h];
      pc[h] = pn[h];
      const int k = k0 + kStep + lg + h * kRowLanes;
      if (k < total) {
        const int sl = slot_of(k);
        jn[h] = w.cell_atoms[sl];
        pn[h] = w.xs[sl];
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int k = k0 + lg + h * kRowLanes;
      const int j = jc[h];
      bool take = false;
      if (k < total && j != i) {
        double x = pc[h].x - pi.x, y = pc[h].y - pi.y, z = pc[h].z - pi.z;
        mic(b, x, y, z);
        take = dist2(x, y, z) < rv2;
      }
      const unsigned gm = (unsigned)(__ballot(take) >> (gw * kRowLanes)) & ((1u << kRowLanes) - 1u);
      if (take) {
        const int at = found + __popc(gm & ((1u << lg) - 1u));
        if (at < w.MV) {
          row[at] = j;
          if (gid) key[at] = gid[j];
        }
      }
      found += __popc(gm);
    }
  }
  __syncthreads();
  const int n = min(found, w.MV);
  if (gid) {
    for (int k = lg; k < n; k += kRowLanes) {
      const uint64_t kv = key[k];
      const int v = row[k];
      int r = 0;
      for (int t = 0; t < n; ++t) {
        const uint64_t kt = key[t];
        r += (kt < kv || (kt == kv && row[t] < v)) ? 1 : 0;
      }
      sorted[r] = v;
    }
  } else {
    // four entries per LDS read (row is 16-byte aligned: MV is a multiple of 4)
    const int4* row4 = reinterpret_cast<const int4*>(row);
    const int n4 = n >> 2;
    for (int k = lg; k < n; k += kRowLanes) {
      const int v = row[k];
      int r = 0;
      for (int t = 0; t < n4; ++t) {
        const int4 x = row4[t];
        r += (x.x < v) + (x.y < v) + (x.z < v) + (x.w < v);
      }
      for (int t = n4 * 4; t < n; ++t) r += row[t] < v ? 1 : 0;
      sorted[r] = v;
    }
  }
  __syncthreads();
  if (active) {
    for (int k = lg; k < n; k += kRowLanes) w.vl[(size_t)i * w.MV + k] = sorted[k];
    if (lg == 0) {
      w.vn[i] = found;
      if (found > w.MV) atomicMax(&w.flags[kFlagVerletOverflow], found);
    }
  }
  __syncthreads();  // the next pass reuses the LDS
  }
}

// Position of the centre inside each neighbour's row (binary search on the
// row key); -1 when the neighbour has no row (outside the centre range).
// Rows are symmetric (the same FP64 test from both ends), so when both ends
// are centres and no image map is given, only the lower slot searches: it
// finds q (i in j's row) and writes the reverse entries of both directed
// edges, i -> j and j -> i; every entry is still written exactly once.  A
// lane searches kRevBatch edges of its centre in lockstep.
constexpr int kRevBatch = 4;

__global__ __launch_bounds__(256) void k_verlet_reverse(const int* __restrict__ go, int c1, int c2,
                                                        const uint64_t* __restrict__ gid, DevWork w)
{
  const int go_word = *go;
  if (!go_word) return;
  const bool listed = rebuild_listed(w.map);
  const int ncent = listed ? w.map.req_off[2 * w.map.n_req] : c2 - c1;
  const int lg = threadIdx.x % kRowLanes;
  for (int kc = (blockIdx.x * blockDim.x + threadIdx.x) / kRowLanes; kc < ncent;
       kc += gridDim.x * (blockDim.x / kRowLanes)) {
  const int i = listed ? w.map.lists[kc] : c1 + kc;
  if (i < c1 || i >= c2 || (!listed && !slot_rebuilt(w.map, go_word, i))) continue;
  const int n = min(w.vn[i], w.MV);
  const int* ri = w.vl + (size_t)i * w.MV;
  const ImageMap& im = w.map;
  for (int p0 = lg; p0 < n; p0 += kRevBatch * kRowLanes) {
    int t[kRevBatch], lo[kRevBatch], hi[kRevBatch], q[kRevBatch], ov[kRevBatch];
    bool own[kRevBatch];
    const int* ro[kRevBatch];
#pragma unroll
    for (int u = 0; u < kRevBatch; ++u) {
      const int p = p0 + u * kRowLanes;
      // the reverse of i -> j is found in the row of j; with an image map, j
      // stands for its owner o = owner[j] shifted by S, and the reverse entry
      // is o -> (the image of i shifted by -S), one of i's own ghost slots
      int o = -1, tt = i;
      if (p < n) {
        const int j = ri[p];
        o = j;
        if (im.owner) {
          o = im.owner[j];
          const int c = im.image[j];
          if (c != image_code(0, 0, 0)) {
            // i's ghosts are stored in increasing image code
            const int want = image_neg(c);
            int lo = im.gstart[i], hi = im.gstart[i + 1] - 1;
            tt = -1;
            while (lo <= hi) {
              const int mid = (lo + hi) >> 1;
              const int v = im.image[mid];
              if (v == want) {
                tt = mid;
                break;
              }
              if (v < want) lo = mid + 1;
              else hi = mid - 1;
            }
          }
        }
      }
      const bool centre = o >= c1 && o < c2;
      const bool mine = im.owner || !centre || o > i;
      own[u] = p < n && mine;
      const bool search = own[u] && tt >= 0 && centre;
      t[u] = tt;
      ov[u] = o;
      q[u] = -1;
      lo[u] = 0;
      hi[u] = search ? min(w.vn[o], w.MV) - 1 : -1;
      ro[u] = search ? w.vl + (size_t)o * w.MV : ri;  // an in-bounds row for idle probes
    }
    bool more = true;
    while (more) {
      int v[kRevBatch];
#pragma unroll
      for (int u = 0; u < kRevBatch; ++u) v[u] = ro[u][lo[u] <= hi[u] ? (lo[u] + hi[u]) >> 1 : 0];
      more = false;
#pragma unroll
      for (int u = 0; u < kRevBatch; ++u) {
        if (lo[u] <= hi[u]) {
          const int mid = (lo[u] + hi[u]) >> 1;
          if (v[u] == t[u]) {
            q[u] = mid;
            lo[u] = hi[u] + 1;
          } else if (key_less(gid, v[u], t[u])) {
            lo[u] = mid + 1;
          } else {
            hi[u] = mid - 1;
          }
        }
        more = more || lo[u] <= hi[u];
      }
    }
#pragma unroll
    for (int u = 0; u < kRevBatch; ++u) {
      const int p = p0 + u * kRowLanes;
      if (own[u]) {
        w.vrev[(size_t)i * w.MV + p] = q[u];
        if (!im.owner && q[u] >= 0) w.vrev[(size_t)ov[u] * w.MV + q[u]] = p;
      }
    }
  }
  }
}

inline int nblk(int n) { return (n + kBlock - 1) / kBlock; }
// workgroups of 256 lanes of a rebuild kernel, which strides over its
// centres or slots: 2,048 of them fill 256 CUs at 32 waves each, and a
// launch that finds no rebuild asked returns in about a microsecond
constexpr int kListBlocks = 2048;

// LDS of one centre of the build: its row buffers and candidate ranges
size_t build_lds_per_centre(int mv, bool keyed) { return build_row_lds(mv, keyed) + kRangeLds; }

// centres per build workgroup: 16 groups of kRowLanes lanes unless the LDS
// would exceed 64 KiB
int build_groups(int mv, bool keyed)
{
  const size_t per = build_lds_per_centre(mv, keyed);
  int groups = 256 / kRowLanes;
  while (groups > 1 && per * groups > 64 * 1024) groups >>= 1;
  return groups;
}

}  // namespace

hipError_t launch_request_rebuild(DevWork& w, hipStream_t s)
{
  k_set_flag<<<1, 1, 0, s>>>(w.flags + kFlagRebuild, kRebuildAsked | kRebuildAll);
  return hipGetLastError();
}

hipError_t launch_clear_ints(int* p, long long n, hipStream_t s)
{
  if (n > 0) k_clear_ints<<<(int)std::min<long long>((n + kBlock - 1) / kBlock, 1024), kBlock, 0, s>>>(p, n);
  return hipGetLastError();
}

hipError_t launch_pack(const DevBox& b, int N, int T, const double* pos, const int* type, bool forced, DevWork& w,
                       hipStream_t s, double half_skin)
{
  const double half = half_skin > 0.0 ? half_skin : 0.5 * kSkin;
  if (N > 0)
    k_pack<<<nblk(N), kBlock, 0, s>>>(b, N, T, pos, type, w.x0, half * half, forced ? 1 : 0, w.xq, w.flags, w.map);
  else if (forced)
    k_set_flag<<<1, 1, 0, s>>>(w.flags + kFlagRebuild, kRebuildAsked | kRebuildAll);
  return hipGetLastError();
}

hipError_t launch_rebuild(const DevBox& b, int N, int c1, int c2, double rv2, const uint64_t* gid, DevWork& w,
                          hipStream_t s, const int* x0_ranges, bool capped)
{
  const int* go = w.flags + kFlagRebuild;
  const int ncell = b.nc[0] * b.nc[1] * b.nc[2];
  const int r[4] = {x0_ranges ? x0_ranges[0] : 0, x0_ranges ? x0_ranges[1] : N, x0_ranges ? x0_ranges[2] : 0,
                    x0_ranges ? x0_ranges[3] : 0};
  // the kernels stride over the centres and slots (with compacted lists,
  // the listed ones, counted on the device), so the grids may be capped
  const bool listed = rebuild_listed(w.map);
  if (listed) {
    k_list_scan<<<1, 1024, 0, s>>>(go, w.map);
    k_list_fill<<<w.map.n_req, kBlock, 0, s>>>(go, w.map);
  }
  const bool cap = capped || listed;
  const int slot_blocks = cap ? std::min(nblk(N), kListBlocks) : nblk(N);
  if (N > 0)
    k_cell_count<<<slot_blocks, kBlock, 0, s>>>(go, b, N, w.xq, w.cell_of, w.cell_rank, w.cell_count, w.x0, r[0],
                                                r[1], r[2], r[3], w.flags, w.map);
  if (ncell <= kScanOneBlock) {
    k_scan_cells<<<1, 1024, 0, s>>>(go, ncell, w.cell_count, w.cell_start);
  } else {
    const int tiles = (ncell + 4095) / 4096;
    int* tile_sum = w.cell_tiles;
    int* tile_off = w.cell_tiles + ((tiles + 4) & ~3);  // 16-byte aligned for the scan's vector stores
    k_scan_tiles<<<tiles, 1024, 0, s>>>(go, ncell, w.cell_count, w.cell_start, tile_sum);
    k_scan_cells<<<1, 1024, 0, s>>>(go, tiles, tile_sum, tile_off);
    k_scan_offsets<<<nblk(ncell), kBlock, 0, s>>>(go, ncell, tile_off, w.cell_start);
  }
  if (N > 0)
    k_cell_fill<<<slot_blocks, kBlock, 0, s>>>(go, N, w.cell_of, w.cell_rank, w.cell_start, w.xq, w.cell_atoms, w.xs,
                                               w.map);
  if (c2 > c1) {
    const bool keyed = gid != nullptr;
    const int groups = build_groups(w.MV, keyed);
    constexpr int rows_per_block = 256 / kRowLanes;
    // build workgroups hold groups * kRowLanes lanes
    int nb = (c2 - c1 + groups - 1) / groups;
    int nr = (c2 - c1 + rows_per_block - 1) / rows_per_block;
    if (cap) {
      nb = std::min(nb, kListBlocks * (rows_per_block / groups));
      nr = std::min(nr, kListBlocks);
    }
    const size_t lds = build_lds_per_centre(w.MV, keyed) * groups;
    if (listed) k_verlet_build<true><<<nb, groups * kRowLanes, lds, s>>>(go, b, c1, c2, groups, w.xq, rv2, gid, w);
    else k_verlet_build<false><<<nb, groups * kRowLanes, lds, s>>>(go, b, c1, c2, groups, w.xq, rv2, gid, w);
    k_verlet_reverse<<<nr, 256, 0, s>>>(go, c1, c2, gid, w);
  }
  // the request marks of this rebuild are spent
  if (w.map.slot_req && w.map.n_req > 0) k_clear_moved<<<nblk(w.map.n_req), kBlock, 0, s>>>(go, w.map);
  return hipGetLastError();
}

}  // namespace dmgx
