// kernels.hip -- HIP kernels of the NEP molecular-dynamics step for gfx950.
//
// Design (CDNA4, wave64):
//  * Verlet list with a 1 A skin, rebuilt only when an atom moved more than
//    skin/2 (neighbor.cu:446-582); rows hold slot indices, sorted ascending
//    (or by global id), with a precomputed reverse position per entry.
//  * Every step a filter pass (16 lanes per centre, ballot compaction) turns
//    each Verlet row into one compact edge row: angular edges first (slots
//    0..na-1, ascending), radial-only edges from the end of the row
//    downward, each edge a float4 (FP64 minimum-image r_ij rounded to float,
//    neighbour slot).  Later kernels stream these rows instead of
//    re-gathering positions.  All rows are row-major ([i*M + slot]).
//  * One fused per-atom kernel does descriptor -> ANN -> adjoints -> angular
//    partial forces: the angular moments s[n][lm] stay in registers, are
//    turned in place into the adjoint dE/ds, and each angular edge's
//    dE_i/dr_ij is written at i's slot.
//  * The gather kernel (8 lanes per atom) forms the radial pair force from
//    both centres' basis-space adjoints, adds the angular partials of both
//    directions (dE_j/dr_ji read from j's row through the reverse index) and
//    ZBL, and writes FP64 force/virial.  No floating-point atomics anywhere:
//    results are deterministic and independent of launch order.
//  * Integrator and thermo in FP64, thermo sums compensated (two-sum).
#include "device_common.hpp"

#include <hip/hip_ext.h>

#include <cmath>

namespace dmgx {

__constant__ float c_c3[(kMaxL + 1) * (kMaxL + 1) - 1];
__constant__ float c_zc[kMaxL + 1][kMaxL + 1][kMaxL + 1];
__constant__ float c_c4[5];
__constant__ float c_c5;

void upload_angular_constants(const AngularConstants& a)
{
  (void)hipMemcpyToSymbol(HIP_SYMBOL(c_c3), a.c3, sizeof(a.c3));
  (void)hipMemcpyToSymbol(HIP_SYMBOL(c_zc), a.zc, sizeof(a.zc));
  (void)hipMemcpyToSymbol(HIP_SYMBOL(c_c4), a.c4, sizeof(a.c4));
  (void)hipMemcpyToSymbol(HIP_SYMBOL(c_c5), &a.c5, sizeof(a.c5));
}

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

// cosine cutoff (nep_utilities.cuh:317-342).  cos and sin of pi d/rc come
// from the hardware sine/cosine (v_sin_f32 / v_cos_f32), whose argument is in
// revolutions: 0.5 d/rc lies in [0, 0.5), far inside their [-256, 256]
// domain, with absolute errors near 1e-7.
__device__ __forceinline__ void cutoff_fc(float rc, float rcinv, float d, float& fc)
{
  fc = d < rc ? 0.5f * __builtin_amdgcn_cosf(0.5f * d * rcinv) + 0.5f : 0.0f;
}
__device__ __forceinline__ void cutoff_fc_fcp(float rc, float rcinv, float d, float& fc, float& fcp)
{
  if (d < rc) {
    const float rev = 0.5f * d * rcinv;
    const float sn = __builtin_amdgcn_sinf(rev), cs = __builtin_amdgcn_cosf(rev);
    fc = 0.5f * cs + 0.5f;
    fcp = -1.57079633f * sn * rcinv;
  } else {
    fc = 0.0f;
    fcp = 0.0f;
  }
}

// Chebyshev radial basis f_k = (T_k(x)+1)/2 fc, x = 2(d/rc-1)^2-1; entries
// k >= kn are zero
template <int K>
__device__ __forceinline__ void cheb(int kn, float rcinv, float d, float fc, float* f)
{
  const float t = d * rcinv - 1.0f;
  const float x = 2.0f * t * t - 1.0f;
  const float hfc = 0.5f * fc;
  float tm2 = 1.0f, tm1 = x;
  f[0] = fc;
  if (K > 1) f[1] = (x + 1.0f) * hfc;
#pragma unroll
  for (int k = 2; k < K; ++k) {
    const float tk = 2.0f * x * tm1 - tm2;
    tm2 = tm1;
    tm1 = tk;
    f[k] = k < kn ? (tk + 1.0f) * hfc : 0.0f;
  }
}

// basis and d/dd (T_k' = k U_{k-1})
template <int K>
__device__ __forceinline__ void cheb_d(int kn, float rcinv, float d, float fc, float fcp, float* f, float* fp)
{
  const float t = d * rcinv - 1.0f;
  const float x = 2.0f * t * t - 1.0f;
  const float dx = 4.0f * t * rcinv;  // dx/dd
  f[0] = fc;
  fp[0] = fcp;
  float tm2 = 1.0f, tm1 = x, um2 = 1.0f, um1 = 2.0f * x;
  if (K > 1) {
    const float v = 0.5f * (x + 1.0f);
    f[1] = v * fc;
    fp[1] = 0.5f * dx * fc + v * fcp;
  }
#pragma unroll
  for (int k = 2; k < K; ++k) {
    const float tk = 2.0f * x * tm1 - tm2;
    const float v = 0.5f * (tk + 1.0f);
    const float der = 0.5f * k * um1 * dx;  // (T_k'/2) dx/dd, T_k' = k U_{k-1}
    f[k] = k < kn ? v * fc : 0.0f;
    fp[k] = k < kn ? der * fc + v * fcp : 0.0f;
    tm2 = tm1;
    tm1 = tk;
    const float uk = 2.0f * x * um1 - um2;
    um2 = um1;
    um1 = uk;
  }
}

// real solid harmonics Y_{L,m}(u), L=1..LM, packed at L*L-1 (+2m-1 re, +2m im)
template <int LM>
__device__ __forceinline__ void harmonics(float ux, float uy, float uz, float* Y)
{
  float zp[LM + 1];
  zp[0] = 1.0f;
#pragma unroll
  for (int k = 1; k <= LM; ++k) zp[k] = zp[k - 1] * uz;
  float pr[LM + 1], pim[LM + 1];
  pr[0] = 1.0f;
  pim[0] = 0.0f;
#pragma unroll
  for (int k = 1; k <= LM; ++k) {
    pr[k] = pr[k - 1] * ux - pim[k - 1] * uy;
    pim[k] = pr[k - 1] * uy + pim[k - 1] * ux;
  }
#pragma unroll
  for (int L = 1; L <= LM; ++L) {
#pragma unroll
    for (int m = 0; m <= L; ++m) {
      float Z = 0.0f;
#pragma unroll
      for (int k = (L - m) & 1; k <= L - m; k += 2) Z += c_zc[L][m][k] * zp[k];
      const int b = L * L - 1;
      if (m == 0) {
        Y[b] = Z;
      } else {
        Y[b + 2 * m - 1] = Z * pr[m];
        Y[b + 2 * m] = Z * pim[m];
      }
    }
  }
}

// Screened-Coulomb ZBL pair energy f(d) and df/dd with the cosine switch
// between r_in and r_out (nep_utilities.cuh:347-433); zp is one pair's
// {1/a, Zi Zj K_C, r_in, r_out, (c_k, d_k) x4} record.
__device__ __forceinline__ void zbl_pair(const float* __restrict__ zp, float d, float dinv, float& f, float& fp)
{
  const float x = d * zp[0];
  float phi = 0.0f, dphi = 0.0f;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float t = zp[4 + 2 * k] * expf(-zp[5 + 2 * k] * x);
    phi += t;
    dphi -= zp[5 + 2 * k] * t;
  }
  const float vv = zp[1] * phi * dinv;
  const float vp = zp[1] * dphi * zp[0] * dinv - vv * dinv;
  const float r1 = zp[2], r2 = zp[3];
  float fc = 0.0f, fcp = 0.0f;
  if (d < r1) {
    fc = 1.0f;
  } else if (d < r2) {
    const float wv = 3.14159265f / (r2 - r1);
    float sn, cs;
    sincosf(wv * (d - r1), &sn, &cs);
    fc = 0.5f * cs + 0.5f;
    fcp = -0.5f * wv * sn;
  }
  f = vv * fc;
  fp = vp * fc + vv * fcp;
}

// Loads the first kp floats of a 16-byte aligned row as float4s into out[0..KP)
// (kp <= KP, both multiples of 4); entries past kp are left as they are.
template <int KP>
__device__ __forceinline__ void load_row4(const float* __restrict__ row, int kp, float* out)
{
  const float4* q = reinterpret_cast<const float4*>(row);
#pragma unroll
  for (int c = 0; c < KP / 4; ++c) {
    if (4 * c < kp) {
      const float4 t = q[c];
      out[4 * c] = t.x;
      out[4 * c + 1] = t.y;
      out[4 * c + 2] = t.z;
      out[4 * c + 3] = t.w;
    }
  }
}

// ---------------------------------------------------------------------------
// per-step typewise filter (nep.cu:503-561) into compact edge rows
// ---------------------------------------------------------------------------

// kFilterLanes lanes per centre walk its Verlet row kFilterLanes entries per
// round; the kept entries of each round are compacted with a wavefront
// ballot, so every store of the kernel is contiguous inside the row.  The
// loads of kFilterBatch rounds (row entries, their reverse positions, then
// the neighbours' packed positions) are issued before any of them is used,
// so a wave waits for two memory latencies per batch instead of two per
// round.  Types were validated when the step's positions were packed; here
// they are only clamped into range.  Pair cutoffs are FP64
// ((rc_i + rc_j) / 2)^2, as the reference's typewise test (nep.cu:503-561).
#ifndef DMGX_FILTER_LANES
#define DMGX_FILTER_LANES 16
#endif
#ifndef DMGX_FILTER_BATCH
#define DMGX_FILTER_BATCH 7
#endif
#ifndef DMGX_GATHER_LANES
#define DMGX_GATHER_LANES 8
#endif
#ifndef DMGX_GATHER_BATCH
#define DMGX_GATHER_BATCH 2
#endif
// the fused descriptor kernel's workgroup (256: three workgroups per CU by
// LDS; 512, which would allow four wavefronts per SIMD, measured no faster)
#ifndef DMGX_DESC_BLOCK
#define DMGX_DESC_BLOCK 256
#endif
constexpr int kDescBlock = DMGX_DESC_BLOCK;
constexpr int kFilterLanes = DMGX_FILTER_LANES;
constexpr int kFilterBatch = DMGX_FILTER_BATCH;
constexpr int kGatherLanes = DMGX_GATHER_LANES;
constexpr int kGatherBatch = DMGX_GATHER_BATCH;  // radial edges per lane per batch (kR <= 9)

// CAND: the rebuild-time split of the rows into candidate rows instead (the
// same pass with every cutoff widened by the skin, writing slot indices
// instead of edge records); it runs only when this step rebuilt the rows.
template <bool ONE, bool MAP, int G, int RB, bool CAND = false>
__global__ __launch_bounds__(kBlock) void k_filter(DevBox b, DevModel m, int c1, int c2, int nblocks, int close,
                                                 DevWork w)
{
  constexpr int PER = kBlock / G;
  static_assert(G >= 1 && G <= 32 && (G & (G - 1)) == 0, "filter lanes: a power of two up to 32");
  constexpr unsigned long long gmask = (1ull << G) - 1ull;
  if (CAND && !w.flags[kFlagRebuild]) return;
  __shared__ double cut2[ONE ? 1 : 2 * kMaxTypes * kMaxTypes];  // [ti][tj] radial, then angular
  const int T = m.T;
  const double widen = CAND ? kSkin : 0.0;
  double rr2 = 0.0, ra2 = 0.0;
  if (ONE) {
    const double rr = (m.rcR[0] + m.rcR[0]) * 0.5 + widen, ra = (m.rcA[0] + m.rcA[0]) * 0.5 + widen;
    rr2 = rr * rr;
    ra2 = ra * ra;
  } else {
    for (int k = threadIdx.x; k < T * T; k += blockDim.x) {
      const int p = k / T, q = k % T;
      const double rr = (m.rcR[p] + m.rcR[q]) * 0.5 + widen, ra = (m.rcA[p] + m.rcA[q]) * 0.5 + widen;
      cut2[k] = rr * rr;
      cut2[T * T + k] = ra * ra;
    }
    __syncthreads();
  }
  // the rebuild kernels of this step ran before this one: close the request
  // (one filter launch per evaluation does; the others of a split one skip)
  if (!CAND && close && blockIdx.x == 0 && threadIdx.x == 0 && w.flags[kFlagRebuild]) {
    w.flags[kFlagRebuild] = 0;
    w.flags[kFlagRebuildCount] += 1;
  }
  const int g = threadIdx.x / G, lg = threadIdx.x % G, gw = g % (64 / G);
  const int i = c1 + xcd_tile(nblocks) * PER + g;
  const bool active = i < c2;
  double4 pi = make_double4(0, 0, 0, 0);
  int ti = 0, nv = 0;
  if (active) {
    pi = w.xq[i];
    ti = ONE ? 0 : safe_type((int)__double_as_longlong(pi.w), T);
    nv = min(w.vn[i], w.MV);
  }
  const size_t vrow = (size_t)i * w.MV, erow = (size_t)i * w.MR, arow = (size_t)i * w.MA;
  const unsigned below = (1u << lg) - 1u;
  int na = 0, nro = 0;
  // the row's angular bits, assembled a 32-bit word at a time by the
  // group's first lane (a chunk of G consecutive entries never straddles
  // a word) and stored once the word is complete
  unsigned* const mrow = CAND ? nullptr : w.amask + (size_t)i * ((w.MV + 31) >> 5);
  unsigned mword = 0u;
  int mcur = 0;
  for (int p0 = 0; __any(p0 < nv); p0 += G * RB) {
    int jj[RB], qq[RB], jo[RB];
    double4 pj[RB];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int p = p0 + r * G + lg;
      jj[r] = -1;
      qq[r] = -1;
      if (p < nv) {
        jj[r] = w.vl[vrow + p];
        qq[r] = w.vrev[vrow + p];
      }
    }
    // with an image map an entry stands for its owner in the edge record
    // and the reverse index; its own slot only gives the position
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      pj[r] = jj[r] >= 0 ? w.xq[jj[r]] : make_double4(0, 0, 0, 0);
      jo[r] = (MAP && jj[r] >= 0) ? w.map.owner[jj[r]] : jj[r];
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const int p = p0 + r * G + lg;
      bool isA = false, isR = false;
      float4 rec = make_float4(0.f, 0.f, 0.f, 0.f);
      if (jj[r] >= 0) {
        double x = pj[r].x - pi.x, y = pj[r].y - pi.y, z = pj[r].z - pi.z;
        mic(b, x, y, z);
        const double d2 = dist2(x, y, z);
        double cr = rr2, ca = ra2;
        if (!ONE) {
          const int tj = safe_type((int)__double_as_longlong(pj[r].w), T);
          cr = cut2[ti * T + tj];
          ca = cut2[T * T + ti * T + tj];
        }
        if (d2 < cr) {
          isA = d2 < ca;
          isR = !isA;
          rec = make_float4((float)x, (float)y, (float)z, __int_as_float(jo[r]));
        }
      }
      const unsigned gA = (unsigned)((__ballot(isA) >> (gw * G)) & gmask);
      const unsigned gR = (unsigned)((__ballot(isR) >> (gw * G)) & gmask);
      // every row entry gets its angular slot or -1: one contiguous store
      const int a = na + __popc(gA & below);
      if (CAND) {
        // candidate rows hold the slot itself (its position is read every
        // step); the reverse index becomes a candidate slot in k_split_reverse
        if (jj[r] >= 0) w.cpos[vrow + p] = (isA && a < w.MA) ? a : -1;
        if (isA && a < w.MA) {
          w.cl[vrow + a] = jj[r];
          w.cq[arow + a] = qq[r] >= 0 ? jo[r] * w.MV + qq[r] : -1;
        }
        if (isR) {
          const int slot = w.MV - 1 - (nro + __popc(gR & below));
          if (slot >= 0) w.cl[vrow + slot] = jj[r];
        }
      } else {
        const int base = p0 + r * G;
        if (lg == 0 && base < nv) {
          if ((base >> 5) != mcur) {
            mrow[mcur] = mword;
            mword = 0u;
            mcur = base >> 5;
          }
          mword |= gA << (base & 31);
        }
        if (isA && a < w.MA) {
          w.edge[erow + a] = rec;
          w.eq[arow + a] = qq[r] >= 0 ? jo[r] * w.MV + qq[r] : -1;
        }
        if (isR) {
          const int slot = w.MR - 1 - (nro + __popc(gR & below));
          if (slot >= 0) w.edge[erow + slot] = rec;
        }
      }
      na += __popc(gA);
      nro += __popc(gR);
    }
  }
  // na + nro <= nv <= MV = MR = MA: the two ends of a row never meet
  if (!CAND && active && lg == 0 && nv > 0) mrow[mcur] = mword;
  if (active && lg == 0) {
    if (CAND) {
      w.nrc[i] = min(na + nro, w.MV);
      w.nac[i] = min(na, w.MA);
    } else {
      w.nr[i] = min(na + nro, w.MR);
      w.na[i] = min(na, w.MA);
    }
  }
}

// Rebuild-time reverse of the candidate rows: an angular candidate i->j's
// Verlet reverse index j*MV + q becomes j*MA + (i's angular candidate slot
// in j's row), or -1 where j has no row in [c1, c2) or i is not one of its
// angular candidates.
template <int G>
__global__ __launch_bounds__(kBlock) void k_split_reverse(int c1, int c2, int nblocks, DevWork w)
{
  if (!w.flags[kFlagRebuild]) return;
  constexpr int PER = kBlock / G;
  const int g = threadIdx.x / G, lg = threadIdx.x % G;
  const int i = c1 + xcd_tile(nblocks) * PER + g;
  if (i >= c2) return;
  const int n = w.nac[i];
  const size_t arow = (size_t)i * w.MA;
  for (int a = lg; a < n; a += G) {
    const int q = w.cq[arow + a];
    int v = -1;
    if (q >= 0) {
      const int j = q / w.MV;
      if (j >= c1 && j < c2) {
        const int c = w.cpos[q];
        if (c >= 0) v = j * w.MA + c;
      }
    }
    w.cq[arow + a] = v;
  }
}

// ---------------------------------------------------------------------------
// fused descriptor + ANN + adjoint + angular partial forces
// ---------------------------------------------------------------------------

// ANN modes of the descriptor kernel
enum : int { kAnnValu = 0, kAnnMfma32 = 1, kAnnMfmaBf16 = 2, kAnnMfmaSplit = 3 };

// multi-type models with at most this many types keep their coefficient
// blocks in LDS
constexpr int kTypesLds = 4;
// multi-type models with at most this many types run their ANN on the
// matrix cores: every type's weights staged in LDS, one GEMM pass per type
// present in the wavefront, each atom taking the pass of its own type
constexpr int kTypesMfma = 2;

template <int NR, int KR, int NA, int KA, int LM, bool Q222, bool Q1111, bool ONE, bool EXACT, int ANN = kAnnValu,
          bool CL = false>
struct Cfg {
  static constexpr int nr = NR, kr = KR, na = NA, ka = KA, lm = LM, ann = ANN;
  static constexpr bool q222 = Q222, q1111 = Q1111, one = ONE, exact = EXACT;
  static constexpr bool cl = CL && EXACT && !ONE;  // coefficient blocks in LDS (T <= kTypesLds)
  // type pairs of the LDS tables: every pair of up to kMaxTypes types, or
  // the types an LDS-table instance serves (kTypesMfma with the matrix-core
  // ANN, which stages every type's weights too)
  static constexpr int ltypes = !cl ? kMaxTypes : (ANN != kAnnValu ? kTypesMfma : kTypesLds);
  static constexpr int nabc = (LM + 1) * (LM + 1) - 1;
  static constexpr int numl = LM + (Q222 ? 1 : 0) + (Q1111 ? 1 : 0);
  static constexpr int dim = NR + NA * numl;
};

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short s16x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16, round to nearest even (inputs are finite)
__device__ __forceinline__ short to_bf16(float x)
{
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

// LDS floats per wavefront of the matrix-core ANN: 64 rows of DMAX+1 (odd
// stride), then 64 energies
template <int DMAX>
constexpr int ann_lds_floats() { return 64 * (DMAX + 1) + 128; }

// ANN of the 64 atoms of a wavefront on the matrix cores (single-type
// model, H <= 64).  Lane l holds atom l's scaled descriptor q[0..D).
//   GEMM 1: Z^T[h][atom] = sum_d W0[h][d] Q[atom][d]   (A = W0, B = Q^T)
//   x = tanh(Z - b0), E = sum_h w1 x, G = w1 (1 - x^2)
//   GEMM 2: Fp[atom][d] = sum_h G[atom][h] W0[h][d]     (A = G, B = W0)
// With the 16x16 output layout (lane l, register r) -> (row 4(l/16)+r,
// column l%16), the GEMM-1 accumulator of lane l holds atom 16nt + l%16 at
// hidden units 16mt + 4(l/16) + r; ordering GEMM 2's contraction index as
// (mt, r) makes those registers its A operand as they are.  Fp rows and the
// energies go back to their lanes through LDS.  BF16 rounds Q, W0 and G to
// bfloat16 (16x16x16 bf16 MFMA, FP32 accumulation); FP32 uses the 16x16x4
// FP32 MFMA (exact products).
// LDS copy of one type's ANN weights for the matrix-core path: W0 rows
// padded to 64 hidden units with an odd row stride (no bank conflicts), then
// b0 and w1 (zero past H)
template <int DMAX>
constexpr int ann_w_stride() { return DMAX + 1; }
// floats of W0 rows: 64 x WS floats, or with SPLIT three bfloat16 parts of
// 64 x WS shorts each (96 x WS floats); then b0 and w1
template <int DMAX, bool SPLIT = false>
constexpr int ann_w_floats() { return (SPLIT ? 96 : 64) * ann_w_stride<DMAX>() + 128; }

// bfloat16 -> float (exact)
__device__ __forceinline__ float from_bf16(short b) { return __uint_as_float(((unsigned)(unsigned short)b) << 16); }

// x = hi + mid + lo + r with hi = bf16(x), mid = bf16(x - hi),
// lo = bf16(x - hi - mid) (each difference exact in FP32), |r| <= 2^-27 |x|
struct Bf16Split {
  short hi, mid, lo;
};
__device__ __forceinline__ Bf16Split split_bf16(float x)
{
  const short hi = to_bf16(x);
  const float r1 = x - from_bf16(hi);
  const short mid = to_bf16(r1);
  return Bf16Split{hi, mid, to_bf16(r1 - from_bf16(mid))};
}

// SPLIT: W0 is stored as its three bfloat16 parts, shorts [64][WS] each
// (hi, then mid, then lo), in place of the float rows
template <int DMAX, bool SPLIT = false>
__device__ __forceinline__ void stage_ann_weights(const DevModel& m, int D, float* __restrict__ ws, int t = 0)
{
  constexpr int WS = ann_w_stride<DMAX>();
  const int H = m.H;
  const float* __restrict__ w0 = m.w0 + (size_t)t * H * D;
  const float* __restrict__ b0 = m.b0 + (size_t)t * H;
  const float* __restrict__ w1 = m.w1 + (size_t)t * H;
  short* __restrict__ whl = reinterpret_cast<short*>(ws);
  for (int k = threadIdx.x; k < 64 * WS; k += blockDim.x) {
    const int h = k / WS, d = k % WS;
    const float x = (h < H && d < D) ? w0[h * D + d] : 0.0f;
    if (SPLIT) {
      const Bf16Split p = split_bf16(x);
      whl[k] = p.hi;
      whl[64 * WS + k] = p.mid;
      whl[128 * WS + k] = p.lo;
    } else {
      ws[k] = x;
    }
  }
  for (int h = threadIdx.x; h < 64; h += blockDim.x) {
    constexpr int bo = (SPLIT ? 96 : 64) * WS;
    ws[bo + h] = h < H ? b0[h] : 0.0f;
    ws[bo + 64 + h] = h < H ? w1[h] : 0.0f;
  }
}


// orders this wavefront's LDS writes before its later LDS reads (each
// wavefront works in its own LDS region, so no workgroup barrier is needed)
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// MT: multi-type (ti is the lane's type, T the number of types <=
// kTypesMfma, ws holds the types' weight blocks one after another): per
// type present in the wavefront one pass over the hidden tiles whose
// activation gradients and energies count only for atoms of that type.
template <int DMAX, int MODE, bool MT = false>
__device__ __forceinline__ void ann_mfma(const DevModel& m, int D, const float* q, float* Fp, float& E,
                                         float* __restrict__ lds, const float* __restrict__ ws, int ti = 0,
                                         int T = 1)
{
  // MODE kAnnMfma32: FP32 operands (16x16x4 FP32 MFMA, exact products);
  // kAnnMfmaBf16: bfloat16 operands; kAnnMfmaSplit: every operand split
  // into three bfloat16 parts (hi + mid + lo) and each product formed from
  // the six part products of relative size above 2^-24 (hi*hi, hi*mid,
  // mid*hi, mid*mid, hi*lo, lo*hi; 16x16x16 bf16 MFMAs, FP32 accumulation,
  // smallest first): FP32 accuracy from bfloat16 products, the "bf16x6"
  // method of Henry, Tang and Heinecke (ARITH 2019)
  constexpr bool BF16 = MODE == kAnnMfmaBf16;
  constexpr bool X3 = MODE == kAnnMfmaSplit;
  constexpr int DS = DMAX + 1;
  constexpr int DT = (DMAX + 15) / 16;
  const int lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
  float* Qs = lds;
  float* Es = lds + 64 * DS;
  int* Ts = reinterpret_cast<int*>(lds + 64 * DS + 64);
#pragma unroll
  for (int d = 0; d < DMAX; ++d) Qs[lane * DS + d] = d < D ? q[d] : 0.0f;
  unsigned present = 1u;
  if (MT) {
    Ts[lane] = ti;
    present = 0u;
    for (int t = 0; t < T; ++t)
      if (__any(ti == t)) present |= 1u << t;
  }
  wave_lds_sync();
  const int H = m.H, HT = (H + 15) >> 4;
  constexpr int WS = ann_w_stride<DMAX>();
#pragma unroll 1
  for (int nt = 0; nt < 4; ++nt) {
    const float* qrow = Qs + (16 * nt + lr) * DS;
    // split descriptor operands of this atom tile, reused by every hidden tile
    s16x4 qh[X3 ? DT : 1], qm[X3 ? DT : 1], ql[X3 ? DT : 1];
    if (X3) {
#pragma unroll
      for (int ks = 0; ks < DT; ++ks)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int d = 16 * ks + 4 * lg + c;
          const Bf16Split p = split_bf16(d < D ? qrow[d] : 0.0f);
          qh[ks][c] = p.hi;
          qm[ks][c] = p.mid;
          ql[ks][c] = p.lo;
        }
    }
    f32x4 facc[DT];
#pragma unroll
    for (int dt = 0; dt < DT; ++dt) facc[dt] = f32x4{0.f, 0.f, 0.f, 0.f};
    float ep = 0.0f;
    const int at = MT ? Ts[16 * nt + lr] : 0;  // type of this lane's atom in the tile
    // the type loop is unrolled over its compile-time bound: a runtime-bounded
    // loop that skips absent types gave wrong accumulators in GEMM 2 here
    // (forces off by 5% with exact energies), the unrolled form does not
#pragma unroll
    for (int t = 0; t < (MT ? kTypesMfma : 1); ++t) {
      if (t >= T || !((present >> t) & 1u)) continue;
      const bool mine = !MT || at == t;
      // [64][WS] floats (or three [64][WS] bfloat16 parts), zero-padded
      const float* __restrict__ W0 = ws + (size_t)t * ann_w_floats<DMAX, X3>();
      const float* __restrict__ b0 = W0 + (X3 ? 96 : 64) * WS;
      const float* __restrict__ w1 = b0 + 64;
      // one tile of 16 hidden units at a time: GEMM 1, activation, and its
      // share of GEMM 2, so that only one accumulator tile is live
#pragma unroll 1
      for (int mt = 0; mt < HT; ++mt) {
        f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
        const int ha = 16 * mt + lr;
        const short* __restrict__ Wh = reinterpret_cast<const short*>(W0);
        const short* __restrict__ Wm = Wh + 64 * WS;
        const short* __restrict__ Wl = Wh + 128 * WS;
        if (X3) {
#pragma unroll
          for (int ks = 0; ks < DT; ++ks) {
            s16x4 ah, am, al;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              // columns past the row stride would read the next row or, past
              // the last row, the float biases as bfloat16 pairs (possibly
              // NaN patterns): they hold zeros
              const int d = 16 * ks + 4 * lg + c;
              ah[c] = d < WS ? Wh[ha * WS + d] : (short)0;
              am[c] = d < WS ? Wm[ha * WS + d] : (short)0;
              al[c] = d < WS ? Wl[ha * WS + d] : (short)0;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, qh[ks], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, ql[ks], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am, qm[ks], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(am, qh[ks], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, qm[ks], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, qh[ks], acc, 0, 0, 0);
          }
        } else if (!BF16) {
#pragma unroll
          for (int ks = 0; ks < (DMAX + 3) / 4; ++ks) {
            const int d = 4 * ks + lg;
            const float bq = d < D ? qrow[d] : 0.0f;
            const float a = W0[ha * WS + d];
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bq, acc, 0, 0, 0);
          }
        } else {
#pragma unroll
          for (int ks = 0; ks < (DMAX + 15) / 16; ++ks) {
            s16x4 bq, a;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              const int d = 16 * ks + 4 * lg + c;
              bq[c] = to_bf16(d < D ? qrow[d] : 0.0f);
              a[c] = to_bf16(W0[ha * WS + d]);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bq, acc, 0, 0, 0);
          }
        }
        float G[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int h = 16 * mt + 4 * lg + r;
          G[r] = 0.0f;
          if (h < H && mine) {
            const float x = tanhf(acc[r] - b0[h]);
            ep += w1[h] * x;
            G[r] = w1[h] * (1.0f - x * x);
          }
        }
        if (X3) {
          s16x4 gh, gm, gl;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const Bf16Split p = split_bf16(G[r]);
            gh[r] = p.hi;
            gm[r] = p.mid;
            gl[r] = p.lo;
          }
#pragma unroll
          for (int dt = 0; dt < DT; ++dt) {
            const int d = 16 * dt + lr;
            s16x4 bh, bm, bl;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              const int h = 16 * mt + 4 * lg + c;
              bh[c] = d < WS ? Wh[h * WS + d] : (short)0;
              bm[c] = d < WS ? Wm[h * WS + d] : (short)0;
              bl[c] = d < WS ? Wl[h * WS + d] : (short)0;
            }
            facc[dt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gl, bh, facc[dt], 0, 0, 0);
            facc[dt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gh, bl, facc[dt], 0, 0, 0);
            facc[dt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gm, bm, facc[dt], 0, 0, 0);
            facc[dt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gm, bh, facc[dt], 0, 0, 0);
            facc[dt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gh, bm, facc[dt], 0, 0, 0);
            facc[dt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(gh, bh, facc[dt], 0, 0, 0);
          }
        } else if (!BF16) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int h = 16 * mt + 4 * lg + r;
#pragma unroll
            for (int dt = 0; dt < DT; ++dt) {
              const int d = 16 * dt + lr;
              const float bw = W0[h * WS + d];
              facc[dt] = __builtin_amdgcn_mfma_f32_16x16x4f32(G[r], bw, facc[dt], 0, 0, 0);
            }
          }
        } else {
          s16x4 a;
#pragma unroll
          for (int r = 0; r < 4; ++r) a[r] = to_bf16(G[r]);
#pragma unroll
          for (int dt = 0; dt < DT; ++dt) {
            const int d = 16 * dt + lr;
            s16x4 bw;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
              const int h = 16 * mt + 4 * lg + c;
              bw[c] = to_bf16(W0[h * WS + d]);
            }
            facc[dt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bw, facc[dt], 0, 0, 0);
          }
        }
      }
    }  // types
    ep += __shfl_xor(ep, 16);
    ep += __shfl_xor(ep, 32);
    if (lg == 0) Es[16 * nt + lr] = ep;
    // this tile's Fp rows replace its (already consumed) Q rows
#pragma unroll
    for (int dt = 0; dt < DT; ++dt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int d = 16 * dt + lr;
        if (d < DMAX) Qs[(16 * nt + 4 * lg + r) * DS + d] = facc[dt][r];
      }
  }
  wave_lds_sync();
#pragma unroll
  for (int d = 0; d < DMAX; ++d) Fp[d] = d < D ? Qs[lane * DS + d] : 0.0f;
  E = Es[lane];
}

// Static LDS of a descriptor instance: pair tables, coefficient blocks, ANN
// weights and the per-wavefront ANN staging.  Instances stay within
// kDescLdsMax per workgroup; where the matrix-core ANN would not, the vector
// ANN runs.
constexpr int kDescLdsMax = 96 * 1024;
template <class C>
constexpr int desc_lds_bytes()
{
  const int pairs = C::ltypes * C::ltypes;
  const int pcut = C::one ? 1 + 2 : 4 * pairs + 4 * pairs;  // float cutoffs, then the FP64 squares
  const int ctab = C::cl ? pairs * (C::nr * C::kr + C::na * C::ka) : 1;
  const int mfma = C::ann != kAnnValu;
  const int annw = mfma ? (C::one ? 1 : kTypesMfma) * ann_w_floats<C::dim, C::ann == kAnnMfmaSplit>() : 1;
  const int annl = mfma ? (kDescBlock / 64) * ann_lds_floats<C::dim>() : 0;
  return 4 * (pcut + ctab + annw + annl);
}

template <int NR, int KR, int NA, int KA, bool Q2, bool Q1, bool ONE, int ANN = kAnnMfma32>
constexpr bool mfma_ann_fits()
{
  // single-type instances reach two wavefronts per SIMD, so they keep two
  // workgroups per CU (64 KiB each); multi-type ones run one either way
  return desc_lds_bytes<Cfg<NR, KR, NA, KA, 4, Q2, Q1, ONE, true, ANN, !ONE>>() <=
         (ONE ? 64 * 1024 : kDescLdsMax);
}

// sum of v over the P consecutive lanes of a lane group (every lane gets it)
template <int P>
__device__ __forceinline__ float group_sum(float v)
{
#pragma unroll
  for (int off = 1; off < P; off <<= 1) v += __shfl_xor(v, off, P);
  return v;
}

// CAND: the edges come from the candidate rows (DevWork::split).  edge, vsa
// and eq are w.edge, w.vsa and w.eq, passed as restrict parameters: with
// candidate rows this kernel writes them, and only through them, so the
// model's tables stay provably unmodified (scalar loads).
// P: lanes per atom.  With P > 1 (systems too small to fill the GPU with one
// lane per atom) a group of P lanes shares an atom: each takes every P-th
// edge and every P-th hidden neuron of the (vector-ALU) ANN, and the partial
// sums are combined across the group, so every lane holds the atom's
// descriptor, moments and derivatives.
template <class C, bool CAND, int P = 1>
__global__ __launch_bounds__(kDescBlock) void k_descriptor(DevBox b, DevModel m, int N, int c1, int c2, int nblocks,
                                                       const int* __restrict__ type, double* __restrict__ pe,
                                                       float4* __restrict__ edge, int* __restrict__ vsa,
                                                       int* __restrict__ eq, int accumulate, DevWork w)
{
  constexpr int NR = C::nr, KR = C::kr, NA = C::na, KA = C::ka, LM = C::lm, NABC = C::nabc, DMAX = C::dim;
  static_assert(P == 1 || (P == 4 && !CAND && C::ann == kAnnValu), "lane groups: the filter's rows, vector ANN");
  const int lr = P == 1 ? 0 : (int)(threadIdx.x & (P - 1));  // lane within the atom's group
  const int i0 = c1 + xcd_tile(nblocks) * (int)(blockDim.x / P) + (int)(threadIdx.x / P);
  const bool active = i0 < c2;
  // the matrix-core ANN needs every lane of the wavefront: lanes past the
  // range carry an empty row of a valid atom and write nothing
  const int i = active ? i0 : c1;
  // multi-type: per type pair the float cutoffs and their inverses (radial,
  // angular; the FP64 pair mean rounded once) and, with C::cl, the
  // coefficient blocks [pair][n][k] (radial, then angular), staged in LDS
  constexpr int kPairs = C::ltypes * C::ltypes;
  static_assert(desc_lds_bytes<C>() <= kDescLdsMax, "descriptor LDS above its bound per workgroup");
  __shared__ float pcut[C::one ? 1 : 4 * kPairs];
  __shared__ double pcut2[C::one ? 1 : 2 * kPairs];  // FP64 squared pair cutoffs (candidate rows)
  __shared__ float ctab[C::cl ? kPairs * (NR * KR + NA * KA) : 1];
  if (!C::one) {
    const int T = m.T;
    for (int k = threadIdx.x; k < T * T; k += blockDim.x) {
      const int p = k / T, q = k % T;
      const double rrd = (m.rcR[p] + m.rcR[q]) * 0.5, rad = (m.rcA[p] + m.rcA[q]) * 0.5;
      const float rr = (float)rrd, ra = (float)rad;
      pcut[k] = rr;
      pcut[kPairs + k] = 1.0f / rr;
      pcut[2 * kPairs + k] = ra;
      pcut[3 * kPairs + k] = 1.0f / ra;
      pcut2[k] = rrd * rrd;
      pcut2[kPairs + k] = rad * rad;
    }
    if (C::cl) {
      const int nrad = T * T * NR * KR, nang = T * T * NA * KA;
      for (int k = threadIdx.x; k < nrad + nang; k += blockDim.x) ctab[k] = k < nrad ? m.cR[k] : m.cA[k - nrad];
    }
    __syncthreads();
  }
  // candidate rows: the rebuild kernels of this step ran before this launch
  // (the only descriptor launch of the evaluation), which closes the request
  if (CAND && blockIdx.x == 0 && threadIdx.x == 0 && w.flags[kFlagRebuild]) {
    w.flags[kFlagRebuild] = 0;
    w.flags[kFlagRebuildCount] += 1;
  }
  if (C::ann == kAnnValu && !active) return;
  __shared__ float ann_w[C::ann != kAnnValu ? (C::one ? 1 : kTypesMfma) * ann_w_floats<DMAX, C::ann == kAnnMfmaSplit>()
                                           : 1];
  if (C::ann != kAnnValu) {
    // every type's weights (at most kTypesMfma types), staged once per workgroup
    for (int t = 0; t < (C::one ? 1 : m.T); ++t)
      stage_ann_weights<DMAX, C::ann == kAnnMfmaSplit>(
          m, C::exact ? DMAX : m.D, ann_w + t * ann_w_floats<DMAX, C::ann == kAnnMfmaSplit>(), t);
    __syncthreads();
  }
  const int nR = C::exact ? NR : m.nR, nA = C::exact ? NA : m.nA;
  // 3-body order: components L <= Lm, then the 4-/5-body blocks
  const int Lm = C::exact ? LM : m.L;
  const int kR = C::exact ? KR : m.kR, kA = C::exact ? KA : m.kA;
  const bool h222 = C::exact ? C::q222 : (m.has222 != 0);
  const bool h1111 = C::exact ? C::q1111 : (m.has1111 != 0);
  const int D = C::exact ? DMAX : m.D;
  const int T = C::one ? 1 : m.T;
  const int ti = C::one ? 0 : safe_type(type[i], T);
  float rcRf[kMaxTypes], rcAf[kMaxTypes];
  // per-type float cutoffs (the FP64 values rounded once)
  if (C::one) {
    rcRf[0] = (float)m.rcR[0];
    rcAf[0] = (float)m.rcA[0];
  }
  int nr = 0, na = 0;
  if (active && !CAND) {
    nr = w.nr[i];
    na = w.na[i];
  }
  // with candidate rows this kernel writes the angular edges it reads back
  const float4* erow = edge + (size_t)i * w.MR;
  float4* __restrict__ frow = w.f12 + (size_t)i * w.MA;
  float q[DMAX];
#pragma unroll
  for (int d = 0; d < DMAX; ++d) q[d] = 0.0f;

  // ---- radial descriptor: q_n = sum_j sum_k c[ti,tj,n,k] f_k(r_ij)
  // The basis values are summed per neighbour type and contracted with the
  // coefficients once per atom (single type: one sum; up to C::ltypes types:
  // one per type); more types contract per edge.
  float Fs[KR];
#pragma unroll
  for (int k = 0; k < KR; ++k) Fs[k] = 0.0f;
  float Ft[C::cl ? C::ltypes : 1][KR];
#pragma unroll
  for (int t = 0; t < (C::cl ? C::ltypes : 1); ++t)
#pragma unroll
    for (int k = 0; k < KR; ++k) Ft[t][k] = 0.0f;
  // the radial loop keeps kAhead edge records in flight (its work per edge
  // is short); the angular loops below load the next record before using
  // the current one and the first record of the next 128-byte line a line
  // ahead
  constexpr int kAhead = 8;
  auto radial_edge = [&](const float4 r, const int tj) {
    const float d2 = r.x * r.x + r.y * r.y + r.z * r.z;
    const float d = d2 * __builtin_amdgcn_rsqf(d2);  // edges have d > 0
    const float rc = C::one ? rcRf[0] : pcut[ti * T + tj];
    const float rcinv = C::one ? 1.0f / rc : pcut[kPairs + ti * T + tj];
    float fc;
    cutoff_fc(rc, rcinv, d, fc);
    float f[KR];
    cheb<KR>(kR, rcinv, d, fc, f);
    if (C::one) {
#pragma unroll
      for (int k = 0; k < KR; ++k)
        if (k < kR) Fs[k] += f[k];
    } else if (C::cl) {
#pragma unroll
      for (int t = 0; t < C::ltypes; ++t) {
        if (t < T) {
          const float wt = tj == t ? 1.0f : 0.0f;
#pragma unroll
          for (int k = 0; k < KR; ++k) Ft[t][k] = fmaf(wt, f[k], Ft[t][k]);
        }
      }
    } else {
      const float* c = m.cR + (size_t)(ti * T + tj) * nR * kR;
#pragma unroll
      for (int n = 0; n < NR; ++n) {
        if (n < nR) {
          float g = 0.0f;
#pragma unroll
          for (int k = 0; k < KR; ++k)
            if (k < kR) g += c[n * kR + k] * f[k];
          q[n] += g;
        }
      }
    }
  };
  if (!CAND) {
    auto ridx = [&](int e) { return e < na ? e : w.MR - 1 - (e - na); };
    float4 ring[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; ++u)
      ring[u] = u * P + lr < nr ? erow[ridx(u * P + lr)] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int e0 = lr; e0 < nr; e0 += kAhead * P)
#pragma unroll
      for (int u = 0; u < kAhead; ++u) {
        const int e = e0 + u * P;
        if (e >= nr) break;
        const float4 r = ring[u];
        if (e + kAhead * P < nr) ring[u] = erow[ridx(e + kAhead * P)];
        radial_edge(r, C::one ? 0 : safe_type(type[__float_as_int(r.w)], T));
      }
    if (P > 1) {
      // the group's partial radial sums
      if (C::one) {
#pragma unroll
        for (int k = 0; k < KR; ++k) Fs[k] = group_sum<P>(Fs[k]);
      } else if (C::cl) {
#pragma unroll
        for (int t = 0; t < (C::cl ? C::ltypes : 1); ++t)
#pragma unroll
          for (int k = 0; k < KR; ++k) Ft[t][k] = group_sum<P>(Ft[t][k]);
      } else {
#pragma unroll
        for (int n = 0; n < NR; ++n) q[n] = group_sum<P>(q[n]);
      }
    }
  } else {
    // candidate rows: the FP64 minimum-image distance of every candidate
    // (the filter's expressions, so the same edges pass), the radial sum over
    // those within the pair's radial cutoff, and the step's angular edges
    // (in candidate order) with their reverse indices and the candidates'
    // slot status for the gather.  Slot indices are read eight candidates
    // ahead, the packed positions four.
    const int nrc = active ? w.nrc[i] : 0, nac = active ? w.nac[i] : 0;
    const size_t vrow = (size_t)i * w.MV, arow = (size_t)i * w.MA;
    const int* __restrict__ crow = w.cl + vrow;
    float4* ew = edge + (size_t)i * w.MR;
    const double4 pi = w.xq[i];
    double cr1 = 0.0, ca1 = 0.0;
    if (C::one) {
      const double rr = (m.rcR[0] + m.rcR[0]) * 0.5, ra = (m.rcA[0] + m.rcA[0]) * 0.5;
      cr1 = rr * rr;
      ca1 = ra * ra;
    }
    auto cidx = [&](int e) { return e < nac ? e : w.MV - 1 - (e - nac); };
    constexpr int kSplitAhead = 4;
    int jr[2 * kSplitAhead];
    double4 pr[kSplitAhead];
#pragma unroll
    for (int u = 0; u < 2 * kSplitAhead; ++u) jr[u] = u < nrc ? crow[cidx(u)] : 0;
#pragma unroll
    for (int u = 0; u < kSplitAhead; ++u) pr[u] = u < nrc ? w.xq[jr[u]] : make_double4(0.0, 0.0, 0.0, 0.0);
    int nin = 0;
    for (int e0 = 0; e0 < nrc; e0 += 2 * kSplitAhead)
#pragma unroll
      for (int u = 0; u < 2 * kSplitAhead; ++u) {
        const int e = e0 + u;
        if (e >= nrc) break;
        const int j = jr[u];
        const double4 pj = pr[u % kSplitAhead];
        if (e + kSplitAhead < nrc) pr[u % kSplitAhead] = w.xq[jr[(u + kSplitAhead) % (2 * kSplitAhead)]];
        if (e + 2 * kSplitAhead < nrc) jr[u] = crow[cidx(e + 2 * kSplitAhead)];
        double x = pj.x - pi.x, y = pj.y - pi.y, z = pj.z - pi.z;
        mic(b, x, y, z);
        const double d2 = dist2(x, y, z);
        int tj = 0;
        double cr = cr1, ca = ca1;
        if (!C::one) {
          tj = safe_type((int)__double_as_longlong(pj.w), T);
          cr = pcut2[ti * T + tj];
          ca = pcut2[kPairs + ti * T + tj];
        }
        const bool in_r = d2 < cr;
        const int jo = w.map.owner ? w.map.owner[j] : j;
        const float4 r = make_float4((float)x, (float)y, (float)z, __int_as_float(jo));
        if (e < nac) {
          const bool in_a = in_r && d2 < ca;
          vsa[arow + e] = in_a ? na : -1;
          if (in_a) {
            ew[na] = r;
            eq[arow + na] = w.cq[arow + e];
            ++na;
          }
        }
        if (in_r) {
          ++nin;
          radial_edge(r, tj);
        }
      }
    if (active) {
      w.nr[i] = nin;
      w.na[i] = na;
    }
  }
  if (C::one) {
#pragma unroll
    for (int n = 0; n < NR; ++n) {
      if (n < nR) {
        float g = 0.0f;
#pragma unroll
        for (int k = 0; k < KR; ++k)
          if (k < kR) g += m.cR[n * kR + k] * Fs[k];
        q[n] = g;
      }
    }
  } else if (C::cl) {
#pragma unroll
    for (int n = 0; n < NR; ++n) {
      float g = 0.0f;
#pragma unroll
      for (int t = 0; t < C::ltypes; ++t) {
        if (t < T) {
          const float* c = ctab + (ti * T + t) * NR * KR + n * KR;
#pragma unroll
          for (int k = 0; k < KR; ++k) g += c[k] * Ft[t][k];
        }
      }
      q[n] = g;
    }
  }

  // ---- angular moments s[n][lm] = sum_j g_n(r_ij) Y_lm(u_ij)
  float s[NA][NABC];
#pragma unroll
  for (int n = 0; n < NA; ++n)
#pragma unroll
    for (int a = 0; a < NABC; ++a) s[n][a] = 0.0f;
  float zpe = 0.0f;
  // each lane takes every P-th edge; the next record is loaded before the
  // current one is used, and the first record of the next 128-byte line
  // (eight of the lane's records) a line ahead
  float4 rn = erow[lr < na ? lr : 0];
  float4 ahead = erow[lr + 8 * P < na ? lr + 8 * P : 0];
  for (int e = lr; e < na; e += P) {
    const float4 r = rn;
    const int en = e + P;
    if (en < na) {
      if ((((en - lr) / P) & 7) == 0) {
        rn = ahead;
        if (en + 8 * P < na) ahead = erow[en + 8 * P];
      } else {
        rn = erow[en];
      }
    }
    const float d2 = r.x * r.x + r.y * r.y + r.z * r.z;
    const float dinv = __builtin_amdgcn_rsqf(d2);  // hardware 1/sqrt (1 ulp); edges have d > 0
    const float d = d2 * dinv;
    float rc;
    int tj = 0;
    if (C::one) {
      rc = rcAf[0];
    } else {
      tj = safe_type(type[__float_as_int(r.w)], T);
      rc = pcut[2 * kPairs + ti * T + tj];
    }
    if (m.zbl) {
      // half of each ZBL pair energy belongs to the centre; ZBL pairs are
      // taken from the angular rows, as the reference's ZBL pass does
      float zf, zfp;
      zbl_pair(m.zblp + (size_t)(ti * m.T + tj) * 12, d, dinv, zf, zfp);
      zpe += 0.5f * zf;
    }
    const float rcinv = C::one ? 1.0f / rc : pcut[3 * kPairs + ti * T + tj];
    float fc;
    cutoff_fc(rc, rcinv, d, fc);
    float f[KA];
    cheb<KA>(kA, rcinv, d, fc, f);
    const float* c = C::cl ? ctab + T * T * NR * KR + (ti * T + tj) * NA * KA
                           : m.cA + (size_t)(ti * T + tj) * nA * kA;
    float g[NA];
#pragma unroll
    for (int n = 0; n < NA; ++n) {
      g[n] = 0.0f;
      if (n < nA) {
#pragma unroll
        for (int k = 0; k < KA; ++k)
          if (k < kA) g[n] += c[n * kA + k] * f[k];
      }
    }
    float Y[NABC];
    harmonics<LM>(r.x * dinv, r.y * dinv, r.z * dinv, Y);
#pragma unroll
    for (int n = 0; n < NA; ++n)
#pragma unroll
      for (int a = 0; a < NABC; ++a) s[n][a] += g[n] * Y[a];
  }
  if (P > 1) {
    // the group's partial moments and ZBL energies
#pragma unroll
    for (int n = 0; n < NA; ++n)
#pragma unroll
      for (int a = 0; a < NABC; ++a) s[n][a] = group_sum<P>(s[n][a]);
    zpe = group_sum<P>(zpe);
  }

  // ---- invariants
#pragma unroll
  for (int n = 0; n < NA; ++n) {
    if (n >= nA) break;
#pragma unroll
    for (int L = 1; L <= LM; ++L) {
      float v = 0.0f;
#pragma unroll
      for (int a = L * L - 1; a < (L + 1) * (L + 1) - 1; ++a) v += c_c3[a] * s[n][a] * s[n][a];
      if (L <= Lm) q[nR + (L - 1) * nA + n] = v;
    }
    int li = Lm;
    if (h222) {
      const float* x = &s[n][3];
      q[nR + li * nA + n] = c_c4[0] * x[0] * x[0] * x[0] + c_c4[1] * x[0] * (x[1] * x[1] + x[2] * x[2]) +
                            c_c4[2] * x[0] * (x[3] * x[3] + x[4] * x[4]) +
                            c_c4[3] * x[3] * (x[2] * x[2] - x[1] * x[1]) + c_c4[4] * x[1] * x[2] * x[4];
      ++li;
    }
    if (h1111) {
      const float r2 = s[n][0] * s[n][0] + s[n][1] * s[n][1] + s[n][2] * s[n][2];
      q[nR + li * nA + n] = c_c5 * r2 * r2;
    }
  }

  // ---- ANN (nep_utilities.cuh:209-310): E = sum_h w1 tanh(w0.q - b0) - bias
#pragma unroll
  for (int d = 0; d < DMAX; ++d)
    if (d < D) q[d] *= m.qs[d];
  float Fp[DMAX];
  float E = 0.0f;
  if (C::ann != kAnnValu) {
    __shared__ float ann_lds[kDescBlock / 64][ann_lds_floats<DMAX>()];
    ann_mfma<DMAX, C::ann, !C::one>(m, D, q, Fp, E, ann_lds[threadIdx.x >> 6], ann_w, ti, T);
  } else {
    const int H = m.H;
    const float* w0 = m.w0 + (size_t)ti * H * D;
    const float* b0 = m.b0 + (size_t)ti * H;
    const float* w1 = m.w1 + (size_t)ti * H;
#pragma unroll
    for (int d = 0; d < DMAX; ++d) Fp[d] = 0.0f;
    for (int h = lr; h < H; h += P) {
      const float* wr = w0 + (size_t)h * D;
      float a = 0.0f;
#pragma unroll
      for (int d = 0; d < DMAX; ++d)
        if (d < D) a += wr[d] * q[d];
      const float x = tanhf(a - b0[h]);
      E += w1[h] * x;
      const float gq = w1[h] * (1.0f - x * x);
#pragma unroll
      for (int d = 0; d < DMAX; ++d)
        if (d < D) Fp[d] += gq * wr[d];
    }
    if (P > 1) {
      // the group's hidden neurons
      E = group_sum<P>(E);
#pragma unroll
      for (int d = 0; d < DMAX; ++d) Fp[d] = group_sum<P>(Fp[d]);
    }
  }
  E -= m.b1[0];
  if (m.version == 5) E -= m.w1x[ti];
#pragma unroll
  for (int d = 0; d < DMAX; ++d)
    if (d < D) Fp[d] *= m.qs[d];
  const double etot = (double)E + (double)zpe;
  if (active && lr == 0) pe[i] = accumulate ? pe[i] + etot : etot;

  // ---- radial adjoint in basis space: bR[i][t][k] = sum_n Fp_n c[ti,t,n,k]
  for (int t = 0; t < T; ++t) {
    const float* c = C::cl ? ctab + (ti * T + t) * NR * KR : m.cR + (size_t)(ti * T + t) * nR * kR;
    float* out = w.bR + ((size_t)i * T + t) * m.kRp;
    if (!active) break;
#pragma unroll
    for (int k = 0; k < ((KR + 3) & ~3); ++k) {
      float v = 0.0f;
      if (k < kR) {
#pragma unroll
        for (int n = 0; n < NR; ++n)
          if (n < nR) v += Fp[n] * c[n * kR + k];
      }
      if (k < m.kRp && (P == 1 || (k & (P - 1)) == lr)) out[k] = v;
    }
  }

  // ---- angular adjoint a[n][lm] = dE/ds[n][lm], written over s
#pragma unroll
  for (int n = 0; n < NA; ++n) {
    if (n >= nA) break;
    float extra[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) extra[a] = 0.0f;
    int li = Lm;
    if (h222) {
      const float F = Fp[nR + li * nA + n];
      const float* x = &s[n][3];
      extra[3] = F * (3.0f * c_c4[0] * x[0] * x[0] + c_c4[1] * (x[1] * x[1] + x[2] * x[2]) +
                      c_c4[2] * (x[3] * x[3] + x[4] * x[4]));
      extra[4] = F * (2.0f * c_c4[1] * x[0] * x[1] - 2.0f * c_c4[3] * x[3] * x[1] + c_c4[4] * x[2] * x[4]);
      extra[5] = F * (2.0f * c_c4[1] * x[0] * x[2] + 2.0f * c_c4[3] * x[3] * x[2] + c_c4[4] * x[1] * x[4]);
      extra[6] = F * (2.0f * c_c4[2] * x[0] * x[3] + c_c4[3] * (x[2] * x[2] - x[1] * x[1]));
      extra[7] = F * (2.0f * c_c4[2] * x[0] * x[4] + c_c4[4] * x[1] * x[2]);
      ++li;
    }
    if (h1111) {
      const float F = Fp[nR + li * nA + n];
      const float r2 = s[n][0] * s[n][0] + s[n][1] * s[n][1] + s[n][2] * s[n][2];
      const float k5 = 4.0f * c_c5 * F * r2;
      extra[0] = k5 * s[n][0];
      extra[1] = k5 * s[n][1];
      extra[2] = k5 * s[n][2];
    }
#pragma unroll
    for (int L = 1; L <= LM; ++L) {
      const float F2 = L <= Lm ? 2.0f * Fp[nR + (L - 1) * nA + n] : 0.0f;  // no adjoint above the order
#pragma unroll
      for (int a = L * L - 1; a < (L + 1) * (L + 1) - 1; ++a) {
        float v = F2 * c_c3[a] * s[n][a];
        if (a < 8) v += extra[a];
        s[n][a] = v;
      }
    }
  }

  // ---- angular partial forces dE_i/dr_ij (the edges of the moments loop)
  rn = erow[lr < na ? lr : 0];
  ahead = erow[lr + 8 * P < na ? lr + 8 * P : 0];
  for (int e = lr; e < na; e += P) {
    const float4 r = rn;
    const int en = e + P;
    if (en < na) {
      if ((((en - lr) / P) & 7) == 0) {
        rn = ahead;
        if (en + 8 * P < na) ahead = erow[en + 8 * P];
      } else {
        rn = erow[en];
      }
    }
    const int j = __float_as_int(r.w);
    const float d2 = r.x * r.x + r.y * r.y + r.z * r.z;
    const float dinv = __builtin_amdgcn_rsqf(d2);  // hardware 1/sqrt (1 ulp); edges have d > 0
    const float d = d2 * dinv;
    const float ux = r.x * dinv, uy = r.y * dinv, uz = r.z * dinv;
    float rc;
    int tj = 0;
    if (C::one) {
      rc = rcAf[0];
    } else {
      tj = safe_type(type[j], T);
      rc = pcut[2 * kPairs + ti * T + tj];
    }
    const float rcinv = C::one ? 1.0f / rc : pcut[3 * kPairs + ti * T + tj];
    float fc, fcp;
    cutoff_fc_fcp(rc, rcinv, d, fc, fcp);
    float f[KA], fp[KA];
    cheb_d<KA>(kA, rcinv, d, fc, fcp, f, fp);
    const float* c = C::cl ? ctab + T * T * NR * KR + (ti * T + tj) * NA * KA
                           : m.cA + (size_t)(ti * T + tj) * nA * kA;
    float g[NA], gp[NA];
#pragma unroll
    for (int n = 0; n < NA; ++n) {
      g[n] = 0.0f;
      gp[n] = 0.0f;
      if (n < nA) {
#pragma unroll
        for (int k = 0; k < KA; ++k)
          if (k < kA) {
            g[n] += c[n * kA + k] * f[k];
            gp[n] += c[n * kA + k] * fp[k];
          }
      }
    }
    // contract the adjoint with the basis: A = sum_n a g_n, B = sum_n a g'_n,
    // then Pr = sum B Y and G = sum A dY/du (Y = Z(uz) Re/Im (ux+i uy)^m)
    float zp[LM + 1];
    zp[0] = 1.0f;
#pragma unroll
    for (int k = 1; k <= LM; ++k) zp[k] = zp[k - 1] * uz;
    float pr[LM + 1], pim[LM + 1];
    pr[0] = 1.0f;
    pim[0] = 0.0f;
#pragma unroll
    for (int k = 1; k <= LM; ++k) {
      pr[k] = pr[k - 1] * ux - pim[k - 1] * uy;
      pim[k] = pr[k - 1] * uy + pim[k - 1] * ux;
    }
    float Pr = 0.0f, Gx = 0.0f, Gy = 0.0f, Gz = 0.0f;
#pragma unroll
    for (int L = 1; L <= LM; ++L) {
#pragma unroll
      for (int mm = 0; mm <= L; ++mm) {
        float Z = 0.0f, dZ = 0.0f;
#pragma unroll
        for (int k = (L - mm) & 1; k <= L - mm; k += 2) {
          Z += c_zc[L][mm][k] * zp[k];
          if (k > 0) dZ += c_zc[L][mm][k] * (float)k * zp[k - 1];
        }
        const int b = L * L - 1;
        if (mm == 0) {
          float A = 0.0f, B = 0.0f;
#pragma unroll
          for (int n = 0; n < NA; ++n) {
            A += s[n][b] * g[n];
            B += s[n][b] * gp[n];
          }
          Pr += B * Z;
          Gz += A * dZ;
        } else {
          const int ir = b + 2 * mm - 1, ii = b + 2 * mm;
          float Ar = 0.0f, Ai = 0.0f, Br = 0.0f, Bi = 0.0f;
#pragma unroll
          for (int n = 0; n < NA; ++n) {
            Ar += s[n][ir] * g[n];
            Ai += s[n][ii] * g[n];
            Br += s[n][ir] * gp[n];
            Bi += s[n][ii] * gp[n];
          }
          Pr += Z * (Br * pr[mm] + Bi * pim[mm]);
          const float mz = (float)mm * Z;
          Gx += mz * (Ar * pr[mm - 1] + Ai * pim[mm - 1]);
          Gy += mz * (Ai * pr[mm - 1] - Ar * pim[mm - 1]);
          Gz += dZ * (Ar * pr[mm] + Ai * pim[mm]);
        }
      }
    }
    const float gu = Gx * ux + Gy * uy + Gz * uz;
    const float4 fo = make_float4(Pr * ux + (Gx - gu * ux) * dinv, Pr * uy + (Gy - gu * uy) * dinv,
                                  Pr * uz + (Gz - gu * uz) * dinv, 0.0f);
    frow[e] = fo;
  }
}

// ---------------------------------------------------------------------------
// gather: radial pair force, angular partials, ZBL -> FP64 force/virial
// ---------------------------------------------------------------------------

// CAND: the radial edges come from the candidate rows (DevWork::split)
template <int KR, bool ONE, int G, bool CAND>
__global__ __launch_bounds__(kBlock) void k_gather(DevBox b, DevModel m, int n1, int n2, int nd1, int nd2,
                                                   int nblocks, const int* __restrict__ type, int N,
                                                   double* __restrict__ force, double* __restrict__ virial,
                                                   int accumulate, DevWork w)
{
  // G lanes per atom share its edges; the per-lane FP64 sums are combined
  // with a butterfly over the group at the end.  Each lane takes its edges
  // RB at a time: the RB edge records are loaded, then the RB neighbour
  // adjoint rows, then all of them are used, so the loads of a batch overlap.
  constexpr int PER = kBlock / G;
  constexpr int RB = KR <= 9 ? kGatherBatch : 2;
  static_assert(G >= 1 && G <= 64 && (G & (G - 1)) == 0, "gather lanes: a power of two");
  const int g = threadIdx.x / G, lg = threadIdx.x % G;
  const int i = n1 + xcd_tile(nblocks) * PER + g;
  const bool active = i < n2;
  const int T = ONE ? 1 : m.T;
  const int kR = m.kR;
  int ti = 0, nr = 0, na = 0, nac = 0;
  double4 pi = make_double4(0.0, 0.0, 0.0, 0.0);
  if (active) {
    ti = ONE ? 0 : safe_type(type[i], T);
    na = w.na[i];
    if (CAND) {
      // candidate rows: the radial loop runs over the candidates and keeps
      // those within the pair's radial cutoff (the descriptor's test)
      nr = w.nrc[i];
      nac = w.nac[i];
      pi = w.xq[i];
    } else {
      nr = w.nr[i];
    }
  }
  const int* __restrict__ crow = w.cl + (size_t)i * w.MV;
  constexpr int KP = (KR + 3) & ~3;  // padded row: KP/4 float4 loads
  const int kRp = m.kRp;
  const float4* __restrict__ erow = w.edge + (size_t)i * w.MR;
  const float4* __restrict__ frow = w.f12 + (size_t)i * w.MA;
  const int* __restrict__ qrow = w.eq + (size_t)i * w.MA;
  const float* __restrict__ bR = w.bR;
  double fx = 0.0, fy = 0.0, fz = 0.0;
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  float bi[KP];
#pragma unroll
  for (int k = 0; k < KP; ++k) bi[k] = 0.0f;
  if (ONE && active) load_row4<KP>(bR + (size_t)i * kRp, kRp, bi);
  const float rcR0 = (float)((m.rcR[0] + m.rcR[0]) * 0.5);
  // radial: F_i += (G_i + G_j) r_ij/d with G_c = sum_k b_c[k] f'_k(d)
  for (int e0 = lg; e0 < nr; e0 += G * RB) {
    float4 rr[RB];
    int jj[RB], tt[RB];
    if (!CAND) {
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int e = e0 + r * G;
        rr[r] = make_float4(0.f, 0.f, 0.f, 0.f);
        jj[r] = -1;
        if (e < nr) {
          rr[r] = erow[e < na ? e : w.MR - 1 - (e - na)];
          jj[r] = __float_as_int(rr[r].w);
        }
      }
    } else {
      int jc[RB];
      double4 pj[RB];
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        const int e = e0 + r * G;
        jc[r] = e < nr ? crow[e < nac ? e : w.MV - 1 - (e - nac)] : -1;
      }
#pragma unroll
      for (int r = 0; r < RB; ++r) pj[r] = jc[r] >= 0 ? w.xq[jc[r]] : make_double4(0.0, 0.0, 0.0, 0.0);
#pragma unroll
      for (int r = 0; r < RB; ++r) {
        rr[r] = make_float4(0.f, 0.f, 0.f, 0.f);
        jj[r] = -1;
        if (jc[r] >= 0) {
          double x = pj[r].x - pi.x, y = pj[r].y - pi.y, z = pj[r].z - pi.z;
          mic(b, x, y, z);
          const double d2 = dist2(x, y, z);
          double rc = (m.rcR[0] + m.rcR[0]) * 0.5;
          if (!ONE) rc = (m.rcR[ti] + m.rcR[safe_type((int)__double_as_longlong(pj[r].w), T)]) * 0.5;
          if (d2 < rc * rc) {
            jj[r] = w.map.owner ? w.map.owner[jc[r]] : jc[r];
            rr[r] = make_float4((float)x, (float)y, (float)z, __int_as_float(jj[r]));
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) tt[r] = (ONE || jj[r] < 0) ? 0 : safe_type(type[jj[r]], T);
    float bj[RB][KP], bo[ONE ? 1 : RB][KP];
#pragma unroll
    for (int r = 0; r < RB; ++r) {
#pragma unroll
      for (int k = 0; k < KP; ++k) bj[r][k] = 0.0f;
      if (jj[r] >= nd1 && jj[r] < nd2) load_row4<KP>(bR + ((size_t)jj[r] * T + ti) * kRp, kRp, bj[r]);
      if (!ONE) {
#pragma unroll
        for (int k = 0; k < KP; ++k) bo[ONE ? 0 : r][k] = 0.0f;
        if (jj[r] >= 0) load_row4<KP>(bR + ((size_t)i * T + tt[r]) * kRp, kRp, bo[ONE ? 0 : r]);
      }
    }
#pragma unroll
    for (int r = 0; r < RB; ++r) {
      const float4 q = rr[r];
      // lanes past the row end carry r = 0: d = 0 and 1/d = 0 there
      const float d2 = q.x * q.x + q.y * q.y + q.z * q.z;
      const float dinv = jj[r] >= 0 ? __builtin_amdgcn_rsqf(d2) : 0.0f;
      const float d = d2 * dinv;
      const float rc = ONE ? rcR0 : (float)((m.rcR[ti] + m.rcR[tt[r]]) * 0.5);
      const float rcinv = 1.0f / rc;
      float fc, fcp;
      cutoff_fc_fcp(rc, rcinv, d, fc, fcp);
      float f[KR], fp[KR];
      cheb_d<KR>(kR, rcinv, d, fc, fcp, f, fp);
      float Gi = 0.0f, Gj = 0.0f;
#pragma unroll
      for (int k = 0; k < KR; ++k) {
        Gi += (ONE ? bi[k] : bo[ONE ? 0 : r][k]) * fp[k];
        Gj += bj[r][k] * fp[k];
      }
      const float sc = (Gi + Gj) * dinv;  // zero past the row end
      fx += (double)(sc * q.x);
      fy += (double)(sc * q.y);
      fz += (double)(sc * q.z);
      // virial_i += r_ij (x) dE_j/dr_ji,  dE_j/dr_ji = -Gj r_ij/d
      const float gj = -Gj * dinv;
      const float fjx = gj * q.x, fjy = gj * q.y, fjz = gj * q.z;
      v[0] += (double)(q.x * fjx);
      v[1] += (double)(q.y * fjy);
      v[2] += (double)(q.z * fjz);
      v[3] += (double)(q.x * fjy);
      v[4] += (double)(q.x * fjz);
      v[5] += (double)(q.y * fjz);
      v[6] += (double)(q.y * fjx);
      v[7] += (double)(q.z * fjx);
      v[8] += (double)(q.z * fjy);
    }
  }
  // angular: F_i += dE_i/dr_ij - dE_j/dr_ji; the second term is read from
  // j's row at the slot where i sits (Verlet reverse position -> angular slot)
  constexpr int AB = 3;
  for (int e0 = lg; e0 < na; e0 += G * AB) {
    float4 rr[AB], aa[AB], bb[AB];
    int jj[AB], qv[AB], e2[AB];
#pragma unroll
    for (int r = 0; r < AB; ++r) {
      const int e = e0 + r * G;
      rr[r] = aa[r] = make_float4(0.f, 0.f, 0.f, 0.f);
      jj[r] = qv[r] = -1;
      if (e < na) {
        rr[r] = erow[e];
        aa[r] = frow[e];
        qv[r] = qrow[e];
        jj[r] = __float_as_int(rr[r].w);
      }
    }
    const int lim = w.rcap * w.MV;
#pragma unroll
    for (int r = 0; r < AB; ++r) {
      e2[r] = -1;
      if (jj[r] >= nd1 && jj[r] < nd2 && qv[r] >= 0 && qv[r] < lim) {
        if (CAND) {
          e2[r] = w.vsa[qv[r]];
        } else {
          // i's Verlet position q in j's row (qv = j MV + q); its angular
          // slot is the number of angular entries of j's row before it
          const int q = qv[r] - jj[r] * w.MV;
          const int mw = (w.MV + 31) >> 5;
          const unsigned* mk = w.amask + (size_t)jj[r] * mw;
          if (mw == 4) {
            // rows of up to 128 entries: the row's four words in one load
            const uint4 m4 = *reinterpret_cast<const uint4*>(mk);
            const unsigned wd[4] = {m4.x, m4.y, m4.z, m4.w};
            const int wq = q >> 5;
            const unsigned word = wd[wq];
            if ((word >> (q & 31)) & 1u) {
              int c = __popc(word & ((1u << (q & 31)) - 1u));
              c += (wq > 0 ? __popc(m4.x) : 0) + (wq > 1 ? __popc(m4.y) : 0) + (wq > 2 ? __popc(m4.z) : 0);
              e2[r] = c;
            }
          } else {
            const unsigned word = mk[q >> 5];
            if ((word >> (q & 31)) & 1u) {
              int c = __popc(word & ((1u << (q & 31)) - 1u));
              for (int k = 0; k < (q >> 5); ++k) c += __popc(mk[k]);
              e2[r] = c;
            }
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < AB; ++r) {
      bb[r] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (e2[r] >= 0 && e2[r] < w.MA) bb[r] = w.f12[(size_t)jj[r] * w.MA + e2[r]];
    }
#pragma unroll
    for (int r = 0; r < AB; ++r) {
      const float4 q = rr[r];
      const bool jdep = jj[r] >= nd1 && jj[r] < nd2;
      float gx = aa[r].x - bb[r].x, gy = aa[r].y - bb[r].y, gz = aa[r].z - bb[r].z;
      float bx = bb[r].x, by = bb[r].y, bz = bb[r].z;
      if (m.zbl && jj[r] >= 0) {
        // each directed pair carries half of the ZBL pair gradient
        // (nep_utilities.cuh:347-433 switch, nep.cu ZBL pass)
        const int tj = ONE ? 0 : safe_type(type[jj[r]], T);
        const float d = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z);
        const float dinv = 1.0f / d;
        float zf, zfp;
        zbl_pair(m.zblp + (size_t)(ti * m.T + tj) * 12, d, dinv, zf, zfp);
        const float h = 0.5f * zfp * dinv;  // dE_i/dr_ij = h r_ij
        const float c = jdep ? 2.0f * h : h;
        gx += c * q.x;
        gy += c * q.y;
        gz += c * q.z;
        if (jdep) {  // dE_j/dr_ji = -h r_ij
          bx -= h * q.x;
          by -= h * q.y;
          bz -= h * q.z;
        }
      }
      fx += (double)gx;
      fy += (double)gy;
      fz += (double)gz;
      v[0] += (double)(q.x * bx);
      v[1] += (double)(q.y * by);
      v[2] += (double)(q.z * bz);
      v[3] += (double)(q.x * by);
      v[4] += (double)(q.x * bz);
      v[5] += (double)(q.y * bz);
      v[6] += (double)(q.y * bx);
      v[7] += (double)(q.z * bx);
      v[8] += (double)(q.z * by);
    }
  }
#pragma unroll
  for (int off = 1; off < G; off <<= 1) {
    fx += __shfl_xor(fx, off, G);
    fy += __shfl_xor(fy, off, G);
    fz += __shfl_xor(fz, off, G);
#pragma unroll
    for (int c = 0; c < 9; ++c) v[c] += __shfl_xor(v[c], off, G);
  }
  if (!active || lg != 0) return;
  if (accumulate) {
    force[i] += fx;
    force[N + i] += fy;
    force[2 * N + i] += fz;
#pragma unroll
    for (int c = 0; c < 9; ++c) virial[(size_t)c * N + i] += v[c];
  } else {
    force[i] = fx;
    force[N + i] = fy;
    force[2 * N + i] = fz;
#pragma unroll
    for (int c = 0; c < 9; ++c) virial[(size_t)c * N + i] = v[c];
  }
}

// ---------------------------------------------------------------------------
// integrator and thermo (FP64)
// ---------------------------------------------------------------------------

// (the first half also adds the drift to `unwrapped` when given:
// update_unwrapped, runtime.cu:423-440)
__global__ void k_velocity_verlet(int first_half, int N, double dt, const double* __restrict__ mass,
                                  double* __restrict__ pos, double* __restrict__ vel,
                                  const double* __restrict__ force, double* __restrict__ unwrapped)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double half = dt * 0.5;
  const double im = 1.0 / mass[i];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double v = vel[c * N + i] + force[c * N + i] * im * half;
    vel[c * N + i] = v;
    if (first_half) {
      const double x0 = pos[c * N + i], x1 = x0 + v * dt;
      pos[c * N + i] = x1;
      if (unwrapped) unwrapped[c * N + i] += x1 - x0;
    }
  }
}

__global__ void k_wrap(DevBox b, int N, double* __restrict__ pos)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double x = pos[i], y = pos[N + i], z = pos[2 * N + i];
  double s[3] = {b.hi[0] * x + b.hi[1] * y + b.hi[2] * z, b.hi[3] * x + b.hi[4] * y + b.hi[5] * z,
                 b.hi[6] * x + b.hi[7] * y + b.hi[8] * z};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (!b.pbc[d]) continue;
    if (s[d] < 0.0) s[d] += 1.0;
    else if (s[d] > 1.0) s[d] -= 1.0;
  }
  pos[i] = b.h[0] * s[0] + b.h[1] * s[1] + b.h[2] * s[2];
  pos[N + i] = b.h[3] * s[0] + b.h[4] * s[1] + b.h[5] * s[2];
  pos[2 * N + i] = b.h[6] * s[0] + b.h[7] * s[1] + b.h[8] * s[2];
}

// First half of a velocity-Verlet step of the MD loop in one kernel: the
// previous step's pending thermostat factor (v *= f), the half kick and
// drift, the periodic wrap, the step's packed (x, y, z, type) record and the
// displacement test against the last rebuild.  The FP64 expressions are
// those of k_scale_by, k_velocity_verlet, k_wrap and the neighbour pack, in
// that order and without contraction, so the results are bit-identical to
// running them as separate kernels.
__global__ __launch_bounds__(kBlock) void k_md_first(DevBox b, int N, int T, double dt,
                                                     const double* __restrict__ mass, double* __restrict__ pos,
                                                     double* __restrict__ vel, const double* __restrict__ force,
                                                     const double* __restrict__ factor,
                                                     const int* __restrict__ type, const double* __restrict__ x0,
                                                     double lim2, int forced, double4* __restrict__ xq,
                                                     int* __restrict__ flags, double* __restrict__ unwrapped)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (forced && i == 0) flags[kFlagRebuild] = kRebuildAsked | kRebuildAll;
  bool moved = false, bad = false;
  if (i < N) {
    const double half = dt * 0.5;
    const double im = 1.0 / mass[i];
    double x[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double v = vel[c * N + i];
      if (factor) v *= factor[0];
      v = v + force[c * N + i] * im * half;
      vel[c * N + i] = v;
      const double p0 = pos[c * N + i];
      x[c] = p0 + v * dt;
      if (unwrapped) unwrapped[c * N + i] += x[c] - p0;
    }
    double sf[3] = {b.hi[0] * x[0] + b.hi[1] * x[1] + b.hi[2] * x[2], b.hi[3] * x[0] + b.hi[4] * x[1] + b.hi[5] * x[2],
                    b.hi[6] * x[0] + b.hi[7] * x[1] + b.hi[8] * x[2]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (!b.pbc[d]) continue;
      if (sf[d] < 0.0) sf[d] += 1.0;
      else if (sf[d] > 1.0) sf[d] -= 1.0;
    }
    const double px = b.h[0] * sf[0] + b.h[1] * sf[1] + b.h[2] * sf[2];
    const double py = b.h[3] * sf[0] + b.h[4] * sf[1] + b.h[5] * sf[2];
    const double pz = b.h[6] * sf[0] + b.h[7] * sf[1] + b.h[8] * sf[2];
    pos[i] = px;
    pos[N + i] = py;
    pos[2 * N + i] = pz;
    const int t = type[i];
    bad = t < 0 || t >= T;
    xq[i] = make_double4(px, py, pz, __longlong_as_double((long long)t));
    if (!forced) {
      double dx = px - x0[i], dy = py - x0[N + i], dz = pz - x0[2 * N + i];
      mic(b, dx, dy, dz);
      moved = !(dist2(dx, dy, dz) <= lim2);
    }
  }
  const bool lead = (threadIdx.x & 63) == 0;
  if (__any(moved) && lead) atomicOr(&flags[kFlagRebuild], kRebuildAsked);
  if (__any(bad) && lead) atomicOr(&flags[kFlagBadType], 1);
}

// error-free two-sum accumulation (s, c) += x
struct KSum {
  double s, c;
};
__device__ __forceinline__ void ksum_add(KSum& k, double x)
{
  const double t = k.s + x;
  const double bp = t - k.s;
  const double err = (k.s - (t - bp)) + (x - bp);
  k.s = t;
  k.c += err;
}
__device__ __forceinline__ void ksum_merge(KSum& a, const KSum& b)
{
  ksum_add(a, b.s);
  a.c += b.c;
}

constexpr int kThermoThreads = 256;

int thermo_blocks(int N)
{
  int nb = (N + kThermoThreads - 1) / kThermoThreads;
  return nb < 1 ? 1 : (nb > 512 ? 512 : nb);
}

// compensated reduction of 8 quantities over a 256-thread workgroup: wave
// shuffles, then the 4 wave results through LDS; the total lands in thread 0
__device__ __forceinline__ void block_reduce8(KSum* acc)
{
#pragma unroll
  for (int q = 0; q < 8; ++q)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      KSum o{__shfl_down(acc[q].s, off), __shfl_down(acc[q].c, off)};
      ksum_merge(acc[q], o);
    }
  __shared__ double sh[kThermoThreads / 64][8][2];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      sh[wave][q][0] = acc[q].s;
      sh[wave][q][1] = acc[q].c;
    }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 8; ++q)
      for (int w = 1; w < kThermoThreads / 64; ++w) ksum_merge(acc[q], KSum{sh[w][q][0], sh[w][q][1]});
}

// velocity factor from this step's normalised temperature, formed where the
// thermo lives (same expressions as Controller::after_force, no contraction)
__device__ __forceinline__ double ensemble_factor(int kind, int N, double coupling, double t_target,
                                                  const double* __restrict__ rand2, double T)
{
#pragma clang fp contract(off)
  double f = 1.0;
  if (kind == 1) {
    f = sqrt(1.0 + coupling * (t_target / T - 1.0));
  } else if (kind == 2) {
    // resamplekin (thermostat_svr.cpp:107-124) with rr and the chi-square sum drawn on the host
    const int ndeg = 3 * N;
    const double kk = T * (ndeg * 8.617343e-5 * 0.5);  // old_KE *= ndeg * K_B * 0.5
    const double sigma = ndeg * 8.617343e-5 * t_target * 0.5;
    const double fac = coupling;  // exp(-1/taut) or 0, formed on the host
    const double rr = rand2[0], sum = rand2[1];
    const double knew = kk + (1.0 - fac) * (sigma * (sum + rr * rr) / ndeg - kk) +
                        2.0 * rr * sqrt(kk * sigma / ndeg * (1.0 - fac) * fac);
    f = sqrt(knew / kk);
  }
  return f;
}

// per-workgroup compensated partial sums of the 8 thermo quantities
// (find_owned_thermo_sums, runtime.cu:442-493)
__global__ __launch_bounds__(kThermoThreads) void k_thermo_partial(int N, const double* __restrict__ mass,
                                                                   const double* __restrict__ pe,
                                                                   const double* __restrict__ vel,
                                                                   const double* __restrict__ vir,
                                                                   double* __restrict__ partial)
{
#pragma clang fp contract(off)
  KSum acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const double m = mass[i];
    const double vx = vel[i], vy = vel[N + i], vz = vel[2 * N + i];
    ksum_add(acc[0], m * (vx * vx + vy * vy + vz * vz));
    ksum_add(acc[1], pe[i]);
    ksum_add(acc[2], vir[i] + m * vx * vx);
    ksum_add(acc[3], vir[N + i] + m * vy * vy);
    ksum_add(acc[4], vir[2 * N + i] + m * vz * vz);
    ksum_add(acc[5], vir[3 * N + i] + m * vx * vy);
    ksum_add(acc[6], vir[4 * N + i] + m * vx * vz);
    ksum_add(acc[7], vir[5 * N + i] + m * vy * vz);
  }
  block_reduce8(acc);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      partial[(blockIdx.x * 8 + q) * 2 + 0] = acc[q].s;
      partial[(blockIdx.x * 8 + q) * 2 + 1] = acc[q].c;
    }
}

// Second half kick of the MD step fused with the per-workgroup thermo
// partial sums of the kicked velocities (the sums of k_thermo_partial).
template <bool THERMO>
__global__ __launch_bounds__(kThermoThreads) void k_md_second(int N, double dt, const double* __restrict__ mass,
                                                              double* __restrict__ vel,
                                                              const double* __restrict__ force,
                                                              const double* __restrict__ pe,
                                                              const double* __restrict__ vir,
                                                              double* __restrict__ partial)
{
#pragma clang fp contract(off)
  KSum acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = {0.0, 0.0};
  const double half = dt * 0.5;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x) {
    const double m = mass[i];
    const double im = 1.0 / m;
    double v3[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double v = vel[c * N + i] + force[c * N + i] * im * half;
      vel[c * N + i] = v;
      v3[c] = v;
    }
    if (THERMO) {
      const double vx = v3[0], vy = v3[1], vz = v3[2];
      ksum_add(acc[0], m * (vx * vx + vy * vy + vz * vz));
      ksum_add(acc[1], pe[i]);
      ksum_add(acc[2], vir[i] + m * vx * vx);
      ksum_add(acc[3], vir[N + i] + m * vy * vy);
      ksum_add(acc[4], vir[2 * N + i] + m * vz * vz);
      ksum_add(acc[5], vir[3 * N + i] + m * vx * vy);
      ksum_add(acc[6], vir[4 * N + i] + m * vx * vz);
      ksum_add(acc[7], vir[5 * N + i] + m * vy * vz);
    }
  }
  if (THERMO) {
    block_reduce8(acc);
    if (threadIdx.x == 0)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        partial[(blockIdx.x * 8 + q) * 2 + 0] = acc[q].s;
        partial[(blockIdx.x * 8 + q) * 2 + 1] = acc[q].c;
      }
  }
}

// normalize_global_thermo (runtime_internal.hpp:704-716): T = sum m v^2/(3 N kB), stress/V
// With step_index, the step's target temperature and draws are
// targets[k] and rand2[2k .. 2k+1] for k = *step_index, which the kernel then
// advances: the same launch serves every step of a chunk (a replayed graph).
__global__ __launch_bounds__(kThermoThreads) void k_thermo_final(int nb, int N, double volume,
                                                                 const double* __restrict__ partial,
                                                                 double* __restrict__ thermo, int kind,
                                                                 double coupling, double t_target,
                                                                 const double* __restrict__ rand2,
                                                                 double* __restrict__ factor,
                                                                 const double* __restrict__ targets,
                                                                 int* __restrict__ step_index)
{
  KSum acc[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) acc[q] = {0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int q = 0; q < 8; ++q) ksum_merge(acc[q], KSum{partial[(b * 8 + q) * 2], partial[(b * 8 + q) * 2 + 1]});
  block_reduce8(acc);
  if (threadIdx.x == 0) {
    if (step_index) {
      const int k = *step_index;
      t_target = targets[k];
      rand2 += 2 * k;
      *step_index = k + 1;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      double v = acc[q].s + acc[q].c;
      if (q == 0) v /= 3.0 * N * 8.617343e-5;
      else if (q >= 2) v /= volume;
      thermo[q] = v;
      if (q == 0 && factor) factor[0] = ensemble_factor(kind, N, coupling, t_target, rand2, v);
    }
  }
}

__global__ void k_chunk_start(double* __restrict__ factor, int* __restrict__ step_index)
{
  factor[0] = 1.0;
  *step_index = 0;
}

__global__ void k_ensemble_factor(int kind, int N, double coupling, double t_target,
                                  const double* __restrict__ rand2, const double* __restrict__ thermo,
                                  double* __restrict__ factor)
{
  factor[0] = ensemble_factor(kind, N, coupling, t_target, rand2, thermo[0]);
}

__global__ void k_scale_by(int n, const double* __restrict__ factor, double* __restrict__ data)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) data[i] *= factor[0];
}

__global__ void k_scale(int n, double factor, double* __restrict__ data)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) data[i] *= factor;
}

__global__ void k_check_scaled(int n, double factor, const double* __restrict__ data, int* __restrict__ flag)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !isfinite(data[i] * factor)) atomicOr(flag, 1);
}

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------

static inline int nblk(int n) { return (n + kBlock - 1) / kBlock; }

// lanes per centre (8, 16 or 32) and the Verlet entries each lane loads
// per pass (7, 7 or 4)
// closes a rebuild request as the filter does, for an evaluation whose
// first filter launch has no centres
__global__ void k_close_request(int* __restrict__ flags)
{
  if (flags[kFlagRebuild]) {
    flags[kFlagRebuild] = 0;
    flags[kFlagRebuildCount] += 1;
  }
}

template <int G, int RB, bool CAND = false>
static void run_filter(const DevBox& b, const DevModel& m, int c1, int c2, int close, DevWork& w, hipStream_t s)
{
  constexpr int per = kBlock / G;
  const int nb = (c2 - c1 + per - 1) / per;
  if (w.map.owner) {
    if (m.T == 1) k_filter<true, true, G, RB, CAND><<<xcd_grid(nb), kBlock, 0, s>>>(b, m, c1, c2, nb, close, w);
    else k_filter<false, true, G, RB, CAND><<<xcd_grid(nb), kBlock, 0, s>>>(b, m, c1, c2, nb, close, w);
  } else {
    if (m.T == 1) k_filter<true, false, G, RB, CAND><<<xcd_grid(nb), kBlock, 0, s>>>(b, m, c1, c2, nb, close, w);
    else k_filter<false, false, G, RB, CAND><<<xcd_grid(nb), kBlock, 0, s>>>(b, m, c1, c2, nb, close, w);
  }
  if (CAND) k_split_reverse<G><<<xcd_grid(nb), kBlock, 0, s>>>(c1, c2, nb, w);
}

hipError_t launch_split(const DevBox& b, const DevModel& m, int c1, int c2, DevWork& w, hipStream_t s, int lanes)
{
  if (!w.split || !w.cl) return hipErrorInvalidValue;
  if (c2 > c1) {
    if (lanes == 8) run_filter<8, kFilterBatch, true>(b, m, c1, c2, 0, w, s);
    else if (lanes == 32) run_filter<32, 4, true>(b, m, c1, c2, 0, w, s);
    else run_filter<kFilterLanes, kFilterBatch, true>(b, m, c1, c2, 0, w, s);
  } else {
    k_close_request<<<1, 1, 0, s>>>(w.flags);  // no descriptor launch closes it
  }
  return hipGetLastError();
}

hipError_t launch_filter(const DevBox& b, const DevModel& m, int c1, int c2, DevWork& w, hipStream_t s, int lanes,
                         bool close)
{
  if (c2 > c1) {
    const int cl = close ? 1 : 0;
    if (lanes == 8) run_filter<8, kFilterBatch>(b, m, c1, c2, cl, w, s);
    else if (lanes == 32) run_filter<32, 4>(b, m, c1, c2, cl, w, s);
    else run_filter<kFilterLanes, kFilterBatch>(b, m, c1, c2, cl, w, s);
  } else if (close) {
    k_close_request<<<1, 1, 0, s>>>(w.flags);
  }
  return hipGetLastError();
}

template <class C>
static void run_descriptor(const DevBox& b, const DevModel& m, int N, int c1, int c2, const int* type, double* pe,
                           bool acc, DevWork& w, hipStream_t s, hipEvent_t e0, hipEvent_t e1, int lanes = 1)
{
  if constexpr (C::ann == kAnnValu) {
    if (lanes == 4 && !w.split) {
      constexpr int per = kDescBlock / 4;
      const int nb = (c2 - c1 + per - 1) / per;
      auto k = k_descriptor<C, false, 4>;
      if (e0 || e1)
        hipExtLaunchKernelGGL(k, dim3(xcd_grid(nb)), dim3(kDescBlock), 0, s, e0, e1, 0, b, m, N, c1, c2, nb, type,
                              pe, w.edge, w.vsa, w.eq, acc ? 1 : 0, w);
      else
        k<<<xcd_grid(nb), kDescBlock, 0, s>>>(b, m, N, c1, c2, nb, type, pe, w.edge, w.vsa, w.eq, acc ? 1 : 0, w);
      return;
    }
  }
  const int nb = (c2 - c1 + kDescBlock - 1) / kDescBlock;
  auto k = w.split ? k_descriptor<C, true> : k_descriptor<C, false>;
  if (e0 || e1)
    hipExtLaunchKernelGGL(k, dim3(xcd_grid(nb)), dim3(kDescBlock), 0, s, e0, e1, 0, b, m, N, c1, c2, nb, type, pe,
                          w.edge, w.vsa, w.eq, acc ? 1 : 0, w);
  else
    k<<<xcd_grid(nb), kDescBlock, 0, s>>>(b, m, N, c1, c2, nb, type, pe, w.edge, w.vsa, w.eq, acc ? 1 : 0, w);
}

// single-type exact configuration with the ANN mode picked at run time:
// matrix cores for H <= 64 (FP32 or BF16 operands), vector ALUs otherwise
template <int NR, int KR, int NA, int KA, bool Q2, bool Q1>
static void run_single(const DevBox& b, const DevModel& m, int N, int c1, int c2, const int* type, double* pe, bool acc, int ann,
                       DevWork& w, hipStream_t s, hipEvent_t e0, hipEvent_t e1, int lanes)
{
  const int mode = (m.H <= 64 && !(lanes == 4 && !w.split)) ? ann : kAnnValu;
  if constexpr (mfma_ann_fits<NR, KR, NA, KA, Q2, Q1, true>()) {
    if (mode == kAnnMfma32) {
      run_descriptor<Cfg<NR, KR, NA, KA, 4, Q2, Q1, true, true, kAnnMfma32>>(b, m, N, c1, c2, type, pe, acc, w, s, e0,
                                                                            e1);
      return;
    }
    if (mode == kAnnMfmaBf16) {
      run_descriptor<Cfg<NR, KR, NA, KA, 4, Q2, Q1, true, true, kAnnMfmaBf16>>(b, m, N, c1, c2, type, pe, acc, w, s, e0,
                                                                              e1);
      return;
    }
    if constexpr (mfma_ann_fits<NR, KR, NA, KA, Q2, Q1, true, kAnnMfmaSplit>()) {
      if (mode == kAnnMfmaSplit) {
        run_descriptor<Cfg<NR, KR, NA, KA, 4, Q2, Q1, true, true, kAnnMfmaSplit>>(b, m, N, c1, c2, type, pe, acc, w, s,
                                                                                  e0, e1);
        return;
      }
    }
  }
  run_descriptor<Cfg<NR, KR, NA, KA, 4, Q2, Q1, true, true, kAnnValu>>(b, m, N, c1, c2, type, pe, acc, w, s, e0, e1,
                                                                      lanes);
}

template <int NR, int KR, int NA, int KA, bool Q2, bool Q1>
static void run_multi(const DevBox& b, const DevModel& m, int N, int c1, int c2, const int* type, double* pe, bool acc, int ann,
                      DevWork& w, hipStream_t s, hipEvent_t e0, hipEvent_t e1, int lanes)
{
  const int mode = (m.H <= 64 && m.T <= kTypesMfma && !(lanes == 4 && !w.split)) ? ann : kAnnValu;
  if constexpr (mfma_ann_fits<NR, KR, NA, KA, Q2, Q1, false>()) {
    if (mode == kAnnMfma32) {
      run_descriptor<Cfg<NR, KR, NA, KA, 4, Q2, Q1, false, true, kAnnMfma32, true>>(b, m, N, c1, c2, type, pe, acc, w,
                                                                                   s, e0, e1);
      return;
    }
    if (mode == kAnnMfmaBf16) {
      run_descriptor<Cfg<NR, KR, NA, KA, 4, Q2, Q1, false, true, kAnnMfmaBf16, true>>(b, m, N, c1, c2, type, pe, acc,
                                                                                     w, s, e0, e1);
      return;
    }
    if constexpr (mfma_ann_fits<NR, KR, NA, KA, Q2, Q1, false, kAnnMfmaSplit>()) {
      if (mode == kAnnMfmaSplit) {
        run_descriptor<Cfg<NR, KR, NA, KA, 4, Q2, Q1, false, true, kAnnMfmaSplit, true>>(b, m, N, c1, c2, type, pe,
                                                                                         acc, w, s, e0, e1);
        return;
      }
    }
  }
  if (m.T <= kTypesLds)
    run_descriptor<Cfg<NR, KR, NA, KA, 4, Q2, Q1, false, true, kAnnValu, true>>(b, m, N, c1, c2, type, pe, acc, w, s,
                                                                               e0, e1, lanes);
  else
    run_descriptor<Cfg<NR, KR, NA, KA, 4, Q2, Q1, false, true>>(b, m, N, c1, c2, type, pe, acc, w, s, e0, e1, lanes);
}

namespace {
bool model_is(const DevModel& m, int nr, int kr, int na, int ka, int L, bool a, bool b, bool one)
{
  return m.nR == nr && m.kR == kr && m.nA == na && m.kA == ka && m.L == L && (m.has222 != 0) == a &&
         (m.has1111 != 0) == b && (m.T == 1) == one;
}

// sizes of the exactly instantiated configurations (nR, kR, nA, kA, q222,
// q1111, single type); the bounded sizes below come after them
struct ExactCfg {
  int nr, kr, na, ka;
  bool q2, q1, one;
};
constexpr ExactCfg kExact[] = {
    {5, 9, 5, 9, true, true, true},     {7, 9, 5, 9, true, false, true},   {13, 17, 9, 13, true, true, true},
    {7, 11, 7, 11, true, true, false},  {5, 9, 5, 9, true, false, false},  {9, 9, 7, 9, false, false, false},
    {7, 9, 5, 9, true, true, true},     {7, 9, 5, 9, true, true, false},   {9, 13, 7, 13, true, true, true},
    {9, 13, 7, 13, true, true, false}};
constexpr int kBuckets[][4] = {{7, 9, 5, 9}, {9, 13, 7, 13}};
}  // namespace

bool descriptor_bucket(const DevModel& m, int dims[4])
{
  for (const ExactCfg& c : kExact)
    if (model_is(m, c.nr, c.kr, c.na, c.ka, 4, c.q2, c.q1, c.one)) return false;
  if (m.L > 4) return false;
  for (const auto& b : kBuckets)
    if (m.nR <= b[0] && m.kR <= b[1] && m.nA <= b[2] && m.kA <= b[3]) {
      for (int k = 0; k < 4; ++k) dims[k] = b[k];
      return true;
    }
  return false;
}

hipError_t launch_descriptor(const DevBox& b, const DevModel& m, int N, int c1, int c2, const int* type, double* pe,
                             bool acc, int ann, DevWork& w, hipStream_t s, hipEvent_t e0, hipEvent_t e1, int lanes)
{
  if (lanes != 1 && lanes != 4) return hipErrorInvalidValue;
  if (c2 <= c1) return hipSuccess;
  const bool one = m.T == 1;
  const bool q2 = m.has222 != 0, q1 = m.has1111 != 0;
  auto is = [&](int nr, int kr, int na, int ka, int L, bool a, bool b, bool o) {
    return model_is(m, nr, kr, na, ka, L, a, b, o);
  };
  // exact instantiations: the configurations in use, and the bounded sizes
  // that other models are padded to (descriptor_bucket); generic otherwise
  if (is(5, 9, 5, 9, 4, true, true, true))
    run_single<5, 9, 5, 9, true, true>(b, m, N, c1, c2, type, pe, acc, ann, w, s, e0, e1, lanes);
  else if (is(7, 9, 5, 9, 4, true, false, true))
    run_single<7, 9, 5, 9, true, false>(b, m, N, c1, c2, type, pe, acc, ann, w, s, e0, e1, lanes);
  else if (is(13, 17, 9, 13, 4, true, true, true))
    run_descriptor<Cfg<13, 17, 9, 13, 4, true, true, true, true>>(b, m, N, c1, c2, type, pe, acc, w, s, e0, e1, lanes);
  else if (is(7, 11, 7, 11, 4, true, true, false))
    run_multi<7, 11, 7, 11, true, true>(b, m, N, c1, c2, type, pe, acc, ann, w, s, e0, e1, lanes);
  else if (is(5, 9, 5, 9, 4, true, false, false))
    run_multi<5, 9, 5, 9, true, false>(b, m, N, c1, c2, type, pe, acc, ann, w, s, e0, e1, lanes);
  else if (is(9, 9, 7, 9, 4, false, false, false))
    run_multi<9, 9, 7, 9, false, false>(b, m, N, c1, c2, type, pe, acc, ann, w, s, e0, e1, lanes);
  else if (is(7, 9, 5, 9, 4, true, true, true))
    run_single<7, 9, 5, 9, true, true>(b, m, N, c1, c2, type, pe, acc, ann, w, s, e0, e1, lanes);
  else if (is(7, 9, 5, 9, 4, true, true, false))
    run_multi<7, 9, 5, 9, true, true>(b, m, N, c1, c2, type, pe, acc, ann, w, s, e0, e1, lanes);
  else if (is(9, 13, 7, 13, 4, true, true, true))
    run_single<9, 13, 7, 13, true, true>(b, m, N, c1, c2, type, pe, acc, ann, w, s, e0, e1, lanes);
  else if (is(9, 13, 7, 13, 4, true, true, false))
    run_multi<9, 13, 7, 13, true, true>(b, m, N, c1, c2, type, pe, acc, ann, w, s, e0, e1, lanes);
  else if (m.L <= kPadL && m.nR <= kMaxNR && m.kR <= kMaxK && m.nA <= kMaxNA && m.kA <= kMaxK)
    run_descriptor<Cfg<kMaxNR, kMaxK, kMaxNA, kMaxK, kPadL, true, true, false, false>>(b, m, N, c1, c2, type, pe, acc,
                                                                                      w, s, e0, e1, lanes);
  else if (m.L <= kMaxL && m.nR <= kMaxNR && m.kR <= kMaxK && m.nA <= kMaxNA && m.kA <= kMaxK)
    // 3-body orders 5..8 (80 angular moments per radial function): the
    // generic instance at order 8, its per-atom arrays partly in scratch
    run_descriptor<Cfg<kMaxNR, kMaxK, kMaxNA, kMaxK, kMaxL, true, true, false, false>>(b, m, N, c1, c2, type, pe, acc,
                                                                                      w, s, e0, e1, lanes);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int G>
static void run_gather(const DevBox& b, const DevModel& m, int N, int n1, int n2, int nd1, int nd2, const int* type,
                       double* force, double* virial, int a, DevWork& w, hipStream_t s)
{
  constexpr int per = kBlock / G;
  const int nb = (n2 - n1 + per - 1) / per;
  const int g = xcd_grid(nb);
  auto pick = [&](auto k_rows, auto k_cand) { return w.split ? k_cand : k_rows; };
  decltype(&k_gather<9, true, G, false>) k;
  if (m.T == 1 && m.kR == 9) k = pick(k_gather<9, true, G, false>, k_gather<9, true, G, true>);
  else if (m.T == 1) k = pick(k_gather<kMaxK, true, G, false>, k_gather<kMaxK, true, G, true>);
  else if (m.kR <= 11) k = pick(k_gather<11, false, G, false>, k_gather<11, false, G, true>);
  else k = pick(k_gather<kMaxK, false, G, false>, k_gather<kMaxK, false, G, true>);
  k<<<g, kBlock, 0, s>>>(b, m, n1, n2, nd1, nd2, nb, type, N, force, virial, a, w);
}

hipError_t launch_gather(const DevBox& b, const DevModel& m, int N, int n1, int n2, int nd1, int nd2,
                         const int* type, double* force, double* virial, bool acc, DevWork& w, hipStream_t s,
                         int lanes)
{
  if (n2 <= n1) return hipSuccess;
  const int a = acc ? 1 : 0;
  if (lanes == 4) run_gather<4>(b, m, N, n1, n2, nd1, nd2, type, force, virial, a, w, s);
  else if (lanes == 16) run_gather<16>(b, m, N, n1, n2, nd1, nd2, type, force, virial, a, w, s);
  else run_gather<kGatherLanes>(b, m, N, n1, n2, nd1, nd2, type, force, virial, a, w, s);
  return hipGetLastError();
}

hipError_t launch_velocity_verlet(int first_half, int N, double dt, const double* mass, double* pos,
                                  double* vel, const double* force, hipStream_t s, double* unwrapped)
{
  if (N > 0) k_velocity_verlet<<<nblk(N), kBlock, 0, s>>>(first_half, N, dt, mass, pos, vel, force, unwrapped);
  return hipGetLastError();
}

hipError_t launch_wrap(const DevBox& b, int N, double* pos, hipStream_t s)
{
  if (N > 0) k_wrap<<<nblk(N), kBlock, 0, s>>>(b, N, pos);
  return hipGetLastError();
}

hipError_t launch_thermo(int N, double volume, const double* mass, const double* pe, const double* vel,
                         const double* virial, DevWork& w, hipStream_t s)
{
  const int nb = thermo_blocks(N);
  k_thermo_partial<<<nb, kThermoThreads, 0, s>>>(N, mass, pe, vel, virial, w.partial);
  k_thermo_final<<<1, kThermoThreads, 0, s>>>(nb, N, volume, w.partial, w.thermo, 0, 0.0, 0.0, nullptr, nullptr,
                                              nullptr, nullptr);
  return hipGetLastError();
}

hipError_t launch_scale(int n, double factor, double* data, hipStream_t s)
{
  if (n > 0) k_scale<<<nblk(n), kBlock, 0, s>>>(n, factor, data);
  return hipGetLastError();
}

hipError_t launch_ensemble_scale(int kind, int N, double coupling, double t_target, const double* rand2,
                                 double* factor, const double* thermo, double* vel, hipStream_t s)
{
  k_ensemble_factor<<<1, 1, 0, s>>>(kind, N, coupling, t_target, rand2, thermo, factor);
  k_scale_by<<<nblk(3 * N), kBlock, 0, s>>>(3 * N, factor, vel);
  return hipGetLastError();
}

hipError_t launch_md_first(const DevBox& b, int N, int T, double dt, const double* mass, double* pos, double* vel,
                           const double* force, const double* factor, const int* type, bool forced, DevWork& w,
                           hipStream_t s, double* unwrapped, double half_skin)
{
  const double half = half_skin > 0.0 ? half_skin : 0.5 * kSkin;
  if (N > 0)
    k_md_first<<<nblk(N), kBlock, 0, s>>>(b, N, T, dt, mass, pos, vel, force, factor, type, w.x0, half * half,
                                          forced ? 1 : 0, w.xq, w.flags, unwrapped);
  return hipGetLastError();
}

hipError_t launch_md_second(int N, double dt, const double* mass, double* vel, const double* force, const double* pe,
                            const double* virial, bool thermo, DevWork& tw, hipStream_t s)
{
  if (N <= 0) return hipSuccess;
  const int nb = thermo_blocks(N);
  if (thermo)
    k_md_second<true><<<nb, kThermoThreads, 0, s>>>(N, dt, mass, vel, force, pe, virial, tw.partial);
  else
    k_md_second<false><<<nb, kThermoThreads, 0, s>>>(N, dt, mass, vel, force, pe, virial, tw.partial);
  return hipGetLastError();
}

hipError_t launch_thermo_final(int N, double volume, int kind, double coupling, double t_target,
                               const double* rand2, double* factor, DevWork& tw, hipStream_t s,
                               const double* targets, int* step_index)
{
  k_thermo_final<<<1, kThermoThreads, 0, s>>>(thermo_blocks(N), N, volume, tw.partial, tw.thermo, kind, coupling,
                                                t_target, rand2, kind ? factor : nullptr, targets, step_index);
  return hipGetLastError();
}

hipError_t launch_chunk_start(double* factor, int* step_index, hipStream_t s)
{
  k_chunk_start<<<1, 1, 0, s>>>(factor, step_index);
  return hipGetLastError();
}

hipError_t launch_scale_by(int n, const double* factor, double* data, hipStream_t s)
{
  if (n > 0) k_scale_by<<<nblk(n), kBlock, 0, s>>>(n, factor, data);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// small boxes: periodic replicas and the fold of their results
// ---------------------------------------------------------------------------

// Replica k = (ia, ib, ic) of atom i lands at slot k*N + i, shifted by
// ia*a + ib*b + ic*c (the box columns); replica 0 is the original.
__global__ void k_replicate(int N, int r0, int r1, int r2, DevBox b, const double* __restrict__ pos,
                            const int* __restrict__ type, const uint64_t* __restrict__ gid,
                            double* __restrict__ pos2, int* __restrict__ type2, uint64_t* __restrict__ gid2)
{
#pragma clang fp contract(off)
  const int R = r0 * r1 * r2;
  const long long n2 = (long long)N * R;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n2) return;
  const int k = (int)(t / N), i = (int)(t % N);
  const int ia = k % r0, ib = (k / r0) % r1, ic = k / (r0 * r1);
#pragma unroll
  for (int c = 0; c < 3; ++c)
    pos2[c * n2 + t] = pos[c * N + i] + (ia * b.h[3 * c] + ib * b.h[3 * c + 1] + ic * b.h[3 * c + 2]);
  type2[t] = type[i];
  if (gid) gid2[t] = gid[i];
}

// Copies (or adds) the first N entries of each component of the replicated
// outputs (stride n2) into the caller's arrays (stride N).
__global__ void k_fold(int N, long long n2, int accumulate, const double* __restrict__ pe2,
                       const double* __restrict__ f2, const double* __restrict__ v2, double* __restrict__ pe,
                       double* __restrict__ force, double* __restrict__ virial)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (accumulate) {
    pe[i] += pe2[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) force[c * N + i] += f2[c * n2 + i];
#pragma unroll
    for (int c = 0; c < 9; ++c) virial[(size_t)c * N + i] += v2[c * n2 + i];
  } else {
    pe[i] = pe2[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) force[c * N + i] = f2[c * n2 + i];
#pragma unroll
    for (int c = 0; c < 9; ++c) virial[(size_t)c * N + i] = v2[c * n2 + i];
  }
}

hipError_t launch_replicate(int N, const int reps[3], const DevBox& b, const double* pos, const int* type,
                            const uint64_t* gid, double* pos2, int* type2, uint64_t* gid2, hipStream_t s)
{
  const long long n2 = (long long)N * reps[0] * reps[1] * reps[2];
  if (n2 > 0)
    k_replicate<<<(unsigned)((n2 + kBlock - 1) / kBlock), kBlock, 0, s>>>(N, reps[0], reps[1], reps[2], b, pos,
                                                                        type, gid, pos2, type2, gid2);
  return hipGetLastError();
}

hipError_t launch_fold(int N, long long n2, bool accumulate, const double* pe2, const double* f2, const double* v2,
                       double* pe, double* force, double* virial, hipStream_t s)
{
  if (N > 0) k_fold<<<nblk(N), kBlock, 0, s>>>(N, n2, accumulate ? 1 : 0, pe2, f2, v2, pe, force, virial);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// serving loop: masked integrator, ghost refresh, relayout and extraction
// ---------------------------------------------------------------------------

// velocity-Verlet half step of the owned atoms whose request takes an MD
// step this iteration (the FP64 expressions of k_velocity_verlet)
__global__ void k_serve_kick(int first_half, int n, int Nt, const int* __restrict__ atom_req,
                             const int* __restrict__ req_md, const double* __restrict__ req_dt,
                             const double* __restrict__ mass, double* __restrict__ pos, double* __restrict__ vel,
                             const double* __restrict__ force)
{
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = atom_req[i];
  if (!req_md[r]) return;
  const double dt = req_dt[r];
  const double half = dt * 0.5;
  const double im = 1.0 / mass[i];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double v = vel[c * Nt + i] + force[c * Nt + i] * im * half;
    vel[c * Nt + i] = v;
    if (first_half) pos[c * Nt + i] += v * dt;
  }
}

// ghost k (slot n_owned + k) = its owner + a box shift
__global__ void k_serve_ghosts(int n_owned, int n_ghost, int Nt, const int* __restrict__ owner,
                               const double* __restrict__ shift, double* __restrict__ pos)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_ghost) return;
  const int g = n_owned + k, o = owner[k];
#pragma unroll
  for (int c = 0; c < 3; ++c) pos[c * Nt + g] = pos[c * Nt + o] + shift[3 * k + c];
}

// owned atoms of continuing requests move to their new slots (and regions)
// Assembly of a new serving layout, one thread per slot: the slot's request
// is found by binary search over the requests' first slots (both slot ranges
// are in request order and tile [0, n_owned) and [n_owned, Nt)).
__global__ void k_serve_assemble(int R, const ServeSlots* __restrict__ req, ServeLayout from, ServeLayout to,
                                 int* __restrict__ atom_req)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int Nt = to.Nt, no = to.n_owned;
  if (s == 0) to.gstart[no] = Nt;
  if (s >= Nt) return;
  const bool owned = s < no;
  int lo = 0, hi = R - 1;
  while (lo < hi) {  // last request whose range starts at or before s
    const int mid = (lo + hi + 1) >> 1;
    if ((owned ? req[mid].new_o : req[mid].new_g) <= s) lo = mid;
    else hi = mid - 1;
  }
  const ServeSlots q = req[lo];
  if (owned) {
    const int k = s - q.new_o;
    atom_req[s] = lo;
    to.owner[s] = s;
    to.image[s] = image_code(0, 0, 0);
    if (q.old_o >= 0) {
      const int src = q.old_o + k;
      to.type[s] = from.type[src];
      to.mass[s] = from.mass[src];
      to.gstart[s] = q.new_g + (from.gstart[src] - q.old_g);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        to.pos[c * Nt + s] = from.pos[c * from.Nt + src] + q.dtr[c];
        to.vel[c * Nt + s] = from.vel[c * from.Nt + src];
        to.force[c * Nt + s] = from.force[c * from.Nt + src];
        to.home[c * no + s] = from.home[c * from.n_owned + src] + q.dtr[c];
      }
    } else {
      to.type[s] = q.pi[k];
      to.gstart[s] = q.new_g + q.pi[q.n + k];
      to.mass[s] = q.pd[k];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double x = q.pd[q.n + 3 * k + c] + q.dtr[c];
        to.pos[c * Nt + s] = x;
        to.home[c * no + s] = x;
        to.vel[c * Nt + s] = q.pd[4 * q.n + 3 * k + c];
        to.force[c * Nt + s] = q.pd[7 * q.n + 3 * k + c];
      }
    }
  } else {
    const int k = s - q.new_g;
    const int g = s - no;
    if (q.old_g >= 0) {
      const int src = q.old_g + k;
      to.owner[s] = q.new_o + (from.owner[src] - q.old_o);
      to.image[s] = from.image[src];
      to.type[s] = from.type[src];
      to.mass[s] = from.mass[src];
#pragma unroll
      for (int c = 0; c < 3; ++c) to.gshift[3 * g + c] = from.gshift[3 * (src - from.n_owned) + c];
    } else {
      const int a = q.pi[2 * q.n + k];
      to.owner[s] = q.new_o + a;
      to.image[s] = q.pi[2 * q.n + q.ng + k];
      to.type[s] = q.pi[a];
      to.mass[s] = q.pd[a];
#pragma unroll
      for (int c = 0; c < 3; ++c) to.gshift[3 * g + c] = q.pd[10 * q.n + 3 * k + c];
    }
  }
}

// the request of slot s among the first R table rows (last whose range starts
// at or before s), by owned or ghost range
__device__ __forceinline__ int serve_request_of(int R, const ServeSlots* __restrict__ req, int s, bool owned)
{
  int lo = 0, hi = R - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((owned ? req[mid].new_o : req[mid].new_g) <= s) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

constexpr int kCarryLanes = 16;  // lanes per row in k_serve_carry_rows

__global__ void k_serve_carry_rows(int R, const ServeSlots* __restrict__ req, int old_no, int o_keep,
                                   const int* __restrict__ vl0, const int* __restrict__ vrev0,
                                   const int* __restrict__ vn0, DevWork w)
{
  const int s = blockIdx.x * (blockDim.x / kCarryLanes) + threadIdx.x / kCarryLanes;
  const int lane = threadIdx.x % kCarryLanes;
  if (s >= o_keep) return;
  const ServeSlots t = req[serve_request_of(R, req, s, true)];
  const int so = t.old_o + (s - t.new_o);
  const int n = vn0[so];
  if (lane == 0) w.vn[s] = n;
  const int m = min(n, w.MV);
  const size_t ro = (size_t)so * w.MV, rn = (size_t)s * w.MV;
  // every entry of a request's row is one of its own slots (requests are
  // disjoint in space): owned slots and ghosts each keep their offset
  for (int p = lane; p < m; p += kCarryLanes) {
    const int j = vl0[ro + p];
    w.vl[rn + p] = j < old_no ? t.new_o + (j - t.old_o) : t.new_g + (j - t.old_g);
    w.vrev[rn + p] = vrev0[ro + p];
  }
}

__global__ void k_serve_carry_x0(int R, const ServeSlots* __restrict__ req, int old_nt, int new_no, int new_nt,
                                 int o_keep, int g_keep, const double* __restrict__ x0_0, double* __restrict__ x0)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int total = o_keep + (g_keep - new_no);
  if (k >= total) return;
  const bool owned = k < o_keep;
  const int s = owned ? k : new_no + (k - o_keep);
  const ServeSlots t = req[serve_request_of(R, req, s, owned)];
  const int so = owned ? t.old_o + (s - t.new_o) : t.old_g + (s - t.new_g);
#pragma unroll
  for (int c = 0; c < 3; ++c) x0[(size_t)c * new_nt + s] = x0_0[(size_t)c * old_nt + so] + t.dtr[c];
}

__global__ void k_serve_drift(int n_owned, int Nt, const int* __restrict__ atom_req, const double* __restrict__ pos,
                              const double* __restrict__ home, double soft2, double hard2, int* __restrict__ req_flag)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n_owned) return;
  double d2 = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double d = pos[c * Nt + s] - home[c * n_owned + s];
    d2 += d * d;
  }
  if (!(d2 <= soft2)) atomicMax(&req_flag[atom_req[s]], (d2 <= hard2) ? 1 : 2);
}

// finished requests -> their result blocks, one thread per atom; the
// request is found by binary search over the blocks' first atoms
__global__ void k_serve_extract(int R, const ServeOut* __restrict__ req, long long no, int Nt,
                                const double* __restrict__ pos, const double* __restrict__ vel,
                                const double* __restrict__ force, const double* __restrict__ pe,
                                const double* __restrict__ vir, double* __restrict__ out)
{
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= no) return;
  int lo = 0, hi = R - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (req[mid].out / kServeOutDoubles <= k) lo = mid;
    else hi = mid - 1;
  }
  const ServeOut& q = req[lo];
  const int a = (int)(k - q.out / kServeOutDoubles), n = q.n, i = q.slot0 + a;
  double* blk = out + q.out;
  blk[a] = pe[i];
  double x[3], f[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    x[c] = pos[c * Nt + i] - q.trans[c];
    blk[n + c * n + a] = force[c * Nt + i];
    blk[7 * n + c * n + a] = vel[c * Nt + i];
  }
#pragma unroll
  for (int c = 0; c < 9; ++c) blk[10 * n + c * n + a] = vir[(size_t)c * Nt + i];
  // back into the request's box: fractional coordinates, wrapped where periodic
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    f[c] = q.hi[3 * c] * x[0] + q.hi[3 * c + 1] * x[1] + q.hi[3 * c + 2] * x[2];
    if (q.pbc[c]) f[c] -= floor(f[c]);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) blk[4 * n + c * n + a] = q.h[3 * c] * f[0] + q.h[3 * c + 1] * f[1] + q.h[3 * c + 2] * f[2];
}

hipError_t launch_serve_kick(int first_half, int n, int Nt, const int* atom_req, const int* req_md,
                             const double* req_dt, const double* mass, double* pos, double* vel, const double* force,
                             hipStream_t s)
{
  if (n > 0) k_serve_kick<<<nblk(n), kBlock, 0, s>>>(first_half, n, Nt, atom_req, req_md, req_dt, mass, pos, vel, force);
  return hipGetLastError();
}

hipError_t launch_serve_ghosts(int n_owned, int n_ghost, int Nt, const int* owner, const double* shift, double* pos,
                               hipStream_t s)
{
  if (n_ghost > 0) k_serve_ghosts<<<nblk(n_ghost), kBlock, 0, s>>>(n_owned, n_ghost, Nt, owner, shift, pos);
  return hipGetLastError();
}

hipError_t launch_serve_assemble(int R, const ServeSlots* req, const ServeLayout& from, const ServeLayout& to,
                                 int* atom_req, hipStream_t s)
{
  if (R > 0 && to.Nt > 0) k_serve_assemble<<<nblk(to.Nt), kBlock, 0, s>>>(R, req, from, to, atom_req);
  return hipGetLastError();
}

hipError_t launch_serve_carry_rows(int R_keep, const ServeSlots* req, int old_no, int old_nt, int new_no, int new_nt,
                                   int o_keep, int g_keep, const int* vl0, const int* vrev0, const int* vn0,
                                   const double* x0_0, DevWork& w, hipStream_t s)
{
  if (R_keep <= 0) return hipSuccess;
  if (o_keep > 0) {
    constexpr int per = kBlock / kCarryLanes;
    k_serve_carry_rows<<<(o_keep + per - 1) / per, kBlock, 0, s>>>(R_keep, req, old_no, o_keep, vl0, vrev0, vn0, w);
  }
  const int total = o_keep + (g_keep - new_no);
  if (total > 0)
    k_serve_carry_x0<<<nblk(total), kBlock, 0, s>>>(R_keep, req, old_nt, new_no, new_nt, o_keep, g_keep, x0_0, w.x0);
  return hipGetLastError();
}

hipError_t launch_serve_drift(int n_owned, int Nt, const int* atom_req, const double* pos, const double* home,
                              double soft2, double hard2, int* req_flag, hipStream_t s)
{
  if (n_owned > 0)
    k_serve_drift<<<nblk(n_owned), kBlock, 0, s>>>(n_owned, Nt, atom_req, pos, home, soft2, hard2, req_flag);
  return hipGetLastError();
}

hipError_t launch_serve_extract(int R, const ServeOut* req, long long no, int Nt, const double* pos,
                                const double* vel, const double* force, const double* pe, const double* vir,
                                double* out, hipStream_t s)
{
  if (R > 0 && no > 0)
    k_serve_extract<<<(unsigned)((no + kBlock - 1) / kBlock), kBlock, 0, s>>>(R, req, no, Nt, pos, vel, force, pe,
                                                                             vir, out);
  return hipGetLastError();
}

// largest |v|^2: a wavefront maximum, then one atomicMax per wavefront on
// the bit pattern (non-negative doubles order as their bits)
__global__ void k_max_speed2(int N, const double* __restrict__ vel, unsigned long long* __restrict__ out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v2 = 0.0;
  if (i < N) {
    const double x = vel[i], y = vel[N + i], z = vel[2 * N + i];
    v2 = x * x + y * y + z * z;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v2 = fmax(v2, __shfl_xor(v2, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(v2));
}

__global__ void k_clear_u64(unsigned long long* p) { *p = 0ull; }

hipError_t launch_max_speed2(int N, const double* vel, unsigned long long* out, hipStream_t s)
{
  k_clear_u64<<<1, 1, 0, s>>>(out);
  if (N > 0) k_max_speed2<<<nblk(N), kBlock, 0, s>>>(N, vel, out);
  return hipGetLastError();
}

hipError_t launch_check_scaled(int n, double factor, const double* data, int* flag, hipStream_t s)
{
  if (n > 0) k_check_scaled<<<nblk(n), kBlock, 0, s>>>(n, factor, data, flag);
  return hipGetLastError();
}

}  // namespace dmgx
