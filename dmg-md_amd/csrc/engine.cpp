// engine.cpp -- host orchestration of the NEP step (see engine.hpp).
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

namespace dmgx {

namespace {

constexpr double kCoulomb = 14.399645;  // K_C_SP, eV*A (common.cuh:57)
constexpr int kMdChunk = 128;                // steps enqueued between two reads of the status words

// covalent radii for the typewise ZBL cutoff (nep_utilities.cuh:175-186),
// a data table that has to be identical to the reference's
const double kCovalentRadius[94] = {
  0.426667, 0.613333, 1.6,     1.25333, 1.02667, 1.0,     0.946667, 0.84,    0.853333,
  0.893333, 1.86667,  1.66667, 1.50667, 1.38667, 1.46667, 1.36,     1.32,    1.28,
  2.34667,  2.05333,  1.77333, 1.62667, 1.61333, 1.46667, 1.42667,  1.38667, 1.33333,
  1.32,     1.34667,  1.45333, 1.49333, 1.45333, 1.53333, 1.46667,  1.52,    1.56,
  2.52,     2.22667,  1.96,    1.85333, 1.76,    1.65333, 1.53333,  1.50667, 1.50667,
  1.44,     1.53333,  1.64,    1.70667, 1.68,    1.68,    1.64,     1.76,    1.74667,
  2.78667,  2.34667,  2.16,    1.96,    2.10667, 2.09333, 2.08,     2.06667, 2.01333,
  2.02667,  2.01333,  2.0,     1.98667, 1.98667, 1.97333, 2.04,     1.94667, 1.82667,
  1.74667,  1.64,     1.57333, 1.54667, 1.48,    1.49333, 1.50667,  1.76,    1.73333,
  1.73333,  1.81333,  1.74667, 1.84,    1.89333, 2.68,    2.41333,  2.22667, 2.10667,
  2.02667,  2.04,     2.05333, 2.06667};

void check(hipError_t e, const char* what)
{
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// ANN arithmetic of a call: FP32 on the matrix cores unless split-bfloat16
// (32), bfloat16 (4) operands or the vector ALUs (16) are asked for
int ann_mode(unsigned flags)
{
  return (flags & 32u) ? 3 : (flags & 4u) ? 2 : (flags & 16u) ? 0 : 1;
}

template <class T>
void dalloc(T** p, size_t n, const char* what)
{
  if (*p) {
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (n == 0) n = 1;
  if (hipMalloc((void**)p, n * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    throw std::runtime_error(std::string("device allocation failed: ") + what);
  }
}

template <class T>
void dfree(T** p)
{
  if (*p) (void)hipFree(*p);
  *p = nullptr;
}

}  // namespace

void box_inverse(const double* h, double* hi)
{
  const double det = h[0] * (h[4] * h[8] - h[5] * h[7]) + h[1] * (h[5] * h[6] - h[3] * h[8]) +
                     h[2] * (h[3] * h[7] - h[4] * h[6]);
  hi[0] = (h[4] * h[8] - h[5] * h[7]) / det;
  hi[1] = (h[2] * h[7] - h[1] * h[8]) / det;
  hi[2] = (h[1] * h[5] - h[2] * h[4]) / det;
  hi[3] = (h[5] * h[6] - h[3] * h[8]) / det;
  hi[4] = (h[0] * h[8] - h[2] * h[6]) / det;
  hi[5] = (h[2] * h[3] - h[0] * h[5]) / det;
  hi[6] = (h[3] * h[7] - h[4] * h[6]) / det;
  hi[7] = (h[1] * h[6] - h[0] * h[7]) / det;
  hi[8] = (h[0] * h[4] - h[1] * h[3]) / det;
}

double box_volume(const double* h)
{
  return std::fabs(h[0] * (h[4] * h[8] - h[5] * h[7]) + h[1] * (h[5] * h[6] - h[3] * h[8]) +
                   h[2] * (h[3] * h[7] - h[4] * h[6]));
}

// ---------------------------------------------------------------------------
// Profiler
// ---------------------------------------------------------------------------

Profiler::~Profiler()
{
  for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
}

hipEvent_t Profiler::take()
{
  if (used_ == pool_.size()) {
    hipEvent_t e;
    check(hipEventCreate(&e), "event create");
    pool_.push_back(e);
  }
  return pool_[used_++];
}

void Profiler::begin(int kernel, hipStream_t s)
{
  if (mode_ != 1) return;
  Rec r{kernel, take(), nullptr};
  check(hipEventRecord(r.a, s), "event record");
  recs_.push_back(r);
  open_ = (int)recs_.size() - 1;
}

void Profiler::end(hipStream_t s)
{
  if (mode_ != 1 || open_ < 0) return;
  recs_[open_].b = take();
  check(hipEventRecord(recs_[open_].b, s), "event record");
  open_ = -1;
}

bool Profiler::attach(int kernel, hipEvent_t* start, hipEvent_t* stop)
{
  *start = *stop = nullptr;
  if (mode_ != 2 || kernel != kProfDescriptor) return false;
  Rec r{kernel, take(), take()};
  recs_.push_back(r);
  *start = r.a;
  *stop = r.b;
  return true;
}

void Profiler::collect()
{
  for (const Rec& r : recs_) {
    if (!r.b) continue;
    check(hipEventSynchronize(r.b), "event sync");
    float ms = 0.0f;
    check(hipEventElapsedTime(&ms, r.a, r.b), "event time");
    launches_[r.kernel] += 1;
    total_[r.kernel] += ms;
  }
  recs_.clear();
  used_ = 0;
}

void Profiler::read(int kernel, int* launches, double* total_ms)
{
  collect();
  if (kernel < 0 || kernel >= kProfCount) throw std::invalid_argument("unknown kernel id");
  *launches = launches_[kernel];
  *total_ms = total_[kernel];
}

void Profiler::reset()
{
  collect();
  for (int k = 0; k < kProfCount; ++k) {
    launches_[k] = 0;
    total_[k] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// NepEngine
// ---------------------------------------------------------------------------

NepEngine::NepEngine(const std::string& path)
{
  model_.load(path);
  const NepModel& m = model_;
  NepModel::Image im = m.device_image();
  // ZBL pair table [T][T][12]: a_inv, zizj, r_in, r_out, (a_k, b_k) x4
  const int T = m.num_types;
  std::vector<float> zt((size_t)T * T * 12, 0.0f);
  if (m.zbl_enabled) {
    const double uni[8] = {0.18175, 3.1998, 0.50986, 0.94229, 0.28022, 0.4029, 0.02817, 0.20162};
    for (int a = 0; a < T; ++a)
      for (int b = 0; b < T; ++b) {
        float* z = &zt[(size_t)(a * T + b) * 12];
        const int za = m.Z[a], zb = m.Z[b];
        z[0] = (float)((std::pow((double)za, 0.23) + std::pow((double)zb, 0.23)) * 2.134563);
        z[1] = (float)(kCoulomb * za * zb);
        double r1 = m.zbl_rc_inner, r2 = m.zbl_rc_outer;
        double ab[8];
        std::copy(uni, uni + 8, ab);
        if (m.zbl_flexible) {
          const int lo = std::min(a, b), hi = std::max(a, b);
          const int idx = lo * T - (lo * (lo - 1)) / 2 + (hi - lo);
          const double* p = &m.zbl_para[10 * idx];
          r1 = p[0];
          r2 = p[1];
          std::copy(p + 2, p + 10, ab);
        } else if (m.zbl_typewise) {
          r2 = std::min((kCovalentRadius[za - 1] + kCovalentRadius[zb - 1]) * m.zbl_factor, r2);
          r1 = 0.0;
        }
        z[2] = (float)r1;
        z[3] = (float)r2;
        for (int k = 0; k < 8; ++k) z[4 + k] = (float)ab[k];
      }
  }
  const size_t zoff = im.data.size();
  im.data.insert(im.data.end(), zt.begin(), zt.end());
  dalloc(&params_, im.data.size(), "parameters");
  check(hipMemcpy(params_, im.data.data(), im.data.size() * sizeof(float), hipMemcpyHostToDevice),
        "upload parameters");
  upload_angular_constants(derive_angular_constants());
  dm_.T = T;
  dm_.nR = m.nR;
  dm_.nA = m.nA;
  dm_.kR = m.kR;
  dm_.kRp = (m.kR + 3) & ~3;
  dm_.kA = m.kA;
  dm_.L = m.L_max;
  dm_.has222 = m.has222;
  dm_.has1111 = m.has1111;
  dm_.numL = m.num_L;
  dm_.D = m.dim;
  dm_.H = m.H;
  dm_.version = m.version;
  dm_.zbl = m.zbl_enabled ? 1 : 0;
  for (int t = 0; t < kMaxTypes; ++t) {
    dm_.rcR[t] = t < T ? m.rcR[t] : 0.0;
    dm_.rcA[t] = t < T ? m.rcA[t] : 0.0;
  }
  dm_.w0 = params_ + im.off_w0;
  dm_.b0 = params_ + im.off_b0;
  dm_.w1 = params_ + im.off_w1;
  dm_.w1x = params_ + im.off_w1x;
  dm_.b1 = params_ + im.off_b1;
  dm_.cR = params_ + im.off_cR;
  dm_.cA = params_ + im.off_cA;
  dm_.qs = params_ + im.off_qs;
  dm_.zblp = params_ + zoff;
  // The descriptor runs on an exact instance: the model's own sizes, or the
  // bounded sizes it is padded to.  Padding: Chebyshev coefficients beyond
  // the model's orders are zero; the descriptor has 3-body order 4 and both
  // many-body blocks, each model component keeps its ANN input weights and
  // scaler at its place in that layout, and the added components get zero
  // weights, so they change neither the energy nor, with zero adjoints, the
  // forces.  The radial-adjoint pitch stays the model's (the gather reads it).
  dmd_ = dm_;
  int bd[4];
  if (descriptor_bucket(dm_, bd)) {
    const int nR = m.nR, kR = m.kR, nA = m.nA, kA = m.kA, D = m.dim, H = m.H;
    const int NRb = bd[0], KRb = bd[1], NAb = bd[2], KAb = bd[3], Db = NRb + NAb * (kPadL + 2);
    std::vector<int> col(D);
    for (int d = 0; d < nR; ++d) col[d] = d;
    int b = 0;
    for (; b < m.L_max; ++b)
      for (int n = 0; n < nA; ++n) col[nR + b * nA + n] = NRb + b * NAb + n;
    if (m.has222) {
      for (int n = 0; n < nA; ++n) col[nR + b * nA + n] = NRb + kPadL * NAb + n;
      ++b;
    }
    if (m.has1111)
      for (int n = 0; n < nA; ++n) col[nR + b * nA + n] = NRb + (kPadL + 1) * NAb + n;
    const size_t n_w0 = (size_t)T * H * Db, n_cR = (size_t)T * T * NRb * KRb, n_cA = (size_t)T * T * NAb * KAb;
    std::vector<float> pad(n_w0 + Db + n_cR + n_cA, 0.0f);
    float* pw0 = pad.data();
    float* pqs = pw0 + n_w0;
    float* pcR = pqs + Db;
    float* pcA = pcR + n_cR;
    for (int t = 0; t < T; ++t)
      for (int h = 0; h < H; ++h)
        for (int d = 0; d < D; ++d)
          pw0[((size_t)t * H + h) * Db + col[d]] = im.data[im.off_w0 + ((size_t)t * H + h) * D + d];
    for (int d = 0; d < Db; ++d) pqs[d] = 1.0f;
    for (int d = 0; d < D; ++d) pqs[col[d]] = im.data[im.off_qs + d];
    for (int pr = 0; pr < T * T; ++pr) {
      for (int n = 0; n < nR; ++n)
        for (int k = 0; k < kR; ++k)
          pcR[((size_t)pr * NRb + n) * KRb + k] = im.data[im.off_cR + ((size_t)pr * nR + n) * kR + k];
      for (int n = 0; n < nA; ++n)
        for (int k = 0; k < kA; ++k)
          pcA[((size_t)pr * NAb + n) * KAb + k] = im.data[im.off_cA + ((size_t)pr * nA + n) * kA + k];
    }
    dalloc(&params_pad_, pad.size(), "padded descriptor parameters");
    check(hipMemcpy(params_pad_, pad.data(), pad.size() * sizeof(float), hipMemcpyHostToDevice),
          "upload padded parameters");
    dmd_.nR = NRb;
    dmd_.kR = KRb;
    dmd_.nA = NAb;
    dmd_.kA = KAb;
    dmd_.L = kPadL;
    dmd_.has222 = 1;
    dmd_.has1111 = 1;
    dmd_.numL = kPadL + 2;
    dmd_.D = Db;
    dmd_.w0 = params_pad_;
    dmd_.qs = params_pad_ + n_w0;
    dmd_.cR = params_pad_ + n_w0 + Db;
    dmd_.cA = params_pad_ + n_w0 + Db + n_cR;
  }
  check(hipHostMalloc((void**)&hflags_, kNumFlags * sizeof(int), hipHostMallocDefault), "pinned flags");
  // the overlapped filter's stream and its fork/join events (made here, not
  // at first use, which may fall inside a caller's graph capture)
  check(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking), "side stream");
  check(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "fork event");
  check(hipEventCreateWithFlags(&join_, hipEventDisableTiming), "join event");
  dalloc(&w_.flags, kNumFlags, "flags");
  check(hipMemset(w_.flags, 0, kNumFlags * sizeof(int)), "clear flags");
}

NepEngine::~NepEngine()
{
  if (side_) (void)hipStreamSynchronize(side_);  // an overlapped filter may still use the workspace
  free_all();
  dfree(&params_);
  dfree(&params_pad_);
  dfree(&sb_pos_);
  dfree(&sb_type_);
  dfree(&sb_gid_);
  dfree(&sb_pe_);
  dfree(&sb_f_);
  dfree(&sb_vir_);
  dfree(&w_.flags);
  if (hflags_) (void)hipHostFree(hflags_);
  if (side_) (void)hipStreamDestroy(side_);
  if (fork_) (void)hipEventDestroy(fork_);
  if (join_) (void)hipEventDestroy(join_);
}

void NepEngine::free_all()
{
  dfree(&w_.xq);
  dfree(&w_.cell_of);
  dfree(&w_.cell_rank);
  dfree(&w_.cell_atoms);
  dfree(&w_.xs);
  dfree(&w_.vn);
  dfree(&w_.vl);
  dfree(&w_.vrev);
  dfree(&w_.x0);
  dfree(&w_.nr);
  dfree(&w_.na);
  dfree(&w_.edge);
  dfree(&w_.eq);
  dfree(&w_.vsa);
  dfree(&w_.amask);
  dfree(&w_.bR);
  dfree(&w_.f12);
  dfree(&w_.cl);
  dfree(&w_.cpos);
  dfree(&w_.cq);
  dfree(&w_.nrc);
  dfree(&w_.nac);
  dfree(&w_.cell_count);
  dfree(&w_.cell_start);
  dfree(&w_.cell_tiles);
  w_.cap = w_.rcap = 0;
  w_.ncell_cap = 0;
  have_list_ = false;
}

// Row capacity in a layout: a multiple of 4 (16-byte row chunks in the
// Verlet build's sort), or a power of two; layout -1 takes the power of two
// where it costs at most 1.25x.
int NepEngine::padded_rows(int mv, int layout) const
{
  const int r4 = (mv + 3) & ~3;
  int p2 = 4;
  while (p2 < mv) p2 <<= 1;
  if (layout < 0) layout = (p2 <= 1.25 * mv) ? 1 : 0;
  return layout == 1 ? p2 : r4;
}

// device bytes per row entry: Verlet slot, reverse index, edge record,
// edge reverse, angular partial force and an angular bit; and with
// candidate rows the candidate slot, its angular slot, its position and
// reverse index
constexpr double kRowEntryBytes = 4 + 4 + 16 + 4 + 16 + 4.0 / 32.0;
constexpr double kSplitEntryBytes = 4 + 4 + 4 + 4;

void NepEngine::allocate(int capacity, int mv, int centres)
{
  mv_base_ = mv;
  mv = padded_rows(mv, tune_.row_layout);
  const size_t cap = (size_t)std::max(capacity, 1);
  // centres with rows: by default every slot
  const size_t rc = (size_t)std::max(1, std::min(capacity, centres < 0 ? capacity : centres));
  const size_t rows = rc * (size_t)mv;
  if (rows >= (size_t)INT32_MAX)
    throw std::invalid_argument("centres x Verlet row capacity must stay below 2^31");
  dalloc(&w_.xq, cap, "packed positions");
  dalloc(&w_.cell_of, cap, "cell_of");
  dalloc(&w_.cell_rank, cap, "cell rank");
  dalloc(&w_.cell_atoms, cap, "cell_atoms");
  dalloc(&w_.xs, cap, "cell-ordered positions");
  dalloc(&w_.x0, 3 * cap, "x0");
  dalloc(&w_.vn, rc, "vn");
  dalloc(&w_.vl, rows, "verlet rows");
  dalloc(&w_.vrev, rows, "verlet reverse");
  dalloc(&w_.nr, rc, "nr");
  dalloc(&w_.na, rc, "na");
  dalloc(&w_.edge, rows, "edges");
  dalloc(&w_.eq, rows, "edge reverse");
  dalloc(&w_.amask, rc * (size_t)((mv + 31) / 32), "angular bits");
  dfree(&w_.vsa);  // candidate rows only
  dalloc(&w_.bR, rc * model_.num_types * (size_t)((model_.kR + 3) & ~3), "radial adjoint");
  dalloc(&w_.f12, rows, "angular partials");
  w_.cap = (int)cap;
  w_.rcap = (int)rc;
  w_.MV = mv;
  tune_.row_capacity = mv;
  w_.MR = mv;
  w_.MA = mv;
  have_list_ = false;
  if (tune_.split) allocate_split();
}

void NepEngine::allocate_split()
{
  const size_t cap = (size_t)std::max(w_.rcap, 1), rows = cap * (size_t)w_.MV;
  dalloc(&w_.cl, rows, "candidate rows");
  dalloc(&w_.vsa, rows, "candidate->angular");
  dalloc(&w_.cpos, rows, "verlet->angular candidate");
  dalloc(&w_.cq, rows, "candidate reverse");
  dalloc(&w_.nrc, cap, "candidates");
  dalloc(&w_.nac, cap, "angular candidates");
  w_.split = 1;
  have_list_ = false;  // the candidate rows are made by the next rebuild
}

void NepEngine::set_desc_lanes(int lanes)
{
  if (lanes != 0 && lanes != 1 && lanes != 4) throw std::invalid_argument("descriptor lanes per atom are 0, 1 or 4");
  tune_.desc_lanes = lanes;
  tune_.tuned = true;
}

void NepEngine::set_split(int split)
{
  if (split != 0 && split != 1) throw std::invalid_argument("split is 0 (per-step filter) or 1 (rebuild-time rows)");
  tune_.split = split;
  if (split) {
    if (w_.cap > 0 && w_.MV > 0) allocate_split();
  } else {
    w_.split = 0;
    dfree(&w_.vsa);
    dfree(&w_.cl);
    dfree(&w_.cpos);
    dfree(&w_.cq);
    dfree(&w_.nrc);
    dfree(&w_.nac);
  }
  box_changed_ = true;
  tune_.tuned = true;
}

void NepEngine::reserve(int capacity, int centres)
{
  if (capacity < 0) throw std::invalid_argument("negative capacity");
  if (centres < 0) centres = capacity;
  centres = std::min(centres, capacity);
  if (capacity <= w_.cap && centres <= w_.rcap && w_.MV > 0) return;
  const int mv = std::max(mv_base_, (int)std::ceil(model_.MN_R * std::pow((model_.rcR_max + kSkin) / model_.rcR_max, 3)));
  const int cap = std::max(capacity, w_.cap);
  allocate(cap, std::max(mv, 8), std::min(cap, std::max(centres, w_.rcap)));
}

void NepEngine::ensure_cells(int ncell)
{
  if (ncell <= w_.ncell_cap) return;
  dalloc(&w_.cell_count, (size_t)ncell, "cell counts");
  dalloc(&w_.cell_start, (size_t)ncell + 1, "cell starts");
  dalloc(&w_.cell_tiles, 2 * ((size_t)ncell / 4096 + 2) + 8, "cell scan tiles");
  check(hipMemset(w_.cell_count, 0, sizeof(int) * ncell), "clear cell counts");
  w_.ncell_cap = ncell;
}

DevBox NepEngine::make_devbox(const Box& box) const
{
  DevBox b{};
  std::memcpy(b.h, box.h, sizeof(b.h));
  box_inverse(box.h, b.hi);
  for (int d = 0; d < 3; ++d) b.pbc[d] = box.pbc[d] ? 1 : 0;
  const double* h = box.h;
  b.ortho = (h[1] == 0 && h[2] == 0 && h[3] == 0 && h[5] == 0 && h[6] == 0 && h[7] == 0) ? 1 : 0;
  const double vol = box_volume(h);
  if (!(vol > 0.0) || !std::isfinite(vol)) throw std::invalid_argument("box has non-positive volume");
  const double a[3] = {h[0], h[3], h[6]}, bb[3] = {h[1], h[4], h[7]}, c[3] = {h[2], h[5], h[8]};
  const double* pairs[3][2] = {{bb, c}, {c, a}, {a, bb}};
  const double edge = 0.5 * (model_.rcR_max + kSkin);
  for (int d = 0; d < 3; ++d) {
    const double* u = pairs[d][0];
    const double* v = pairs[d][1];
    const double x = u[1] * v[2] - u[2] * v[1], y = u[2] * v[0] - u[0] * v[2], z = u[0] * v[1] - u[1] * v[0];
    b.thick[d] = vol / std::sqrt(x * x + y * y + z * z);
    b.nc[d] = std::max(1, (int)std::floor(b.thick[d] / edge));
  }
  return b;
}

NepEngine::Plan NepEngine::plan(const Box& box, int N, int n1, int n2, int nd1, int nd2, const int* type,
                                const double* pos, const uint64_t* gid, double* pe, double* force, double* virial,
                                unsigned flags, const ImageMap* map)
{
  Plan p;
  if (N < 0 || n1 < 0 || n2 < n1 || n2 > N || nd1 < 0 || nd2 < nd1 || nd2 > N)
    throw std::invalid_argument("centre range is outside [0, N)");
  if (n1 < nd1 || n2 > nd2) throw std::invalid_argument("force centres must be dependency centres");
  if (N > 0 && (!type || !pos || !pe || !force || !virial)) throw std::invalid_argument("null array");
  const bool by_gid = (flags & 8u) != 0;
  if (by_gid && !gid) throw std::invalid_argument("DMGX_SORT_BY_GID needs global_id");
  if (N == 0 || nd2 == nd1) return p;
  const ImageMap im = map ? *map : ImageMap{};
  if (im.owner && (!im.image || !im.gstart)) throw std::invalid_argument("incomplete image map");
  if (im.owner && (n1 != 0 || nd1 != 0 || gid))
    throw std::invalid_argument("an image map needs the owned slots first and rows ordered by slot");
  if (im.slot_req && (!im.owner || !im.req_moved || im.n_owned != nd2 || im.n_req <= 0))
    throw std::invalid_argument("a request map needs an image map, its moved words and the owned slots as centres");
  if ((im.req_range || im.req_off || im.lists) && !(im.slot_req && im.req_range && im.req_off && im.lists))
    throw std::invalid_argument("compacted rebuild lists need a request map, its ranges, offsets and lists");
  p.box = make_devbox(box);
  for (int d = 0; d < 3; ++d)
    if (p.box.pbc[d] && p.box.thick[d] <= 2.5 * (model_.rcR_max + kSkin))
      throw std::invalid_argument(
          "box thinner than 2.5 (rc + skin) in a periodic direction");
  const long long ncell = (long long)p.box.nc[0] * p.box.nc[1] * p.box.nc[2];
  if (ncell > (1LL << 28)) throw std::invalid_argument("cell grid too large");
  // geometric growth (a serving batch grows request by request); rows (for
  // the slots below nd2, the ones that can be centres) stay addressable by
  // int (rcap * MV < 2^31)
  if (N > w_.cap || nd2 > w_.rcap || w_.MV == 0) {
    const long long lim = (long long)(INT32_MAX - 1) / std::max(w_.MV, 128);
    const long long cap = N > w_.cap ? std::max<long long>(N, 2LL * w_.cap) : w_.cap;
    const long long rc = nd2 > w_.rcap ? std::max<long long>(nd2, std::min<long long>(2LL * w_.rcap, lim)) : w_.rcap;
    reserve((int)std::min<long long>(cap, INT32_MAX - 1), (int)rc);
  }
  if (ncell > w_.ncell_cap)
    ensure_cells((int)std::max<long long>(ncell, std::min<long long>(2LL * w_.ncell_cap, 1LL << 28)));
  p.forced = !have_list_ || box_changed_ || N != last_N_ || nd1 != last_nd1_ || nd2 != last_nd2_ ||
             by_gid != last_gid_sort_ || (flags & 2u) || std::memcmp(box.h, last_h_, sizeof(last_h_)) != 0 ||
             std::memcmp(box.pbc, last_pbc_, sizeof(last_pbc_)) != 0 || im.owner != last_map_.owner ||
             im.image != last_map_.image || im.gstart != last_map_.gstart;
  if (p.forced) scale_since_build_ = 1.0;  // this evaluation rebuilds from the current positions
  p.empty = false;
  w_.map = im;
  if (w_.split) {
    // candidate rows are remade for every centre at a rebuild: no request map
    w_.map.slot_req = nullptr;
    w_.map.req_moved = nullptr;
    w_.map.req_range = nullptr;
    w_.map.req_off = nullptr;
    w_.map.lists = nullptr;
    w_.map.n_req = 0;
  }
  last_map_ = im;
  have_list_ = true;
  box_changed_ = false;
  last_N_ = N;
  last_nd1_ = nd1;
  last_nd2_ = nd2;
  last_gid_sort_ = by_gid;
  std::memcpy(last_h_, box.h, sizeof(last_h_));
  std::memcpy(last_pbc_, box.pbc, sizeof(last_pbc_));
  return p;
}

void NepEngine::launch_force(const Plan& p, int N, int n1, int n2, int nd1, int nd2, const int* type,
                             const uint64_t* gid, double* pe, double* force, double* virial, unsigned flags,
                             hipStream_t s)
{
  if (p.empty) return;
  const bool by_gid = (flags & 8u) != 0;
  const double rv = model_.rcR_max + kSkin;
  prof_.begin(kProfRebuild, s);
  ++evaluations_;
  check(launch_rebuild(p.box, N, nd1, nd2, rv * rv, by_gid ? gid : nullptr, w_, s, nullptr, rare_rebuilds_),
        "rebuild");
  if (w_.split) check(launch_split(p.box, dm_, nd1, nd2, w_, s, tune_.filter_lanes), "candidate rows");
  prof_.end(s);
  const bool acc = (flags & 1u) != 0;
  // ANN: FP32 on the matrix cores unless BF16 operands or the vector path are asked for
  const int ann = ann_mode(flags);
  launch_filter_descriptor(p.box, N, nd1, nd2, type, pe, acc, ann, true, s, true);
  prof_.begin(kProfGather, s);
  check(launch_gather(p.box, dm_, N, n1, n2, nd1, nd2, type, force, virial, acc, w_, s, tune_.gather_lanes),
        "gather");
  prof_.end(s);
  last_box_ = p.box;
  last_N_stats_ = N;
  last_c1_ = nd1;
  last_c2_ = nd2;
}

void NepEngine::adopt_carried(const Box& box, int N, int n_owned, int c_keep, int g_keep, const int* type,
                              const double* pos, double* pe, double* force, double* virial, const ImageMap& map,
                              hipStream_t s)
{
  if (N > w_.cap || n_owned > w_.rcap) throw std::logic_error("adopt_carried: workspace smaller than the batch");
  if (c_keep < 0 || c_keep > n_owned || g_keep < n_owned || g_keep > N)
    throw std::invalid_argument("adopt_carried: carried ranges outside the batch");
  const Plan p = plan(box, N, 0, n_owned, 0, n_owned, type, pos, nullptr, pe, force, virial, 0u, &map);
  if (p.empty) return;
  const double rv = model_.rcR_max + kSkin;
  const int ranges[4] = {c_keep, n_owned, g_keep, N};
  check(launch_pack(p.box, N, model_.num_types, pos, type, true, w_, s, half_skin()), "pack");
  check(launch_rebuild(p.box, N, c_keep, n_owned, rv * rv, nullptr, w_, s, ranges), "rebuild new centres");
  check(launch_filter(p.box, dm_, 0, 0, w_, s, tune_.filter_lanes, true), "close rebuild request");
  scale_since_build_ = 1.0;
  last_small_ = false;
  last_box_ = p.box;
  last_N_stats_ = N;
  last_c1_ = 0;
  last_c2_ = n_owned;
}

void NepEngine::launch_key(std::vector<long long>& k) const
{
  const DevWork& w = w_;
  auto p = [&](const void* q) { k.push_back((long long)(uintptr_t)q); };
  for (long long v : {(long long)w.cap, (long long)w.rcap, (long long)w.MV, (long long)w.MR, (long long)w.MA,
                      (long long)w.ncell_cap, (long long)w.split, (long long)w.map.n_owned, (long long)w.map.n_req})
    k.push_back(v);
  for (const void* q : {(const void*)w.xq, (const void*)w.cell_of, (const void*)w.cell_rank,
                        (const void*)w.cell_count, (const void*)w.cell_start, (const void*)w.cell_atoms,
                        (const void*)w.xs, (const void*)w.cell_tiles, (const void*)w.vn, (const void*)w.vl,
                        (const void*)w.vrev, (const void*)w.x0, (const void*)w.nr, (const void*)w.na,
                        (const void*)w.edge, (const void*)w.eq, (const void*)w.vsa, (const void*)w.amask,
                        (const void*)w.bR,
                        (const void*)w.f12, (const void*)w.cl, (const void*)w.cpos, (const void*)w.cq,
                        (const void*)w.nrc, (const void*)w.nac, (const void*)w.map.owner, (const void*)w.map.image,
                        (const void*)w.map.gstart, (const void*)w.map.slot_req, (const void*)w.map.req_moved,
                        (const void*)w.map.req_range, (const void*)w.map.req_off, (const void*)w.map.lists,
                        (const void*)w.flags, (const void*)w.partial, (const void*)w.thermo})
    p(q);
  for (int v : {tune_.filter_lanes, tune_.gather_lanes, tune_.row_layout, tune_.overlap, tune_.split,
                tune_.desc_lanes})
    k.push_back(v);
}

bool NepEngine::small_box(const Box& box, int reps[3]) const
{
  const DevBox b = make_devbox(box);
  const double need = 2.5 * (model_.rcR_max + kSkin);
  bool any = false;
  for (int d = 0; d < 3; ++d) {
    reps[d] = 1;
    if (b.pbc[d] && b.thick[d] <= need) {
      reps[d] = (int)std::floor(need / b.thick[d]) + 1;
      any = true;
    }
  }
  return any;
}

void NepEngine::enqueue_small(const Box& box, const int reps[3], int N, const int* type, const double* pos,
                              const uint64_t* gid, double* pe, double* force, double* virial, unsigned flags,
                              hipStream_t s)
{
  const long long R = (long long)reps[0] * reps[1] * reps[2];
  const long long n2 = (long long)N * R;
  if (n2 > (1LL << 30)) throw std::invalid_argument("small box: too many periodic replicas");
  if (n2 > sb_cap_) {
    dalloc(&sb_pos_, 3 * (size_t)n2, "replica positions");
    dalloc(&sb_type_, (size_t)n2, "replica types");
    dalloc(&sb_gid_, (size_t)n2, "replica ids");
    dalloc(&sb_pe_, (size_t)n2, "replica potential");
    dalloc(&sb_f_, 3 * (size_t)n2, "replica forces");
    dalloc(&sb_vir_, 9 * (size_t)n2, "replica virial");
    sb_cap_ = n2;
  }
  Box big = box;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) big.h[3 * r + c] *= reps[c];
  const bool by_gid = (flags & 8u) != 0;
  check(launch_replicate(N, reps, make_devbox(box), pos, type, by_gid ? gid : nullptr, sb_pos_, sb_type_, sb_gid_, s),
        "replicate");
  const int M = (int)n2;
  const uint64_t* gid2 = by_gid ? sb_gid_ : nullptr;
  const unsigned f2 = flags & ~1u;  // the replicas' outputs are overwritten, then folded
  const Plan p = plan(big, M, 0, N, 0, M, sb_type_, sb_pos_, gid2, sb_pe_, sb_f_, sb_vir_, f2);
  last_small_ = true;
  prof_.begin(kProfCheck, s);
  check(launch_pack(p.box, M, model_.num_types, sb_pos_, sb_type_, p.forced, w_, s), "pack / displacement check");
  prof_.end(s);
  launch_force(p, M, 0, N, 0, M, sb_type_, gid2, sb_pe_, sb_f_, sb_vir_, f2, s);
  check(launch_fold(N, n2, (flags & 1u) != 0, sb_pe_, sb_f_, sb_vir_, pe, force, virial, s), "fold replicas");
}

double NepEngine::half_skin() const
{
  return 0.5 * (kSkin - std::fabs(1.0 - scale_since_build_) * (model_.rcR_max + kSkin));
}

void NepEngine::notify_box_scaled(double scale, hipStream_t stream)
{
  if (!have_list_ || last_small_ || last_N_ <= 0) {
    box_changed_ = true;
    return;
  }
  for (double& v : last_h_) v *= scale;  // the same products the caller forms for its box
  check(launch_scale(3 * last_N_, scale, w_.x0, stream), "rescale displacement references");
  scale_since_build_ *= scale;
  if (std::fabs(1.0 - scale_since_build_) * (model_.rcR_max + kSkin) > 0.5 * kSkin) box_changed_ = true;
}

void NepEngine::enqueue(const Box& box, int N, int n1, int n2, int nd1, int nd2, const int* type,
                        const double* pos, const uint64_t* gid, double* pe, double* force, double* virial,
                        unsigned flags, hipStream_t s, const ImageMap* map)
{
  int reps[3];
  if (N > 0 && nd2 > nd1 && small_box(box, reps)) {
    if (map && map->owner) throw std::invalid_argument("an image map needs a box without thin periodic directions");
    if (n1 != 0 || nd1 != 0 || n2 != N || nd2 != N)
      throw std::invalid_argument("a periodic box thinner than 2.5 (rc + skin) needs all atoms as centres");
    if (!type || !pos || !pe || !force || !virial) throw std::invalid_argument("null array");
    if ((flags & 8u) && !gid) throw std::invalid_argument("DMGX_SORT_BY_GID needs global_id");
    enqueue_small(box, reps, N, type, pos, gid, pe, force, virial, flags, s);
    return;
  }
  const Plan p = plan(box, N, n1, n2, nd1, nd2, type, pos, gid, pe, force, virial, flags, map);
  if (p.empty) return;
  last_small_ = false;
  // rebuild decision on the device: forced by the host, or any atom moved
  // more than half the skin left
  prof_.begin(kProfCheck, s);
  check(launch_pack(p.box, N, model_.num_types, pos, type, p.forced, w_, s, half_skin()),
        "pack / displacement check");
  prof_.end(s);
  launch_force(p, N, n1, n2, nd1, nd2, type, gid, pe, force, virial, flags, s);
}

const int* NepEngine::read_flags(hipStream_t s)
{
  check(hipMemcpyAsync(hflags_, w_.flags, kNumFlags * sizeof(int), hipMemcpyDeviceToHost, s), "read flags");
  check(hipStreamSynchronize(s), "synchronise");
  return hflags_;
}

bool NepEngine::settle(hipStream_t s)
{
  read_flags(s);
  if (hflags_[kFlagRebuildCount] != seen_rebuilds_) {  // a rebuild ran: fresh references
    seen_rebuilds_ = hflags_[kFlagRebuildCount];
    scale_since_build_ = 1.0;
  }
  // the rebuild rate over at least 64 evaluations picks the rebuild grids
  // (both give the same rows); the count restarts with a new allocation
  const long long de = evaluations_ - rate_evaluations_;
  const int dr = hflags_[kFlagRebuildCount] - rate_rebuilds_;
  if (de >= 64 || dr < 0) {
    if (dr >= 0) rare_rebuilds_ = 16LL * dr < de;
    rate_evaluations_ = evaluations_;
    rate_rebuilds_ = hflags_[kFlagRebuildCount];
  }
  if (hflags_[kFlagBadType]) {
    check(hipMemsetAsync(w_.flags + kFlagBadType, 0, sizeof(int), s), "clear flag");
    throw std::invalid_argument("an atom type lies outside [0, num_types)");
  }
  if (hflags_[kFlagOutOfBox]) {
    check(hipMemsetAsync(w_.flags + kFlagOutOfBox, 0, sizeof(int), s), "clear flag");
    box_changed_ = true;
    throw std::invalid_argument("a position lies outside the box in a non-periodic direction");
  }
  if (hflags_[kFlagVerletOverflow] > w_.MV) {
    const int mv = (int)std::ceil(hflags_[kFlagVerletOverflow] * 1.2) + 8;
    if (mv > 4096) throw std::runtime_error("Verlet row overflow: more than 4096 neighbours within rc + skin");
    check(hipMemsetAsync(w_.flags + kFlagVerletOverflow, 0, sizeof(int), s), "clear flag");
    check(hipStreamSynchronize(s), "synchronise");
    allocate(w_.cap, mv, w_.rcap);  // drops the list: the next enqueue rebuilds
    return true;
  }
  return false;
}

void NepEngine::compute(const Box& box, int N, int n1, int n2, int nd1, int nd2, const int* type,
                        const double* pos, const uint64_t* gid, double* pe, double* force, double* virial,
                        unsigned flags, hipStream_t s, const ImageMap* map)
{
  for (int attempt = 0; attempt < 4; ++attempt) {
    enqueue(box, N, n1, n2, nd1, nd2, type, pos, gid, pe, force, virial, flags, s, map);
    if (!settle(s)) return;
    if (flags & 1u)
      throw std::runtime_error(
          "Verlet row overflow: accumulated outputs are invalid (workspace grown; clear them and call again)");
  }
  throw std::runtime_error("Verlet rows kept overflowing");
}

// Smallest centre range split for the overlapped schedule: each half keeps
// at least this many centres (below it the launches are too short to gain)
constexpr int kOverlapMinCentres = 16384;
// Below this many centres one lane per atom leaves most SIMDs idle (256 CUs x
// 4 SIMDs x 64 lanes): the untuned descriptor then takes four lanes per atom
constexpr int kLaneGroupCentres = 65536;

int NepEngine::desc_lanes(int n, int ann) const
{
  if (w_.split || (ann != 1 && ann != 0)) return 1;  // candidate rows, or BF16 operands asked for
  if (tune_.desc_lanes) return tune_.desc_lanes;
  return n < kLaneGroupCentres ? 4 : 1;
}

void NepEngine::launch_filter_descriptor(const DevBox& b, int N, int c1, int c2, const int* type, double* pe,
                                         bool acc, int ann, bool close, hipStream_t s, bool profile)
{
  const int n = c2 - c1;
  // per-kernel event timing (profiling mode 1) keeps the sequential order,
  // so that each kernel group is timed alone; candidate rows have no filter
  // pass (the descriptor, a single launch, closes the rebuild request)
  const bool split = !w_.split && tune_.overlap && n >= 2 * kOverlapMinCentres &&
                     !(profile && prof_.on() && !prof_.attach_mode());
  if (!split) {
    if (!w_.split) {
      if (profile) prof_.begin(kProfFilter, s);
      check(launch_filter(b, dm_, c1, c2, w_, s, tune_.filter_lanes, close), "filter");
      if (profile) prof_.end(s);
    }
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    const bool attached = profile && prof_.attach(kProfDescriptor, &ev0, &ev1);
    if (profile && !attached) prof_.begin(kProfDescriptor, s);
    check(launch_descriptor(b, dmd_, N, c1, c2, type, pe, acc, ann, w_, s, ev0, ev1, desc_lanes(n, ann)),
          "descriptor");
    if (profile && !attached) prof_.end(s);
    return;
  }
  // halves on workgroup boundaries of the descriptor (256 centres)
  const int mid = c1 + ((n / 2 + 255) & ~255);
  check(launch_filter(b, dm_, c1, mid, w_, s, tune_.filter_lanes, close), "filter");
  // the side stream starts after the first filter (and the rebuild before
  // it): the second filter overlaps the first descriptor
  check(hipEventRecord(fork_, s), "fork");
  check(hipStreamWaitEvent(side_, fork_, 0), "fork wait");
  check(launch_filter(b, dm_, mid, c2, w_, side_, tune_.filter_lanes, false), "filter");
  check(hipEventRecord(join_, side_), "join");
  for (int h = 0; h < 2; ++h) {
    if (h == 1) check(hipStreamWaitEvent(s, join_, 0), "join wait");
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (profile) prof_.attach(kProfDescriptor, &ev0, &ev1);
    check(launch_descriptor(b, dmd_, N, h ? mid : c1, h ? c2 : mid, type, pe, acc, ann, w_, s, ev0, ev1,
                            desc_lanes(n, ann)),
          "descriptor");
  }
}

void NepEngine::launch_step_kernels(const DevBox& b, int N, int nc, const int* type, double* pe, double* force,
                                    double* virial, unsigned flags, hipStream_t s)
{
  const int ann = ann_mode(flags);
  launch_filter_descriptor(b, N, 0, nc, type, pe, false, ann, true, s, false);
  check(launch_gather(b, dm_, N, 0, nc, 0, nc, type, force, virial, false, w_, s, tune_.gather_lanes), "gather");
}

const NepEngine::Tuning& NepEngine::autotune(const Box& box, int N, int nc, const int* type, const double* pos,
                                              double* pe, double* force, double* virial, unsigned flags,
                                              hipStream_t s, const ImageMap* map)
{
  int reps[3];
  if (N <= 0 || nc <= 0 || nc > N || small_box(box, reps)) return tune_;  // small boxes keep the defaults
  flags &= ~1u;
  // the filter's shapes first, then the candidate rows against them
  if (w_.split) set_split(0);
  compute(box, N, 0, nc, 0, nc, type, pos, nullptr, pe, force, virial, flags, s, map);
  flags &= ~2u;
  // rows, edges, adjoints and partials are current: the filter rewrites its
  // outputs identically and the gather overwrites the forces, so every
  // candidate can be timed on the same inputs
  const DevBox b = make_devbox(box);
  // each candidate: one warm-up launch, then kReps launches, each between
  // two consecutive events; its time is the median launch
  constexpr int kReps = 7;
  hipEvent_t ev[kReps + 1];
  for (hipEvent_t& e : ev) check(hipEventCreate(&e), "tune event");
  auto time_median = [&](const std::function<void()>& launch) {
    launch();
    check(hipEventRecord(ev[0], s), "tune event");
    for (int k = 0; k < kReps; ++k) {
      launch();
      check(hipEventRecord(ev[k + 1], s), "tune event");
    }
    check(hipEventSynchronize(ev[kReps]), "tune event");
    float ms[kReps];
    for (int k = 0; k < kReps; ++k) check(hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]), "tune event");
    std::sort(ms, ms + kReps);
    return (double)ms[kReps / 2];
  };
  // the device busy before the first candidate is timed (its clocks up, and
  // freshly allocated rows touched)
  auto warm = [&] {
    for (int k = 0; k < 10; ++k) launch_step_kernels(b, N, nc, type, pe, force, virial, flags, s);
  };
  warm();
  // lanes of the filter and gather in the current row layout, then the
  // whole force sequence with them
  auto tune_lanes = [&](int layout) {
    const int fl[3] = {8, 16, 32}, gl[3] = {4, 8, 16};
    int best = 1;
    for (int k = 0; k < 3; ++k) {
      tune_.filter_ms[k] = time_median([&] { check(launch_filter(b, dm_, 0, nc, w_, s, fl[k]), "tune filter"); });
      if (tune_.filter_ms[k] < tune_.filter_ms[best]) best = k;
    }
    tune_.filter_lanes = fl[best];
    best = 1;
    for (int k = 0; k < 3; ++k) {
      tune_.gather_ms[k] = time_median([&] {
        check(launch_gather(b, dm_, N, 0, nc, 0, nc, type, force, virial, false, w_, s, gl[k]), "tune gather");
      });
      if (tune_.gather_ms[k] < tune_.gather_ms[best]) best = k;
    }
    tune_.gather_lanes = gl[best];
    tune_.force_ms[layout] = time_median([&] { launch_step_kernels(b, N, nc, type, pe, force, virial, flags, s); });
  };
  const int layout0 = (w_.MV == padded_rows(mv_base_, 1)) ? 1 : 0;
  tune_.row_layout = layout0;
  tune_lanes(layout0);
  // the other row layout, where it differs and its rows fit comfortably
  const int other = 1 - layout0;
  const int mv_other = padded_rows(mv_base_, other);
  size_t free_b = 0, total_b = 0;
  check(hipMemGetInfo(&free_b, &total_b), "memory info");
  const double extra = (double)w_.rcap * std::max(0, mv_other - w_.MV) * kRowEntryBytes;
  if (mv_other != w_.MV && (double)w_.rcap * mv_other < (double)INT32_MAX && extra < 0.5 * (double)free_b) {
    const Tuning first = tune_;
    tune_.row_layout = other;
    bool ok = true;
    try {
      allocate(w_.cap, mv_base_, w_.rcap);
      compute(box, N, 0, nc, 0, nc, type, pos, nullptr, pe, force, virial, flags | 2u, s, map);
    } catch (const std::runtime_error&) {
      ok = false;  // no room for the other layout's rows: keep the first
      (void)hipGetLastError();
    }
    if (ok) {
      warm();
      tune_lanes(other);
    }
    // wider rows cost memory and every rebuild more, which the force
    // kernels' timing does not see: they must win by 3% to be kept
    const double bar = mv_other > first.row_capacity ? 0.97 : 1.0;
    if (!ok || !(tune_.force_ms[other] < bar * first.force_ms[layout0])) {
      const double t_other = ok ? tune_.force_ms[other] : 0.0;
      tune_ = first;
      tune_.force_ms[other] = t_other;
      allocate(w_.cap, mv_base_, w_.rcap);
      // fresh rows: the remaining candidates are timed on this layout's rows
      compute(box, N, 0, nc, 0, nc, type, pos, nullptr, pe, force, virial, flags | 2u, s, map);
    }
  }
  // the overlapped filter/descriptor schedule, where the centres split
  if (nc >= 2 * kOverlapMinCentres) {
    double t[2];
    for (int o = 0; o < 2; ++o) {
      tune_.overlap = o;
      t[o] = time_median([&] { launch_step_kernels(b, N, nc, type, pe, force, virial, flags, s); });
      tune_.overlap_ms[o] = t[o];
    }
    tune_.overlap = t[1] < t[0] ? 1 : 0;
  }
  // one or four lanes per atom in the descriptor (FP32 requests only)
  {
    const int ann = ann_mode(flags);
    if (ann == 0 || ann == 1) {
      double t[2];
      for (int k = 0; k < 2; ++k) {
        tune_.desc_lanes = k ? 4 : 1;
        t[k] = time_median([&] { launch_step_kernels(b, N, nc, type, pe, force, virial, flags, s); });
        tune_.desc_ms[k] = t[k];
      }
      tune_.desc_lanes = t[1] < t[0] ? 4 : 1;
    }
  }
  // candidate rows split at rebuild time, with every gather shape, where
  // their rows fit comfortably; kept if the step's force kernels are faster
  {
    double t_filter = tune_.force_ms[tune_.row_layout];
    if (tune_.overlap_ms[tune_.overlap] > 0.0) t_filter = std::min(t_filter, tune_.overlap_ms[tune_.overlap]);
    const int dl = tune_.desc_lanes == 4 ? 1 : 0;
    if (tune_.desc_ms[dl] > 0.0) t_filter = std::min(t_filter, tune_.desc_ms[dl]);
    tune_.split_ms[0] = t_filter;
    tune_.split_ms[1] = 0.0;
    size_t free_s = 0, total_s = 0;
    check(hipMemGetInfo(&free_s, &total_s), "memory info");
    const double need = (double)w_.rcap * (w_.MV * kSplitEntryBytes + 8.0);
    if (need < 0.5 * (double)free_s) {
      const Tuning first = tune_;
      bool ok = true;
      try {
        set_split(1);
        compute(box, N, 0, nc, 0, nc, type, pos, nullptr, pe, force, virial, flags | 2u, s, map);
      } catch (const std::runtime_error&) {
        ok = false;  // no room for the candidate rows
        (void)hipGetLastError();
      }
      double best = 0.0;
      int best_g = tune_.gather_lanes;
      if (ok) {
        const int gl[3] = {4, 8, 16};
        for (int g : gl) {
          tune_.gather_lanes = g;
          const double t = time_median([&] { launch_step_kernels(b, N, nc, type, pe, force, virial, flags, s); });
          if (best == 0.0 || t < best) {
            best = t;
            best_g = g;
          }
        }
      }
      if (ok && best < t_filter) {
        tune_.gather_lanes = best_g;
        tune_.split_ms[1] = best;
      } else {
        set_split(0);
        tune_ = first;
        tune_.split_ms[1] = best;
      }
    }
  }
  for (hipEvent_t& e : ev) (void)hipEventDestroy(e);
  tune_.tuned = true;
  // the outputs as a step with the chosen shapes writes them (a rebuild
  // where the rows were re-allocated)
  compute(box, N, 0, nc, 0, nc, type, pos, nullptr, pe, force, virial, flags, s, map);
  check(hipStreamSynchronize(s), "tune");
  return tune_;
}

void NepEngine::set_tiles(int filter_lanes, int gather_lanes)
{
  if (filter_lanes != 8 && filter_lanes != 16 && filter_lanes != 32)
    throw std::invalid_argument("filter lanes per centre are 8, 16 or 32");
  if (gather_lanes != 4 && gather_lanes != 8 && gather_lanes != 16)
    throw std::invalid_argument("gather lanes per atom are 4, 8 or 16");
  tune_.filter_lanes = filter_lanes;
  tune_.gather_lanes = gather_lanes;
  tune_.tuned = true;
}

namespace {

// Hands out 256-byte aligned pieces of one buffer; with a null base it only
// counts the bytes the pieces need.
struct Carver {
  char* base;
  size_t used = 0;
  template <class T>
  T* take(size_t n)
  {
    used = (used + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
    used += std::max<size_t>(n, 1) * sizeof(T);
    return p;
  }
};

// The workspace of compute_in for cap slots, rows of mv entries and ncell
// cells, carved from `base` (or only measured when base is null).
DevWork carve_work(char* base, int cap, int mv, long long ncell, int T, int kRp, bool split, size_t* bytes)
{
  Carver c{base};
  DevWork w{};
  const size_t n = (size_t)cap, rows = n * (size_t)mv;
  w.flags = c.take<int>(kNumFlags);
  w.cell_count = c.take<int>((size_t)ncell);
  w.cell_start = c.take<int>((size_t)ncell + 1);
  w.cell_tiles = c.take<int>(2 * ((size_t)ncell / 4096 + 2) + 8);
  w.xq = c.take<double4>(n);
  w.xs = c.take<double4>(n);
  w.cell_of = c.take<int>(n);
  w.cell_rank = c.take<int>(n);
  w.cell_atoms = c.take<int>(n);
  w.vn = c.take<int>(n);
  w.nr = c.take<int>(n);
  w.na = c.take<int>(n);
  w.x0 = c.take<double>(3 * n);
  w.bR = c.take<float>(n * (size_t)T * (size_t)kRp);
  w.vl = c.take<int>(rows);
  w.vrev = c.take<int>(rows);
  w.eq = c.take<int>(rows);
  w.amask = c.take<unsigned>(n * (size_t)((mv + 31) / 32));
  w.edge = c.take<float4>(rows);
  w.f12 = c.take<float4>(rows);
  if (split) {
    w.split = 1;
    w.cl = c.take<int>(rows);
    w.vsa = c.take<int>(rows);
    w.cpos = c.take<int>(rows);
    w.cq = c.take<int>(rows);
    w.nrc = c.take<int>(n);
    w.nac = c.take<int>(n);
  }
  w.cap = w.rcap = cap;
  w.MV = w.MR = w.MA = mv;
  w.ncell_cap = (int)ncell;
  *bytes = c.used;
  return w;
}

}  // namespace

void NepEngine::compute_in(void* ws, size_t ws_bytes, const Box& box, int N, int n1, int n2, int nd1, int nd2,
                           const int* type, const double* pos, const uint64_t* gid, double* pe, double* force,
                           double* virial, unsigned flags, hipStream_t s)
{
  if (N < 0 || n1 < 0 || n2 < n1 || n2 > N || nd1 < 0 || nd2 < nd1 || nd2 > N)
    throw std::invalid_argument("centre range is outside [0, N)");
  if (n1 < nd1 || n2 > nd2) throw std::invalid_argument("force centres must be dependency centres");
  if (N > 0 && (!type || !pos || !pe || !force || !virial)) throw std::invalid_argument("null array");
  const bool by_gid = (flags & 8u) != 0;
  if (by_gid && !gid) throw std::invalid_argument("DMGX_SORT_BY_GID needs global_id");
  if (!ws || (reinterpret_cast<uintptr_t>(ws) & 255)) throw std::invalid_argument("workspace not 256-byte aligned");
  if (N == 0 || nd2 == nd1) return;
  int reps[3];
  if (small_box(box, reps))
    throw std::invalid_argument(
        "a periodic box thinner than 2.5 (rc + skin) is evaluated on a replicated supercell, which the fixed "
        "workspace does not hold; use dmgx_nep_compute");
  const DevBox b = make_devbox(box);
  const long long ncell = (long long)b.nc[0] * b.nc[1] * b.nc[2];
  if (ncell > (1LL << 28)) throw std::invalid_argument("cell grid too large");
  // rows as the allocating path sizes them before any overflow has grown them
  const int mv_need = std::max(8, (int)std::ceil(model_.MN_R * std::pow((model_.rcR_max + kSkin) / model_.rcR_max, 3)));
  const int mv = padded_rows(mv_need, tune_.row_layout);
  const int kRp = (model_.kR + 3) & ~3;
  size_t need = 0;
  const bool split = tune_.split != 0;
  carve_work(nullptr, N, mv, ncell, model_.num_types, kRp, split, &need);
  if (need > ws_bytes) {
    size_t one = 0, none = 0;
    carve_work(nullptr, 1, mv, ncell, model_.num_types, kRp, split, &one);
    carve_work(nullptr, 0, mv, ncell, model_.num_types, kRp, split, &none);
    const size_t per_atom = std::max<size_t>(one - none, 1);
    throw std::invalid_argument("the system needs " + std::to_string(need) + " bytes of workspace (" +
                                std::to_string(N) + " slots, rows of " + std::to_string(mv) + ", " +
                                std::to_string(ncell) + " cells); the fixed workspace holds " +
                                std::to_string(ws_bytes) + " (about " +
                                std::to_string(ws_bytes > none ? (ws_bytes - none) / per_atom : 0) +
                                " slots of this model in this box)");
  }
  DevWork w = carve_work(static_cast<char*>(ws), N, mv, ncell, model_.num_types, kRp, split, &need);
  // the workspace keeps the previous call's contents: clear the status words
  // and the cell counts (the cell scan expects them at zero) with a kernel
  check(launch_clear_ints(w.flags, kNumFlags, s), "clear status words");
  check(launch_clear_ints(w.cell_count, ncell, s), "clear cell counts");
  check(launch_pack(b, N, model_.num_types, pos, type, true, w, s), "pack");
  const double rv = model_.rcR_max + kSkin;
  check(launch_rebuild(b, N, nd1, nd2, rv * rv, by_gid ? gid : nullptr, w, s), "rebuild");
  if (split) check(launch_split(b, dm_, nd1, nd2, w, s, tune_.filter_lanes), "candidate rows");
  else check(launch_filter(b, dm_, nd1, nd2, w, s, tune_.filter_lanes), "filter");
  const bool acc = (flags & 1u) != 0;
  const int ann = ann_mode(flags);
  const int lanes = (split || (ann != 1 && ann != 0)) ? 1
                    : tune_.desc_lanes ? tune_.desc_lanes : (nd2 - nd1 < kLaneGroupCentres ? 4 : 1);
  check(launch_descriptor(b, dmd_, N, nd1, nd2, type, pe, acc, ann, w, s, nullptr, nullptr, lanes), "descriptor");
  check(launch_gather(b, dm_, N, n1, n2, nd1, nd2, type, force, virial, acc, w, s, tune_.gather_lanes), "gather");
  check(hipMemcpyAsync(hflags_, w.flags, kNumFlags * sizeof(int), hipMemcpyDeviceToHost, s), "read status words");
  check(hipStreamSynchronize(s), "synchronise");
  if (hflags_[kFlagBadType]) throw std::invalid_argument("an atom type lies outside [0, num_types)");
  if (hflags_[kFlagOutOfBox])
    throw std::invalid_argument("a position lies outside the box in a non-periodic direction");
  if (hflags_[kFlagVerletOverflow] > mv)
    throw std::runtime_error("Verlet row overflow: " + std::to_string(hflags_[kFlagVerletOverflow]) +
                             " neighbours within rc + skin, the fixed workspace's rows hold " + std::to_string(mv) +
                             " (MN_radial of the model); outputs are invalid");
}

double NepEngine::bytes_per_slot(bool centre) const
{
  // packed and cell-ordered positions, cell index, rank and entry, reference
  const double slot = 2 * sizeof(double4) + 3 * sizeof(int) + 3 * sizeof(double);
  if (!centre) return slot;
  int mv = w_.MV;
  if (mv <= 0)
    mv = padded_rows(std::max(8, (int)std::ceil(model_.MN_R * std::pow((model_.rcR_max + kSkin) / model_.rcR_max, 3))),
                     tune_.row_layout);
  // row lengths, radial adjoints and the rows
  double per = 3 * sizeof(int) + (double)model_.num_types * ((model_.kR + 3) & ~3) * sizeof(float);
  double entry = kRowEntryBytes;
  if (tune_.split) {
    per += 2 * sizeof(int);
    entry += kSplitEntryBytes;
  }
  return slot + per + entry * mv;
}

void NepEngine::stats(int* max_radial, int* max_angular, int* max_verlet, int* rebuilds)
{
  const int n = last_c2_ - last_c1_;
  int mr = 0, ma = 0, mv = 0;
  check(hipDeviceSynchronize(), "sync");
  if (n > 0 && have_list_) {
    std::vector<int> a(n), b(n), c(n);
    check(hipMemcpy(a.data(), w_.nr + last_c1_, n * sizeof(int), hipMemcpyDeviceToHost), "copy");
    check(hipMemcpy(b.data(), w_.na + last_c1_, n * sizeof(int), hipMemcpyDeviceToHost), "copy");
    check(hipMemcpy(c.data(), w_.vn + last_c1_, n * sizeof(int), hipMemcpyDeviceToHost), "copy");
    mr = *std::max_element(a.begin(), a.end());
    ma = *std::max_element(b.begin(), b.end());
    mv = *std::max_element(c.begin(), c.end());
  }
  int f[kNumFlags];
  check(hipMemcpy(f, w_.flags, sizeof(f), hipMemcpyDeviceToHost), "copy");
  if (max_radial) *max_radial = mr;
  if (max_angular) *max_angular = ma;
  if (max_verlet) *max_verlet = mv;
  if (rebuilds) *rebuilds = f[kFlagRebuildCount];
}

void NepEngine::totals(long long* sum_radial, long long* sum_angular, long long* sum_verlet, int* centres)
{
  const int n = last_c2_ - last_c1_;
  long long a = 0, b = 0, c = 0;
  if (n > 0 && have_list_) {
    std::vector<int> r(n), g(n), v(n);
    check(hipDeviceSynchronize(), "sync");
    check(hipMemcpy(r.data(), w_.nr + last_c1_, n * sizeof(int), hipMemcpyDeviceToHost), "copy");
    check(hipMemcpy(g.data(), w_.na + last_c1_, n * sizeof(int), hipMemcpyDeviceToHost), "copy");
    check(hipMemcpy(v.data(), w_.vn + last_c1_, n * sizeof(int), hipMemcpyDeviceToHost), "copy");
    for (int k = 0; k < n; ++k) {
      a += r[k];
      b += g[k];
      c += v[k];
    }
  }
  if (sum_radial) *sum_radial = a;
  if (sum_angular) *sum_angular = b;
  if (sum_verlet) *sum_verlet = c;
  if (centres) *centres = n;
}

void NepEngine::copy_verlet_rows(int c1, int c2, int max_nb, int* counts, int* rows)
{
  if (!have_list_) throw std::runtime_error("no neighbour list yet");
  if (c1 < last_c1_ || c2 > last_c2_ || c2 < c1) throw std::invalid_argument("rows outside the centre range");
  const int N = last_N_stats_;
  check(hipDeviceSynchronize(), "sync");
  std::vector<int> vn(N), vl((size_t)w_.MV * N);
  check(hipMemcpy(vn.data(), w_.vn, N * sizeof(int), hipMemcpyDeviceToHost), "copy");
  check(hipMemcpy(vl.data(), w_.vl, vl.size() * sizeof(int), hipMemcpyDeviceToHost), "copy");
  for (int i = c1; i < c2; ++i) {
    const int n = std::min(vn[i], w_.MV);
    counts[i - c1] = vn[i];
    if (n > max_nb) throw std::invalid_argument("max_nb too small");
    std::copy(vl.begin() + (size_t)i * w_.MV, vl.begin() + (size_t)i * w_.MV + n, rows + (size_t)(i - c1) * max_nb);
  }
}

void NepEngine::copy_pair_rows(int which, int c1, int c2, int max_nb, int* counts, int* rows)
{
  if (!have_list_) throw std::runtime_error("no neighbour list yet");
  if (c1 < last_c1_ || c2 > last_c2_ || c2 < c1) throw std::invalid_argument("rows outside the centre range");
  const int N = last_N_stats_;
  // candidate rows keep the step's angular rows (written by the descriptor)
  // but no radial rows: for those the filter pass runs once on the last
  // evaluation's positions
  if (w_.split && which == 0)
    check(launch_filter(last_box_, dm_, last_c1_, last_c2_, w_, nullptr, tune_.filter_lanes, false), "filter");
  check(hipDeviceSynchronize(), "sync");
  std::vector<int> nr(N), na(N);
  std::vector<float4> e((size_t)w_.MR * N);
  check(hipMemcpy(nr.data(), w_.nr, N * sizeof(int), hipMemcpyDeviceToHost), "copy");
  check(hipMemcpy(na.data(), w_.na, N * sizeof(int), hipMemcpyDeviceToHost), "copy");
  check(hipMemcpy(e.data(), w_.edge, e.size() * sizeof(float4), hipMemcpyDeviceToHost), "copy");
  for (int i = c1; i < c2; ++i) {
    const int n = which == 0 ? nr[i] : na[i];
    counts[i - c1] = n;
    if (n > max_nb) throw std::invalid_argument("max_nb too small");
    std::vector<int> r(n);
    for (int k = 0; k < n; ++k) {
      const int slot = k < na[i] ? k : w_.MR - 1 - (k - na[i]);
      float w = e[(size_t)i * w_.MR + slot].w;
      int j;
      std::memcpy(&j, &w, sizeof(int));
      r[k] = j;
    }
    std::sort(r.begin(), r.end());
    std::copy(r.begin(), r.end(), rows + (size_t)(i - c1) * max_nb);
  }
}

// ---------------------------------------------------------------------------
// MdState
// ---------------------------------------------------------------------------

MdState::MdState(NepEngine* nep, const Box& box, int N, const int* type, const double* mass, const double* pos,
                 const double* vel)
    : nep_(nep), box_(box), N_(N)
{
  if (N <= 0) throw std::invalid_argument("MD needs at least one atom");
  for (int i = 0; i < N; ++i) {
    if (type[i] < 0 || type[i] >= nep->model().num_types) throw std::invalid_argument("atom type out of range");
    if (!(mass[i] > 0.0)) throw std::invalid_argument("masses must be positive");
  }
  check(hipStreamCreate(&stream_), "stream");
  dalloc(&pos_, 3 * (size_t)N, "positions");
  dalloc(&vel_, 3 * (size_t)N, "velocities");
  dalloc(&force_, 3 * (size_t)N, "forces");
  dalloc(&pe_, (size_t)N, "potential");
  dalloc(&vir_, 9 * (size_t)N, "virial");
  dalloc(&mass_, (size_t)N, "mass");
  dalloc(&type_, (size_t)N, "type");
  dalloc(&tw_.partial, (size_t)thermo_blocks(N) * 16, "thermo partials");
  dalloc(&tw_.thermo, 8, "thermo");
  dalloc(&vmax2_, 1, "max speed");
  check(hipMemcpy(pos_, pos, 3 * sizeof(double) * N, hipMemcpyHostToDevice), "upload");
  check(hipMemcpy(vel_, vel, 3 * sizeof(double) * N, hipMemcpyHostToDevice), "upload");
  check(hipMemcpy(mass_, mass, sizeof(double) * N, hipMemcpyHostToDevice), "upload");
  check(hipMemcpy(type_, type, sizeof(int) * N, hipMemcpyHostToDevice), "upload");
  check(hipMemset(force_, 0, 3 * sizeof(double) * N), "clear");
  check(hipMemset(pe_, 0, sizeof(double) * N), "clear");
  check(hipMemset(vir_, 0, 9 * sizeof(double) * N), "clear");
  nep_->reserve(N);
  order_.resize(N);
  for (int i = 0; i < N; ++i) order_[i] = i;
}

namespace {
// the low 21 bits of v spread to every third bit (Morton / Z-order interleave)
unsigned long long morton3(long long v)
{
  unsigned long long x = (unsigned long long)v & 0x1fffffULL;
  x = (x | x << 32) & 0x1f00000000ffffULL;
  x = (x | x << 16) & 0x1f0000ff0000ffULL;
  x = (x | x << 8) & 0x100f00f00f00f00fULL;
  x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
  x = (x | x << 2) & 0x1249249249249249ULL;
  return x;
}

// a[c·N + i] <- a[c·N + perm[i]] for each of the K components of an SoA array
template <class T>
void permute_soa(T* dev, int N, int K, const std::vector<int>& perm, hipStream_t s)
{
  std::vector<T> src((size_t)K * N), dst((size_t)K * N);
  check(hipMemcpyAsync(src.data(), dev, sizeof(T) * src.size(), hipMemcpyDeviceToHost, s), "reorder download");
  check(hipStreamSynchronize(s), "reorder download");
  for (int c = 0; c < K; ++c)
    for (int i = 0; i < N; ++i) dst[(size_t)c * N + i] = src[(size_t)c * N + perm[i]];
  check(hipMemcpyAsync(dev, dst.data(), sizeof(T) * dst.size(), hipMemcpyHostToDevice, s), "reorder upload");
  check(hipStreamSynchronize(s), "reorder upload");
}
}  // namespace

void MdState::reorder()
{
  wrap();
  std::vector<double> x(3 * (size_t)N_);
  check(hipMemcpyAsync(x.data(), pos_, sizeof(double) * x.size(), hipMemcpyDeviceToHost, stream_), "download");
  check(hipStreamSynchronize(stream_), "download");
  double hi[9];
  box_inverse(box_.h, hi);
  // cells per direction: thickness / (half the Verlet range), at least one
  const double* h = box_.h;
  const double a[3] = {h[0], h[3], h[6]}, b[3] = {h[1], h[4], h[7]}, c[3] = {h[2], h[5], h[8]};
  auto cross = [](const double* u, const double* v, double* w) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
  };
  double bc[3], ca[3], ab[3];
  cross(b, c, bc);
  cross(c, a, ca);
  cross(a, b, ab);
  const double vol = std::fabs(box_volume(h));
  const double* nrm[3] = {bc, ca, ab};
  const double edge = 0.5 * (nep_->model().rcR_max + kSkin);
  long long nc[3];
  for (int d = 0; d < 3; ++d) {
    const double thick = vol / std::sqrt(nrm[d][0] * nrm[d][0] + nrm[d][1] * nrm[d][1] + nrm[d][2] * nrm[d][2]);
    nc[d] = std::max(1LL, (long long)std::floor(thick / edge));
  }
  std::vector<unsigned long long> key(N_);
  for (int i = 0; i < N_; ++i) {
    long long cidx[3];
    for (int d = 0; d < 3; ++d) {
      double s = hi[3 * d] * x[i] + hi[3 * d + 1] * x[N_ + i] + hi[3 * d + 2] * x[2 * (size_t)N_ + i];
      s -= std::floor(s);  // open directions: a fraction of the current extent is enough for locality
      cidx[d] = std::min(nc[d] - 1, std::max(0LL, (long long)std::floor(s * nc[d])));
    }
    key[i] = morton3(cidx[0]) | (morton3(cidx[1]) << 1) | (morton3(cidx[2]) << 2);
  }
  std::vector<int> perm(N_);
  for (int i = 0; i < N_; ++i) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](int p, int q) { return key[p] < key[q]; });
  permute_soa(pos_, N_, 3, perm, stream_);
  permute_soa(vel_, N_, 3, perm, stream_);
  permute_soa(force_, N_, 3, perm, stream_);
  permute_soa(pe_, N_, 1, perm, stream_);
  permute_soa(vir_, N_, 9, perm, stream_);
  permute_soa(mass_, N_, 1, perm, stream_);
  permute_soa(type_, N_, 1, perm, stream_);
  if (unwrapped_) permute_soa(unwrapped_, N_, 3, perm, stream_);
  std::vector<int> o(N_);
  for (int i = 0; i < N_; ++i) o[i] = order_[perm[i]];
  order_.swap(o);
  // the Verlet rows hold storage indices: rebuild them at the next force
  nep_->notify_box_changed();
}

MdState::~MdState()
{
  if (stream_) (void)hipStreamSynchronize(stream_);
  dfree(&pos_);
  dfree(&vel_);
  dfree(&force_);
  dfree(&pe_);
  dfree(&vir_);
  dfree(&mass_);
  dfree(&type_);
  dfree(&tw_.partial);
  dfree(&tw_.thermo);
  dfree(&vmax2_);
  dfree(&unwrapped_);
  dfree(&ck_unwrapped_);
  dfree(&ck_pos_);
  dfree(&ck_vel_);
  dfree(&ck_force_);
  dfree(&ck_pe_);
  dfree(&ck_vir_);
  drop_step_graph();
  dfree(&kstep_);
  dfree(&rand_);
  dfree(&factor_);
  if (hrand_) (void)hipHostFree(hrand_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void MdState::arrays(double** pos, double** vel, double** force, double** pe, double** vir, double** mass,
                     int** type)
{
  if (pos) *pos = pos_;
  if (vel) *vel = vel_;
  if (force) *force = force_;
  if (pe) *pe = pe_;
  if (vir) *vir = vir_;
  if (mass) *mass = mass_;
  if (type) *type = type_;
}

void MdState::wrap()
{
  DevBox b{};
  std::memcpy(b.h, box_.h, sizeof(b.h));
  box_inverse(box_.h, b.hi);
  for (int d = 0; d < 3; ++d) b.pbc[d] = box_.pbc[d];
  check(launch_wrap(b, N_, pos_, stream_), "wrap");
}

void MdState::force(unsigned flags)
{
  ++force_calls_;
  wrap();
  nep_->compute(box_, N_, 0, N_, 0, N_, type_, pos_, nullptr, pe_, force_, vir_, flags & ~1u, stream_);
}

void MdState::enqueue_thermo()
{
  Profiler& p = nep_->profiler();
  p.begin(kProfThermo, stream_);
  check(launch_thermo(N_, box_volume(box_.h), mass_, pe_, vel_, vir_, tw_, stream_), "thermo");
  p.end(stream_);
}

void MdState::thermo(double* out)
{
  enqueue_thermo();
  check(hipMemcpyAsync(out, tw_.thermo, 8 * sizeof(double), hipMemcpyDeviceToHost, stream_), "thermo copy");
  check(hipStreamSynchronize(stream_), "thermo");
}

void MdState::vv(int first_half, double dt)
{
  Profiler& p = nep_->profiler();
  p.begin(kProfIntegrate, stream_);
  check(launch_velocity_verlet(first_half, N_, dt, mass_, pos_, vel_, force_, stream_, unwrapped_),
        "velocity Verlet");
  p.end(stream_);
}

const NepEngine::Tuning& MdState::autotune(unsigned flags)
{
  wrap();
  const NepEngine::Tuning& t = nep_->autotune(box_, N_, N_, type_, pos_, pe_, force_, vir_, flags, stream_);
  have_force_ = true;
  return t;
}

void MdState::drop_step_graph()
{
  if (step_exec_) (void)hipGraphExecDestroy(step_exec_);
  if (step_graph_) (void)hipGraphDestroy(step_graph_);
  step_exec_ = nullptr;
  step_graph_ = nullptr;
  step_key_.clear();
}

void MdState::step_graph(const NepEngine::Plan& plan, double dt, int kind, double coupling, double volume,
                         unsigned flags, bool thermo_on)
{
  std::vector<long long> key;
  nep_->launch_key(key);
  auto bits = [](double v) {
    long long b;
    std::memcpy(&b, &v, sizeof(b));
    return b;
  };
  for (long long v : {(long long)N_, (long long)flags, (long long)kind, (long long)thermo_on, bits(dt), bits(coupling),
                      bits(volume), bits(nep_->half_skin())})
    key.push_back(v);
  for (double v : box_.h) key.push_back(bits(v));
  for (int v : box_.pbc) key.push_back(v);
  for (const void* q : {(const void*)pos_, (const void*)vel_, (const void*)force_, (const void*)pe_,
                        (const void*)vir_, (const void*)mass_, (const void*)type_, (const void*)unwrapped_,
                        (const void*)factor_, (const void*)rand_, (const void*)kstep_, (const void*)tw_.partial,
                        (const void*)tw_.thermo})
    key.push_back((long long)(uintptr_t)q);
  if (step_exec_ && key == step_key_) return;
  drop_step_graph();
  check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "step capture");
  hipError_t e = launch_md_first(plan.box, N_, nep_->model().num_types, dt, mass_, pos_, vel_, force_,
                                 kind != 0 ? factor_ : nullptr, type_, false, nep_->work(), stream_, unwrapped_,
                                 nep_->half_skin());
  if (e == hipSuccess) {
    try {
      nep_->launch_force(plan, N_, 0, N_, 0, N_, type_, nullptr, pe_, force_, vir_, flags & ~1u, stream_);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    e = launch_md_second(N_, dt, mass_, vel_, force_, pe_, vir_, thermo_on, tw_, stream_);
  }
  if (e == hipSuccess && thermo_on)
    e = launch_thermo_final(N_, volume, kind, coupling, 0.0, rand_, factor_, tw_, stream_, rand_ + 2 * kMdChunk,
                            kind != 0 ? kstep_ : nullptr);
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(stream_, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    check(e != hipSuccess ? e : e2, "step capture");
  }
  step_graph_ = g;
  check(hipGraphInstantiate(&step_exec_, step_graph_, nullptr, nullptr, 0), "step graph instantiate");
  step_key_ = std::move(key);
}

void MdState::enable_unwrapped()
{
  if (unwrapped_) return;
  dalloc(&unwrapped_, 3 * (size_t)N_, "unwrapped positions");
  check(hipMemcpyAsync(unwrapped_, pos_, 3 * (size_t)N_ * sizeof(double), hipMemcpyDeviceToDevice, stream_),
        "unwrapped positions");
  check(hipStreamSynchronize(stream_), "unwrapped positions");
}

void MdState::checkpoint(bool save)
{
  struct {
    double* live;
    double** copy;
    size_t n;
  } buf[6] = {{pos_, &ck_pos_, 3}, {vel_, &ck_vel_, 3}, {force_, &ck_force_, 3}, {pe_, &ck_pe_, 1},
              {vir_, &ck_vir_, 9}, {unwrapped_, &ck_unwrapped_, 3}};
  for (auto& e : buf) {
    if (!e.live) continue;
    const size_t bytes = e.n * (size_t)N_ * sizeof(double);
    if (!*e.copy) dalloc(e.copy, e.n * (size_t)N_, "checkpoint");
    if (save) check(hipMemcpyAsync(*e.copy, e.live, bytes, hipMemcpyDeviceToDevice, stream_), "checkpoint");
    else check(hipMemcpyAsync(e.live, *e.copy, bytes, hipMemcpyDeviceToDevice, stream_), "restore");
  }
}

void MdState::run(int steps, double dt_fs, const EnsembleConfig& ens, unsigned flags, double* thermo_out, int every,
                  const Output& on_output)
{
  RunOptions opt;
  opt.every = every;
  opt.on_output = on_output;
  run(steps, dt_fs, ens, flags, thermo_out, opt);
}

void MdState::run(int steps, double dt_fs, const EnsembleConfig& ens, unsigned flags, double* thermo_out,
                  const RunOptions& opt)
{
  if (steps < 0) throw std::invalid_argument("negative step count");
  if (opt.max_distance < 0.0 || !std::isfinite(opt.max_distance))
    throw std::invalid_argument("maximum distance per step should be > 0");
  if (opt.correct_every < 0) throw std::invalid_argument("negative correction interval");
  const double dt = dt_fs / kTimeUnit;
  const int every = opt.every;
  const Output& on_output = opt.on_output;
  const bool adaptive = opt.max_distance > 0.0;
  const bool correct = opt.correct_every > 0 && opt.on_correct;
  Controller ctl(ens, N_, dt);
  elapsed_ = 0.0;
  double th[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (!have_force_) {
    force(flags);
    have_force_ = true;
  }
  // NHC needs the pre-force kinetic energy on the host before every force
  // (scalar chain state): it stays on the synchronous path, as do replicated
  // small boxes, the adaptive step and velocity corrections.  NPT runs the
  // device-resident step below one step per chunk, with a single
  // synchronisation per step for the barostat's pressure.
  int reps[3];
  if (ctl.needs_pre_force_ke() || nep_->small_box(box_, reps) || adaptive || correct) {
    for (int step = 0; step < steps; ++step) {
      if (correct && step % opt.correct_every == 0) {
        check(hipStreamSynchronize(stream_), "velocity correction");
        opt.on_correct(step);
      }
      double dts = dt;
      if (adaptive) {
        // the largest speed at the step's start (its bits: non-negative
        // doubles order as their bit patterns)
        check(launch_max_speed2(N_, vel_, vmax2_, stream_), "max speed");
        unsigned long long bits = 0;
        check(hipMemcpyAsync(&bits, vmax2_, sizeof(bits), hipMemcpyDeviceToHost, stream_), "max speed");
        check(hipStreamSynchronize(stream_), "max speed");
        double v2;
        std::memcpy(&v2, &bits, sizeof(v2));
        if (v2 != 0.0) {
          const double limited = opt.max_distance / std::sqrt(v2);
          if (limited < dt) dts = limited;
        }
      }
      elapsed_ += dts;
      if (ctl.needs_pre_force_ke()) {
        thermo(th);
        const double ke = 0.5 * th[0] * 3.0 * N_ * kBoltzmann;
        if (auto f = ctl.before_force(dts, ke, step, steps))
          check(launch_scale(3 * N_, *f, vel_, stream_), "scale velocity");
      }
      vv(1, dts);
      force(flags);
      vv(0, dts);
      thermo(th);
      if (auto f = ctl.after_force(th, step, steps)) check(launch_scale(3 * N_, *f, vel_, stream_), "scale velocity");
      if (ctl.has_barostat()) {
        const double sc = ctl.barostat_scale(th, box_volume(box_.h), step, steps);
        if (sc != 1.0) {
          for (double& v : box_.h) v *= sc;
          check(launch_scale(3 * N_, sc, pos_, stream_), "scale positions");
          nep_->notify_box_scaled(sc, stream_);
        }
      }
      if (on_output && ((every > 0 && (step + 1) % every == 0) || step + 1 == steps)) {
        check(hipStreamSynchronize(stream_), "output");
        on_output(step + 1, th);
      }
    }
    if (steps == 0) thermo(th);
    if (thermo_out) std::memcpy(thermo_out, th, sizeof(th));
    return;
  }
  // Device-resident chunks: the rebuild decision and the thermostat factor
  // are formed on the device, so a chunk is enqueued without a host
  // synchronisation; the status flags are read once per chunk and a Verlet
  // overflow rolls the chunk back to its checkpoint.
  int kind = 0;
  double coupling = 0.0;
  if (ens.kind == Ensemble::nvt_ber || ens.kind == Ensemble::npt_ber) {
    coupling = 1.0 / ens.t_coupling;
    kind = coupling > 1.0e-5 ? 1 : 0;
  } else if (ens.kind == Ensemble::nvt_bdp || ens.kind == Ensemble::npt_scr) {
    kind = 2;
    coupling = ens.t_coupling > 0.1 ? std::exp(-1.0 / ens.t_coupling) : 0.0;
  }
  const bool baro = ctl.has_barostat();
  const bool thermo_on = kind != 0 || baro;
  constexpr int kChunk = kMdChunk;
  // device and pinned: the thermostat draws of a chunk, then its target
  // temperatures (pinned: then one step's thermo)
  if (!rand_) dalloc(&rand_, 3 * kChunk, "thermostat draws");
  if (!factor_) dalloc(&factor_, 1, "thermostat factor");
  if (!kstep_) dalloc(&kstep_, 1, "chunk step");
  if (!hrand_)
    check(hipHostMalloc((void**)&hrand_, (3 * kChunk + 8) * sizeof(double), hipHostMallocDefault), "pinned");
  double* hth = hrand_ + 3 * kChunk;
  double* targets = rand_ + 2 * kChunk;
  // one captured step, replayed, when asked for (not under the per-kernel
  // profiler, whose events are host work per launch, nor with the filter on
  // a side stream)
  const bool graphs = step_graph_on_ && !baro && !nep_->profiler().on() && nep_->tuning().overlap == 0;
  for (int c0 = 0; c0 < steps;) {
    // a barostat step needs the step's pressure on the host: chunks of one
    int n = baro ? 1 : std::min(kChunk, steps - c0);
    if (on_output && every > 0) n = std::min(n, every - c0 % every);
    const Rng saved = ctl.rng();
    checkpoint(true);
    bool redo = false;
    for (int attempt = 0; attempt < 4; ++attempt) {
      if (kind != 0) {
        // the staging buffer is free: every chunk (and every attempt) ends
        // with the status read's synchronisation, after its upload
        for (int k = 0; k < n; ++k) {
          if (kind == 2) ctl.draw_bdp(&hrand_[2 * k], &hrand_[2 * k + 1]);
          hrand_[2 * kChunk + k] = ctl.target_temperature(c0 + k, steps);
        }
        check(hipMemcpyAsync(rand_, hrand_, 3 * kChunk * sizeof(double), hipMemcpyHostToDevice, stream_),
              "draws");
        // no pending factor at the chunk's first step; its first draws
        check(launch_chunk_start(factor_, kstep_, stream_), "chunk start");
      }
      // per step: fused first half (pending factor, kick, drift, wrap, pack,
      // displacement test), force, fused second half (+ thermo partials),
      // and the thermo reduction that also forms the next factor
      const double volume = box_volume(box_.h);
      double* const fac = kind != 0 ? factor_ : nullptr;
      if (graphs && n > 0) {
        const NepEngine::Plan plan =
            nep_->plan(box_, N_, 0, N_, 0, N_, type_, pos_, nullptr, pe_, force_, vir_, flags & ~1u);
        if (plan.forced) check(launch_request_rebuild(nep_->work(), stream_), "rebuild request");
        step_graph(plan, dt, kind, coupling, volume, flags, thermo_on);
        for (int k = 0; k < n; ++k) check(hipGraphLaunch(step_exec_, stream_), "step graph");
      }
      for (int k = 0; k < n && !graphs; ++k) {
        const NepEngine::Plan plan =
            nep_->plan(box_, N_, 0, N_, 0, N_, type_, pos_, nullptr, pe_, force_, vir_, flags & ~1u);
        Profiler& p = nep_->profiler();
        p.begin(kProfIntegrate, stream_);
        // the displacement test honours what box rescales have used of the skin
        check(launch_md_first(plan.box, N_, nep_->model().num_types, dt, mass_, pos_, vel_, force_, fac, type_,
                              plan.forced, nep_->work(), stream_, unwrapped_, nep_->half_skin()),
              "velocity Verlet (first half)");
        p.end(stream_);
        nep_->launch_force(plan, N_, 0, N_, 0, N_, type_, nullptr, pe_, force_, vir_, flags & ~1u, stream_);
        p.begin(kProfIntegrate, stream_);
        check(launch_md_second(N_, dt, mass_, vel_, force_, pe_, vir_, thermo_on, tw_, stream_),
              "velocity Verlet (second half)");
        p.end(stream_);
        if (thermo_on) {
          p.begin(kProfThermo, stream_);
          check(launch_thermo_final(N_, volume, kind, coupling, 0.0, rand_, factor_, tw_, stream_, targets,
                                    kind != 0 ? kstep_ : nullptr),
                "thermostat");
          p.end(stream_);
        }
      }
      const bool pending = kind != 0 && n > 0;
      if (pending) check(launch_scale_by(3 * N_, factor_, vel_, stream_), "thermostat scale");
      // a barostat step's thermo (before the thermostat rescale) comes back
      // with the status words, in the same synchronisation
      if (baro)
        check(hipMemcpyAsync(hth, tw_.thermo, 8 * sizeof(double), hipMemcpyDeviceToHost, stream_), "thermo copy");
      redo = nep_->settle(stream_);
      if (!redo) break;
      checkpoint(false);
      ctl.set_rng(saved);
    }
    if (redo) throw std::runtime_error("Verlet rows kept overflowing");
    if (baro) {
      // the isotropic box step from the pressure before the thermostat
      // rescale (ensemble_control.cpp:262-304): box and positions scaled,
      // the Verlet list kept while the scale leaves room in the skin
      std::memcpy(th, hth, sizeof(th));
      const double sc = ctl.barostat_scale(th, box_volume(box_.h), c0, steps);
      if (sc != 1.0) {
        for (double& v : box_.h) v *= sc;
        check(launch_scale(3 * N_, sc, pos_, stream_), "scale positions");
        nep_->notify_box_scaled(sc, stream_);
      }
    }
    force_calls_ += n;
    c0 += n;
    elapsed_ += dt * n;
    if (on_output && ((every > 0 && c0 % every == 0) || c0 == steps)) {
      // the last step's thermo before its rescale (NVE: nothing rescales)
      double tho[8];
      if (baro) {
        std::memcpy(tho, th, sizeof(tho));
        check(hipStreamSynchronize(stream_), "output");
      } else {
        if (kind == 0) enqueue_thermo();
        check(hipMemcpyAsync(tho, tw_.thermo, sizeof(tho), hipMemcpyDeviceToHost, stream_), "thermo copy");
        check(hipStreamSynchronize(stream_), "output");
      }
      on_output(c0, tho);
    }
  }
  // a barostat run reports its last step's thermo before the rescales, as
  // the per-step path does; the others the final state's
  if (!baro || steps == 0) thermo(th);
  if (thermo_out) std::memcpy(thermo_out, th, sizeof(th));
}

}  // namespace dmgx
