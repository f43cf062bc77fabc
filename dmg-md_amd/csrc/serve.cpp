// serve.cpp -- the serving loop (see serve.hpp).
#include "serve.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>

namespace dmgx {

namespace {

constexpr int kTuneMinAtoms = 65536;  // owned atoms of the batch that triggers the autotuner

void check(hipError_t e, const char* what)
{
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

double now_s()
{
  using C = std::chrono::steady_clock;
  return std::chrono::duration<double>(C::now().time_since_epoch()).count();
}

// fractional coordinates of x under h^-1
void to_frac(const double* hi, const double* x, double* s)
{
  for (int d = 0; d < 3; ++d) s[d] = hi[3 * d] * x[0] + hi[3 * d + 1] * x[1] + hi[3 * d + 2] * x[2];
}

void to_cart(const double* h, const double* s, double* x)
{
  for (int d = 0; d < 3; ++d) x[d] = h[3 * d] * s[0] + h[3 * d + 1] * s[1] + h[3 * d + 2] * s[2];
}

// perpendicular widths of the cell spanned by the columns of h
void thickness(const double* h, double* t)
{
  const double a[3] = {h[0], h[3], h[6]}, b[3] = {h[1], h[4], h[7]}, c[3] = {h[2], h[5], h[8]};
  const double* pairs[3][2] = {{b, c}, {c, a}, {a, b}};
  const double vol = std::fabs(box_volume(h));
  for (int d = 0; d < 3; ++d) {
    const double* u = pairs[d][0];
    const double* v = pairs[d][1];
    const double x = u[1] * v[2] - u[2] * v[1], y = u[2] * v[0] - u[0] * v[2], z = u[0] * v[1] - u[1] * v[0];
    t[d] = vol / std::sqrt(x * x + y * y + z * z);
  }
}

}  // namespace

template <class T>
void Server::DBuf<T>::need(long long n, const char* what)
{
  n = std::max<long long>(n, 1);
  if (n <= cap && p) return;
  // grows by half again at least: a freed block synchronises the device,
  // and a growing batch would otherwise reallocate at every relayout
  const long long want = std::max<long long>(n, cap + cap / 2);
  release();
  if (hipMalloc((void**)&p, (size_t)want * sizeof(T)) == hipSuccess) {
    cap = want;
    return;
  }
  (void)hipGetLastError();
  if (want == n || hipMalloc((void**)&p, (size_t)n * sizeof(T)) != hipSuccess) {
    p = nullptr;
    throw std::runtime_error(std::string("device allocation failed: ") + what);
  }
  cap = n;
}

template <class T>
void Server::DBuf<T>::release()
{
  if (p) (void)hipFree(p);
  p = nullptr;
  cap = 0;
}

// Device buffers for request packs, reused by size class (powers of two from
// 4 KiB): freeing device memory would synchronise the device under the
// serving loop, so buffers are only freed with the server.
// A pool miss allocates a slab of kSlabBytes (or one block, if larger) and
// splits it into blocks of the class: allocation calls serialise with the
// serving loop's, so they are kept few.
constexpr size_t kSlabBytes = size_t(32) << 20;

void* Server::pool_get(std::multimap<size_t, void*>& pool, std::vector<void*>& slabs, bool pinned, size_t bytes,
                       size_t* cap)
{
  size_t c = 4096;
  while (c < bytes) c <<= 1;
  *cap = c;
  {
    std::lock_guard<std::mutex> lk(pack_mu_);
    auto it = pool.find(c);
    if (it != pool.end()) {
      void* p = it->second;
      pool.erase(it);
      return p;
    }
  }
  const size_t slab = std::max(c, kSlabBytes);
  void* p = nullptr;
  if (pinned) {
    if (hipHostMalloc(&p, slab, hipHostMallocDefault) != hipSuccess)
      throw std::runtime_error("out of pinned host memory for a request upload");
  } else if (hipMalloc(&p, slab) != hipSuccess) {
    throw std::runtime_error("out of device memory for a request upload");
  }
  std::lock_guard<std::mutex> lk(pack_mu_);
  slabs.push_back(p);
  for (size_t off = c; off + c <= slab; off += c) pool.emplace(c, static_cast<char*>(p) + off);
  return p;
}

void* Server::pack_get(size_t bytes, size_t* cap) { return pool_get(pack_free_, pack_slabs_, false, bytes, cap); }

void Server::pack_put(void* p, size_t cap)
{
  if (!p) return;
  std::lock_guard<std::mutex> lk(pack_mu_);
  pack_free_.emplace(cap, p);
}

// pinned host buffers for building packs, reused the same way
void* Server::pin_get(size_t bytes, size_t* cap) { return pool_get(pin_free_, pin_slabs_, true, bytes, cap); }

struct Server::Request {
  long long id = 0;
  int model = 0;
  Box box{};
  double hi[9] = {};
  int n = 0;
  int ng = 0;  // periodic images within rc + skin + margin of the box
  int steps = 0;
  double dt = 0;
  long long bytes = 0;  // device bytes while in flight (slots x the model's bytes per slot)
  // the request's atoms and images on the device until its first layout
  // (ServeSlots pack layout), from the server's pack pool
  void* pack = nullptr;
  size_t pack_cap = 0;
  void* hpack = nullptr;  // its pinned host copy, until the upload is consumed
  size_t hpack_cap = 0;
  hipEvent_t ready = nullptr;  // the upload's completion on the upload stream
  double lo[3] = {0, 0, 0}, hi_ext[3] = {0, 0, 0};  // extent of owned + ghost positions
  // progress
  int state = 0;  // 0: needs its first force evaluation, 1: MD steps, 2: done
  int step = 0;
  bool live = true;
  double t_submit = 0, t_done = 0;
  std::string error;
  // placement in the current layout of its model's lane
  double trans[3] = {0, 0, 0};
  int slot0 = -1, gslot0 = -1;  // first owned / ghost slot in the layout
  bool placed = false;
  // types and masses as submitted (a new image shell is packed from them)
  std::vector<int> types;
  std::vector<double> masses;
  // results: block of kServeOutDoubles n doubles (kernels.hpp)
  std::shared_ptr<ResultBlock> res;
  long long res_off = 0;
  long long footprint() const { return (long long)n + ng; }
};

// One model's batch: its engine, stream, layout arrays and captured force
// launches, driven by its own thread.  `active` (the layout's requests, in
// layout order) changes only on that thread and under mu_; `incoming`,
// `dirty` and `slots_now` are read and written under mu_.
struct Server::Lane {
  NepEngine* nep = nullptr;
  hipStream_t stream = nullptr;
  std::thread thread;
  std::vector<std::shared_ptr<Request>> active;
  std::vector<std::shared_ptr<Request>> incoming;  // admitted, placed by the next layout
  bool dirty = false;  // membership changed since the last layout
  long long seen_gen = -1;  // admit_gen_ when this lane last ran admission
  long long slots_now = 0;  // slots of the current layout
  bool any_md = false;  // requests taking an MD step in the frame being evaluated
  int n_owned = 0, n_ghost = 0, n_total = 0;
  Box batch_box{};
  DBuf<double> pos, vel, force, pe, vir, mass;
  DBuf<int> type;
  DBuf<int> owner;     // slot -> owned slot it is an image of (itself for owned slots)
  DBuf<int> image;     // slot -> image code (image_code(0, 0, 0) for owned slots)
  DBuf<int> gstart;    // owned slot -> first ghost slot of its images [n_owned + 1]
  DBuf<double> g_shift;  // ghost k -> box shift [3]
  // the next layout's per-slot arrays, assembled on the device and swapped in
  DBuf<double> pos2, vel2, force2, mass2, g_shift2;
  DBuf<double> home, home2;  // owned slots' positions where their request's shell was built
  // the previous layout's Verlet rows (and the displacement references of its
  // slots), copied out of the engine at a relayout for the continuing requests
  DBuf<int> vl0, vrev0, vn0;
  DBuf<double> x00;
  DBuf<int> req_flag;        // active request -> 1 once an atom drifted past the shell's budget
  DBuf<int> req_moved;       // layout request -> 1 once an atom moved skin/2 since its rows were built
  DBuf<int4> req_range;      // layout request -> first owned slot, owned count, first ghost slot, ghost count
  DBuf<int> req_off, lists;  // a rebuild's compacted centre and slot lists of the requests it remakes
  int n_req = 0;             // requests of the current layout
  int* h_drift = nullptr;    // pinned copy of req_flag, read after each frame
  int h_drift_cap = 0;
  DBuf<int> type2, owner2, image2, gstart2;
  DBuf<int> atom_req;    // owned slot -> active request index
  DBuf<int> req_md;      // active request -> 1 when it takes an MD step this iteration
  DBuf<double> req_dt;   // active request -> dt (natural units)
  DBuf<ServeSlots> slots;  // layout table (one row per active request)
  DBuf<ServeOut> outq;     // finished requests (extract table)
  DBuf<double> out;        // finished requests' result blocks
  char* h_stage = nullptr;  // pinned staging for this lane's uploads
  size_t h_stage_bytes = 0;
  bool forced_rebuild = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;

  void stage(size_t bytes)
  {
    if (bytes <= h_stage_bytes) return;
    if (h_stage) (void)hipHostFree(h_stage);
    h_stage = nullptr;
    bytes = std::max(bytes, h_stage_bytes * 2);
    h_stage_bytes = 0;
    check(hipHostMalloc((void**)&h_stage, bytes, hipHostMallocDefault), "pinned staging");
    h_stage_bytes = bytes;
  }
  bool live() const
  {
    for (auto& r : active)
      if (r->live) return true;
    return false;
  }
  ImageMap map() const
  {
    ImageMap m;
    m.owner = owner.p;
    m.image = image.p;
    m.gstart = gstart.p;
    if (n_req > 0) {
      // requests are disjoint: a rebuild remakes only the moved ones' rows
      m.slot_req = atom_req.p;
      m.req_moved = req_moved.p;
      m.n_owned = n_owned;
      m.n_req = n_req;
      m.req_range = req_range.p;
      m.req_off = req_off.p;
      m.lists = lists.p;
    }
    return m;
  }
  void release()
  {
    for (DBuf<double>* b : {&pos, &vel, &force, &pe, &vir, &mass, &pos2, &vel2, &force2, &mass2, &req_dt, &g_shift,
                            &g_shift2, &out, &home, &home2, &x00})
      b->release();
    for (DBuf<int>* b : {&type, &type2, &atom_req, &req_md, &owner, &owner2, &image, &image2, &gstart, &gstart2,
                         &req_flag, &req_moved, &req_off, &lists, &vl0, &vrev0, &vn0})
      b->release();
    if (h_drift) (void)hipHostFree(h_drift);
    h_drift = nullptr;
    h_drift_cap = 0;
    slots.release();
    outq.release();
    req_range.release();
    if (h_stage) (void)hipHostFree(h_stage);
    h_stage = nullptr;
  }
};

// device bytes per slot of a lane's layout: position, velocity and force
// (3 doubles each) and mass, type, owner, image, ghost shift and home
// position, all double-buffered; potential, virial, the atom->request index,
// the result block of a finished atom and the carried displacement
// reference.  (Layout buffers grow by half again when they grow; the default
// budget, half the free memory, leaves the other half for that slack.)
constexpr double kLayoutSlotBytes =
    2 * (9 * 8 + 8 + 3 * 4 + 3 * 8 + 3 * 8) + 8 + 9 * 8 + 4 + kServeOutDoubles * 8 + 3 * 8;

double Server::slot_bytes(int model, bool owned) const
{
  // owned slots are centres: their rows, plus the previous layout's rows and
  // reverse slots copied at a relayout for the continuing requests
  const NepEngine& e = *lanes_[(size_t)model]->nep;
  if (!owned) return e.bytes_per_slot(false) + kLayoutSlotBytes;
  return e.bytes_per_slot(true) + kLayoutSlotBytes + 2.0 * 4.0 * std::max(e.work().MV, 1);
}

void Server::pin_put(void* p, size_t cap)
{
  if (!p) return;
  std::lock_guard<std::mutex> lk(pack_mu_);
  pin_free_.emplace(cap, p);
}

// returns a request's upload buffers and event once the upload is consumed
// (or never will be)
void Server::release_upload(Request& r)
{
  if (r.ready) (void)hipEventDestroy(r.ready);
  r.ready = nullptr;
  pack_put(r.pack, r.pack_cap);
  r.pack = nullptr;
  pin_put(r.hpack, r.hpack_cap);
  r.hpack = nullptr;
}

Server::Server(const std::vector<NepEngine*>& models, const ServeConfig& cfg) : cfg_(cfg)
{
  if (models.empty()) throw std::invalid_argument("a server needs at least one potential");
  for (size_t a = 0; a < models.size(); ++a) {
    if (!models[a]) throw std::invalid_argument("null potential");
    for (size_t b = 0; b < a; ++b)
      if (models[a] == models[b]) throw std::invalid_argument("a potential handle is given twice");
  }
  if (cfg_.max_atoms <= 0 || cfg_.max_requests <= 0) throw std::invalid_argument("bad serving limits");
  if (cfg_.max_atoms >= (1LL << 30)) throw std::invalid_argument("max_atoms must stay below 2^30");
  if (cfg_.hbm_bytes < 0) throw std::invalid_argument("negative HBM budget");
  budget_ = cfg_.hbm_bytes;
  if (budget_ == 0) {
    size_t free_b = 0, total_b = 0;
    check(hipMemGetInfo(&free_b, &total_b), "device memory info");
    budget_ = (long long)(free_b / 2);
  }
  st_.hbm_budget = budget_;
  try {
    for (NepEngine* m : models) {
      auto L = std::make_unique<Lane>();
      L->nep = m;
      lanes_.push_back(std::move(L));
      Lane& l = *lanes_.back();
      // each batch grows request by request: size its engine once, up front,
      // for what the budget allows, instead of reallocating (and
      // synchronising the device) as it grows
      // (with a ghost per owned slot and rows for the owned ones)
      const int k = (int)lanes_.size() - 1;
      const long long fit = (long long)(budget_ / (slot_bytes(k, true) + slot_bytes(k, false)));
      const long long cap = std::min({cfg_.max_atoms, cfg_.reserve_atoms, 2 * fit});
      m->reserve((int)std::max<long long>(cap, 1), (int)std::max<long long>(cap / 2, 1));
      m->reserve_cells(1LL << 23);
      check(hipStreamCreate(&l.stream), "serving stream");
    }
    check(hipStreamCreate(&upload_), "serving upload stream");
  } catch (...) {
    for (auto& L : lanes_) {
      if (L->stream) (void)hipStreamDestroy(L->stream);
      L->release();
    }
    throw;
  }
  for (auto& L : lanes_) {
    Lane* l = L.get();
    l->thread = std::thread([this, l] { lane_worker(*l); });
  }
}

Server::~Server()
{
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  for (auto& L : lanes_)
    if (L->thread.joinable()) L->thread.join();
  for (auto& L : lanes_) {
    if (L->stream) (void)hipStreamSynchronize(L->stream);
    drop_graph(*L);
    L->release();
  }
  if (upload_) (void)hipStreamSynchronize(upload_);
  for (auto& kv : all_) release_upload(*kv.second);  // every request, waiting ones included
  // dropping the requests returns their result blocks to the pinned pool
  waiting_.clear();
  for (auto& L : lanes_) {
    L->active.clear();
    L->incoming.clear();
  }
  all_.clear();
  for (void* p : pack_slabs_) (void)hipFree(p);
  for (void* p : pin_slabs_) (void)hipHostFree(p);
  if (upload_) (void)hipStreamDestroy(upload_);
  for (auto& L : lanes_)
    if (L->stream) (void)hipStreamDestroy(L->stream);
}

void Server::make_pack(Request& rq, const int* type, const double* mass, const double* pos, const double* vel,
                       const double* force, long long limit)
{
  Request* r = &rq;
  const int N = r->n;
  const int model = r->model;
  NepEngine* nep = lanes_[(size_t)model]->nep;
  const Box& box = r->box;
  std::vector<double> wpos(3 * (size_t)N);  // wrapped into the box, SoA
  // wrap into the box and collect the periodic images of the shell
  const double rc = nep->model().rcR_max;
  // a single point never moves: its shell needs no margin for motion
  const double d1 = rc + kSkin + (r->steps > 0 ? cfg_.margin : 0.0);
  double th[3];
  thickness(box.h, th);
  int K[3];
  for (int d = 0; d < 3; ++d) {
    K[d] = box.pbc[d] ? (int)std::ceil(d1 / th[d]) : 0;
    if (K[d] > kImageMax) throw std::invalid_argument("request box too thin for its periodic image shell");
  }
  std::vector<double> frac(3 * (size_t)N);
  for (int i = 0; i < N; ++i) {
    double x[3] = {pos[i], pos[N + i], pos[2 * N + i]}, s[3];
    to_frac(r->hi, x, s);
    for (int d = 0; d < 3; ++d) {
      if (!std::isfinite(s[d])) throw std::invalid_argument("non-finite position");
      if (box.pbc[d]) s[d] -= std::floor(s[d]);
      frac[3 * i + d] = s[d];
    }
    to_cart(box.h, s, x);
    for (int d = 0; d < 3; ++d) wpos[(size_t)d * N + i] = x[d];
  }
  std::vector<int> g_owner, g_image;
  std::vector<double> g_shift;
  // every image within d1 of the box (per face pair: a lower bound of the
  // Euclidean distance, so the shell is a superset of what the rows need);
  // an owned atom's Verlet neighbours at any rebuild lie within
  // rc + skin of the box plus the request's motion, covered by the margin
  // (an atom's images in increasing image code, c then b then a: the
  // reverse search of the Verlet build looks them up by bisection)
  for (int i = 0; i < N; ++i)
    for (int ic = -K[2]; ic <= K[2]; ++ic)
      for (int ib = -K[1]; ib <= K[1]; ++ib)
        for (int ia = -K[0]; ia <= K[0]; ++ia) {
          if (ia == 0 && ib == 0 && ic == 0) continue;
          const int img[3] = {ia, ib, ic};
          double far = 0.0;
          for (int d = 0; d < 3; ++d) {
            const double u = frac[3 * i + d] + img[d];
            const double e = std::max(0.0, std::max(-u, u - 1.0)) * th[d];
            far = std::max(far, e);
          }
          if (far > d1) continue;
          const double sh[3] = {box.h[0] * ia + box.h[1] * ib + box.h[2] * ic,
                                box.h[3] * ia + box.h[4] * ib + box.h[5] * ic,
                                box.h[6] * ia + box.h[7] * ib + box.h[8] * ic};
          g_owner.push_back(i);
          g_image.push_back(image_code(ia, ib, ic));
          g_shift.insert(g_shift.end(), sh, sh + 3);
        }
  if (g_owner.size() >= (1u << 30)) throw std::invalid_argument("request has too many periodic images");
  r->ng = (int)g_owner.size();
  const int NG = r->ng;
  for (int d = 0; d < 3; ++d) {
    r->lo[d] = 1e300;
    r->hi_ext[d] = -1e300;
  }
  auto extend = [&](const double* x) {
    for (int d = 0; d < 3; ++d) {
      r->lo[d] = std::min(r->lo[d], x[d]);
      r->hi_ext[d] = std::max(r->hi_ext[d], x[d]);
    }
  };
  for (int i = 0; i < N; ++i) {
    const double x[3] = {wpos[i], wpos[(size_t)N + i], wpos[2 * (size_t)N + i]};
    extend(x);
  }
  for (int g = 0; g < NG; ++g) {
    const int o = g_owner[g];
    const double x[3] = {wpos[o] + g_shift[3 * g], wpos[(size_t)N + o] + g_shift[3 * g + 1],
                         wpos[2 * (size_t)N + o] + g_shift[3 * g + 2]};
    extend(x);
  }
  const size_t doff = serve_pack_double_offset(N, NG), bytes = serve_pack_bytes(N, NG);
  r->bytes = (long long)std::ceil(N * slot_bytes(model, true) + NG * slot_bytes(model, false)) + (long long)bytes;
  if (r->footprint() > cfg_.max_atoms)
    throw std::invalid_argument("request (with its periodic shell) exceeds the serving atom budget");
  if (r->bytes > limit)
    throw std::invalid_argument("request (with its periodic shell) needs " + std::to_string(r->bytes) +
                                " device bytes, more than the serving HBM budget of " + std::to_string(limit));
  // the pack: built here and uploaded on the submitting thread, so the
  // serving loop only copies it device-to-device when it places the request
  r->hpack = pin_get(bytes, &r->hpack_cap);
  char* host = static_cast<char*>(r->hpack);
  int* pi = reinterpret_cast<int*>(host);
  double* pd = reinterpret_cast<double*>(host + doff);
  std::copy(type, type + N, pi);
  for (int g = 0, a = 0; a < N; ++a) {
    pi[N + a] = g;
    while (g < NG && g_owner[g] == a) ++g;
  }
  std::copy(g_owner.begin(), g_owner.end(), pi + 2 * (size_t)N);
  std::copy(g_image.begin(), g_image.end(), pi + 2 * (size_t)N + NG);
  std::copy(mass, mass + N, pd);
  for (int a = 0; a < N; ++a)
    for (int d = 0; d < 3; ++d) {
      pd[N + 3 * (size_t)a + d] = wpos[(size_t)d * N + a];
      pd[4 * (size_t)N + 3 * (size_t)a + d] = vel ? vel[(size_t)d * N + a] : 0.0;
      pd[7 * (size_t)N + 3 * (size_t)a + d] = force ? force[(size_t)d * N + a] : 0.0;
    }
  std::copy(g_shift.begin(), g_shift.end(), pd + 10 * (size_t)N);
  // asynchronous upload: the lane's stream waits for `ready` before it
  // places the request
  r->pack = pack_get(bytes, &r->pack_cap);
  hipError_t e = hipEventCreateWithFlags(&r->ready, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMemcpyAsync(r->pack, host, bytes, hipMemcpyHostToDevice, upload_);
  if (e == hipSuccess) e = hipEventRecord(r->ready, upload_);
  if (e != hipSuccess) {
    release_upload(*r);
    throw std::runtime_error(std::string("request upload: ") + hipGetErrorString(e));
  }
}

long long Server::submit(int model, const Box& box, int N, const int* type, const double* mass, const double* pos,
                         const double* vel, int steps, double dt_fs)
{
  if (model < 0 || model >= (int)lanes_.size()) throw std::invalid_argument("model index out of range");
  if (N <= 0 || !type || !mass || !pos) throw std::invalid_argument("a request needs atoms, types, masses, positions");
  if (steps < 0) throw std::invalid_argument("negative step count");
  if (steps > 0 && !(dt_fs > 0.0)) throw std::invalid_argument("MD requests need dt_fs > 0");
  const double t_call = now_s();
  NepEngine* nep = lanes_[(size_t)model]->nep;
  auto r = std::make_shared<Request>();
  r->model = model;
  r->box = box;
  const double vol = box_volume(box.h);
  if (!(std::fabs(vol) > 0.0) || !std::isfinite(vol)) throw std::invalid_argument("box has no volume");
  box_inverse(box.h, r->hi);
  r->n = N;
  r->steps = steps;
  r->dt = dt_fs / kTimeUnit;
  const int T = nep->model().num_types;
  for (int i = 0; i < N; ++i) {
    if (type[i] < 0 || type[i] >= T) throw std::invalid_argument("atom type out of range");
    if (!(mass[i] > 0.0)) throw std::invalid_argument("masses must be positive");
  }
  make_pack(*r, type, mass, pos, vel, nullptr, budget_);
  r->types.assign(type, type + N);
  r->masses.assign(mass, mass + N);
  std::lock_guard<std::mutex> lk(mu_);
  if (!error_.empty()) {
    release_upload(*r);
    throw std::runtime_error("serving loop failed: " + error_);
  }
  r->id = next_id_++;
  r->t_submit = t_call;
  if (t_first_ < 0) t_first_ = r->t_submit;
  waiting_.push_back(r);
  all_[r->id] = r;
  st_.submitted++;
  admit_gen_++;
  cv_.notify_all();
  return r->id;
}

bool Server::poll(long long id, bool wait, double* potential, double* force, double* position, double* velocity,
                  double* virial, double* latency_ms)
{
  std::unique_lock<std::mutex> lk(mu_);
  auto it = all_.find(id);
  if (it == all_.end()) throw std::invalid_argument("unknown request id");
  auto r = it->second;
  if (wait) done_cv_.wait(lk, [&] { return r->state == 2 || !error_.empty(); });
  if (r->state != 2) {
    if (!error_.empty()) throw std::runtime_error("serving loop failed: " + error_);
    return false;
  }
  if (!r->error.empty()) throw std::runtime_error(r->error);
  const size_t n = (size_t)r->n;
  const double* blk = static_cast<const double*>(r->res->p) + r->res_off;
  if (potential) std::copy(blk, blk + n, potential);
  if (force) std::copy(blk + n, blk + 4 * n, force);
  if (position) std::copy(blk + 4 * n, blk + 7 * n, position);
  if (velocity) std::copy(blk + 7 * n, blk + 10 * n, velocity);
  if (virial) std::copy(blk + 10 * n, blk + 19 * n, virial);
  if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
  return true;
}

void Server::drain()
{
  std::unique_lock<std::mutex> lk(mu_);
  done_cv_.wait(lk, [&] { return st_.completed == st_.submitted || !error_.empty(); });
  if (!error_.empty()) throw std::runtime_error("serving loop failed: " + error_);
}

void Server::reset_stats()
{
  std::lock_guard<std::mutex> lk(mu_);
  if (!error_.empty()) throw std::runtime_error("serving loop failed: " + error_);
  if (st_.completed != st_.submitted) throw std::invalid_argument("statistics are reset on a drained server only");
  const long long budget = st_.hbm_budget;
  st_ = ServeStats{};
  st_.hbm_budget = budget;
  latencies_.clear();
  t_first_ = -1;
  t_last_ = 0;
}

ServeStats Server::stats()
{
  std::lock_guard<std::mutex> lk(mu_);
  ServeStats s = st_;
  if (!latencies_.empty()) {
    std::vector<double> l = latencies_;
    std::sort(l.begin(), l.end());
    auto q = [&](double f) { return l[std::min(l.size() - 1, (size_t)std::floor(f * (l.size() - 1) + 0.5))]; };
    s.p50_ms = q(0.50) * 1e3;
    s.p99_ms = q(0.99) * 1e3;
    s.max_ms = l.back() * 1e3;
    double sum = 0;
    for (double v : l) sum += v;
    s.mean_ms = sum / l.size() * 1e3;
  }
  if (t_first_ >= 0 && t_last_ > t_first_) {
    s.elapsed_s = t_last_ - t_first_;
    s.frames_per_s = s.frames / s.elapsed_s;
  }
  return s;
}

// ---------------------------------------------------------------------------
// worker
// ---------------------------------------------------------------------------

void Server::lane_worker(Lane& L)
{
  for (;;) {
    bool relay = false;
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || !error_.empty() || L.dirty || L.live() || admit_gen_ != L.seen_gen; });
      if (stop_ || !error_.empty()) return;
      L.seen_gen = admit_gen_;
      admit();
      relay = L.dirty;
      if (!relay && !L.live()) continue;
    }
    try {
      const double t0 = now_s();
      if (relay) {
        std::shared_lock<std::shared_mutex> g = gpu_shared();
        relayout(L);
      }
      if (!L.live()) continue;
      if (needs_tuning(L)) {
        std::unique_lock<std::shared_mutex> g = gpu_exclusive();
        tune(L);
      }
      // a stream capture may not overlap another thread's device-wide
      // synchronisation (a buffer freed while a layout grows): the graph of
      // a new layout is captured with the device to this lane.  With several
      // lanes those stalls cost more than the launches a replay saves
      // (measured: 82-84 k frames/s launched directly against 77-80 k with
      // graphs on the two-model stream), so only a single lane captures.
      if (cfg_.use_graph && lanes_.size() == 1 && !L.graph_exec) {
        drop_graph(L);
        {
          std::unique_lock<std::shared_mutex> g = gpu_exclusive();
          capture_force(L, cfg_.flags & ~3u);
        }
        // the instantiation touches no stream: the other lanes go on meanwhile
        check(hipGraphInstantiate(&L.graph_exec, L.graph, nullptr, nullptr, 0), "serve graph instantiate");
      }
      std::shared_lock<std::shared_mutex> g = gpu_shared();
      const double t1 = now_s();
      // the half-kick and the image-shell check; new shells (and a new
      // layout) before the force evaluation where an atom left a shell's
      // budget
      enqueue_kick(L);
      if (L.any_md) {
        check(hipStreamSynchronize(L.stream), "serve shell check");
        if (needs_shells(L)) {
          const double ts = now_s();
          reshell(L);
          relayout(L);
          upload_md_flags(L);
          std::lock_guard<std::mutex> lk(mu_);
          st_.reshell_s += now_s() - ts;
        }
      }
      enqueue_force(L);
      finish_frame(L);
      const double t2 = now_s();
      retire(L);
      const double t3 = now_s();
      std::lock_guard<std::mutex> lk(mu_);
      st_.relayout_s += t1 - t0;
      st_.iterate_s += t2 - t1;
      st_.retire_s += t3 - t2;
      st_.iterations++;
      long long slots = 0;
      for (auto& l : lanes_) slots += l->slots_now;
      st_.max_slots = std::max(st_.max_slots, slots);
      // a lane whose finished requests hold a quarter of its slots drops
      // them at once (a new layout of its live requests), whether or not
      // requests are waiting: every frame evaluates the whole layout
      if (!L.dirty) {
        long long live = 0, dead = 0;
        for (auto& r : L.active) (r->live ? live : dead) += r->footprint();
        if (live > 0 && 4 * dead >= live + dead) L.dirty = true;
      }
    } catch (const std::exception& e) {
      std::lock_guard<std::mutex> lk(mu_);
      if (error_.empty()) error_ = e.what();
      cv_.notify_all();
      done_cv_.notify_all();
      return;
    }
  }
}

// Turnstile: a lane waiting for the device to itself holds gpu_turn_, so no
// lane starts another shared stretch before it gets its turn.
std::shared_lock<std::shared_mutex> Server::gpu_shared()
{
  { std::lock_guard<std::mutex> t(gpu_turn_); }
  return std::shared_lock<std::shared_mutex>(gpu_mu_);
}

std::unique_lock<std::shared_mutex> Server::gpu_exclusive()
{
  std::lock_guard<std::mutex> t(gpu_turn_);
  return std::unique_lock<std::shared_mutex>(gpu_mu_);
}

bool Server::needs_tuning(Lane& L) const { return !L.nep->tuning().tuned && L.n_owned >= kTuneMinAtoms; }

void Server::tune(Lane& L)
{
  const unsigned flags = cfg_.flags & ~3u;
  const ImageMap map = L.map();
  check(launch_serve_ghosts(L.n_owned, L.n_ghost, L.n_total, L.owner.p + L.n_owned, L.g_shift.p, L.pos.p, L.stream),
        "serve ghosts");
  L.nep->autotune(L.batch_box, L.n_total, L.n_owned, L.type.p, L.pos.p, L.pe.p, L.force.p, L.vir.p,
                  flags | (L.forced_rebuild ? 2u : 0u), L.stream, &map);
  drop_graph(L);
  L.forced_rebuild = false;
}

// Called with mu_ held.  Admits waiting requests in submission order while
// the device bytes of the live requests plus the new ones stay within the
// HBM budget (and the slot and request limits); the first request that does
// not fit stops admission, so large requests are not starved.  A lane takes
// its admitted requests into a new layout when it is empty, when it is
// already due for one (it drops its finished requests), or when those
// requests' slots plus the slots of its finished requests still in the
// layout reach a quarter of the layout, so that a layout (and the Verlet
// rebuild it forces) is not redone for every single arrival; requests a lane
// declines stay waiting without holding budget.
void Server::admit()
{
  long long live_bytes = 0, live_slots = 0;
  int live = 0;
  const size_t nl = lanes_.size();
  std::vector<long long> lane_live(nl, 0), lane_dead(nl, 0);
  for (size_t l = 0; l < nl; ++l) {
    for (auto* list : {&lanes_[l]->active, &lanes_[l]->incoming})
      for (auto& r : *list) {
        if (r->live) {
          live_bytes += r->bytes;
          live_slots += r->footprint();
          lane_live[l] += r->footprint();
          ++live;
        } else {
          lane_dead[l] += r->footprint();
        }
      }
  }
  std::vector<std::vector<size_t>> cand(nl);
  std::vector<long long> add(nl, 0);
  long long add_bytes = 0, add_slots = 0;
  int added = 0;
  for (size_t k = 0; k < waiting_.size(); ++k) {
    const Request& r = *waiting_[k];
    if (live_bytes + add_bytes + r.bytes > budget_ || live_slots + add_slots + r.footprint() > cfg_.max_atoms ||
        live + added + 1 > cfg_.max_requests)
      break;
    add_bytes += r.bytes;
    add_slots += r.footprint();
    ++added;
    cand[(size_t)r.model].push_back(k);
    add[(size_t)r.model] += r.footprint();
  }
  std::vector<char> take(waiting_.size(), 0);
  bool any = false;
  for (size_t l = 0; l < nl; ++l) {
    if (cand[l].empty()) continue;
    Lane& L = *lanes_[l];
    const bool empty = lane_live[l] == 0;
    // a lane already due for a new layout (its finished requests are being
    // dropped) takes whatever was admitted for it at no extra rebuild
    if (!empty && !L.dirty && 4 * (add[l] + lane_dead[l]) < lane_live[l] + lane_dead[l]) continue;
    for (size_t k : cand[l]) {
      L.incoming.push_back(waiting_[k]);
      take[k] = 1;
    }
    L.dirty = true;
    any = true;
  }
  if (!any) return;
  std::deque<std::shared_ptr<Request>> rest;
  for (size_t k = 0; k < waiting_.size(); ++k)
    if (!take[k]) rest.push_back(waiting_[k]);
  waiting_.swap(rest);
  long long in_flight = 0;
  for (auto& L : lanes_)
    for (auto* list : {&L->active, &L->incoming})
      for (auto& r : *list)
        if (r->live) in_flight += r->bytes;
  st_.max_bytes_in_flight = std::max(st_.max_bytes_in_flight, in_flight);
  cv_.notify_all();  // the lanes given requests
}

void Server::relayout(Lane& L)
{
  drop_graph(L);  // the captured launches take the old layout's arrays
  std::vector<std::shared_ptr<Request>> act;
  {
    // finished requests (their results already extracted) leave the layout
    // and the continuing requests come first, so that their Verlet rows
    // can be carried over (the others' rows are built)
    std::lock_guard<std::mutex> lk(mu_);
    std::vector<std::shared_ptr<Request>> keep;
    for (auto& r : L.active)
      if (r->live && r->placed) keep.push_back(r);
    for (auto& r : L.active)
      if (r->live && !r->placed) keep.push_back(r);
    for (auto& r : L.incoming) keep.push_back(r);
    L.incoming.clear();
    L.dirty = false;
    L.active.swap(keep);
    act = L.active;
  }
  // the previous layout's rows, copied out before the engine may grow (only
  // when they are that layout's: a layout that has not been evaluated since
  // an uncarried relayout has none)
  DevWork& ew = L.nep->work();
  const int old_no = L.n_owned, old_nt = L.n_total, mv = ew.MV;
  bool carry = !L.forced_rebuild && old_no > 0 && !ew.split && ew.vl && ew.cap >= old_nt && ew.rcap >= old_no;
  if (carry) {
    const size_t rows = (size_t)old_no * (size_t)mv;
    L.vl0.need((long long)rows, "serve carried rows");
    L.vrev0.need((long long)rows, "serve carried reverse");
    L.vn0.need(old_no, "serve carried row lengths");
    L.x00.need(3LL * old_nt, "serve carried references");
    check(hipMemcpyAsync(L.vl0.p, ew.vl, rows * 4, hipMemcpyDeviceToDevice, L.stream), "serve carry rows");
    check(hipMemcpyAsync(L.vrev0.p, ew.vrev, rows * 4, hipMemcpyDeviceToDevice, L.stream), "serve carry reverse");
    check(hipMemcpyAsync(L.vn0.p, ew.vn, (size_t)old_no * 4, hipMemcpyDeviceToDevice, L.stream), "serve carry lengths");
    check(hipMemcpyAsync(L.x00.p, ew.x0, (size_t)old_nt * 24, hipMemcpyDeviceToDevice, L.stream),
          "serve carry references");
  }
  // placement: every request's extended region in its own cell of a cubic
  // grid of equal slots, separated by more than the Verlet radius
  const double gap = L.nep->model().rcR_max + kSkin + 2.0 + 2.0 * cfg_.margin;
  double slot[3] = {0, 0, 0};
  for (auto& r : act)
    for (int d = 0; d < 3; ++d) slot[d] = std::max(slot[d], r->hi_ext[d] - r->lo[d] + gap);
  const int R = (int)act.size();
  int side = 1;
  while ((long long)side * side * side < R) ++side;
  long long n_owned = 0, n_ghost = 0;
  for (auto& r : act) {
    n_owned += r->n;
    n_ghost += r->ng;
  }
  const long long Nt = n_owned + n_ghost;
  if (Nt >= (1LL << 30)) throw std::runtime_error("serving batch too large");
  if (Nt > L.nep->work().cap || n_owned > L.nep->work().rcap) {
    // geometric growth by a quarter (the rows of the previous layout are
    // already copied); rows for the owned slots only
    check(hipStreamSynchronize(L.stream), "serve carry");
    const DevWork& w = L.nep->work();
    const long long lim = (long long)(INT32_MAX - 1) / std::max(w.MV, 128);
    const long long cap = Nt > w.cap ? std::max<long long>(Nt, w.cap + w.cap / 4) : w.cap;
    const long long rc =
        n_owned > w.rcap ? std::max<long long>(n_owned, std::min<long long>(w.rcap + w.rcap / 4, lim)) : w.rcap;
    L.nep->reserve((int)std::min<long long>(cap, INT32_MAX - 1), (int)rc);
  }
  carry = carry && L.nep->work().MV == mv;
  // the new layout's arrays (double-buffered: the previous layout is the source)
  L.pos2.need(3 * Nt, "serve positions");
  L.vel2.need(3 * Nt, "serve velocities");
  L.force2.need(3 * Nt, "serve forces");
  L.mass2.need(Nt, "serve masses");
  L.type2.need(Nt, "serve types");
  L.owner2.need(Nt, "serve slot owners");
  L.image2.need(Nt, "serve slot images");
  L.gstart2.need(n_owned + 1, "serve ghost ranges");
  L.g_shift2.need(3 * n_ghost, "serve ghost shifts");
  L.pe.need(Nt, "serve potential");
  L.vir.need(9 * Nt, "serve virial");
  L.atom_req.need(n_owned, "serve atom->request");
  L.req_md.need(R, "serve request flags");
  L.req_dt.need(R, "serve request dt");
  L.slots.need(R, "serve layout table");
  L.home2.need(3 * n_owned, "serve home positions");
  L.req_flag.need(R, "serve drift flags");
  L.req_moved.need(R, "serve moved flags");
  L.req_range.need(R, "serve request ranges");
  L.req_off.need(2LL * R + 2, "serve rebuild offsets");
  L.lists.need(n_owned + Nt, "serve rebuild lists");
  if (R > L.h_drift_cap) {
    const int cap = std::max(R, 2 * L.h_drift_cap);
    if (L.h_drift) (void)hipHostFree(L.h_drift);
    L.h_drift = nullptr;
    L.h_drift_cap = 0;
    check(hipHostMalloc((void**)&L.h_drift, (size_t)cap * sizeof(int), hipHostMallocDefault), "pinned drift flags");
    L.h_drift_cap = cap;
  }
  // host side: one table row per request; new requests bring their packs
  std::vector<ServeSlots> tab((size_t)R);
  std::vector<int4> ranges((size_t)std::max(R, 1));
  std::vector<double> h_dt((size_t)std::max(R, 1));
  std::vector<Request*> consumed;
  long long o = 0, gk = n_owned;
  int r_keep = 0;
  long long o_keep = 0, g_keep = n_owned;
  for (int q = 0; q < R; ++q) {
    Request& r = *act[q];
    const int g3[3] = {q % side, (q / side) % side, q / (side * side)};
    double tnew[3];
    for (int d = 0; d < 3; ++d) tnew[d] = g3[d] * slot[d] + 0.5 * gap - r.lo[d];
    h_dt[q] = r.dt;
    ServeSlots& t = tab[q];
    t.new_o = (int)o;
    t.new_g = (int)gk;
    ranges[q] = make_int4((int)o, r.n, (int)gk, r.ng);
    t.n = r.n;
    t.ng = r.ng;
    t.pi = nullptr;
    t.pd = nullptr;
    if (r.placed) {
      t.old_o = r.slot0;
      t.old_g = r.gslot0;
      for (int d = 0; d < 3; ++d) t.dtr[d] = tnew[d] - r.trans[d];
      ++r_keep;
      o_keep += r.n;
      g_keep += r.ng;
    } else {
      if (!r.pack) throw std::logic_error("serving: a new request has no pack");
      t.old_o = t.old_g = -1;
      check(hipStreamWaitEvent(L.stream, r.ready, 0), "serve wait for upload");
      t.pi = static_cast<const int*>(r.pack);
      t.pd = reinterpret_cast<const double*>(static_cast<const char*>(r.pack) + serve_pack_double_offset(r.n, r.ng));
      for (int d = 0; d < 3; ++d) t.dtr[d] = tnew[d];
      consumed.push_back(&r);
    }
    r.slot0 = (int)o;
    r.gslot0 = (int)gk;
    for (int d = 0; d < 3; ++d) r.trans[d] = tnew[d];
    r.placed = true;
    o += r.n;
    gk += r.ng;
  }
  // uploads through the lane's pinned staging buffer (its stream is idle here)
  const size_t bytes = tab.size() * sizeof(ServeSlots) + h_dt.size() * 8 + ranges.size() * sizeof(int4) + 4 * 16;
  L.stage(bytes);
  char* hs = L.h_stage;
  size_t off = 0;
  auto up = [&](void* dev, const void* src, size_t n) {
    if (n == 0) return;
    off = (off + 15) & ~size_t(15);
    std::memcpy(hs + off, src, n);
    check(hipMemcpyAsync(dev, hs + off, n, hipMemcpyHostToDevice, L.stream), "serve upload");
    off += n;
  };
  up(L.slots.p, tab.data(), tab.size() * sizeof(ServeSlots));
  up(L.req_range.p, ranges.data(), (size_t)R * sizeof(int4));
  up(L.req_dt.p, h_dt.data(), h_dt.size() * 8);
  ServeLayout from, to;
  from.Nt = L.n_total;
  from.n_owned = L.n_owned;
  from.pos = L.pos.p;
  from.vel = L.vel.p;
  from.force = L.force.p;
  from.mass = L.mass.p;
  from.gshift = L.g_shift.p;
  from.type = L.type.p;
  from.owner = L.owner.p;
  from.image = L.image.p;
  from.gstart = L.gstart.p;
  from.home = L.home.p;
  to.Nt = (int)Nt;
  to.n_owned = (int)n_owned;
  to.pos = L.pos2.p;
  to.vel = L.vel2.p;
  to.force = L.force2.p;
  to.mass = L.mass2.p;
  to.gshift = L.g_shift2.p;
  to.type = L.type2.p;
  to.owner = L.owner2.p;
  to.image = L.image2.p;
  to.gstart = L.gstart2.p;
  to.home = L.home2.p;
  check(launch_serve_assemble(R, L.slots.p, from, to, L.atom_req.p, L.stream), "serve assemble");
  check(launch_clear_ints(L.req_flag.p, R, L.stream), "serve drift flags");
  check(launch_clear_ints(L.req_moved.p, R, L.stream), "serve moved flags");
  if (carry)
    check(launch_serve_carry_rows(r_keep, L.slots.p, old_no, old_nt, (int)n_owned, (int)Nt, (int)o_keep, (int)g_keep,
                                  L.vl0.p, L.vrev0.p, L.vn0.p, L.x00.p, L.nep->work(), L.stream),
          "serve carry rows");
  check(hipStreamSynchronize(L.stream), "serve relayout");
  for (Request* r : consumed) release_upload(*r);
  std::swap(L.pos, L.pos2);
  std::swap(L.vel, L.vel2);
  std::swap(L.force, L.force2);
  std::swap(L.mass, L.mass2);
  std::swap(L.type, L.type2);
  std::swap(L.owner, L.owner2);
  std::swap(L.image, L.image2);
  std::swap(L.gstart, L.gstart2);
  std::swap(L.g_shift, L.g_shift2);
  std::swap(L.home, L.home2);
  if (R > 0) std::memset(L.h_drift, 0, (size_t)R * sizeof(int));
  L.n_owned = (int)n_owned;
  L.n_ghost = (int)n_ghost;
  L.n_total = (int)Nt;
  L.n_req = R;
  std::memset(&L.batch_box, 0, sizeof(L.batch_box));
  for (int d = 0; d < 3; ++d) L.batch_box.h[4 * d] = side * slot[d];
  L.batch_box.pbc[0] = L.batch_box.pbc[1] = L.batch_box.pbc[2] = 0;
  L.forced_rebuild = true;
  if (carry && Nt > 0) {
    // the continuing requests keep their rows; only the others' are built
    const ImageMap map = L.map();
    check(launch_serve_ghosts(L.n_owned, L.n_ghost, L.n_total, L.owner.p + L.n_owned, L.g_shift.p, L.pos.p,
                              L.stream),
          "serve ghosts");
    L.nep->adopt_carried(L.batch_box, L.n_total, L.n_owned, (int)o_keep, (int)g_keep, L.type.p, L.pos.p, L.pe.p,
                         L.force.p, L.vir.p, map, L.stream);
    L.forced_rebuild = false;
  }
  std::lock_guard<std::mutex> lk(mu_);
  L.slots_now = Nt;
  st_.relayouts++;
  if (!L.forced_rebuild) st_.rows_carried++;
}

// Uploads which of the lane's requests take an MD step this iteration.
void Server::upload_md_flags(Lane& L)
{
  const int R = (int)L.active.size();
  std::vector<int> md((size_t)std::max(R, 1), 0);
  L.any_md = false;
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (int q = 0; q < R; ++q) {
      md[q] = (L.active[q]->live && L.active[q]->state == 1) ? 1 : 0;
      L.any_md = L.any_md || md[q];
    }
  }
  L.stage(md.size() * 4);
  std::memcpy(L.h_stage, md.data(), md.size() * 4);
  check(hipMemcpyAsync(L.req_md.p, L.h_stage, md.size() * 4, hipMemcpyHostToDevice, L.stream), "serve flags");
}

// Enqueues the first part of a lane's frame on its stream, without waiting:
// the first half-kick of the requests that take an MD step and the check of
// their drift against the image shells (flags copied back).
void Server::enqueue_kick(Lane& L)
{
  const int R = (int)L.active.size();
  upload_md_flags(L);
  if (L.any_md) {
    check(launch_serve_kick(1, L.n_owned, L.n_total, L.atom_req.p, L.req_md.p, L.req_dt.p, L.mass.p, L.pos.p, L.vel.p,
                            L.force.p, L.stream),
          "serve kick 1");
    // the image shell holds every image within rc + skin + margin of the
    // box at its build: exact while no atom is more than margin / 2 from
    // where it was then.  A request past that bound at this step's positions
    // gets a new shell before the force evaluation, with every MD request
    // past a quarter of the margin (one relayout for the wave).
    const double soft = 0.25 * cfg_.margin, hard = 0.5 * cfg_.margin;
    check(launch_serve_drift(L.n_owned, L.n_total, L.atom_req.p, L.pos.p, L.home.p, soft * soft, hard * hard,
                             L.req_flag.p, L.stream),
          "serve drift");
    check(hipMemcpyAsync(L.h_drift, L.req_flag.p, (size_t)R * sizeof(int), hipMemcpyDeviceToHost, L.stream),
          "serve drift flags");
  }
}

// Enqueues the force launches of a lane's frame (the captured graph, or the
// eager launches), without waiting.
void Server::enqueue_force(Lane& L)
{
  const unsigned flags = cfg_.flags & ~3u;
  const ImageMap map = L.map();
  if (L.graph_exec) {
    if (L.forced_rebuild) check(launch_request_rebuild(L.nep->work(), L.stream), "serve rebuild request");
    check(hipGraphLaunch(L.graph_exec, L.stream), "serve force graph");
  } else {
    // no graph for this layout yet (it is captured at the start of the
    // lane's next iteration): the same launches, enqueued directly
    check(launch_serve_ghosts(L.n_owned, L.n_ghost, L.n_total, L.owner.p + L.n_owned, L.g_shift.p, L.pos.p, L.stream),
          "serve ghosts");
    L.nep->enqueue(L.batch_box, L.n_total, 0, L.n_owned, 0, L.n_owned, L.type.p, L.pos.p, nullptr, L.pe.p, L.force.p,
                   L.vir.p, flags | (L.forced_rebuild ? 2u : 0u), L.stream, &map);
  }
}

// Waits for a lane's frame; a Verlet overflow (the engine has grown its rows)
// redoes the force evaluation (enqueued directly; the graph is captured again
// at the next iteration); then the second half-kick is enqueued.
void Server::finish_frame(Lane& L)
{
  const unsigned flags = cfg_.flags & ~3u;
  const ImageMap map = L.map();
  bool done = !L.nep->settle(L.stream);
  for (int attempt = 0; attempt < 4 && !done; ++attempt) {
    drop_graph(L);
    L.nep->enqueue(L.batch_box, L.n_total, 0, L.n_owned, 0, L.n_owned, L.type.p, L.pos.p, nullptr, L.pe.p, L.force.p,
                   L.vir.p, flags | 2u, L.stream, &map);
    done = !L.nep->settle(L.stream);
  }
  if (!done) throw std::runtime_error("serving: Verlet rows kept overflowing");
  L.forced_rebuild = false;
  if (L.any_md)
    check(launch_serve_kick(0, L.n_owned, L.n_total, L.atom_req.p, L.req_md.p, L.req_dt.p, L.mass.p, L.pos.p, L.vel.p,
                            L.force.p, L.stream),
          "serve kick 2");
}

// Captures the force launches of a lane's layout: the plan (host checks,
// workspace sizing) is made first, outside the capture; the pack never forces
// a rebuild itself (a forced rebuild is requested on the device before a
// replay).  The captured kernels take the layout's arrays, so the graph lives
// until the next relayout or workspace reallocation.  Records L.graph; the
// caller instantiates it.
void Server::capture_force(Lane& L, unsigned flags)
{
  const ImageMap map = L.map();
  const NepEngine::Plan p = L.nep->plan(L.batch_box, L.n_total, 0, L.n_owned, 0, L.n_owned, L.type.p, L.pos.p,
                                        nullptr, L.pe.p, L.force.p, L.vir.p, flags, &map);
  if (p.forced) L.forced_rebuild = true;
  check(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal), "serve capture");
  hipError_t e =
      launch_serve_ghosts(L.n_owned, L.n_ghost, L.n_total, L.owner.p + L.n_owned, L.g_shift.p, L.pos.p, L.stream);
  if (e == hipSuccess)
    e = launch_pack(p.box, L.n_total, L.nep->model().num_types, L.pos.p, L.type.p, false, L.nep->work(), L.stream);
  if (e == hipSuccess) {
    try {
      L.nep->launch_force(p, L.n_total, 0, L.n_owned, 0, L.n_owned, L.type.p, nullptr, L.pe.p, L.force.p, L.vir.p,
                          flags, L.stream);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(L.stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
  }
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(L.stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    check(e != hipSuccess ? e : e2, "serve capture");
  }
  L.graph = g;
}

void Server::drop_graph(Lane& L)
{
  if (L.graph_exec) (void)hipGraphExecDestroy(L.graph_exec);
  if (L.graph) (void)hipGraphDestroy(L.graph);
  L.graph_exec = nullptr;
  L.graph = nullptr;
}

// host threads that pack new image shells
constexpr size_t kReshellThreads = 8;

bool Server::needs_shells(Lane& L)
{
  if (!L.any_md || L.n_owned == 0) return false;
  std::lock_guard<std::mutex> lk(mu_);
  for (size_t q = 0; q < L.active.size(); ++q) {
    const Request* r = L.active[q].get();
    if (L.h_drift[q] == 2 && r->live && r->state == 1 && r->placed) return true;
  }
  return false;
}

void Server::reshell(Lane& L)
{
  std::vector<Request*> todo;
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (size_t q = 0; q < L.active.size(); ++q) {
      Request* r = L.active[q].get();
      if (L.h_drift[q] && r->live && r->state == 1 && r->placed) todo.push_back(r);
    }
  }
  // the requests' current state as a finished request's result block
  // (positions wrapped into their boxes, velocities, forces), one extraction
  // for all of them; then new shells from it
  long long no = 0;
  std::vector<ServeOut> tab(todo.size());
  for (size_t q = 0; q < todo.size(); ++q) {
    const Request& r = *todo[q];
    ServeOut& t = tab[q];
    t.slot0 = r.slot0;
    t.n = r.n;
    t.out = (long long)kServeOutDoubles * no;
    for (int d = 0; d < 3; ++d) {
      t.pbc[d] = r.box.pbc[d] ? 1 : 0;
      t.trans[d] = r.trans[d];
    }
    std::memcpy(t.h, r.box.h, sizeof(t.h));
    std::memcpy(t.hi, r.hi, sizeof(t.hi));
    no += r.n;
  }
  const size_t out_doubles = (size_t)kServeOutDoubles * (size_t)no;
  L.outq.need((long long)tab.size(), "serve out table");
  L.out.need((long long)out_doubles, "serve results");
  L.stage(tab.size() * sizeof(ServeOut));
  std::memcpy(L.h_stage, tab.data(), tab.size() * sizeof(ServeOut));
  check(hipMemcpyAsync(L.outq.p, L.h_stage, tab.size() * sizeof(ServeOut), hipMemcpyHostToDevice, L.stream),
        "serve reshell table");
  check(launch_serve_extract((int)tab.size(), L.outq.p, no, L.n_total, L.pos.p, L.vel.p, L.force.p, L.pe.p, L.vir.p,
                             L.out.p, L.stream),
        "serve reshell extract");
  size_t cap = 0;
  double* blk = static_cast<double*>(pin_get(out_doubles * 8, &cap));
  check(hipMemcpyAsync(blk, L.out.p, out_doubles * 8, hipMemcpyDeviceToHost, L.stream), "serve reshell state");
  check(hipStreamSynchronize(L.stream), "serve reshell state");
  // the packs on up to kReshellThreads host threads (each request's own)
  std::atomic<size_t> next{0};
  std::string failure;
  std::mutex fail_mu;
  auto work = [&] {
    for (size_t q; (q = next.fetch_add(1)) < todo.size();) {
      Request* r = todo[q];
      const size_t n = (size_t)r->n;
      const double* b = blk + tab[q].out;
      try {
        // block: potential [n], force [3n], position [3n], velocity [3n], virial
        make_pack(*r, r->types.data(), r->masses.data(), b + 4 * n, b + 7 * n, b + n,
                  std::numeric_limits<long long>::max());
      } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(fail_mu);
        if (failure.empty()) failure = e.what();
      }
      r->placed = false;
    }
  };
  const size_t nt = std::min<size_t>(kReshellThreads, todo.size());
  std::vector<std::thread> pool;
  for (size_t k = 1; k < nt; ++k) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  pin_put(blk, cap);
  if (!failure.empty()) throw std::runtime_error("serving: new image shell: " + failure);
  std::lock_guard<std::mutex> lk(mu_);
  st_.reshells += (long long)todo.size();
}

void Server::retire(Lane& L)
{
  // advance the request states; collect the ones that finished
  std::vector<Request*> fin;
  long long frames = 0, atom_frames = 0;
  std::unique_lock<std::mutex> slk(mu_);
  for (auto& rp : L.active) {
    Request& r = *rp;
    if (!r.live) continue;
    if (r.state == 0) {
      if (r.steps == 0) {
        fin.push_back(&r);
        frames += 1;
        atom_frames += r.n;
      } else {
        r.state = 1;
      }
    } else if (r.state == 1) {
      ++r.step;
      frames += 1;
      atom_frames += r.n;
      if (r.step >= r.steps) fin.push_back(&r);
    }
  }
  slk.unlock();
  if (!fin.empty()) {
    long long no = 0;
    std::vector<ServeOut> tab(fin.size());
    for (size_t q = 0; q < fin.size(); ++q) {
      const Request& r = *fin[q];
      ServeOut& t = tab[q];
      t.slot0 = r.slot0;
      t.n = r.n;
      t.out = (long long)kServeOutDoubles * no;
      for (int d = 0; d < 3; ++d) {
        t.pbc[d] = r.box.pbc[d] ? 1 : 0;
        t.trans[d] = r.trans[d];
      }
      std::memcpy(t.h, r.box.h, sizeof(t.h));
      std::memcpy(t.hi, r.hi, sizeof(t.hi));
      no += r.n;
    }
    const size_t out_doubles = (size_t)kServeOutDoubles * (size_t)no;
    L.outq.need((long long)tab.size(), "serve out table");
    L.out.need((long long)out_doubles, "serve results");
    check(hipStreamSynchronize(L.stream), "serve step");  // the staging buffer is free again
    L.stage(tab.size() * sizeof(ServeOut));
    std::memcpy(L.h_stage, tab.data(), tab.size() * sizeof(ServeOut));
    check(hipMemcpyAsync(L.outq.p, L.h_stage, tab.size() * sizeof(ServeOut), hipMemcpyHostToDevice, L.stream),
          "serve out upload");
    check(launch_serve_extract((int)tab.size(), L.outq.p, no, L.n_total, L.pos.p, L.vel.p, L.force.p, L.pe.p,
                               L.vir.p, L.out.p, L.stream),
          "serve extract");
    // results go straight into a pinned block that the requests keep
    auto blk = std::make_shared<ResultBlock>();
    blk->srv = this;
    blk->p = pin_get(out_doubles * 8, &blk->cap);
    check(hipMemcpyAsync(blk->p, L.out.p, out_doubles * 8, hipMemcpyDeviceToHost, L.stream), "serve results");
    check(hipStreamSynchronize(L.stream), "serve results");
    const double t = now_s();
    std::lock_guard<std::mutex> lk(mu_);
    for (size_t q = 0; q < fin.size(); ++q) {
      Request* r = fin[q];
      r->res = blk;
      r->res_off = tab[q].out;
      r->state = 2;
      r->live = false;
      r->t_done = t;
      latencies_.push_back(r->t_done - r->t_submit);
      st_.completed++;
    }
    st_.frames += frames;
    st_.atom_frames += atom_frames;
    t_last_ = t;
    admit_gen_++;  // their budget is free
    cv_.notify_all();
    done_cv_.notify_all();
    // a finished request leaves the layout at the lane's next relayout; an
    // empty lane drops its layout now
    if (!L.live()) {
      L.active.clear();
      L.n_owned = L.n_ghost = L.n_total = L.n_req = 0;
      L.slots_now = 0;
      drop_graph(L);
    }
    return;
  }
  std::lock_guard<std::mutex> lk(mu_);
  st_.frames += frames;
  st_.atom_frames += atom_frames;
}

}  // namespace dmgx
