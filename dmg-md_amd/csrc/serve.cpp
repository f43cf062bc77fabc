// serve.cpp -- the serving loop (see serve.hpp).
#include "serve.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
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
  release();
  if (hipMalloc((void**)&p, (size_t)n * sizeof(T)) != hipSuccess) {
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

void Server::stage(size_t bytes)
{
  if (bytes <= h_stage_bytes_) return;
  if (h_stage_) (void)hipHostFree(h_stage_);
  h_stage_ = nullptr;
  h_stage_bytes_ = 0;
  check(hipHostMalloc((void**)&h_stage_, bytes, hipHostMallocDefault), "pinned staging");
  h_stage_bytes_ = bytes;
}

// Device buffers for request packs, reused by size class (powers of two from
// 4 KiB): freeing device memory would synchronise the device under the
// serving loop, so buffers are only freed with the server.
void* Server::pack_get(size_t bytes, size_t* cap)
{
  size_t c = 4096;
  while (c < bytes) c <<= 1;
  {
    std::lock_guard<std::mutex> lk(pack_mu_);
    auto it = pack_free_.find(c);
    if (it != pack_free_.end()) {
      void* p = it->second;
      pack_free_.erase(it);
      *cap = c;
      return p;
    }
  }
  void* p = nullptr;
  if (hipMalloc(&p, c) != hipSuccess) throw std::runtime_error("out of device memory for a request upload");
  *cap = c;
  return p;
}

void Server::pack_put(void* p, size_t cap)
{
  if (!p) return;
  std::lock_guard<std::mutex> lk(pack_mu_);
  pack_free_.emplace(cap, p);
}

// pinned host buffers for building packs, reused the same way
void* Server::pin_get(size_t bytes, size_t* cap)
{
  size_t c = 4096;
  while (c < bytes) c <<= 1;
  {
    std::lock_guard<std::mutex> lk(pack_mu_);
    auto it = pin_free_.find(c);
    if (it != pin_free_.end()) {
      void* p = it->second;
      pin_free_.erase(it);
      *cap = c;
      return p;
    }
  }
  void* p = nullptr;
  if (hipHostMalloc(&p, c, hipHostMallocDefault) != hipSuccess)
    throw std::runtime_error("out of pinned host memory for a request upload");
  *cap = c;
  return p;
}

struct Server::Request {
  long long id = 0;
  Box box{};
  double hi[9] = {};
  int n = 0;
  int ng = 0;  // periodic images within rc + skin + margin of the box
  int steps = 0;
  double dt = 0;
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
  // placement in the current layout
  double trans[3] = {0, 0, 0};
  int slot0 = -1, gslot0 = -1;  // first owned / ghost slot in the layout
  bool placed = false;
  // results: block [potential n | force 3n | position 3n | velocity 3n]
  std::shared_ptr<ResultBlock> res;
  long long res_off = 0;
  long long ghosts() const { return ng; }
  long long footprint() const { return n + ghosts(); }
};

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

Server::Server(NepEngine* nep, const ServeConfig& cfg) : nep_(nep), cfg_(cfg)
{
  if (!nep) throw std::invalid_argument("null potential");
  if (cfg_.max_atoms <= 0 || cfg_.max_requests <= 0) throw std::invalid_argument("bad serving limits");
  if (cfg_.max_atoms >= (1LL << 30)) throw std::invalid_argument("max_atoms must stay below 2^30");
  // the batch grows request by request: size the engine once, up front,
  // instead of reallocating (and synchronising the device) as it grows
  nep_->reserve((int)std::min(cfg_.max_atoms, cfg_.reserve_atoms));
  nep_->reserve_cells(1LL << 23);
  check(hipStreamCreate(&stream_), "serving stream");
  check(hipStreamCreate(&upload_), "serving upload stream");
  thread_ = std::thread([this] { worker(); });
}

Server::~Server()
{
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_.notify_all();
  if (thread_.joinable()) thread_.join();
  if (stream_) (void)hipStreamSynchronize(stream_);
  drop_graph();
  for (DBuf<double>* b : {&pos_, &vel_, &force_, &pe_, &vir_, &mass_, &pos2_, &vel2_, &force2_, &mass2_, &req_dt_,
                          &g_shift_, &g_shift2_, &out_})
    b->release();
  for (DBuf<int>* b : {&type_, &type2_, &atom_req_, &req_md_, &owner_, &owner2_, &image_, &image2_, &gstart_,
                       &gstart2_})
    b->release();
  slots_.release();
  outq_.release();
  if (upload_) (void)hipStreamSynchronize(upload_);
  for (auto& kv : all_) release_upload(*kv.second);  // every request, waiting ones included
  // dropping the requests returns their result blocks to the pinned pool
  waiting_.clear();
  active_.clear();
  all_.clear();
  for (auto& kv : pack_free_) (void)hipFree(kv.second);
  for (auto& kv : pin_free_) (void)hipHostFree(kv.second);
  if (upload_) (void)hipStreamDestroy(upload_);
  if (h_stage_) (void)hipHostFree(h_stage_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

long long Server::submit(const Box& box, int N, const int* type, const double* mass, const double* pos,
                         const double* vel, int steps, double dt_fs)
{
  if (N <= 0 || !type || !mass || !pos) throw std::invalid_argument("a request needs atoms, types, masses, positions");
  if (steps < 0) throw std::invalid_argument("negative step count");
  if (steps > 0 && !(dt_fs > 0.0)) throw std::invalid_argument("MD requests need dt_fs > 0");
  const double t_call = now_s();
  auto r = std::make_shared<Request>();
  r->box = box;
  const double vol = box_volume(box.h);
  if (!(std::fabs(vol) > 0.0) || !std::isfinite(vol)) throw std::invalid_argument("box has no volume");
  box_inverse(box.h, r->hi);
  r->n = N;
  r->steps = steps;
  r->dt = dt_fs / kTimeUnit;
  const int T = nep_->model().num_types;
  for (int i = 0; i < N; ++i) {
    if (type[i] < 0 || type[i] >= T) throw std::invalid_argument("atom type out of range");
    if (!(mass[i] > 0.0)) throw std::invalid_argument("masses must be positive");
  }
  std::vector<double> wpos(3 * (size_t)N);  // wrapped into the box, SoA
  // wrap into the box and collect the periodic images of the shell
  const double rc = nep_->model().rcR_max;
  const double d1 = rc + kSkin + cfg_.margin;
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
  for (int i = 0; i < N; ++i)
    for (int ia = -K[0]; ia <= K[0]; ++ia)
      for (int ib = -K[1]; ib <= K[1]; ++ib)
        for (int ic = -K[2]; ic <= K[2]; ++ic) {
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
  if (r->footprint() > cfg_.max_atoms)
    throw std::invalid_argument("request (with its periodic shell) exceeds the serving atom budget");
  // the pack: built here and uploaded on the submitting thread, so the
  // serving loop only copies it device-to-device when it places the request
  const size_t doff = serve_pack_double_offset(N, NG), bytes = serve_pack_bytes(N, NG);
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
    }
  std::copy(g_shift.begin(), g_shift.end(), pd + 7 * (size_t)N);
  // asynchronous upload: the serving stream waits for `ready` before it
  // places the request
  r->pack = pack_get(bytes, &r->pack_cap);
  hipError_t e = hipEventCreateWithFlags(&r->ready, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMemcpyAsync(r->pack, host, bytes, hipMemcpyHostToDevice, upload_);
  if (e == hipSuccess) e = hipEventRecord(r->ready, upload_);
  if (e != hipSuccess) {
    release_upload(*r);
    throw std::runtime_error(std::string("request upload: ") + hipGetErrorString(e));
  }
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
  cv_.notify_all();
  return r->id;
}

bool Server::poll(long long id, bool wait, double* potential, double* force, double* position, double* velocity,
                  double* latency_ms)
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
  if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
  return true;
}

void Server::drain()
{
  std::unique_lock<std::mutex> lk(mu_);
  done_cv_.wait(lk, [&] { return st_.completed == st_.submitted || !error_.empty(); });
  if (!error_.empty()) throw std::runtime_error("serving loop failed: " + error_);
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

void Server::worker()
{
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] {
        if (stop_ || !error_.empty()) return true;
        if (!waiting_.empty()) return true;
        for (auto& r : active_)
          if (r->live) return true;
        return false;
      });
      if (stop_ || !error_.empty()) return;
      admit();
    }
    try {
      const double t0 = now_s();
      if (dirty_) relayout();
      bool any = false;
      for (auto& r : active_) any = any || r->live;
      const double t1 = now_s();
      if (any) iterate();
      const double t2 = now_s();
      if (any) retire();
      const double t3 = now_s();
      std::lock_guard<std::mutex> lk(mu_);
      st_.relayout_s += t1 - t0;
      st_.iterate_s += t2 - t1;
      st_.retire_s += t3 - t2;
      st_.max_slots = std::max<long long>(st_.max_slots, n_total_);
    } catch (const std::exception& e) {
      std::lock_guard<std::mutex> lk(mu_);
      error_ = e.what();
      done_cv_.notify_all();
      return;
    }
  }
}

// Called with mu_ held.  Admits waiting requests in FIFO order while the
// batch (live requests + new ones) fits the budget; a new layout is made when
// the batch is empty, or when the atoms that would be admitted plus those of
// finished requests still in the layout reach a quarter of the layout, so
// that the layout (and the Verlet rebuild it forces) is not redone for every
// single arrival.
bool Server::admit()
{
  long long live_atoms = 0, dead_atoms = 0;
  int live = 0;
  for (auto& r : active_) {
    if (r->live) {
      live_atoms += r->footprint();
      ++live;
    } else {
      dead_atoms += r->footprint();
    }
  }
  long long add_atoms = 0;
  size_t k = 0;
  for (; k < waiting_.size(); ++k) {
    const long long f = waiting_[k]->footprint();
    if (live_atoms + add_atoms + f > cfg_.max_atoms || live + (int)k + 1 > cfg_.max_requests) break;
    add_atoms += f;
  }
  if (k == 0) return false;
  const long long layout_atoms = live_atoms + dead_atoms;
  const bool empty = live == 0;
  if (!empty && 4 * (add_atoms + dead_atoms) < layout_atoms) return false;
  std::vector<std::shared_ptr<Request>> next;
  for (auto& r : active_)
    if (r->live) next.push_back(r);
  for (size_t j = 0; j < k; ++j) next.push_back(waiting_[j]);
  waiting_.erase(waiting_.begin(), waiting_.begin() + (long)k);
  active_.swap(next);
  dirty_ = true;
  return true;
}

void Server::relayout()
{
  drop_graph();  // the captured launches take the old layout's arrays
  // placement: every request's extended region in its own cell of a cubic
  // grid of equal slots, separated by more than the Verlet radius
  const double gap = nep_->model().rcR_max + kSkin + 2.0 + 2.0 * cfg_.margin;
  double slot[3] = {0, 0, 0};
  for (auto& r : active_)
    for (int d = 0; d < 3; ++d) slot[d] = std::max(slot[d], r->hi_ext[d] - r->lo[d] + gap);
  const int R = (int)active_.size();
  int side = 1;
  while ((long long)side * side * side < R) ++side;
  long long n_owned = 0, n_ghost = 0;
  for (auto& r : active_) {
    n_owned += r->n;
    n_ghost += r->ghosts();
  }
  const long long Nt = n_owned + n_ghost;
  if (Nt >= (1LL << 30)) throw std::runtime_error("serving batch too large");
  // the new layout's arrays (double-buffered: the previous layout is the source)
  pos2_.need(3 * Nt, "serve positions");
  vel2_.need(3 * Nt, "serve velocities");
  force2_.need(3 * Nt, "serve forces");
  mass2_.need(Nt, "serve masses");
  type2_.need(Nt, "serve types");
  owner2_.need(Nt, "serve slot owners");
  image2_.need(Nt, "serve slot images");
  gstart2_.need(n_owned + 1, "serve ghost ranges");
  g_shift2_.need(3 * n_ghost, "serve ghost shifts");
  pe_.need(Nt, "serve potential");
  vir_.need(9 * Nt, "serve virial");
  atom_req_.need(n_owned, "serve atom->request");
  req_md_.need(R, "serve request flags");
  req_dt_.need(R, "serve request dt");
  slots_.need(R, "serve layout table");
  // host side: one table row per request; new requests bring their packs
  std::vector<ServeSlots> tab((size_t)R);
  std::vector<double> h_dt((size_t)std::max(R, 1));
  std::vector<Request*> consumed;
  long long o = 0, gk = n_owned;
  for (int q = 0; q < R; ++q) {
    Request& r = *active_[q];
    const int g3[3] = {q % side, (q / side) % side, q / (side * side)};
    double tnew[3];
    for (int d = 0; d < 3; ++d) tnew[d] = g3[d] * slot[d] + 0.5 * gap - r.lo[d];
    h_dt[q] = r.dt;
    ServeSlots& t = tab[q];
    t.new_o = (int)o;
    t.new_g = (int)gk;
    t.n = r.n;
    t.ng = r.ng;
    t.pi = nullptr;
    t.pd = nullptr;
    if (r.placed) {
      t.old_o = r.slot0;
      t.old_g = r.gslot0;
      for (int d = 0; d < 3; ++d) t.dtr[d] = tnew[d] - r.trans[d];
    } else {
      if (!r.pack) throw std::logic_error("serving: a new request has no pack");
      t.old_o = t.old_g = -1;
      check(hipStreamWaitEvent(stream_, r.ready, 0), "serve wait for upload");
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
  // uploads through one pinned staging buffer (the stream is idle here)
  const size_t bytes = tab.size() * sizeof(ServeSlots) + h_dt.size() * 8 + 4 * 16;
  stage(bytes);
  char* hs = h_stage_;
  size_t off = 0;
  auto up = [&](void* dev, const void* src, size_t n) {
    if (n == 0) return;
    off = (off + 15) & ~size_t(15);
    std::memcpy(hs + off, src, n);
    check(hipMemcpyAsync(dev, hs + off, n, hipMemcpyHostToDevice, stream_), "serve upload");
    off += n;
  };
  up(slots_.p, tab.data(), tab.size() * sizeof(ServeSlots));
  up(req_dt_.p, h_dt.data(), h_dt.size() * 8);
  ServeLayout from, to;
  from.Nt = n_total_;
  from.n_owned = n_owned_;
  from.pos = pos_.p;
  from.vel = vel_.p;
  from.force = force_.p;
  from.mass = mass_.p;
  from.gshift = g_shift_.p;
  from.type = type_.p;
  from.owner = owner_.p;
  from.image = image_.p;
  from.gstart = gstart_.p;
  to.Nt = (int)Nt;
  to.n_owned = (int)n_owned;
  to.pos = pos2_.p;
  to.vel = vel2_.p;
  to.force = force2_.p;
  to.mass = mass2_.p;
  to.gshift = g_shift2_.p;
  to.type = type2_.p;
  to.owner = owner2_.p;
  to.image = image2_.p;
  to.gstart = gstart2_.p;
  check(launch_serve_assemble(R, slots_.p, from, to, atom_req_.p, stream_), "serve assemble");
  check(hipStreamSynchronize(stream_), "serve relayout");
  for (Request* r : consumed) release_upload(*r);
  std::swap(pos_, pos2_);
  std::swap(vel_, vel2_);
  std::swap(force_, force2_);
  std::swap(mass_, mass2_);
  std::swap(type_, type2_);
  std::swap(owner_, owner2_);
  std::swap(image_, image2_);
  std::swap(gstart_, gstart2_);
  std::swap(g_shift_, g_shift2_);
  n_owned_ = (int)n_owned;
  n_ghost_ = (int)n_ghost;
  n_total_ = (int)Nt;
  std::memset(&batch_box_, 0, sizeof(batch_box_));
  for (int d = 0; d < 3; ++d) batch_box_.h[4 * d] = side * slot[d];
  batch_box_.pbc[0] = batch_box_.pbc[1] = batch_box_.pbc[2] = 0;
  forced_rebuild_ = true;
  dirty_ = false;
  {
    std::lock_guard<std::mutex> lk(mu_);
    st_.relayouts++;
  }
}

void Server::iterate()
{
  const int R = (int)active_.size();
  // which requests take an MD step this iteration
  std::vector<int> md((size_t)std::max(R, 1), 0);
  bool any_md = false;
  for (int q = 0; q < R; ++q) {
    md[q] = (active_[q]->live && active_[q]->state == 1) ? 1 : 0;
    any_md = any_md || md[q];
  }
  stage(md.size() * 4);
  std::memcpy(h_stage_, md.data(), md.size() * 4);
  check(hipMemcpyAsync(req_md_.p, h_stage_, md.size() * 4, hipMemcpyHostToDevice, stream_), "serve flags");
  if (any_md)
    check(launch_serve_kick(1, n_owned_, n_total_, atom_req_.p, req_md_.p, req_dt_.p, mass_.p, pos_.p, vel_.p,
                            force_.p, stream_),
          "serve kick 1");
  const unsigned flags = cfg_.flags & ~3u;
  if (!nep_->tuning().tuned && n_owned_ >= kTuneMinAtoms) {
    // tile shapes picked once, on the first batch large enough to tell them apart
    ImageMap map;
    map.owner = owner_.p;
    map.image = image_.p;
    map.gstart = gstart_.p;
    check(launch_serve_ghosts(n_owned_, n_ghost_, n_total_, owner_.p + n_owned_, g_shift_.p, pos_.p, stream_),
          "serve ghosts");
    nep_->autotune(batch_box_, n_total_, n_owned_, type_.p, pos_.p, pe_.p, force_.p, vir_.p,
                   flags | (forced_rebuild_ ? 2u : 0u), stream_, &map);
    drop_graph();
    forced_rebuild_ = false;
  }
  if (cfg_.use_graph) {
    // one graph launch per iteration; a Verlet overflow reallocates the
    // engine's rows, so the graph is captured again and replayed
    bool done = false;
    for (int attempt = 0; attempt < 4 && !done; ++attempt) {
      if (!graph_exec_) capture_force(flags);
      if (forced_rebuild_) check(launch_request_rebuild(nep_->work(), stream_), "serve rebuild request");
      check(hipGraphLaunch(graph_exec_, stream_), "serve force graph");
      done = !nep_->settle(stream_);
      if (!done) {
        drop_graph();
        forced_rebuild_ = true;
      }
    }
    if (!done) throw std::runtime_error("serving: Verlet rows kept overflowing");
  } else {
    check(launch_serve_ghosts(n_owned_, n_ghost_, n_total_, owner_.p + n_owned_, g_shift_.p, pos_.p, stream_),
          "serve ghosts");
    ImageMap map;
    map.owner = owner_.p;
    map.image = image_.p;
    map.gstart = gstart_.p;
    nep_->compute(batch_box_, n_total_, 0, n_owned_, 0, n_owned_, type_.p, pos_.p, nullptr, pe_.p, force_.p,
                  vir_.p, flags | (forced_rebuild_ ? 2u : 0u), stream_, &map);
  }
  forced_rebuild_ = false;
  if (any_md)
    check(launch_serve_kick(0, n_owned_, n_total_, atom_req_.p, req_md_.p, req_dt_.p, mass_.p, pos_.p, vel_.p,
                            force_.p, stream_),
          "serve kick 2");
  std::lock_guard<std::mutex> lk(mu_);
  st_.iterations++;
}

// Captures the force launches of the current layout: the plan (host checks,
// workspace sizing) is made first, outside the capture; the pack never forces
// a rebuild itself (a forced rebuild is requested on the device before a
// replay).  The captured kernels take the layout's arrays, so the graph lives
// until the next relayout or workspace reallocation.
void Server::capture_force(unsigned flags)
{
  ImageMap map;
  map.owner = owner_.p;
  map.image = image_.p;
  map.gstart = gstart_.p;
  const NepEngine::Plan p = nep_->plan(batch_box_, n_total_, 0, n_owned_, 0, n_owned_, type_.p, pos_.p, nullptr,
                                       pe_.p, force_.p, vir_.p, flags, &map);
  if (p.forced) forced_rebuild_ = true;
  check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "serve capture");
  hipError_t e = launch_serve_ghosts(n_owned_, n_ghost_, n_total_, owner_.p + n_owned_, g_shift_.p, pos_.p, stream_);
  if (e == hipSuccess)
    e = launch_pack(p.box, n_total_, nep_->model().num_types, pos_.p, type_.p, false, nep_->work(), stream_);
  if (e == hipSuccess) {
    try {
      nep_->launch_force(p, n_total_, 0, n_owned_, 0, n_owned_, type_.p, nullptr, pe_.p, force_.p, vir_.p, flags,
                         stream_);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(stream_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
  }
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(stream_, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    check(e != hipSuccess ? e : e2, "serve capture");
  }
  graph_ = g;
  check(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0), "serve graph instantiate");
}

void Server::drop_graph()
{
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graph_exec_ = nullptr;
  graph_ = nullptr;
}

void Server::retire()
{
  // advance the request states; collect the ones that finished
  std::vector<Request*> fin;
  long long frames = 0, atom_frames = 0;
  std::unique_lock<std::mutex> slk(mu_);
  for (auto& rp : active_) {
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
      t.out = 10 * no;
      for (int d = 0; d < 3; ++d) {
        t.pbc[d] = r.box.pbc[d] ? 1 : 0;
        t.trans[d] = r.trans[d];
      }
      std::memcpy(t.h, r.box.h, sizeof(t.h));
      std::memcpy(t.hi, r.hi, sizeof(t.hi));
      no += r.n;
    }
    outq_.need((long long)tab.size(), "serve out table");
    out_.need(10 * no, "serve results");
    check(hipStreamSynchronize(stream_), "serve step");  // the staging buffer is free again
    stage(tab.size() * sizeof(ServeOut));
    std::memcpy(h_stage_, tab.data(), tab.size() * sizeof(ServeOut));
    check(hipMemcpyAsync(outq_.p, h_stage_, tab.size() * sizeof(ServeOut), hipMemcpyHostToDevice, stream_),
          "serve out upload");
    check(launch_serve_extract((int)tab.size(), outq_.p, no, n_total_, pos_.p, vel_.p, force_.p, pe_.p, out_.p,
                               stream_),
          "serve extract");
    // results go straight into a pinned block that the requests keep
    auto blk = std::make_shared<ResultBlock>();
    blk->srv = this;
    blk->p = pin_get(10 * (size_t)no * 8, &blk->cap);
    check(hipMemcpyAsync(blk->p, out_.p, 10 * (size_t)no * 8, hipMemcpyDeviceToHost, stream_), "serve results");
    check(hipStreamSynchronize(stream_), "serve results");
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
    done_cv_.notify_all();
    // a finished request leaves the layout at the next relayout
    bool any_live = false;
    for (auto& r : active_) any_live = any_live || r->live;
    if (!any_live) {
      active_.clear();
      dirty_ = true;
    }
    return;
  }
  std::lock_guard<std::mutex> lk(mu_);
  st_.frames += frames;
  st_.atom_frames += atom_frames;
}

}  // namespace dmgx
