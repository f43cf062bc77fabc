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

struct Server::Request {
  long long id = 0;
  Box box{};
  double hi[9] = {};
  int n = 0;
  std::vector<int> type;
  std::vector<double> mass, pos, vel;  // pos wrapped into the box, SoA
  int steps = 0;
  double dt = 0;
  // periodic images within rc + skin + margin of the box, grouped by owner:
  // owner (local index), image code and Cartesian shift
  std::vector<int> g_owner, g_image;
  std::vector<double> g_shift;
  double lo[3] = {0, 0, 0}, hi_ext[3] = {0, 0, 0};  // extent of owned + ghost positions
  // progress
  int state = 0;  // 0: needs its first force evaluation, 1: MD steps, 2: done
  int step = 0;
  bool live = true;
  double t_submit = 0, t_done = 0;
  std::string error;
  // placement in the current layout
  double trans[3] = {0, 0, 0};
  int slot0 = -1;
  bool placed = false;
  // results
  std::vector<double> r_pe, r_force, r_pos, r_vel;
  long long ghosts() const { return (long long)g_owner.size(); }
  long long footprint() const { return n + ghosts(); }
};

Server::Server(NepEngine* nep, const ServeConfig& cfg) : nep_(nep), cfg_(cfg)
{
  if (!nep) throw std::invalid_argument("null potential");
  if (cfg_.max_atoms <= 0 || cfg_.max_requests <= 0) throw std::invalid_argument("bad serving limits");
  if (cfg_.max_atoms >= (1LL << 30)) throw std::invalid_argument("max_atoms must stay below 2^30");
  check(hipStreamCreate(&stream_), "serving stream");
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
  for (DBuf<double>* b : {&pos_, &vel_, &force_, &pe_, &vir_, &mass_, &pos2_, &vel2_, &force2_, &req_dt_, &g_shift_,
                          &remap_dt_, &fresh_, &out_shift_, &out_})
    b->release();
  for (DBuf<int>* b : {&type_, &atom_req_, &req_md_, &owner_, &image_, &gstart_, &remap_src_, &fresh_dst_, &out_slot_})
    b->release();
  if (h_stage_) (void)hipHostFree(h_stage_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

long long Server::submit(const Box& box, int N, const int* type, const double* mass, const double* pos,
                         const double* vel, int steps, double dt_fs)
{
  if (N <= 0 || !type || !mass || !pos) throw std::invalid_argument("a request needs atoms, types, masses, positions");
  if (steps < 0) throw std::invalid_argument("negative step count");
  if (steps > 0 && !(dt_fs > 0.0)) throw std::invalid_argument("MD requests need dt_fs > 0");
  auto r = std::make_shared<Request>();
  r->box = box;
  const double vol = box_volume(box.h);
  if (!(std::fabs(vol) > 0.0) || !std::isfinite(vol)) throw std::invalid_argument("box has no volume");
  box_inverse(box.h, r->hi);
  r->n = N;
  r->steps = steps;
  r->dt = dt_fs / kTimeUnit;
  r->type.assign(type, type + N);
  r->mass.assign(mass, mass + N);
  const int T = nep_->model().num_types;
  for (int i = 0; i < N; ++i) {
    if (type[i] < 0 || type[i] >= T) throw std::invalid_argument("atom type out of range");
    if (!(mass[i] > 0.0)) throw std::invalid_argument("masses must be positive");
  }
  r->pos.resize(3 * (size_t)N);
  r->vel.assign(3 * (size_t)N, 0.0);
  if (vel) std::copy(vel, vel + 3 * (size_t)N, r->vel.begin());
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
    for (int d = 0; d < 3; ++d) r->pos[(size_t)d * N + i] = x[d];
  }
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
          r->g_owner.push_back(i);
          r->g_image.push_back(image_code(ia, ib, ic));
          r->g_shift.insert(r->g_shift.end(), sh, sh + 3);
        }
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
    const double x[3] = {r->pos[i], r->pos[(size_t)N + i], r->pos[2 * (size_t)N + i]};
    extend(x);
  }
  for (size_t g = 0; g < r->g_owner.size(); ++g) {
    const int o = r->g_owner[g];
    const double x[3] = {r->pos[o] + r->g_shift[3 * g], r->pos[(size_t)N + o] + r->g_shift[3 * g + 1],
                         r->pos[2 * (size_t)N + o] + r->g_shift[3 * g + 2]};
    extend(x);
  }
  if (r->footprint() > cfg_.max_atoms)
    throw std::invalid_argument("request (with its periodic shell) exceeds the serving atom budget");
  std::lock_guard<std::mutex> lk(mu_);
  if (!error_.empty()) throw std::runtime_error("serving loop failed: " + error_);
  r->id = next_id_++;
  r->t_submit = now_s();
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
  if (potential) std::copy(r->r_pe.begin(), r->r_pe.end(), potential);
  if (force) std::copy(r->r_force.begin(), r->r_force.end(), force);
  if (position) std::copy(r->r_pos.begin(), r->r_pos.end(), position);
  if (velocity) std::copy(r->r_vel.begin(), r->r_vel.end(), velocity);
  if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
  (void)n;
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
  const int old_Nt = n_total_;
  // device arrays; position, velocity and force are double-buffered across the remap
  pos2_.need(3 * Nt, "serve positions");
  vel2_.need(3 * Nt, "serve velocities");
  force2_.need(3 * Nt, "serve forces");
  pe_.need(Nt, "serve potential");
  vir_.need(9 * Nt, "serve virial");
  mass_.need(Nt, "serve masses");
  type_.need(Nt, "serve types");
  atom_req_.need(n_owned, "serve atom->request");
  remap_src_.need(n_owned, "serve remap");
  remap_dt_.need(3 * n_owned, "serve remap shift");
  req_md_.need(R, "serve request flags");
  req_dt_.need(R, "serve request dt");
  owner_.need(Nt, "serve slot owners");
  image_.need(Nt, "serve slot images");
  gstart_.need(n_owned + 1, "serve ghost ranges");
  g_shift_.need(3 * n_ghost, "serve ghost shifts");
  // host side of the layout
  std::vector<int> h_type((size_t)Nt), h_req((size_t)n_owned), h_src((size_t)n_owned), h_own((size_t)Nt),
      h_img((size_t)Nt, image_code(0, 0, 0)), h_gst((size_t)n_owned + 1);
  std::vector<double> h_mass((size_t)Nt), h_dtr(3 * (size_t)n_owned, 0.0), h_gsh(3 * (size_t)n_ghost);
  std::vector<int> h_fdst;
  std::vector<double> h_fresh;
  std::vector<double> h_dt((size_t)std::max(R, 1));
  long long o = 0, gk = n_owned;
  for (int q = 0; q < R; ++q) {
    Request& r = *active_[q];
    const int gx = q % side, gy = (q / side) % side, gz = q / (side * side);
    double tnew[3];
    const int g3[3] = {gx, gy, gz};
    for (int d = 0; d < 3; ++d) tnew[d] = g3[d] * slot[d] + 0.5 * gap - r.lo[d];
    h_dt[q] = r.dt;
    for (int k = 0; k < r.n; ++k) {
      const long long i = o + k;
      h_own[i] = (int)i;
      h_type[i] = r.type[k];
      h_mass[i] = r.mass[k];
      h_req[i] = q;
      if (r.placed) {
        h_src[i] = r.slot0 + k;
        for (int d = 0; d < 3; ++d) h_dtr[3 * i + d] = tnew[d] - r.trans[d];
      } else {
        h_src[i] = -1;
        h_fdst.push_back((int)i);
        for (int d = 0; d < 3; ++d) h_fresh.push_back(r.pos[(size_t)d * r.n + k] + tnew[d]);
        for (int d = 0; d < 3; ++d) h_fresh.push_back(r.vel[(size_t)d * r.n + k]);
      }
    }
    // ghosts in owner order: owned slot o + a owns [h_gst[o + a], h_gst[o + a + 1])
    long long g = 0;
    for (int a = 0; a < r.n; ++a) {
      h_gst[o + a] = (int)gk;
      for (; g < r.ghosts() && r.g_owner[g] == a; ++g, ++gk) {
        const long long k = gk - n_owned;
        h_own[gk] = (int)(o + a);
        h_img[gk] = r.g_image[g];
        for (int d = 0; d < 3; ++d) h_gsh[3 * k + d] = r.g_shift[3 * g + d];
        h_type[gk] = r.type[a];
        h_mass[gk] = r.mass[a];
      }
    }
    r.slot0 = (int)o;
    for (int d = 0; d < 3; ++d) r.trans[d] = tnew[d];
    r.placed = true;
    o += r.n;
  }
  h_gst[n_owned] = (int)gk;
  const long long nf = (long long)h_fdst.size();
  fresh_dst_.need(nf, "serve fresh slots");
  fresh_.need(6 * nf, "serve fresh atoms");
  // uploads through one pinned staging buffer (the stream is idle here)
  const size_t bytes = h_type.size() * 4 + h_mass.size() * 8 + h_req.size() * 4 + h_src.size() * 4 +
                       h_dtr.size() * 8 + h_own.size() * 4 + h_img.size() * 4 + h_gst.size() * 4 +
                       h_gsh.size() * 8 + h_fdst.size() * 4 + h_fresh.size() * 8 + h_dt.size() * 8 + 64 * 16;
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
  up(type_.p, h_type.data(), h_type.size() * 4);
  up(mass_.p, h_mass.data(), h_mass.size() * 8);
  up(atom_req_.p, h_req.data(), h_req.size() * 4);
  up(remap_src_.p, h_src.data(), h_src.size() * 4);
  up(remap_dt_.p, h_dtr.data(), h_dtr.size() * 8);
  up(owner_.p, h_own.data(), h_own.size() * 4);
  up(image_.p, h_img.data(), h_img.size() * 4);
  up(gstart_.p, h_gst.data(), h_gst.size() * 4);
  up(g_shift_.p, h_gsh.data(), h_gsh.size() * 8);
  up(fresh_dst_.p, h_fdst.data(), h_fdst.size() * 4);
  up(fresh_.p, h_fresh.data(), h_fresh.size() * 8);
  up(req_dt_.p, h_dt.data(), h_dt.size() * 8);
  // continuing atoms move from the old layout, new atoms come from staging
  if (old_Nt > 0 && pos_.p)
    check(launch_serve_remap((int)n_owned, (int)Nt, old_Nt, remap_src_.p, remap_dt_.p, pos_.p, vel_.p, force_.p,
                             pos2_.p, vel2_.p, force2_.p, stream_),
          "serve remap");
  check(launch_serve_fresh((int)nf, (int)Nt, fresh_dst_.p, fresh_.p, pos2_.p, vel2_.p, force2_.p, stream_),
        "serve fresh");
  check(hipStreamSynchronize(stream_), "serve relayout");
  std::swap(pos_, pos2_);
  std::swap(vel_, vel2_);
  std::swap(force_, force2_);
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
  check(launch_serve_ghosts(n_owned_, n_ghost_, n_total_, owner_.p + n_owned_, g_shift_.p, pos_.p, stream_),
        "serve ghosts");
  unsigned flags = cfg_.flags & ~1u;
  if (forced_rebuild_) flags |= 2u;
  ImageMap map;
  map.owner = owner_.p;
  map.image = image_.p;
  map.gstart = gstart_.p;
  nep_->compute(batch_box_, n_total_, 0, n_owned_, 0, n_owned_, type_.p, pos_.p, nullptr, pe_.p, force_.p,
                vir_.p, flags, stream_, &map);
  forced_rebuild_ = false;
  if (any_md)
    check(launch_serve_kick(0, n_owned_, n_total_, atom_req_.p, req_md_.p, req_dt_.p, mass_.p, pos_.p, vel_.p,
                            force_.p, stream_),
          "serve kick 2");
  std::lock_guard<std::mutex> lk(mu_);
  st_.iterations++;
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
    for (Request* r : fin) no += r->n;
    out_slot_.need(no, "serve out slots");
    out_shift_.need(3 * no, "serve out shifts");
    out_.need(10 * no, "serve results");
    std::vector<int> slots((size_t)no);
    std::vector<double> shifts(3 * (size_t)no);
    long long k = 0;
    for (Request* r : fin)
      for (int a = 0; a < r->n; ++a, ++k) {
        slots[k] = r->slot0 + a;
        for (int d = 0; d < 3; ++d) shifts[3 * k + d] = r->trans[d];
      }
    const size_t o2 = (slots.size() * 4 + 15) & ~size_t(15);
    const size_t o3 = (o2 + shifts.size() * 8 + 15) & ~size_t(15);
    check(hipStreamSynchronize(stream_), "serve step");  // the staging buffer is free again
    stage(o3 + 10 * (size_t)no * 8);
    char* hs = h_stage_;
    std::memcpy(hs, slots.data(), slots.size() * 4);
    std::memcpy(hs + o2, shifts.data(), shifts.size() * 8);
    check(hipMemcpyAsync(out_slot_.p, hs, slots.size() * 4, hipMemcpyHostToDevice, stream_), "serve out upload");
    check(hipMemcpyAsync(out_shift_.p, hs + o2, shifts.size() * 8, hipMemcpyHostToDevice, stream_),
          "serve out upload");
    check(launch_serve_extract((int)no, n_total_, out_slot_.p, out_shift_.p, pos_.p, vel_.p, force_.p, pe_.p, out_.p,
                               stream_),
          "serve extract");
    check(hipMemcpyAsync(hs + o3, out_.p, 10 * (size_t)no * 8, hipMemcpyDeviceToHost, stream_), "serve results");
    check(hipStreamSynchronize(stream_), "serve results");
    const double* res = reinterpret_cast<const double*>(hs + o3);
    const double t = now_s();
    std::lock_guard<std::mutex> lk(mu_);
    k = 0;
    for (Request* r : fin) {
      const size_t n = (size_t)r->n;
      r->r_pe.resize(n);
      r->r_force.resize(3 * n);
      r->r_pos.resize(3 * n);
      r->r_vel.resize(3 * n);
      for (size_t a = 0; a < n; ++a, ++k) {
        const double* v = res + 10 * k;
        // positions back into the request's own box
        double x[3] = {v[0], v[1], v[2]}, s[3];
        to_frac(r->hi, x, s);
        for (int d = 0; d < 3; ++d)
          if (r->box.pbc[d]) s[d] -= std::floor(s[d]);
        to_cart(r->box.h, s, x);
        for (int d = 0; d < 3; ++d) {
          r->r_pos[d * n + a] = x[d];
          r->r_vel[d * n + a] = v[3 + d];
          r->r_force[d * n + a] = v[6 + d];
        }
        r->r_pe[a] = v[9];
      }
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
