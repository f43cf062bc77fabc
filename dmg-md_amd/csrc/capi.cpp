// capi.cpp -- the C ABI declared in include/dmgx.h.
//
// Every entry point catches C++ exceptions and HIP failures and returns a
// DMGX_E* status with the message kept per thread (dmgx_last_error); nothing
// crosses the ABI as an exception and nothing exits the process (the
// reference exits on parse errors, nep.cu:139-147; here the caller decides).
#include "dmgx.h"

#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "engine.hpp"
#include "ensemble.hpp"
#include "frontend.hpp"
#include "noalloc.hpp"
#include "serve.hpp"

struct dmgx_nep {
  dmgx::NepEngine engine;
  explicit dmgx_nep(const char* path) : engine(path) {}
};

struct dmgx_md {
  dmgx::MdState state;
  template <class... A>
  explicit dmgx_md(A&&... a) : state(std::forward<A>(a)...) {}
};

struct dmgx_server {
  dmgx::Server server;
  dmgx_server(const std::vector<dmgx::NepEngine*>& models, const dmgx::ServeConfig& cfg) : server(models, cfg) {}
};

struct dmgx_rng {
  dmgx::Rng rng;
  explicit dmgx_rng(uint32_t seed) : rng(seed) {}
};

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg)
{
  g_error = msg;
  return code;
}

template <class F>
int guarded(F&& f)
{
  try {
    g_error.clear();
    f();
    return DMGX_OK;
  } catch (const std::invalid_argument& e) {
    return fail(DMGX_EINVAL, e.what());
  } catch (const std::bad_alloc& e) {
    return fail(DMGX_ENOMEM, e.what());
  } catch (const std::runtime_error& e) {
    const std::string m = e.what();
    if (m.find("allocation failed") != std::string::npos) return fail(DMGX_ENOMEM, m);
    if (m.find("overflow") != std::string::npos) return fail(DMGX_EOVERFLOW, m);
    return fail(DMGX_EDEVICE, m);
  } catch (const std::exception& e) {
    return fail(DMGX_EDEVICE, e.what());
  } catch (...) {
    return fail(DMGX_EDEVICE, "unknown error");
  }
}

dmgx::Box to_box(const dmgx_box* b)
{
  if (!b) throw std::invalid_argument("null box");
  dmgx::Box o;
  std::memcpy(o.h, b->h, sizeof(o.h));
  std::memcpy(o.pbc, b->pbc, sizeof(o.pbc));
  return o;
}

}  // namespace

extern "C" {

const char* dmgx_last_error(void) { return g_error.c_str(); }

int dmgx_nep_create(const char* path, dmgx_nep** out)
{
  return guarded([&] {
    if (!path || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    *out = new dmgx_nep(path);
  });
}

void dmgx_nep_destroy(dmgx_nep* nep) { delete nep; }

int dmgx_nep_get_info(const dmgx_nep* nep, dmgx_nep_info* out)
{
  return guarded([&] {
    if (!nep || !out) throw std::invalid_argument("null argument");
    const dmgx::NepModel& m = nep->engine.model();
    std::memset(out, 0, sizeof(*out));
    out->version = m.version;
    out->num_types = m.num_types;
    out->n_max_radial = m.nR - 1;
    out->n_max_angular = m.nA - 1;
    out->basis_size_radial = m.kR - 1;
    out->basis_size_angular = m.kA - 1;
    out->l_max = m.L_max;
    out->num_L = m.num_L;
    out->dim = m.dim;
    out->num_neurons = m.H;
    out->zbl_enabled = m.zbl_enabled;
    out->zbl_flexible = m.zbl_flexible;
    out->MN_radial = m.MN_R;
    out->MN_angular = m.MN_A;
    out->rc_radial_max = m.rcR_max;
    out->rc_angular_max = m.rcA_max;
    for (int t = 0; t < m.num_types && t < 16; ++t) out->atomic_number[t] = m.Z[t];
  });
}

int dmgx_nep_reserve(dmgx_nep* nep, int capacity)
{
  return guarded([&] {
    if (!nep) throw std::invalid_argument("null handle");
    nep->engine.reserve(capacity);
  });
}

int dmgx_nep_compute(dmgx_nep* nep, const dmgx_box* box, int N, int n1, int n2, int nd1, int nd2,
                     const int* type, const double* position, const uint64_t* global_id, double* potential,
                     double* force, double* virial, unsigned flags, void* stream)
{
  return guarded([&] {
    if (!nep) throw std::invalid_argument("null handle");
    nep->engine.compute(to_box(box), N, n1, n2, nd1, nd2, type, position, global_id, potential, force, virial,
                        flags, (hipStream_t)stream);
  });
}

int dmgx_nep_compute_noalloc(dmgx_nep* nep, const dmgx_box* box, int N, int n1, int n2, int nd1, int nd2,
                             const int* type, const double* position, const uint64_t* global_id,
                             double* potential, double* force, double* virial, unsigned flags, void* stream)
{
  return guarded([&] {
    if (!nep) throw std::invalid_argument("null handle");
    void* ws = nullptr;
    size_t bytes = 0;
    const hipError_t e = dmgx::fixed_workspace(&ws, &bytes);
    if (e != hipSuccess) throw std::runtime_error(std::string("fixed workspace: ") + hipGetErrorString(e));
    nep->engine.compute_in(ws, bytes, to_box(box), N, n1, n2, nd1, nd2, type, position, global_id, potential, force,
                           virial, flags, (hipStream_t)stream);
  });
}

int dmgx_nep_notify_box_changed(dmgx_nep* nep)
{
  return guarded([&] {
    if (!nep) throw std::invalid_argument("null handle");
    nep->engine.notify_box_changed();
  });
}

int dmgx_nep_neighbor_stats(dmgx_nep* nep, int* max_radial, int* max_angular, int* max_verlet, int* rebuilds)
{
  return guarded([&] {
    if (!nep) throw std::invalid_argument("null handle");
    nep->engine.stats(max_radial, max_angular, max_verlet, rebuilds);
  });
}

int dmgx_nep_neighbor_totals(dmgx_nep* nep, long long* sum_radial, long long* sum_angular, long long* sum_verlet,
                             int* num_centres)
{
  return guarded([&] {
    if (!nep) throw std::invalid_argument("null handle");
    nep->engine.totals(sum_radial, sum_angular, sum_verlet, num_centres);
  });
}

int dmgx_nep_set_profiling(dmgx_nep* nep, int enable)
{
  return guarded([&] {
    if (!nep) throw std::invalid_argument("null handle");
    if (enable < 0 || enable > 2) throw std::invalid_argument("profiling mode is 0, 1 or 2");
    nep->engine.profiler().enable(enable);
  });
}

namespace {
void fill_tuning(const dmgx::NepEngine::Tuning& t, dmgx_tuning* out)
{
  if (!out) return;
  out->filter_lanes = t.filter_lanes;
  out->gather_lanes = t.gather_lanes;
  for (int k = 0; k < 3; ++k) {
    out->filter_ms[k] = t.filter_ms[k];
    out->gather_ms[k] = t.gather_ms[k];
  }
  out->tuned = t.tuned ? 1 : 0;
  out->row_layout = t.row_layout;
  out->row_capacity = t.row_capacity;
  out->force_ms[0] = t.force_ms[0];
  out->force_ms[1] = t.force_ms[1];
  out->overlap = t.overlap;
  out->overlap_ms[0] = t.overlap_ms[0];
  out->overlap_ms[1] = t.overlap_ms[1];
  out->split = t.split;
  out->split_ms[0] = t.split_ms[0];
  out->split_ms[1] = t.split_ms[1];
  out->desc_lanes = t.desc_lanes;
  out->desc_ms[0] = t.desc_ms[0];
  out->desc_ms[1] = t.desc_ms[1];
}
}  // namespace

int dmgx_nep_get_tuning(const dmgx_nep* nep, dmgx_tuning* out)
{
  return guarded([&] {
    if (!nep || !out) throw std::invalid_argument("null argument");
    fill_tuning(nep->engine.tuning(), out);
  });
}

int dmgx_nep_set_tiles(dmgx_nep* nep, int filter_lanes, int gather_lanes)
{
  return guarded([&] {
    if (!nep) throw std::invalid_argument("null handle");
    nep->engine.set_tiles(filter_lanes, gather_lanes);
  });
}

int dmgx_nep_set_split(dmgx_nep* nep, int split)
{
  return guarded([&] {
    if (!nep) throw std::invalid_argument("null handle");
    nep->engine.set_split(split);
  });
}

int dmgx_nep_set_desc_lanes(dmgx_nep* nep, int lanes)
{
  return guarded([&] {
    if (!nep) throw std::invalid_argument("null handle");
    nep->engine.set_desc_lanes(lanes);
  });
}

int dmgx_nep_autotune(dmgx_nep* nep, const dmgx_box* box, int N, const int* type, const double* position,
                      double* potential, double* force, double* virial, unsigned flags, void* stream,
                      dmgx_tuning* out)
{
  return guarded([&] {
    if (!nep || !box || !type || !position || !potential || !force || !virial)
      throw std::invalid_argument("null argument");
    if (N <= 0) throw std::invalid_argument("N must be positive");
    fill_tuning(nep->engine.autotune(to_box(box), N, N, type, position, potential, force, virial, flags,
                                     (hipStream_t)stream),
                out);
  });
}

int dmgx_nep_profile_read(dmgx_nep* nep, int kernel, int* launches, double* total_ms)
{
  return guarded([&] {
    if (!nep || !launches || !total_ms) throw std::invalid_argument("null argument");
    nep->engine.profiler().read(kernel, launches, total_ms);
  });
}

int dmgx_nep_profile_reset(dmgx_nep* nep)
{
  return guarded([&] {
    if (!nep) throw std::invalid_argument("null handle");
    nep->engine.profiler().reset();
  });
}

int dmgx_nep_copy_verlet_rows(dmgx_nep* nep, int c1, int c2, int max_nb, int* counts, int* rows)
{
  return guarded([&] {
    if (!nep || !counts || !rows) throw std::invalid_argument("null argument");
    nep->engine.copy_verlet_rows(c1, c2, max_nb, counts, rows);
  });
}

int dmgx_nep_copy_pair_rows(dmgx_nep* nep, int which, int c1, int c2, int max_nb, int* counts, int* rows)
{
  return guarded([&] {
    if (!nep || !counts || !rows) throw std::invalid_argument("null argument");
    if (which != 0 && which != 1) throw std::invalid_argument("which must be 0 (radial) or 1 (angular)");
    nep->engine.copy_pair_rows(which, c1, c2, max_nb, counts, rows);
  });
}

int dmgx_velocity_verlet(int first_half, int N, double dt, const double* mass, double* position,
                         double* velocity, const double* force, void* stream)
{
  return guarded([&] {
    if (N < 0) throw std::invalid_argument("negative N");
    if (N > 0 && (!mass || !position || !velocity || !force)) throw std::invalid_argument("null array");
    const hipError_t e =
        dmgx::launch_velocity_verlet(first_half, N, dt, mass, position, velocity, force, (hipStream_t)stream);
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
  });
}

int dmgx_wrap_positions(const dmgx_box* box, int N, double* position, void* stream)
{
  return guarded([&] {
    dmgx::Box b = to_box(box);
    if (N < 0 || (N > 0 && !position)) throw std::invalid_argument("bad arguments");
    dmgx::DevBox d{};
    std::memcpy(d.h, b.h, sizeof(d.h));
    dmgx::box_inverse(b.h, d.hi);
    for (int k = 0; k < 3; ++k) d.pbc[k] = b.pbc[k] ? 1 : 0;
    const hipError_t e = dmgx::launch_wrap(d, N, position, (hipStream_t)stream);
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
  });
}

int dmgx_thermo(const dmgx_box* box, int N, const double* mass, const double* potential, const double* velocity,
                const double* virial, double* thermo_host, void* stream)
{
  return guarded([&] {
    dmgx::Box b = to_box(box);
    if (N <= 0 || !mass || !potential || !velocity || !virial || !thermo_host)
      throw std::invalid_argument("bad arguments");
    dmgx::DevWork w{};
    hipStream_t s = (hipStream_t)stream;
    if (hipMalloc((void**)&w.partial, sizeof(double) * 16 * dmgx::thermo_blocks(N)) != hipSuccess ||
        hipMalloc((void**)&w.thermo, sizeof(double) * 8) != hipSuccess) {
      if (w.partial) (void)hipFree(w.partial);
      throw std::runtime_error("device allocation failed: thermo scratch");
    }
    hipError_t e = dmgx::launch_thermo(N, dmgx::box_volume(b.h), mass, potential, velocity, virial, w, s);
    if (e == hipSuccess) e = hipMemcpyAsync(thermo_host, w.thermo, 8 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(w.partial);
    (void)hipFree(w.thermo);
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
  });
}

int dmgx_scale_velocity(int N, double factor, double* velocity, void* stream)
{
  return guarded([&] {
    if (N < 0 || (N > 0 && !velocity)) throw std::invalid_argument("bad arguments");
    const hipError_t e = dmgx::launch_scale(3 * N, factor, velocity, (hipStream_t)stream);
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
  });
}

int dmgx_scale_positions(int N, double factor, double* position, void* stream)
{
  return guarded([&] {
    if (N < 0 || (N > 0 && !position)) throw std::invalid_argument("bad arguments");
    if (!std::isfinite(factor) || factor <= 0.0) throw std::invalid_argument("scale factor must be finite and > 0");
    hipStream_t s = (hipStream_t)stream;
    int* flag = nullptr;
    if (hipMalloc((void**)&flag, sizeof(int)) != hipSuccess) throw std::runtime_error("device allocation failed");
    int h = 0;
    hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), s);
    if (e == hipSuccess) e = dmgx::launch_check_scaled(3 * N, factor, position, flag, s);
    if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && !h) e = dmgx::launch_scale(3 * N, factor, position, s);
    (void)hipFree(flag);
    if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));
    if (h) throw std::invalid_argument("scaled position is not finite");
  });
}

int dmgx_rng_create(uint32_t seed, dmgx_rng** out)
{
  return guarded([&] {
    if (!out) throw std::invalid_argument("null argument");
    *out = new dmgx_rng(seed);
  });
}

void dmgx_rng_destroy(dmgx_rng* rng) { delete rng; }

double dmgx_gasdev(dmgx_rng* rng) { return dmgx::gasdev(rng->rng); }

double dmgx_gamdev(int ia, dmgx_rng* rng)
{
  try {
    return dmgx::gamdev(ia, rng->rng);
  } catch (const std::exception& e) {
    g_error = e.what();
    return std::nan("");
  }
}

double dmgx_resamplekin_sumnoises(int nn, dmgx_rng* rng) { return dmgx::resamplekin_sumnoises(nn, rng->rng); }

double dmgx_resamplekin(double kk, double sigma, int ndeg, double taut, dmgx_rng* rng)
{
  return dmgx::resamplekin(kk, sigma, ndeg, taut, rng->rng);
}

int dmgx_nhc_init(int atom_count, double temperature, double damping, double dt, dmgx_nhc* out)
{
  return guarded([&] {
    if (!out) throw std::invalid_argument("null argument");
    dmgx::NhcChain c = dmgx::make_nhc_chain(atom_count, temperature, damping, dt);
    for (int i = 0; i < 4; ++i) {
      out->position[i] = c.position[i];
      out->momentum[i] = c.momentum[i];
      out->mass[i] = c.mass[i];
    }
  });
}

double dmgx_nhc_half_step(dmgx_nhc* s, double ek2, double kT, double dN, double dt_half)
{
  dmgx::NhcChain c;
  for (int i = 0; i < 4; ++i) {
    c.position[i] = s->position[i];
    c.momentum[i] = s->momentum[i];
    c.mass[i] = s->mass[i];
  }
  const double f = dmgx::nhc_half_step(c, ek2, kT, dN, dt_half);
  for (int i = 0; i < 4; ++i) {
    s->position[i] = c.position[i];
    s->momentum[i] = c.momentum[i];
    s->mass[i] = c.mass[i];
  }
  return f;
}

double dmgx_npt_scr_scale(double p_target, double p_coupling, double p_inst, double t_target, double volume,
                          double gaussian)
{
  return dmgx::npt_scr_scale(p_target, p_coupling, p_inst, t_target, volume, gaussian);
}

int dmgx_md_create(dmgx_nep* nep, const dmgx_box* box, int N, const int* type_host, const double* mass_host,
                   const double* position_host, const double* velocity_host, dmgx_md** out)
{
  return guarded([&] {
    if (!nep || !out || !type_host || !mass_host || !position_host || !velocity_host)
      throw std::invalid_argument("null argument");
    *out = nullptr;
    *out = new dmgx_md(&nep->engine, to_box(box), N, type_host, mass_host, position_host, velocity_host);
  });
}

void dmgx_md_destroy(dmgx_md* md) { delete md; }

int dmgx_md_run(dmgx_md* md, int steps, double dt_fs, const dmgx_ensemble* ens, unsigned flags,
                double* thermo_host)
{
  return guarded([&] {
    if (!md || !ens) throw std::invalid_argument("null argument");
    dmgx::EnsembleConfig c;
    if (ens->kind < 0 || ens->kind > 5) throw std::invalid_argument("unknown ensemble kind");
    c.kind = (dmgx::Ensemble)ens->kind;
    c.t_init = ens->t_init;
    c.t_final = ens->t_final;
    c.t_coupling = ens->t_coupling;
    c.p_target_gpa = ens->p_target_gpa;
    c.bulk_modulus_gpa = ens->bulk_modulus_gpa;
    c.p_coupling = ens->p_coupling;
    c.seed = ens->seed;
    md->state.run(steps, dt_fs, c, flags, thermo_host);
  });
}

int dmgx_md_arrays(dmgx_md* md, double** position, double** velocity, double** force, double** potential,
                   double** virial, double** mass, int** type)
{
  return guarded([&] {
    if (!md) throw std::invalid_argument("null handle");
    md->state.arrays(position, velocity, force, potential, virial, mass, type);
  });
}

int dmgx_md_box(dmgx_md* md, dmgx_box* out)
{
  return guarded([&] {
    if (!md || !out) throw std::invalid_argument("null argument");
    std::memcpy(out->h, md->state.box().h, sizeof(out->h));
    std::memcpy(out->pbc, md->state.box().pbc, sizeof(out->pbc));
  });
}

int dmgx_md_reorder(dmgx_md* md, int* order_host)
{
  return guarded([&] {
    if (!md) throw std::invalid_argument("null handle");
    md->state.reorder();
    if (order_host) std::memcpy(order_host, md->state.order().data(), sizeof(int) * md->state.order().size());
  });
}

int dmgx_md_autotune(dmgx_md* md, unsigned flags, dmgx_tuning* out)
{
  return guarded([&] {
    if (!md) throw std::invalid_argument("null handle");
    fill_tuning(md->state.autotune(flags), out);
  });
}

int dmgx_md_set_step_graph(dmgx_md* md, int on)
{
  return guarded([&] {
    if (!md) throw std::invalid_argument("null handle");
    md->state.set_step_graph(on != 0);
  });
}

int dmgx_server_create(dmgx_nep* const* models, int num_models, const dmgx_server_config* cfg,
                       dmgx_server** out)
{
  return guarded([&] {
    if (!models || !cfg || !out) throw std::invalid_argument("null argument");
    if (num_models <= 0 || num_models > 64) throw std::invalid_argument("num_models must lie in [1, 64]");
    *out = nullptr;
    std::vector<dmgx::NepEngine*> engines;
    for (int m = 0; m < num_models; ++m) {
      if (!models[m]) throw std::invalid_argument("null potential handle");
      engines.push_back(&models[m]->engine);
    }
    dmgx::ServeConfig c;
    c.max_atoms = cfg->max_atoms;
    c.max_requests = cfg->max_requests;
    c.flags = cfg->flags;
    c.margin = cfg->margin;
    c.use_graph = cfg->use_graph != 0;
    c.hbm_bytes = cfg->hbm_bytes;
    if (!(c.margin >= 0.0)) throw std::invalid_argument("negative shell margin");
    *out = new dmgx_server(engines, c);
  });
}

void dmgx_server_destroy(dmgx_server* server) { delete server; }

int dmgx_server_submit(dmgx_server* server, int model, const dmgx_box* box, int N, const int* type_host,
                       const double* mass_host, const double* position_host, const double* velocity_host, int steps,
                       double dt_fs, long long* id_out)
{
  return guarded([&] {
    if (!server || !box || !id_out) throw std::invalid_argument("null argument");
    *id_out = server->server.submit(model, to_box(box), N, type_host, mass_host, position_host, velocity_host, steps,
                                    dt_fs);
  });
}

int dmgx_server_poll(dmgx_server* server, long long id, int wait, double* potential_host, double* force_host,
                     double* position_host, double* velocity_host, double* virial_host, double* latency_ms,
                     int* done)
{
  return guarded([&] {
    if (!server || !done) throw std::invalid_argument("null argument");
    *done = server->server.poll(id, wait != 0, potential_host, force_host, position_host, velocity_host, virial_host,
                                latency_ms)
                ? 1
                : 0;
  });
}

int dmgx_server_drain(dmgx_server* server)
{
  return guarded([&] {
    if (!server) throw std::invalid_argument("null handle");
    server->server.drain();
  });
}

int dmgx_server_reset_stats(dmgx_server* server)
{
  return guarded([&] {
    if (!server) throw std::invalid_argument("null handle");
    server->server.reset_stats();
  });
}

int dmgx_server_get_stats(dmgx_server* server, dmgx_server_stats* out)
{
  return guarded([&] {
    if (!server || !out) throw std::invalid_argument("null argument");
    const dmgx::ServeStats s = server->server.stats();
    out->submitted = s.submitted;
    out->completed = s.completed;
    out->frames = s.frames;
    out->atom_frames = s.atom_frames;
    out->p50_ms = s.p50_ms;
    out->p99_ms = s.p99_ms;
    out->mean_ms = s.mean_ms;
    out->max_ms = s.max_ms;
    out->elapsed_s = s.elapsed_s;
    out->frames_per_s = s.frames_per_s;
    out->iterations = s.iterations;
    out->relayouts = s.relayouts;
    out->max_slots = s.max_slots;
    out->relayout_s = s.relayout_s;
    out->iterate_s = s.iterate_s;
    out->retire_s = s.retire_s;
    out->hbm_budget = s.hbm_budget;
    out->max_bytes_in_flight = s.max_bytes_in_flight;
    out->reshells = s.reshells;
    out->rows_carried = s.rows_carried;
    out->reshell_s = s.reshell_s;
  });
}

int dmgx_run_directory(const char* directory, const char* run_file, const char* model_file)
{
  return guarded([&] {
    if (!directory) throw std::invalid_argument("null directory");
    dmgx::run_directory(directory, run_file ? run_file : "run.in", model_file ? model_file : "model.xyz");
  });
}

}  // extern "C"
