/*
 * dmgx.h -- C ABI of the MI355X-native NEP molecular-dynamics step.
 *
 * The library (libdmgx.so, built from dmg-md_amd/csrc) is a drop-in for one
 * GPU's share of the reference's NEP force evaluator and integrator.  Every
 * entry point below names the reference interface it replaces (file:line in
 * the reference tree).  Conventions:
 *   - every function returns 0 on success or a negative DMGX_E* code; the
 *     message of the last failure on the calling thread is dmgx_last_error();
 *     nothing here exits the process or throws across the ABI;
 *   - array arguments are caller-owned DEVICE pointers (hipMalloc'd) unless
 *     the name ends in _host; per-atom arrays are structure-of-arrays with
 *     stride N exactly as the reference lays them out (nep.cu:1282-1292):
 *       position/force double[3N] x|y|z, potential double[N],
 *       virial double[9N] xx yy zz xy xz yz yx zx zy, type int[N] (0-based
 *       index into the nep.txt symbol list), global_id uint64[N];
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream);
 *     every launch of a call goes to that stream;
 *   - one dmgx_nep handle per host thread; calls on one handle are not
 *     reentrant (the reference is single-threaded too, nep.cu:1311).
 */
#ifndef DMGX_H
#define DMGX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMGX_OK 0
#define DMGX_EINVAL -1     /* bad argument or input file */
#define DMGX_ENOMEM -2     /* device allocation failed */
#define DMGX_EDEVICE -3    /* HIP runtime / kernel launch error */
#define DMGX_EOVERFLOW -4  /* a neighbour row exceeded its capacity */
#define DMGX_EUNSUPPORTED -5

/* Simulation box: h is the reference's cpu_h[0..8] (row-major 3x3 matrix whose
 * COLUMNS are the cell vectors a, b, c; box.cuh:53-70), pbc 1 = periodic. */
typedef struct dmgx_box {
  double h[9];
  int pbc[3];
} dmgx_box;

typedef struct dmgx_nep dmgx_nep;

/* Per-model dimensions, the params() view of the reference (nep.cuh:80-137). */
typedef struct dmgx_nep_info {
  int version;            /* 4 or 5 */
  int num_types;
  int n_max_radial, n_max_angular;
  int basis_size_radial, basis_size_angular;
  int l_max, num_L, dim, num_neurons;
  int zbl_enabled, zbl_flexible;
  int MN_radial, MN_angular; /* enlarged x1.25 as nep.cu:267-268 */
  double rc_radial_max, rc_angular_max;
  int atomic_number[16];     /* first 16 types */
} dmgx_nep_info;

/* Flags of dmgx_nep_compute */
#define DMGX_ACCUMULATE 1u     /* add into pe/force/virial (the reference's +=
                                  contract, nep.cu:745,869); default overwrite */
#define DMGX_FORCE_REBUILD 2u  /* rebuild the Verlet list regardless of skin */
#define DMGX_ANN_BF16 4u       /* ANN with bf16 operands on the matrix cores, fp32
                                  accumulation (single-type models, <= 64 neurons;
                                  other models keep the fp32 ANN) */
#define DMGX_SORT_BY_GID 8u    /* order rows by global_id (domain rows,
                                  neighbor.cuh:164-196); needs global_id */
#define DMGX_ANN_VALU 16u      /* fp32 ANN on the vector ALUs; the default fp32 ANN
                                  runs on the matrix cores where the model allows */
#define DMGX_ANN_BF16_SPLIT 32u /* ANN on the matrix cores with every operand split into
                                  three bf16 parts (hi + mid + lo) and each product formed
                                  from the six part products above 2^-24 relative, fp32
                                  accumulation: fp32 accuracy from bf16 matrix-core
                                  products (same models as DMGX_ANN_BF16) */

/* NEP::NEP(const char*) -- parse nep.txt once and upload the parameters
 * (nep.cu:135-415).  Supports nep4, nep4_zbl, nep5, nep5_zbl with 3-body
 * orders l_max 1..8 and the q_222 / q_1111 many-body invariants. */
int dmgx_nep_create(const char* nep_txt_path, dmgx_nep** out);

/* ~NEP (nep.cu:455) */
void dmgx_nep_destroy(dmgx_nep* nep);

/* NEP::params()/zbl_params() view (nep.cuh:80-137) */
int dmgx_nep_get_info(const dmgx_nep* nep, dmgx_nep_info* out);

/* NEP::allocate_workspace(int) (nep.cu:432-451): size every per-atom workspace
 * for up to `capacity` local slots.  compute() grows it on demand. */
int dmgx_nep_reserve(dmgx_nep* nep, int capacity);

/* Potential::compute (potential.cuh:82-88) and NEP::compute_domain
 * (nep.cuh:178-189) in one call.
 *   [0,N)      every local slot is a neighbour candidate (the SoA stride);
 *   [nd1,nd2)  dependency centres: Verlet rows, descriptors, Fp, partials;
 *   [n1,n2)    force centres: force, virial (and ZBL energy).
 * The plain large-box call is n1=nd1=0, n2=nd2=N.  global_id may be NULL
 * unless DMGX_SORT_BY_GID is set.  Positions must lie inside the box in
 * non-periodic directions.  Outputs are overwritten unless DMGX_ACCUMULATE. */
int dmgx_nep_compute(dmgx_nep* nep, const dmgx_box* box, int N, int n1, int n2, int nd1, int nd2,
                     const int* type, const double* position, const uint64_t* global_id,
                     double* potential, double* force, double* virial, unsigned flags,
                     void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc): the same
 * evaluation as dmgx_nep_compute (same arguments, flags and results), with
 * every per-atom workspace carved from a fixed 128 MiB device array that the
 * library holds on each GPU until the process ends.  It calls no hipMalloc,
 * hipFree or allocating library routine.  Limits and differences:
 *   - the system must fit the fixed workspace (about 17,000 slots of a model
 *     with rows of 152 entries); larger systems get DMGX_EINVAL with the
 *     sizes in the message;
 *   - Verlet rows hold the capacity MN_radial implies (no growth): a longer
 *     row gives DMGX_EOVERFLOW and invalid outputs;
 *   - periodic boxes thinner than 2.5 (rc + skin) give DMGX_EINVAL;
 *   - the Verlet list is rebuilt on every call, and the handle's own list
 *     and workspace are left as they were;
 *   - the stream is synchronised (the status words are read back);
 *   - every call on one GPU shares that GPU's array: calls that overlap on
 *     one GPU (other threads or streams) race. */
int dmgx_nep_compute_noalloc(dmgx_nep* nep, const dmgx_box* box, int N, int n1, int n2, int nd1, int nd2,
                             const int* type, const double* position, const uint64_t* global_id,
                             double* potential, double* force, double* virial, unsigned flags,
                             void* stream);

/* NEP::notify_box_changed (nep.cuh:170, neighbor.cuh:224): the next compute
 * rebuilds the Verlet list. */
int dmgx_nep_notify_box_changed(dmgx_nep* nep);

/* Neighbour-list statistics of the last compute (the neighbor.out record of
 * nep.cu:1156-1176): max radial/angular/Verlet row lengths, rebuild count. */
int dmgx_nep_neighbor_stats(dmgx_nep* nep, int* max_radial, int* max_angular, int* max_verlet,
                            int* num_rebuilds);
/* Summed row lengths over the dependency centres of the last compute (mean
 * neighbour counts for throughput accounting; no reference counterpart). */
int dmgx_nep_neighbor_totals(dmgx_nep* nep, long long* sum_radial, long long* sum_angular,
                             long long* sum_verlet, int* num_centres);

/* Per-kernel timing with HIP events (no counterpart in the reference; the
 * measurement harness uses it).  Kernel ids: 0 filter, 1 fused
 * descriptor/ANN/angular-partial, 2 gather, 3 Verlet rebuild (all its
 * kernels), 4 displacement check, 5 velocity Verlet, 6 thermo.  enable: 0
 * off; 1 every kernel group, events recorded on the launch stream around it
 * (each record separates the kernels it sits between by a few microseconds);
 * 2 the fused descriptor kernel only, its events attached to its own dispatch
 * (hipExtLaunchKernelGGL), which leaves the stream undisturbed.  profile_read
 * synchronises and returns the launch count and summed milliseconds since
 * the last reset. */
int dmgx_nep_set_profiling(dmgx_nep* nep, int enable);
int dmgx_nep_profile_read(dmgx_nep* nep, int kernel, int* launches, double* total_ms);
int dmgx_nep_profile_reset(dmgx_nep* nep);

/* Tile shapes of the per-step kernels (no counterpart in the reference, whose
 * launch shapes are fixed): lanes per centre of the radial/angular filter
 * (8, 16 or 32) and lanes per atom of the force gather (4, 8 or 16), with the
 * median milliseconds each candidate took when the autotuner timed it (0
 * before), and the row layout: every per-centre row (Verlet, edge,
 * partial-force rows) holds row_capacity entries, the capacity needed
 * rounded up to a multiple of 4 (layout 0) or to a power of two (layout 1;
 * -1 before tuning, when the power of two is taken where it costs at most
 * 1.25x), with the filter + descriptor + gather time of each layout the
 * autotuner timed.  Defaults 16 and 8; tuned = 1 once the autotuner has run.
 * Boxes small enough to be replicated keep the defaults. */
typedef struct dmgx_tuning {
  int filter_lanes, gather_lanes;
  double filter_ms[3]; /* candidates 8, 16, 32 */
  double gather_ms[3]; /* candidates 4, 8, 16 */
  int tuned;
  int row_layout, row_capacity;
  double force_ms[2]; /* layouts 0, 1 */
  int overlap;          /* 1: the filter of the second half of the centres runs on a side
                           stream beside the descriptor of the first half (timed only where
                           each half holds at least 16,384 centres) */
  double overlap_ms[2]; /* filter + descriptor + gather without / with the overlap */
  int split;            /* 1: the rows are split by type pair at rebuild time (every Verlet
                           entry within its pair's rc + skin) and the step has no filter
                           pass: the descriptor and gather form the distances themselves */
  double split_ms[2];   /* the step's force kernels with the filter / with the split rows */
  int desc_lanes;       /* lanes per atom of the descriptor: 1, 4 (a group of four lanes shares
                           an atom's edges and hidden neurons; FP32 ANN on the vector ALUs),
                           0 before tuning (four below 65,536 centres, else one); BF16 ANN
                           requests keep one */
  double desc_ms[2];    /* the step's force kernels with 1 / 4 lanes per atom */
} dmgx_tuning;
int dmgx_nep_get_tuning(const dmgx_nep* nep, dmgx_tuning* out);
/* Sets the shapes directly (DMGX_EINVAL for values other than the
 * candidates); a server bound to the handle keeps them (no autotuning). */
int dmgx_nep_set_tiles(dmgx_nep* nep, int filter_lanes, int gather_lanes);
/* Selects where the typewise radial/angular split of the Verlet rows happens:
 * 0 every step (the filter pass of nep.cu:503-561), 1 once per Verlet
 * rebuild, with the distance test at the point of use.  The next evaluation
 * rebuilds the rows; a server bound to the handle keeps the choice. */
int dmgx_nep_set_split(dmgx_nep* nep, int split);
/* Sets the descriptor's lanes per atom (0: by system size, 1 or 4; no
 * counterpart in the reference, whose launch shape is fixed); a server bound
 * to the handle keeps the choice. */
int dmgx_nep_set_desc_lanes(dmgx_nep* nep, int lanes);
/* Times every candidate on this system (positions in device memory, as
 * dmgx_nep_compute) and keeps the fastest for all later evaluations of this
 * handle; afterwards pe/force/virial hold the system's values as the chosen
 * shapes compute them.  Synchronises the stream. */
int dmgx_nep_autotune(dmgx_nep* nep, const dmgx_box* box, int N, const int* type,
                      const double* position, double* potential, double* force, double* virial,
                      unsigned flags, void* stream, dmgx_tuning* out);

/* Copies the current Verlet rows of centres [c1,c2) to host memory in the
 * order the device stores them: ascending by slot, or by (global id, slot)
 * when the last compute set DMGX_SORT_BY_GID (test access to the row
 * contract of neighbor.cu:141-218 and neighbor.cuh:133-196):
 * counts_host[c2-c1], rows_host[(c2-c1)*max_nb]. */
int dmgx_nep_copy_verlet_rows(dmgx_nep* nep, int c1, int c2, int max_nb, int* counts_host,
                              int* rows_host);
/* Same for the per-step radial (which=0) or angular (which=1) rows. */
int dmgx_nep_copy_pair_rows(dmgx_nep* nep, int which, int c1, int c2, int max_nb,
                            int* counts_host, int* rows_host);

/* ---------------------------------------------------------------------- */
/* Integrator and thermodynamics (FP64)                                    */
/* ---------------------------------------------------------------------- */

/* velocity_verlet kernel (runtime.cu:392-421): v += f/m * dt/2, and on the
 * first half x += v*dt.  dt in natural units (fs / 10.18051). */
int dmgx_velocity_verlet(int first_half, int N, double dt, const double* mass, double* position,
                         double* velocity, const double* force, void* stream);

/* wrap_positions (runtime_internal.hpp:648-681): fold into the box through
 * fractional coordinates. */
int dmgx_wrap_positions(const dmgx_box* box, int N, double* position, void* stream);

/* find_owned_thermo_sums + normalize_global_thermo (runtime.cu:442-535,
 * runtime_internal.hpp:704-716) with compensated FP64 sums:
 * thermo_host[8] = T, sum pe, (sum virial_ab + m v_a v_b)/V for xx yy zz xy
 * xz yz.  Synchronises the stream. */
int dmgx_thermo(const dmgx_box* box, int N, const double* mass, const double* potential,
                const double* velocity, const double* virial, double* thermo_host, void* stream);

/* scale_owned_velocity (runtime.cu:495-510) */
int dmgx_scale_velocity(int N, double factor, double* velocity, void* stream);

/* preflight_scaled_positions + scale_positions (box_control.cu:22-46,128-174):
 * isotropic x <- factor*x; returns DMGX_EINVAL if a scaled position is not
 * finite (positions are left unchanged in that case). */
int dmgx_scale_positions(int N, double factor, double* position, void* stream);

/* ---------------------------------------------------------------------- */
/* Ensemble scalars (host FP64; bit-compatible with the reference)         */
/* ---------------------------------------------------------------------- */

typedef struct dmgx_rng dmgx_rng; /* std::mt19937 stream of SvrState */

/* SvrState(seed) (thermostat_svr.cpp:29) */
int dmgx_rng_create(uint32_t seed, dmgx_rng** out);
void dmgx_rng_destroy(dmgx_rng* rng);
/* gasdev, gamdev, resamplekin_sumnoises, resamplekin (thermostat_svr.cpp:31-124) */
double dmgx_gasdev(dmgx_rng* rng);
double dmgx_gamdev(int ia, dmgx_rng* rng);
double dmgx_resamplekin_sumnoises(int nn, dmgx_rng* rng);
double dmgx_resamplekin(double kk, double sigma, int ndeg, double taut, dmgx_rng* rng);

/* make_nhc_state / nhc_half_step (thermostat_nhc.cpp:30-117); chain state is
 * position[4], momentum[4], mass[4]. */
typedef struct dmgx_nhc {
  double position[4], momentum[4], mass[4];
} dmgx_nhc;
int dmgx_nhc_init(int atom_count, double temperature, double damping, double dt, dmgx_nhc* out);
double dmgx_nhc_half_step(dmgx_nhc* state, double ek2, double kT, double dN, double dt_half);

/* npt_scr_isotropic_scale (npt_scr.cpp:21-34) */
double dmgx_npt_scr_scale(double target_pressure, double pressure_coupling,
                          double pressure_instant, double target_temperature, double volume,
                          double gaussian);

/* ---------------------------------------------------------------------- */
/* MD segment (runtime.cu:680-849 run_segment, single GPU, no MPI)          */
/* ---------------------------------------------------------------------- */

typedef struct dmgx_md dmgx_md;

#define DMGX_ENS_NVE 0
#define DMGX_ENS_NVT_BER 1
#define DMGX_ENS_NVT_BDP 2
#define DMGX_ENS_NVT_NHC 3
#define DMGX_ENS_NPT_BER 4
#define DMGX_ENS_NPT_SCR 5

typedef struct dmgx_ensemble {
  int kind;               /* DMGX_ENS_* */
  double t_init, t_final; /* K */
  double t_coupling;      /* Tdamp in steps (run.in ensemble line) */
  double p_target_gpa, bulk_modulus_gpa, p_coupling; /* NPT */
  uint32_t seed;          /* BDP/SCR RNG seed (default 12345678) */
} dmgx_ensemble;

/* Creates an MD state bound to a NEP handle; copies the caller's host arrays
 * (positions/velocities in A and A/natural-time, masses in amu) into device
 * memory owned by the state. */
int dmgx_md_create(dmgx_nep* nep, const dmgx_box* box, int N, const int* type_host,
                   const double* mass_host, const double* position_host,
                   const double* velocity_host, dmgx_md** out);
void dmgx_md_destroy(dmgx_md* md);
/* Runs `steps` velocity-Verlet steps of dt_fs femtoseconds under `ens`,
 * evaluating forces on the device; thermo of the last step -> thermo_host[8]
 * (may be NULL).  flags as dmgx_nep_compute (DMGX_ANN_BF16). */
int dmgx_md_run(dmgx_md* md, int steps, double dt_fs, const dmgx_ensemble* ens, unsigned flags,
                double* thermo_host);
/* Device pointers of the state (position, velocity, force, potential, virial,
 * mass, type) for zero-copy inspection. */
int dmgx_md_arrays(dmgx_md* md, double** position, double** velocity, double** force,
                   double** potential, double** virial, double** mass, int** type);
int dmgx_md_box(dmgx_md* md, dmgx_box* box_out);
/* Stores the state's atoms in cell order (cells of half the Verlet range),
 * once, so that neighbours sit near each other in memory; every per-atom
 * array, including dmgx_md_arrays', is permuted alike. order_host (N ints,
 * may be NULL) receives, for each stored atom, its index in the arrays given
 * to dmgx_md_create (composed over repeated calls). No counterpart in the
 * reference, which keeps the input order: it is for inputs whose order is
 * spatially random (random insertion, shuffled files), where the neighbour
 * gathers miss the caches. */
int dmgx_md_reorder(dmgx_md* md, int* order_host);
/* dmgx_nep_autotune on the state's current positions (wrapped into the box
 * first); the state's forces are those positions' forces afterwards. */
int dmgx_md_autotune(dmgx_md* md, unsigned flags, dmgx_tuning* out);
/* on != 0: the device-resident MD loop (NVE and the NVT thermostats; not a
 * barostat, not under the per-kernel profiler) replays one captured step,
 * a hipGraph of the step's launches, instead of launching each kernel;
 * bit-identical results.  Off by default: on a 4,096-atom system the step
 * is bound by its kernels, not their launches, and the replay measured
 * slower.  No counterpart in the reference (a per-kernel launch loop). */
int dmgx_md_set_step_graph(dmgx_md* md, int on);

/* ---------------------------------------------------------------------- */
/* Serving loop: continuous batching of independent periodic systems.      */
/* No reference counterpart (the reference runs one system per process).    */
/* ---------------------------------------------------------------------- */

typedef struct dmgx_server dmgx_server;

typedef struct dmgx_server_config {
  long long max_atoms;   /* slots in flight over all models: owned atoms + periodic-image shells */
  int max_requests;      /* requests in flight */
  unsigned flags;        /* DMGX_ANN_BF16 / DMGX_ANN_VALU */
  double margin;         /* extra shell thickness (A) for an MD request's motion: a request
                            whose atoms move margin / 2 gets a new image shell */
  int use_graph;         /* nonzero: the force launches of a batch layout are captured once
                            into a hipGraph and replayed every iteration (a server of one
                            model; with several, frames are launched directly: a capture
                            needs the device to itself, which stalls the other models) */
  long long hbm_bytes;   /* admission budget: device bytes of the requests in flight, each
                            request counting (atoms + image shell) x its model's bytes per
                            slot (engine workspace at the model's row capacity and the batch
                            layout's arrays) plus its upload; 0 = half the device memory
                            free when the server starts */
} dmgx_server_config;

typedef struct dmgx_server_stats {
  long long submitted, completed;
  long long frames;      /* single points + MD steps served */
  long long atom_frames; /* the same, weighted by the request's atom count */
  double p50_ms, p99_ms, mean_ms, max_ms; /* submit-to-result latency */
  double elapsed_s;      /* first submit to last completion */
  double frames_per_s;
  long long iterations, relayouts;
  long long max_slots;   /* largest batch over all models: owned atoms + image ghosts */
  double relayout_s, iterate_s, retire_s; /* worker wall time per phase */
  long long hbm_budget;          /* the admission budget in bytes */
  long long max_bytes_in_flight; /* the most device bytes admitted at once */
  long long reshells;            /* MD requests given a new image shell after their atoms drifted
                                    a quarter of the margin from where the shell was built */
  long long rows_carried;        /* relayouts whose continuing requests kept their Verlet rows
                                    (only the new and re-shelled requests' rows were built) */
  double reshell_s;              /* worker wall time spent building new image shells */
} dmgx_server_stats;

/* Starts a server (one worker thread; one batch and one stream per model) on
 * num_models NEP handles; a request names its model by index into `models`.
 * The handles must not be used elsewhere while the server lives.  Every
 * model's batch is evaluated each iteration, the batches' launches
 * overlapping on the device. */
int dmgx_server_create(dmgx_nep* const* models, int num_models, const dmgx_server_config* cfg,
                       dmgx_server** out);
/* Stops the worker (pending requests are dropped) and frees the server. */
void dmgx_server_destroy(dmgx_server* server);
/* Submits a periodic system (host arrays, copied) for model `model`: steps = 0
 * is a single point (energy, forces, virial), steps > 0 that many NVE
 * velocity-Verlet steps of dt_fs.  velocity may be NULL (zero).  A request
 * whose device bytes exceed the whole budget is refused with DMGX_EINVAL. */
int dmgx_server_submit(dmgx_server* server, int model, const dmgx_box* box, int N, const int* type_host,
                       const double* mass_host, const double* position_host,
                       const double* velocity_host, int steps, double dt_fs, long long* id_out);
/* *done = 1 once the request finished; then its results are copied to the
 * host arrays given (each may be NULL): potential[N], force[3N],
 * position[3N] (wrapped into its box), velocity[3N], virial[9N] (per atom,
 * xx yy zz xy xz yz yx zx zy, as dmgx_nep_compute), latency in ms.
 * wait != 0 blocks until it finishes. */
int dmgx_server_poll(dmgx_server* server, long long id, int wait, double* potential_host,
                     double* force_host, double* position_host, double* velocity_host,
                     double* virial_host, double* latency_ms, int* done);
/* Blocks until every submitted request has finished. */
int dmgx_server_drain(dmgx_server* server);
int dmgx_server_get_stats(dmgx_server* server, dmgx_server_stats* out);
/* Zeroes the statistics (counters, latencies, time window) of a drained
 * server, keeping its device buffers and pools: the next stream is measured
 * on a warmed server.  No counterpart in the reference (serving loop). */
int dmgx_server_reset_stats(dmgx_server* server);

/* ---------------------------------------------------------------------- */
/* Input front end (dmgmd_main.cpp + run_m1_replicated, runtime.cu)         */
/* ---------------------------------------------------------------------- */

/* Runs a run.in / model.xyz input directory: commands potential, velocity,
 * time_step, ensemble (nve, nvt_ber, nvt_bdp, nvt_nhc, npt_ber, npt_scr),
 * dump_thermo, dump_xyz, dump_restart, run (run_parser.cpp:158-315); writes
 * thermo.out, the dump_xyz files, restart.xyz and neighbor.out into the
 * directory in the reference's formats (runtime.cu:1148-1313).  run_file and
 * model_file default to "run.in" and "model.xyz" when NULL. */
int dmgx_run_directory(const char* directory, const char* run_file, const char* model_file);

/* Message of the last failure on this thread ("" if none). */
const char* dmgx_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* DMGX_H */
