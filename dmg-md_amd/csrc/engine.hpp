// engine.hpp -- host orchestration of the NEP step on one GPU.
//
// NepEngine mirrors the reference NEP object (src/gpumd_compat/nep.cuh):
// parse once, size workspaces, compute(box, ranges, SoA arrays) with a
// Verlet list reused across calls until an atom moved more than skin/2 or the
// box changed (neighbor.cu:537-582).  MdState mirrors run_segment
// (src/runtime.cu:680-849) for one GPU without MPI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "ensemble.hpp"
#include "kernels.hpp"
#include "nep_model.hpp"

struct dmgx_box;

namespace dmgx {

struct Box {
  double h[9];
  int pbc[3];
};

// Per-kernel HIP-event timing on the launch stream (off unless enabled).
enum ProfKernel { kProfFilter = 0, kProfDescriptor, kProfGather, kProfRebuild, kProfCheck, kProfIntegrate,
                  kProfThermo, kProfCount };

class Profiler {
 public:
  ~Profiler();
  // 0 off, 1 events recorded around every kernel group, 2 the fused
  // descriptor kernel only, with events attached to its dispatch
  void enable(int mode) { mode_ = mode; }
  bool on() const { return mode_ != 0; }
  bool attach_mode() const { return mode_ == 2; }
  // mode 1: event pair recorded around a launch (group)
  void begin(int kernel, hipStream_t s);
  void end(hipStream_t s);
  // mode 2: an event pair for the dispatch of `kernel` itself, if it is timed
  bool attach(int kernel, hipEvent_t* start, hipEvent_t* stop);
  // synchronises, folds recorded pairs into the totals and clears them
  void collect();
  void read(int kernel, int* launches, double* total_ms);
  void reset();

 private:
  int mode_ = 0;
  struct Rec {
    int kernel;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs_;
  std::vector<hipEvent_t> pool_;
  size_t used_ = 0;
  int open_ = -1;
  int launches_[kProfCount] = {};
  double total_[kProfCount] = {};
  hipEvent_t take();
};

class NepEngine {
 public:
  explicit NepEngine(const std::string& path);
  ~NepEngine();
  NepEngine(const NepEngine&) = delete;
  NepEngine& operator=(const NepEngine&) = delete;

  const NepModel& model() const { return model_; }
  // workspace for `capacity` slots, of which the first `centres` (default:
  // all) can be centres with rows
  void reserve(int capacity, int centres = -1);
  // sizes the cell grid arrays for at least ncell cells
  void reserve_cells(long long ncell) { ensure_cells((int)std::min<long long>(ncell, 1LL << 28)); }
  // Throws std::invalid_argument / std::runtime_error; see dmgx_nep_compute.
  // `map` (optional) marks the slots past the owned ones as periodic images
  // of owned slots (ImageMap); the serving loop's batches use it.
  void compute(const Box& box, int N, int n1, int n2, int nd1, int nd2, const int* type, const double* pos,
               const uint64_t* gid, double* pe, double* force, double* virial, unsigned flags,
               hipStream_t stream, const ImageMap* map = nullptr);
  // Enqueues one evaluation without any host synchronisation (the Verlet
  // rebuild decision is taken on the device).  settle() then synchronises,
  // raises on invalid input, and returns true when a Verlet row overflowed:
  // the workspace has been grown and the enqueued results must be redone.
  void enqueue(const Box& box, int N, int n1, int n2, int nd1, int nd2, const int* type, const double* pos,
               const uint64_t* gid, double* pe, double* force, double* virial, unsigned flags,
               hipStream_t stream, const ImageMap* map = nullptr);
  bool settle(hipStream_t stream);
  // enqueue() in two halves for callers that pack the step's positions
  // themselves (the MD loop fuses the pack into its integrator kernel):
  // plan() validates and takes the host-side rebuild decision; launch_force()
  // enqueues the rebuild, filter, descriptor and gather kernels.
  struct Plan {
    DevBox box;
    bool forced = false;
    bool empty = true;
  };
  Plan plan(const Box& box, int N, int n1, int n2, int nd1, int nd2, const int* type, const double* pos,
            const uint64_t* gid, double* pe, double* force, double* virial, unsigned flags,
            const ImageMap* map = nullptr);
  void launch_force(const Plan& p, int N, int n1, int n2, int nd1, int nd2, const int* type, const uint64_t* gid,
                    double* pe, double* force, double* virial, unsigned flags, hipStream_t stream);
  DevWork& work() { return w_; }
  const DevWork& work() const { return w_; }
  // What captured launches of this engine depend on: every workspace
  // pointer and size and the tile shapes (a graph is valid while they stay)
  void launch_key(std::vector<long long>& key) const;
  // Serving relayouts: adopts a batch of N slots (centres [0, n_owned)) whose
  // first centres [0, c_keep) had their Verlet rows, and their slots their
  // displacement references, carried over from the previous batch by the
  // caller (continuing requests, translated).  Packs the positions and
  // builds the rows of centres [c_keep, n_owned) only, resetting the
  // references of the slots [c_keep, n_owned) and [g_keep, N); the list is
  // then valid for this box, N and map (the device's skin check still asks
  // for a full rebuild once a carried atom has moved half the skin).  The
  // workspace must already hold N slots (reserve) with unchanged rows.
  void adopt_carried(const Box& box, int N, int n_owned, int c_keep, int g_keep, const int* type, const double* pos,
                     double* pe, double* force, double* virial, const ImageMap& map, hipStream_t stream);
  // Tile shapes picked by autotune(): lanes per centre of the filter (8, 16,
  // 32) and per atom of the gather (4, 8, 16), and the row layout: the
  // capacity of every per-centre row (Verlet, edge, partial-force rows)
  // rounded up to a multiple of 4 (layout 0) or to a power of two (layout
  // 1).  Before tuning the power of two is used where it costs at most 1.25x
  // the rows.  autotune() evaluates the given system once (overwriting
  // pe/force/virial), times every lane candidate of the two kernels on it
  // (one warm-up launch, then the median of seven launches, each between two
  // HIP events), then the filter + descriptor + gather sequence in each row
  // layout (re-allocating and rebuilding the rows for the other one), keeps
  // the fastest and leaves the outputs as an evaluation with the chosen
  // shapes writes them.
  struct Tuning {
    int filter_lanes = 16, gather_lanes = 8;
    double filter_ms[3] = {0, 0, 0}, gather_ms[3] = {0, 0, 0};  // per candidate, in the order above
    int row_layout = -1;                // -1 before tuning (the default rule), else 0 or 1
    int row_capacity = 0;               // entries per row now allocated
    double force_ms[2] = {0, 0};        // filter + descriptor + gather per row layout (0: not timed)
    // 1: the filter of the second half of the centres runs on a side stream
    // while the fused descriptor kernel works on the first half (the
    // memory-bound filter overlaps the VALU-bound descriptor)
    int overlap = 0;
    double overlap_ms[2] = {0, 0};      // filter + descriptor + gather without / with the overlap
    // 1: the rows are split by type pair at rebuild time into candidate rows
    // (every Verlet entry within its pair's rc + skin) and the step has no
    // filter pass: the descriptor and the gather form the distances from the
    // packed positions, and the descriptor writes the step's angular edges
    int split = 0;
    double split_ms[2] = {0, 0};        // the step's force kernels with the filter / with candidate rows
    // lanes per atom of the descriptor: 1, 4 (a group of four lanes shares an
    // atom's edges and hidden neurons, the ANN on the vector ALUs in FP32), or
    // 0 before tuning: four below kLaneGroupCentres centres, else one.  BF16
    // and split-BF16 ANN requests keep one lane per atom.
    int desc_lanes = 0;
    double desc_ms[2] = {0, 0};         // the step's force kernels with 1 / 4 lanes per atom
    bool tuned = false;
  };
  const Tuning& tuning() const { return tune_; }
  // Sets the shapes directly (rejects other values); marks them tuned, so
  // a server does not re-tune them
  void set_tiles(int filter_lanes, int gather_lanes);
  // Selects the per-step filter (0) or the rebuild-time split (1); the next
  // evaluation rebuilds the rows
  void set_split(int split);
  // Sets the descriptor's lanes per atom: 0 (by system size), 1 or 4
  void set_desc_lanes(int lanes);
  // (centres [0, nc); `map` as for compute)
  const Tuning& autotune(const Box& box, int N, int nc, const int* type, const double* pos, double* pe,
                         double* force, double* virial, unsigned flags, hipStream_t stream,
                         const ImageMap* map = nullptr);
  // A periodic direction thinner than 2.5 (rc + skin) is evaluated on a
  // replicated supercell; reps[d] is the replication per box vector.
  bool small_box(const Box& box, int reps[3]) const;
  const int* read_flags(hipStream_t stream);
  void notify_box_changed() { box_changed_ = true; }
  // Box and positions were rescaled isotropically by `scale` (a barostat
  // step; the caller scales both).  The Verlet list is kept: the skin
  // check's references are rescaled on `stream` and its budget shrinks by
  // |1 - accumulated scale| (rc + skin), so a pair beyond rc + skin at the
  // build stays beyond rc while no atom's non-affine displacement exceeds
  // half of what is left; once the scale has used half the skin (or for a
  // replicated small box) the next evaluation rebuilds.
  void notify_box_scaled(double scale, hipStream_t stream);
  bool rescaled_since_build() const { return scale_since_build_ != 1.0; }
  // the displacement since the last Verlet build that requests a rebuild:
  // half the skin, less what isotropic box rescales have used of it
  double half_skin() const;
  // One evaluation whose per-atom workspace (packed positions, cell grid,
  // Verlet rows, edge rows, adjoints, partial forces, status words) is carved
  // out of `ws` (ws_bytes long, 256-byte aligned, device memory) instead of
  // the handle's own allocations: nothing is allocated or freed.  The Verlet
  // list is built afresh (the workspace keeps no list between calls), the
  // handle's own list and workspace are left untouched, and the stream is
  // synchronised to read the status words.  Throws std::invalid_argument if
  // the system needs more than ws_bytes or is a replicated small box, and
  // std::runtime_error ("overflow") if a Verlet row exceeds the capacity the
  // model's MN_radial implies.
  void compute_in(void* ws, size_t ws_bytes, const Box& box, int N, int n1, int n2, int nd1, int nd2,
                  const int* type, const double* pos, const uint64_t* gid, double* pe, double* force,
                  double* virial, unsigned flags, hipStream_t stream);
  // Device bytes of the workspace per slot that can be a centre (with its
  // rows at the current row capacity; before the first allocation, the
  // capacity the model's MN_radial implies), or per slot that cannot
  double bytes_per_slot(bool centre = true) const;
  void stats(int* max_radial, int* max_angular, int* max_verlet, int* rebuilds);
  void totals(long long* sum_radial, long long* sum_angular, long long* sum_verlet, int* centres);
  void copy_verlet_rows(int c1, int c2, int max_nb, int* counts, int* rows);
  void copy_pair_rows(int which, int c1, int c2, int max_nb, int* counts, int* rows);
  Profiler& profiler() { return prof_; }

 private:
  DevBox make_devbox(const Box& box) const;
  void enqueue_small(const Box& box, const int reps[3], int N, const int* type, const double* pos,
                     const uint64_t* gid, double* pe, double* force, double* virial, unsigned flags,
                     hipStream_t stream);
  void ensure_cells(int ncell);
  void allocate(int capacity, int mv, int centres = -1);
  void allocate_split();
  void free_all();

  NepModel model_;
  DevModel dm_{};
  float* params_ = nullptr;
  float* params_pad_ = nullptr;  // descriptor parameters padded to a bounded instance (when used)
  DevModel dmd_{};               // the model as the descriptor kernel sees it
  DevWork w_{};
  int* hflags_ = nullptr;  // pinned host mirror of w_.flags
  bool have_list_ = false, box_changed_ = true, last_gid_sort_ = false;
  int last_N_ = -1, last_nd1_ = -1, last_nd2_ = -1;
  double last_h_[9] = {};
  int last_pbc_[3] = {};
  ImageMap last_map_;
  Tuning tune_;
  int mv_base_ = 0;  // row capacity asked for, before the layout's rounding
  double scale_since_build_ = 1.0;  // isotropic box scale accumulated since the last Verlet build
  int seen_rebuilds_ = 0;           // device rebuild count when last read
  // rebuild rate: evaluations enqueued, and the evaluation and rebuild
  // counts where the rate was last taken; with fewer than one rebuild per
  // 16 evaluations the rebuild kernels run on capped grids
  long long evaluations_ = 0, rate_evaluations_ = 0;
  int rate_rebuilds_ = 0;
  bool rare_rebuilds_ = false;
  bool last_small_ = false;         // the last evaluation ran on a replicated small box
  int padded_rows(int mv, int layout) const;
  void launch_step_kernels(const DevBox& b, int N, int nc, const int* type, double* pe, double* force,
                           double* virial, unsigned flags, hipStream_t s);
  // filter + fused descriptor of centres [c1, c2), sequential or overlapped
  // (tune_.overlap); `close` as launch_filter's; ev0/ev1 attached to the
  // (first) descriptor dispatch when given
  void launch_filter_descriptor(const DevBox& b, int N, int c1, int c2, const int* type, double* pe, bool acc,
                                int ann, bool close, hipStream_t s, bool profile);
  // lanes per atom of the descriptor for n centres in ANN mode ann
  int desc_lanes(int n, int ann) const;
  DevBox last_box_{};  // box of the last evaluation (the filter of copy_pair_rows)
  hipStream_t side_ = nullptr;  // the overlapped filter's stream
  hipEvent_t fork_ = nullptr, join_ = nullptr;
  int last_N_stats_ = 0, last_c1_ = 0, last_c2_ = 0;
  Profiler prof_;
  // replicated small-box system (atoms x replicas)
  long long sb_cap_ = 0;
  double *sb_pos_ = nullptr, *sb_pe_ = nullptr, *sb_f_ = nullptr, *sb_vir_ = nullptr;
  int* sb_type_ = nullptr;
  uint64_t* sb_gid_ = nullptr;
};

class MdState {
 public:
  MdState(NepEngine* nep, const Box& box, int N, const int* type, const double* mass, const double* pos,
          const double* vel);
  ~MdState();
  MdState(const MdState&) = delete;
  MdState& operator=(const MdState&) = delete;
  // Every `every` steps (every > 0) and after the last step, `on_output`
  // gets the number of steps done and that step's thermo as computed before
  // the thermostat rescale (the state itself is after it), the point where
  // the reference writes its thermo rows and dumps (runtime.cu:790-840).
  using Output = std::function<void(int steps_done, const double* thermo)>;
  void run(int steps, double dt_fs, const EnsembleConfig& ens, unsigned flags, double* thermo_out, int every = 0,
           const Output& on_output = Output());
  // Options of a run beyond the ensemble: output as above; an adaptive step,
  // each step's dt being min(dt, max_distance / max |v|) over the velocities
  // at the step's start (runtime.cu:594-618); and `on_correct(step)` called
  // before every step whose index is a multiple of correct_every, with the
  // stream idle (the reference's correct_velocity trigger, runtime.cu:749-756).
  // Either of the last two runs the loop step by step.
  struct RunOptions {
    int every = 0;
    Output on_output;
    double max_distance = 0.0;  // A; 0 keeps the step fixed
    int correct_every = 0;
    std::function<void(int step)> on_correct;
  };
  void run(int steps, double dt_fs, const EnsembleConfig& ens, unsigned flags, double* thermo_out,
           const RunOptions& opt);
  // natural time units advanced by the last run (sum of its steps' dt)
  double elapsed() const { return elapsed_; }
  // Tracks unwrapped positions from now on (starting from the current
  // positions; a second call keeps the running values): every drift is also
  // added to them (AtomsState::enable_unwrapped, update_unwrapped,
  // runtime.cu:80-90,423-440).  unwrapped() is null until enabled.
  void enable_unwrapped();
  // NepEngine::autotune on this system's current positions (the forces are
  // those positions' forces afterwards)
  const NepEngine::Tuning& autotune(unsigned flags);
  const double* unwrapped() const { return unwrapped_; }
  // Stores the atoms in cell order (cells of half the Verlet range in Morton
  // order, input order kept within a cell), once, so that neighbours sit near each
  // other in memory. Every per-atom array is permuted alike, forces included;
  // order()[i] is the index, in the arrays given at construction, of the atom
  // now stored at i (composed over repeated calls).
  void reorder();
  const std::vector<int>& order() const { return order_; }
  // the device-resident loop replays one captured step (off by default)
  void set_step_graph(bool on)
  {
    step_graph_on_ = on;
    if (!on) drop_step_graph();
  }
  long long force_calls() const { return force_calls_; }
  void arrays(double** pos, double** vel, double** force, double** pe, double** vir, double** mass, int** type);
  const Box& box() const { return box_; }

 private:
  void force(unsigned flags);
  void wrap();
  void vv(int first_half, double dt);
  void enqueue_thermo();
  void thermo(double* out);
  void checkpoint(bool save);
  NepEngine* nep_;
  Box box_;
  int N_;
  hipStream_t stream_ = nullptr;
  double *pos_ = nullptr, *vel_ = nullptr, *force_ = nullptr, *pe_ = nullptr, *vir_ = nullptr, *mass_ = nullptr;
  int* type_ = nullptr;
  DevWork tw_{};  // thermo scratch only
  bool have_force_ = false;
  double *ck_pos_ = nullptr, *ck_vel_ = nullptr, *ck_force_ = nullptr, *ck_pe_ = nullptr, *ck_vir_ = nullptr;
  double* rand_ = nullptr;    // per-step thermostat draws, then target temperatures, of a chunk (device)
  double* hrand_ = nullptr;   // pinned staging of the draws and targets
  int* kstep_ = nullptr;      // device: the step of the chunk the thermostat launch takes
  // small systems: one step's launches captured and replayed per step
  bool step_graph_on_ = false;
  hipGraph_t step_graph_ = nullptr;
  hipGraphExec_t step_exec_ = nullptr;
  std::vector<long long> step_key_;
  void drop_step_graph();
  // captures (or keeps, while nothing it depends on changed) one step of the
  // device-resident loop: first half, force, second half, thermostat
  void step_graph(const NepEngine::Plan& plan, double dt, int kind, double coupling, double volume, unsigned flags,
                  bool thermo_on);
  double* factor_ = nullptr;  // device velocity factor
  long long force_calls_ = 0;
  unsigned long long* vmax2_ = nullptr;  // device: bits of the largest |v|^2 (adaptive step)
  double* unwrapped_ = nullptr;          // device [3N] when enabled
  double* ck_unwrapped_ = nullptr;
  double elapsed_ = 0.0;
  std::vector<int> order_;  // stored index -> construction index
};

double box_volume(const double* h);
void box_inverse(const double* h, double* hi);

}  // namespace dmgx
