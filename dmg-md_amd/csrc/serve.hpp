// serve.hpp -- the serving loop: many independent periodic systems (requests)
// evaluated together on one GPU with continuous batching.
//
// No reference counterpart: the reference runs one system per process.  A
// request is a periodic box with N atoms, evaluated with one of the server's
// models, either as a single point (steps = 0: energy, forces, virial) or as
// `steps` NVE velocity-Verlet steps.  At submit the request gets a shell of
// the periodic images of its atoms that lie within rc + skin + margin of its
// box.  The active requests of one model are placed in disjoint, translated
// regions of one non-periodic batch box (gaps wider than the Verlet radius),
// laid out as
//     [owned atoms of every request | their image ghosts, grouped by owner]
// and the batch is evaluated with an ImageMap: owned atoms are the only
// force and descriptor centres, and a ghost in a row stands for its owner
// (the owner's descriptor, radial adjoint and angular partials are the
// image's by periodicity), so each request gets exactly its periodic result
// while ghosts cost positions only.  Ghost positions are refreshed from their
// owners (+ box shift) every step.  The shell stays exact while no atom is
// more than margin / 2 from where it was when the shell was built: after each
// half-kick the device checks that, and an MD request past it gets a new
// shell (its atoms wrapped into its box, its images recomputed, velocities
// and forces carried) before that step's force evaluation, together with the
// MD requests past a quarter of the margin.
//
// Each model has its own batch ("lane": layout, engine workspace, stream and
// captured force graph).  A worker thread owns the GPU work: each iteration
// admits waiting requests in submission order while the device bytes of the
// requests in flight stay within the HBM budget, re-lays out the lanes whose
// membership changed, enqueues every lane's frame (one MD step, or the first
// force evaluation) on its own stream so the lanes overlap on the device,
// then settles each lane and retires the finished requests, recording
// submit-to-result latency.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace dmgx {

struct ServeConfig {
  long long max_atoms = 4000000;  // owned + ghost slots in flight, all models
  int max_requests = 4096;
  unsigned flags = 0;             // force flags (ANN mode)
  double margin = 2.0;            // extra shell thickness (A) for an MD request's motion between new shells
  long long hbm_bytes = 0;        // device bytes of the requests in flight; 0: half the free memory at start
  long long reserve_atoms = 1LL << 21;  // engine workspace sized up front per model (capped by the budget)
  bool use_graph = true;          // replay the force launches of a layout as one hipGraph (one model only)
};

struct ServeStats {
  long long submitted = 0, completed = 0, frames = 0, atom_frames = 0;
  double p50_ms = 0, p99_ms = 0, mean_ms = 0, max_ms = 0, elapsed_s = 0, frames_per_s = 0;
  long long iterations = 0, relayouts = 0;
  long long max_slots = 0;
  double relayout_s = 0, iterate_s = 0, retire_s = 0;
  long long hbm_budget = 0, max_bytes_in_flight = 0;
  long long reshells = 0;  // MD requests given a new image shell after drifting
  long long rows_carried = 0;  // relayouts whose continuing requests kept their Verlet rows
  double reshell_s = 0;        // worker wall time building new image shells
};

// The Server owns its NepEngines' GPU work while it lives: the engines must
// not be used by other threads meanwhile.  Requests overlapping on the GPU
// are the point; the results of each are those of its periodic system alone.
class Server {
 public:
  Server(const std::vector<NepEngine*>& models, const ServeConfig& cfg);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // host arrays are copied; position/velocity SoA [3*N]; velocity may be null
  long long submit(int model, const Box& box, int N, const int* type, const double* mass, const double* pos,
                   const double* vel, int steps, double dt_fs);
  // true when done; then copies results (each may be null): potential[N],
  // force[3N], position[3N] (wrapped into the request's box), velocity[3N],
  // virial[9N]
  bool poll(long long id, bool wait, double* potential, double* force, double* position, double* velocity,
            double* virial, double* latency_ms);
  ServeStats stats();
  // zeroes the counters, latencies and time window (a drained server only):
  // a warmed server's buffers and pools are kept for the next measurement
  void reset_stats();
  // blocks until every submitted request is done
  void drain();
  // device bytes one slot (owned atom or image ghost) of a model's batch
  // takes: engine workspace (rows of the model's capacity) and the layout's
  // double-buffered per-slot arrays
  double slot_bytes(int model, bool owned) const;

 private:
  struct Request;
  struct Lane;
  // one thread per lane: its frames, new shells, relayouts and retirements,
  // so that one lane's host work overlaps the others' device work
  void lane_worker(Lane& L);
  void admit();
  // the lane's tile shapes, once, on its first batch large enough to tell
  // them apart, with the device to itself
  bool needs_tuning(Lane& L) const;
  void tune(Lane& L);
  void relayout(Lane& L);
  void upload_md_flags(Lane& L);
  void enqueue_kick(Lane& L);
  void enqueue_force(Lane& L);
  void finish_frame(Lane& L);
  void retire(Lane& L);
  // the request's image shell and upload pack from its positions (wrapped
  // into its box here), velocities and forces, SoA [3n]; sets its ghost
  // count, extent and device bytes (device bytes above `limit` throw)
  void make_pack(Request& r, const int* type, const double* mass, const double* pos, const double* vel,
                 const double* force, long long limit);
  // after the frame's half-kick (the lane's stream idle): whether an MD
  // request of the lane has an atom past the image shell's budget
  bool needs_shells(Lane& L);
  // new shells for those requests and the other MD requests past a quarter
  // of the margin, from their current state (a new layout places them)
  void reshell(Lane& L);

  ServeConfig cfg_;
  std::vector<std::unique_ptr<Lane>> lanes_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  // lanes' iterations hold it shared; tuning and captures hold it
  // exclusively, entering through the turnstile gpu_turn_
  std::shared_mutex gpu_mu_;
  std::mutex gpu_turn_;
  std::shared_lock<std::shared_mutex> gpu_shared();
  std::unique_lock<std::shared_mutex> gpu_exclusive();
  long long admit_gen_ = 0;  // bumped when admission may make progress (a submit, a completion)
  bool stop_ = false;
  std::string error_;  // first worker failure, reported by poll/drain/stats
  long long next_id_ = 1;
  long long budget_ = 0;  // HBM bytes of requests in flight
  std::deque<std::shared_ptr<Request>> waiting_;
  std::map<long long, std::shared_ptr<Request>> all_;
  double t_first_ = -1, t_last_ = 0;
  std::vector<double> latencies_;
  ServeStats st_;

  // device buffer with its capacity (elements)
  template <class T>
  struct DBuf {
    T* p = nullptr;
    long long cap = 0;
    void need(long long n, const char* what);
    void release();
  };

  hipStream_t upload_ = nullptr;  // request uploads (submitting threads)
  std::mutex pack_mu_;
  std::multimap<size_t, void*> pack_free_;  // reusable device pack buffers by capacity
  std::multimap<size_t, void*> pin_free_;   // reusable pinned host buffers by capacity
  std::vector<void*> pack_slabs_, pin_slabs_;  // the allocations the buffers are cut from
  void* pool_get(std::multimap<size_t, void*>& pool, std::vector<void*>& slabs, bool pinned, size_t bytes,
                 size_t* cap);
  void* pack_get(size_t bytes, size_t* cap);
  void pack_put(void* p, size_t cap);
  void* pin_get(size_t bytes, size_t* cap);
  void pin_put(void* p, size_t cap);
  void release_upload(Request& r);
  // a pinned block of finished requests' results, shared by those requests
  // until they are dropped with the server
  struct ResultBlock {
    Server* srv = nullptr;
    void* p = nullptr;
    size_t cap = 0;
    ~ResultBlock() { srv->pin_put(p, cap); }
  };
  void capture_force(Lane& L, unsigned flags);
  void drop_graph(Lane& L);
};

}  // namespace dmgx
