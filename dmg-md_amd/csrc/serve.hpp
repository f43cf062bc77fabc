// serve.hpp -- the serving loop: many independent periodic systems (requests)
// evaluated together on one GPU with continuous batching.
//
// No reference counterpart: the reference runs one system per process.  A
// request is a periodic box with N atoms and either a single-point force
// evaluation (steps = 0) or `steps` NVE velocity-Verlet steps.  At submit
// the request gets a shell of the periodic images of its atoms that lie
// within rc + skin + margin of its box.  Active requests are placed in
// disjoint, translated regions of one non-periodic batch box (gaps wider than
// the Verlet radius), laid out as
//     [owned atoms of every request | their image ghosts, grouped by owner]
// and the batch is evaluated with an ImageMap: owned atoms are the only
// force and descriptor centres, and a ghost in a row stands for its owner
// (the owner's descriptor, radial adjoint and angular partials are the
// image's by periodicity), so each request gets exactly its periodic result
// while ghosts cost positions only.  Ghost positions are refreshed from their
// owners (+ box shift) every step.
//
// A worker thread owns the GPU work: each iteration admits waiting requests
// (FIFO) while the batch stays within the HBM budget (atoms incl. ghosts),
// re-lays the batch out when its membership changed, advances every active
// request by one frame (its first force evaluation, or one MD step), and
// retires the finished ones, recording submit-to-result latency.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"

namespace dmgx {

struct ServeConfig {
  long long max_atoms = 4000000;  // owned + ghost atoms in flight
  int max_requests = 4096;
  unsigned flags = 0;             // force flags (ANN mode)
  double margin = 1.0;            // extra shell thickness (A) for motion during a request
  long long reserve_atoms = 1LL << 21;  // engine workspace sized up front (capped by max_atoms)
  bool use_graph = true;          // replay the force launches of a layout as one hipGraph
};

struct ServeStats {
  long long submitted = 0, completed = 0, frames = 0, atom_frames = 0;
  double p50_ms = 0, p99_ms = 0, mean_ms = 0, max_ms = 0, elapsed_s = 0, frames_per_s = 0;
  long long iterations = 0, relayouts = 0;
  long long max_slots = 0;
  double relayout_s = 0, iterate_s = 0, retire_s = 0;
};

// The Server owns the NepEngine's GPU work while it lives: the engine must not
// be used by other threads meanwhile.  Requests overlapping on the GPU are the
// point; the results of each are those of its periodic system alone.
class Server {
 public:
  Server(NepEngine* nep, const ServeConfig& cfg);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // host arrays are copied; position/velocity SoA [3*N]; velocity may be null
  long long submit(const Box& box, int N, const int* type, const double* mass, const double* pos,
                   const double* vel, int steps, double dt_fs);
  // true when done; then copies results (each may be null): potential[N],
  // force[3N], position[3N] (wrapped into the request's box), velocity[3N]
  bool poll(long long id, bool wait, double* potential, double* force, double* position, double* velocity,
            double* latency_ms);
  ServeStats stats();
  // blocks until every submitted request is done
  void drain();

 private:
  struct Request;
  void worker();
  bool admit();
  void relayout();
  void iterate();
  void retire();

  NepEngine* nep_;
  ServeConfig cfg_;
  hipStream_t stream_ = nullptr;
  std::thread thread_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  bool stop_ = false;
  std::string error_;  // first worker failure, reported by poll/drain/stats
  long long next_id_ = 1;
  std::deque<std::shared_ptr<Request>> waiting_;
  std::map<long long, std::shared_ptr<Request>> all_;
  std::vector<std::shared_ptr<Request>> active_;  // worker only
  bool dirty_ = true;                               // membership changed since the last layout
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

  // batch layout (worker only)
  int n_owned_ = 0, n_ghost_ = 0, n_total_ = 0;
  Box batch_box_{};
  DBuf<double> pos_, vel_, force_, pe_, vir_, mass_;
  DBuf<int> type_;
  DBuf<int> owner_;        // slot -> owned slot it is an image of (itself for owned slots)
  DBuf<int> image_;        // slot -> image code (image_code(0, 0, 0) for owned slots)
  DBuf<int> gstart_;       // owned slot -> first ghost slot of its images [n_owned + 1]
  DBuf<double> g_shift_;   // ghost k -> box shift [3]
  // the next layout's per-slot arrays, assembled on the device and swapped in
  DBuf<double> pos2_, vel2_, force2_, mass2_, g_shift2_;
  DBuf<int> type2_, owner2_, image2_, gstart2_;
  DBuf<int> atom_req_;     // owned slot -> active request index
  DBuf<int> req_md_;       // active request -> 1 when it takes an MD step this iteration
  DBuf<double> req_dt_;    // active request -> dt (natural units)
  DBuf<ServeSlots> slots_; // layout table (one row per active request)
  DBuf<ServeOut> outq_;    // finished requests (extract table)
  DBuf<double> out_;       // finished requests' result blocks (10 doubles per atom)
  hipStream_t upload_ = nullptr;  // request uploads (submitting threads)
  std::mutex pack_mu_;
  std::multimap<size_t, void*> pack_free_;  // reusable device pack buffers by capacity
  std::multimap<size_t, void*> pin_free_;   // reusable pinned host buffers by capacity
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
  char* h_stage_ = nullptr;  // pinned staging for uploads and results
  size_t h_stage_bytes_ = 0;
  void stage(size_t bytes);
  bool forced_rebuild_ = true;
  // the force launches of the current layout (ghost refresh, pack, Verlet
  // rebuild, filter, descriptor, gather) captured once and replayed
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  void capture_force(unsigned flags);
  void drop_graph();
};

}  // namespace dmgx
