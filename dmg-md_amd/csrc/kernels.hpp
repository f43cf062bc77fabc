// kernels.hpp -- device-side structures and launch wrappers of the NEP step.
//
// All kernels live in kernels.hip (one translation unit, so the __constant__
// angular tables are shared); host code calls the launch_* wrappers below.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "nep_model.hpp"

namespace dmgx {

constexpr int kBlock = 256;  // threads per workgroup for per-atom kernels
constexpr double kSkin = 1.0;  // Verlet skin (neighbor.cuh:259)

// Box with its inverse, as the reference's Box (box.cuh:53-70) plus the
// cell grid of the Verlet build.
struct DevBox {
  double h[9];
  double hi[9];
  double thick[3];
  int pbc[3];
  int ortho;
  int nc[3];   // cells per direction
};

// Model parameters as kernel arguments; tables live in device memory.
struct DevModel {
  int T, nR, nA, kR, kA, L, has222, has1111, numL, D, H, version;
  int kRp;  // kR rounded up to a multiple of 4 (row pitch of the radial adjoint)
  int zbl;
  double rcR[kMaxTypes], rcA[kMaxTypes];  // FP64 per-type cutoffs (predicate)
  const float* w0;    // [T][H][D]
  const float* b0;    // [T][H]
  const float* w1;    // [T][H]
  const float* w1x;   // [T] NEP5 typewise output bias
  const float* b1;    // [1]
  const float* cR;    // [T][T][nR][kR]
  const float* cA;    // [T][T][nA][kA]
  const float* qs;    // [D]
  const float* zblp;  // [T][T][12] a_inv, zizj, r_in, r_out, (a_k, b_k) x4
};

// Periodic-image map of a batch whose ghost slots are images of owned slots
// (the serving loop's periodic requests).  owner[j] is the owned slot that j
// is an image of (owner[j] == j for an owned slot), image[j] the lattice shift
// (a, b, c) of that image as image_code(a, b, c), and the ghosts of owned slot
// i occupy slots [gstart[i], gstart[i+1]), in increasing image code.  With a
// map, a row entry j stands
// for its owner everywhere except for its position: the owner's descriptor,
// radial adjoint and angular partials are the image's, so ghosts need no rows
// and no descriptors of their own.
struct ImageMap {
  const int* owner = nullptr;   // [N]
  const int* image = nullptr;   // [N]
  const int* gstart = nullptr;  // [owned + 1], absolute slots
  // Optional request map of a batch of disjoint systems (every row holds
  // only slots of its own system): the request of each owned slot and one
  // "moved" word per request.  With it, a rebuild not requested for every
  // centre remakes only the rows (and displacement references) of the
  // requests with a slot that moved more than skin/2.
  const int* slot_req = nullptr;  // [n_owned]
  int* req_moved = nullptr;       // [n_req]
  int n_owned = 0, n_req = 0;
  // Optional with the request map: each request's slots (first owned slot,
  // owned count, first ghost slot, ghost count) and the buffers a rebuild
  // compacts its work into, so that it visits only the requests it remakes:
  // req_off [2 n_req + 2] (each request's offset in the centre and slot
  // lists, then the two totals) and lists [n_owned + N] (centres, then slots)
  const int4* req_range = nullptr;
  int* req_off = nullptr;
  int* lists = nullptr;
};
// whether a rebuild works from compacted lists (host launches and kernels
// decide alike)
__host__ __device__ inline bool rebuild_listed(const ImageMap& m)
{
  return m.slot_req && m.req_moved && m.req_range && m.req_off && m.lists && m.n_req > 0;
}
// flags[kFlagRebuild] bits: a rebuild is requested; it covers every centre
constexpr int kRebuildAsked = 1, kRebuildAll = 2;
// whether slot i's rows (or reference position) are remade by the rebuild
// whose request word is go
__device__ __forceinline__ bool slot_rebuilt(const ImageMap& m, int go, int i)
{
  if (!m.slot_req || (go & kRebuildAll)) return true;
  return m.req_moved[m.slot_req[i < m.n_owned ? i : m.owner[i]]] != 0;
}
constexpr int kImageMax = 15;  // |a|, |b|, |c| <= kImageMax
__host__ __device__ inline int image_code(int a, int b, int c)
{
  return (a + 16) | ((b + 16) << 5) | ((c + 16) << 10);
}
// the code of the opposite shift (-a, -b, -c)
__host__ __device__ inline int image_neg(int code)
{
  return (32 - (code & 31)) | ((32 - ((code >> 5) & 31)) << 5) | ((32 - ((code >> 10) & 31)) << 10);
}

// Per-atom / per-edge workspaces (device pointers), sized for `cap` slots;
// the per-centre arrays (rows, row lengths, radial adjoints) for the first
// `rcap` slots, the ones that can be centres.  Row arrays are row-major:
// entry p of centre i lives at [i*M + p].
// (Every field is part of NepEngine::launch_key: a captured launch takes
// the struct by value.)
struct DevWork {
  int cap = 0;
  int rcap = 0;
  int MV = 0, MR = 0, MA = 0;  // row capacities: Verlet, radial, angular
  int ncell_cap = 0;
  double4* xq = nullptr;       // [cap] per-step (x, y, z, type bits)
  int* cell_of = nullptr;      // [cap]
  int* cell_rank = nullptr;    // [cap] rank of the atom inside its cell
  int* cell_count = nullptr;   // [ncell] zero between rebuilds
  int* cell_start = nullptr;   // [ncell+1]
  int* cell_atoms = nullptr;   // [cap]
  double4* xs = nullptr;       // [cap] xq in cell_atoms order (rebuild only)
  int* cell_tiles = nullptr;   // [2 (ncell/4096 + 2) + 8] tile sums, then (16-byte aligned) offsets, of the multi-block scan
  int* vn = nullptr;           // [cap] Verlet row length
  int* vl = nullptr;           // [cap*MV] Verlet rows
  int* vrev = nullptr;         // [cap*MV] position of the centre in the neighbour's row
  double* x0 = nullptr;        // [3*cap] positions at the last rebuild
  int* nr = nullptr;           // [cap] radial row length
  int* na = nullptr;           // [cap] angular row length
  float4* edge = nullptr;      // [cap*MR] (rx, ry, rz, bits(j)): angular edges at
                               // 0..na-1, radial-only edges at MR-1 downward
  int* eq = nullptr;           // [cap*MA] of an angular edge i->j: j*MV + (position of i in
                               // j's Verlet row), whose angular bit gives i's slot in j's
                               // angular row; -1 if j has no row
  int* vsa = nullptr;          // [cap*MV] with candidate rows only: candidate slot -> angular slot, -1 if none
  unsigned* amask = nullptr;   // [cap*ceil(MV/32)] per-step filter: bit p of row i set when Verlet
                               // entry p is angular (its angular slot: the set bits below p)
  float* bR = nullptr;         // [cap*T*kRp] radial adjoint in basis space
  float4* f12 = nullptr;       // [cap*MA] dE_i/dr_ij of angular edges
  // Rows split by type pair at rebuild time (split != 0): the step has no
  // filter pass; the descriptor and the gather read candidate rows, form the
  // FP64 distances themselves, and the descriptor writes the step's angular
  // edges (edge, eq, na), its slot status (vsa, indexed by candidate slot)
  // and the radial count nr.
  int split = 0;
  int* cl = nullptr;           // [cap*MV] candidate slots: every Verlet entry within its pair's
                               // rc + skin, angular candidates (rcA + skin) at 0..nac-1,
                               // radial-only ones at MV-1 downward
  int* cpos = nullptr;         // [cap*MV] Verlet position -> angular candidate slot, -1 if none
  int* cq = nullptr;           // [cap*MA] of an angular candidate i->j: j*MA + (i's angular
                               // candidate slot in j's row), an index into vsa; -1 if none
  int* nrc = nullptr;          // [cap] candidates
  int* nac = nullptr;          // [cap] angular candidates
  ImageMap map;                // periodic-image map of the current call (empty: none)
  int* flags = nullptr;        // [kNumFlags] device status words
  double* partial = nullptr;   // thermo block partials
  double* thermo = nullptr;    // [8]
};

// Status words in DevWork::flags
enum : int {
  kFlagVerletOverflow = 0,  // max Verlet row length seen
  kFlagRadialOverflow = 1,  // max radial row length seen
  kFlagAngularOverflow = 2, // max angular row length seen
  kFlagMoved = 3,           // an atom moved more than skin/2
  kFlagOutOfBox = 4,        // a position lies outside the box in a non-periodic direction
  kFlagNonFinite = 5,       // non-finite scaled position
  kFlagBadType = 6,         // a type index outside [0, num_types)
  kFlagRebuild = 7,         // a Verlet rebuild is requested for this step
  kFlagRebuildCount = 8,    // rebuilds done since the flags were allocated
  kNumFlags = 16,
};

void upload_angular_constants(const AngularConstants& a);

// Neighbour build (neighbor.hip).  launch_pack packs (x, y, z, type) for
// the step, validates types and, unless `forced`, sets flags[kFlagRebuild]
// when an atom moved more than skin/2 (and, with a request map, marks the
// atom's request); `forced` asks for every centre unconditionally.
// launch_rebuild enqueues the whole cell list + Verlet build; every kernel of
// it returns at once unless flags[kFlagRebuild] is set.  The request is
// closed (and counted) by the filter kernel that follows.
hipError_t launch_request_rebuild(DevWork& w, hipStream_t s);
// p[0..n) = 0 with a kernel (no memset node, so callers may capture it)
hipError_t launch_clear_ints(int* p, long long n, hipStream_t s);
// half_skin: the displacement since the last build that requests a rebuild
// (default half the skin)
hipError_t launch_pack(const DevBox& b, int N, int T, const double* pos, const int* type, bool forced, DevWork& w,
                       hipStream_t s, double half_skin = -1.0);
// x0_ranges (optional): {a1, b1, a2, b2}, the slots whose displacement
// references this rebuild resets (default every slot); the others keep
// theirs (rows carried over from an earlier build)
// capped: grids of at most 2,048 workgroups striding over the centres and
// slots, for callers that rebuild rarely (an empty launch then costs about a
// microsecond; a rebuild takes about 1.3x as long as on full grids); a
// rebuild from compacted lists always uses capped grids
hipError_t launch_rebuild(const DevBox& b, int N, int c1, int c2, double rv2, const uint64_t* gid, DevWork& w,
                          hipStream_t s, const int* x0_ranges = nullptr, bool capped = false);

// Force evaluation (per step)
// lanes: lanes per centre, 8, 16 (default) or 32
// close: this launch closes the step's rebuild request (exactly one filter
// launch of an evaluation does)
hipError_t launch_filter(const DevBox& b, const DevModel& m, int c1, int c2, DevWork& w, hipStream_t s,
                         int lanes = 16, bool close = true);
// Rebuild-time split of the Verlet rows of centres [c1, c2) into candidate
// rows (w.split mode): both kernels return at once unless a rebuild was
// requested this step; the descriptor launch closes the request.
hipError_t launch_split(const DevBox& b, const DevModel& m, int c1, int c2, DevWork& w, hipStream_t s,
                        int lanes = 16);
// Sizes (nR, kR, nA, kA) of the exactly instantiated descriptor that a model
// without an exact instance of its own is padded to (zero coefficients and
// ANN weights on the added entries, 3-body order 4 with both many-body
// blocks); false where the model has an exact instance, or exceeds every
// padded size and runs on the generic instance.
bool descriptor_bucket(const DevModel& m, int dims[4]);
// ann: 0 FP32 on the vector ALUs, 1 FP32 on the matrix cores, 2 BF16 operands,
// 3 split-BF16 operands (hi + lo, three MFMAs per product)
// on the matrix cores; modes 1 and 2 apply to models of at most 2 types with
// at most 64 neurons (every size but the generic fallback), the vector path
// otherwise.
// start/stop (optional) are recorded with the kernel's own dispatch
// With w.split the descriptor forms its edges from the candidate rows (b is
// the box of the step) and closes the step's rebuild request.  lanes: lanes
// per atom, 1 or 4 (four: every model on the vector-ALU ANN in FP32, with the
// filter's rows; for systems too small to fill the GPU one lane per atom).
hipError_t launch_descriptor(const DevBox& b, const DevModel& m, int N, int c1, int c2, const int* type, double* pe,
                             bool accumulate, int ann, DevWork& w, hipStream_t s, hipEvent_t start = nullptr,
                             hipEvent_t stop = nullptr, int lanes = 1);
// lanes: lanes per atom, 4, 8 (default) or 16
hipError_t launch_gather(const DevBox& b, const DevModel& m, int N, int n1, int n2, int nd1, int nd2,
                         const int* type, double* force, double* virial, bool accumulate, DevWork& w, hipStream_t s,
                         int lanes = 8);

// Integrator / thermo
// first half: kick + drift (and the drift added to `unwrapped` when given)
hipError_t launch_velocity_verlet(int first_half, int N, double dt, const double* mass, double* pos,
                                  double* vel, const double* force, hipStream_t s, double* unwrapped = nullptr);
hipError_t launch_wrap(const DevBox& b, int N, double* pos, hipStream_t s);
hipError_t launch_thermo(int N, double volume, const double* mass, const double* pe, const double* vel,
                         const double* virial, DevWork& w, hipStream_t s);
hipError_t launch_scale(int n, double factor, double* data, hipStream_t s);
// out[0] = bit pattern of max_i |v_i|^2 (velocities SoA [3N])
hipError_t launch_max_speed2(int N, const double* vel, unsigned long long* out, hipStream_t s);
// Device-side velocity factor from the thermo of this step (no host read):
// kind 1 Berendsen sqrt(1 + c (T_t/T - 1)); kind 2 Bussi with the step's
// pre-drawn Gaussian rr and chi-square sum R (the same expression and
// rounding as the host resamplekin).  Writes factor[0] and rescales v.
hipError_t launch_ensemble_scale(int kind, int N, double coupling, double t_target, const double* rand2,
                                 double* factor, const double* thermo, double* vel, hipStream_t s);
hipError_t launch_check_scaled(int n, double factor, const double* data, int* flag, hipStream_t s);

// Fused MD-loop kernels.  launch_md_first: v *= factor[0] (if factor), half
// kick, drift, wrap, pack (x, y, z, type) into w.xq and the displacement test
// (or a forced rebuild request).  launch_md_second: half kick, plus the thermo
// partial sums into tw.partial when `thermo`.  launch_thermo_final: tw.thermo
// from the partials and, when kind != 0, the velocity factor of that
// thermostat into factor[0] (see launch_ensemble_scale for kind).
hipError_t launch_md_first(const DevBox& b, int N, int T, double dt, const double* mass, double* pos, double* vel,
                           const double* force, const double* factor, const int* type, bool forced, DevWork& w,
                           hipStream_t s, double* unwrapped = nullptr, double half_skin = -1.0);
hipError_t launch_md_second(int N, double dt, const double* mass, double* vel, const double* force, const double* pe,
                            const double* virial, bool thermo, DevWork& tw, hipStream_t s);
// With step_index (device), the step's target and draws are targets[k] and
// rand2[2k..] for k = *step_index, advanced by the kernel (t_target unused).
hipError_t launch_thermo_final(int N, double volume, int kind, double coupling, double t_target,
                               const double* rand2, double* factor, DevWork& tw, hipStream_t s,
                               const double* targets = nullptr, int* step_index = nullptr);
// factor[0] = 1 and *step_index = 0: a chunk's first step applies no
// pending factor and takes the chunk's first target and draws
hipError_t launch_chunk_start(double* factor, int* step_index, hipStream_t s);
hipError_t launch_scale_by(int n, const double* factor, double* data, hipStream_t s);

// Small boxes: replicate N atoms into reps[0]*reps[1]*reps[2] periodic copies
// (copy k = ia + r0*(ib + r1*ic) at slots k*N..k*N+N-1, shifted by
// ia*a + ib*b + ic*c), and fold the first N results of the copies back.
hipError_t launch_replicate(int N, const int reps[3], const DevBox& b, const double* pos, const int* type,
                            const uint64_t* gid, double* pos2, int* type2, uint64_t* gid2, hipStream_t s);
hipError_t launch_fold(int N, long long n2, bool accumulate, const double* pe2, const double* f2, const double* v2,
                       double* pe, double* force, double* virial, hipStream_t s);

// Serving loop (arrays of the batch have stride Nt = owned + ghost slots):
// masked half kick of owned atoms whose request takes an MD step; ghost
// positions = owner + shift; assembly of a new layout; packing of finished
// atoms' (position - shift, velocity, force, potential).
hipError_t launch_serve_kick(int first_half, int n, int Nt, const int* atom_req, const int* req_md,
                             const double* req_dt, const double* mass, double* pos, double* vel, const double* force,
                             hipStream_t s);
hipError_t launch_serve_ghosts(int n_owned, int n_ghost, int Nt, const int* owner, const double* shift, double* pos,
                               hipStream_t s);

// One request of a new serving layout: where its owned atoms and image
// ghosts go, and where they come from: the previous layout for a continuing
// request, or the request's own upload (a "pack") for a newly admitted one.
// A pack holds, for n atoms and ng ghosts, the ints type[n], first ghost
// (request-local)[n], ghost owner (request-local)[ng], image code[ng], then
// the doubles mass[n], position[3n] and velocity[3n] (per atom x, y, z, in
// the request's own frame) and ghost shift[3ng].
struct ServeSlots {
  int new_o, new_g;   // first owned / ghost slot in the new layout
  int old_o, old_g;   // first owned / ghost slot in the previous layout, -1 if new
  int n, ng;          // owned atoms, ghosts
  const int* pi;      // a new request's pack: ints
  const double* pd;   // a new request's pack: doubles
  double dtr[3];      // translation added to the positions
};
// pack layout: ints (types [n], first ghost of each atom [n], ghost owners
// [ng], ghost image codes [ng]) then, at an 8-byte boundary, doubles (masses
// [n], positions [3n], velocities [3n], forces [3n], ghost shifts [3ng], the
// per-atom ones atom-major)
inline size_t serve_pack_ints(int n, int ng) { return 2 * (size_t)n + 2 * (size_t)ng; }
inline size_t serve_pack_doubles(int n, int ng) { return 10 * (size_t)n + 3 * (size_t)ng; }
inline size_t serve_pack_double_offset(int n, int ng) { return (serve_pack_ints(n, ng) * 4 + 7) & ~size_t(7); }
inline size_t serve_pack_bytes(int n, int ng)
{
  return serve_pack_double_offset(n, ng) + 8 * serve_pack_doubles(n, ng);
}
// The per-slot arrays of one layout.  gshift is per ghost ([3*(slot - n_owned)]).
struct ServeLayout {
  int Nt = 0, n_owned = 0;
  double *pos = nullptr, *vel = nullptr, *force = nullptr, *mass = nullptr, *gshift = nullptr;
  double* home = nullptr;  // owned slots' positions where their request's image shell was built [3*n_owned]
  int *type = nullptr, *owner = nullptr, *image = nullptr, *gstart = nullptr;
};
// Writes every slot of `to` (R requests in slot order; ghost positions are
// left to launch_serve_ghosts) and atom_req[owned slot] = request index.
hipError_t launch_serve_assemble(int R, const ServeSlots* req, const ServeLayout& from, const ServeLayout& to,
                                 int* atom_req, hipStream_t s);
// req_flag[request] raised to 1 for every request with an owned atom farther
// than sqrt(soft2) from its home position, to 2 beyond sqrt(hard2) (flags are
// only ever raised here)
hipError_t launch_serve_drift(int n_owned, int Nt, const int* atom_req, const double* pos, const double* home,
                              double soft2, double hard2, int* req_flag, hipStream_t s);
// The Verlet rows of a relayout's continuing requests (the first R_keep
// rows of the layout table) moved to their new slots: vl, vrev and vn of
// their owned centres from the copies vl0, vrev0, vn0 (the previous
// layout's, old_no owned slots), every entry mapped through its request's
// owned and ghost slot offsets; x0 of all their slots from x0_0 (stride
// old_nt) plus the translation, in w.x0 (stride new_nt).
// Their owned slots are [0, o_keep), their ghosts [new_no, g_keep).
hipError_t launch_serve_carry_rows(int R_keep, const ServeSlots* req, int old_no, int old_nt, int new_no, int new_nt,
                                   int o_keep, int g_keep, const int* vl0, const int* vrev0, const int* vn0,
                                   const double* x0_0, DevWork& w, hipStream_t s);
// One finished request for launch_serve_extract: its owned slots, the
// offset of its block in the output, its placement translation and its box.
constexpr int kServeOutDoubles = 19;  // result doubles per atom
struct ServeOut {
  int slot0, n;
  long long out;  // first double of its block of kServeOutDoubles n
  int pbc[3];
  double trans[3];
  double h[9], hi[9];
};
// Writes, per finished request, the block [potential n | force 3n | position
// 3n | velocity 3n | virial 9n] (SoA, position wrapped into the request's own
// box, virial in the order xx yy zz xy xz yz yx zx zy).
hipError_t launch_serve_extract(int R, const ServeOut* req, long long no, int Nt, const double* pos,
                                const double* vel, const double* force, const double* pe, const double* vir,
                                double* out, hipStream_t s);

// Number of thermo block partials launch_thermo uses for N atoms.
int thermo_blocks(int N);

}  // namespace dmgx
