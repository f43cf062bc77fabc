/*
 * nep_oracle.h -- FP64 CPU restatement of the NEP4/NEP5 (+ZBL) energy, force
 * and virial, used ONLY as test infrastructure (the checker).  Nothing in the
 * product library links or calls this code; tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load it through ctypes.
 *
 * Reference behaviour restated (file:line into the reference tree):
 *   parser            src/gpumd_compat/nep.cu:135-415 (GPUMD nep.txt grammar)
 *   descriptor + ANN  src/gpumd_compat/nep.cu:569-748,
 *                     third_party/lammps-matpl-overlay/files/src/nep_cpu.cpp:918-1064
 *   radial force      nep.cu:756-881, nep_cpu.cpp:1066-1162
 *   angular force     nep.cu:890-991 + potential.cu:81-208, nep_cpu.cpp:1164-1278
 *   ZBL               nep.cu:996-1111, nep_cpu.cpp:1280-1381
 *   MIC               src/gpumd_compat/box.cuh:77-172
 *   neighbor rows     src/gpumd_compat/neighbor.cu:141-218 (predicate d^2 < rc^2)
 *   step kernels      src/runtime.cu:392-535, src/runtime_internal.hpp:648-716
 *
 * Pinning: tests/test_oracle_*.py check every function here against the
 * reference's golden per-atom outputs (tests/baseline/goldens), against the
 * compiled reference CPU evaluator in oracle/_ref (nep_cpu.cpp), against
 * finite differences of the oracle energy, rotation/permutation invariance,
 * and against closed forms (Legendre / Wigner-3j constants, ZBL textbook form).
 */
#ifndef DMGX_NEP_ORACLE_H
#define DMGX_NEP_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct nepo_model nepo_model;

/* Parse a GPUMD-format nep.txt.  Returns 0 on success, nonzero with a message
 * in err on failure (never exits). */
int nepo_load(const char* path, nepo_model** out, char* err, int errlen);
void nepo_free(nepo_model* m);

/* Integer/real descriptors of the model:
 * ints[0..11] = version, num_types, n_max_R, n_max_A, basis_R, basis_A,
 *               L_max, num_L, dim, num_neurons, zbl_enabled, zbl_flexible
 * reals[0..3] = rc_R_max, rc_A_max, zbl_rc_inner, zbl_rc_outer */
void nepo_info(const nepo_model* m, int* ints, double* reals);

/* When round_to_float != 0 every parsed parameter is rounded to float (the
 * storage precision of the GPU reference, nep.cu:388) before FP64 use. */
void nepo_set_param_rounding(nepo_model* m, int round_to_float);

/* Energy/force/virial of one system.
 *   h[9]   box matrix, row-major with the cell vectors a,b,c as COLUMNS
 *          (h[0]=ax h[1]=bx h[2]=cx h[3]=ay ...), the reference's cpu_h[0..8]
 *   pbc[3] 1 = periodic
 *   centres: [nd1,nd2) get descriptors, [n1,n2) get forces/virials; every
 *          atom in [0,N) is a neighbour candidate (the domain contract of
 *          NEP::compute_domain, nep.cu:1298-1838).  n1=nd1=0, n2=nd2=N is the
 *          plain NEP::compute.
 *   pos    SoA x|y|z (3N); force SoA (3N); virial SoA (9N) in the order
 *          xx yy zz xy xz yz yx zx zy; pe (N).  Outputs are overwritten.
 *   q_out  optional N*dim scaled descriptors (row per atom); fp_out optional
 *          N*dim dE/dq (already multiplied by q_scaler).
 * Returns 0, or nonzero on invalid input. */
int nepo_compute(const nepo_model* m, const double* h, const int* pbc, int N,
                 int n1, int n2, int nd1, int nd2, const int* type,
                 const double* pos, double* pe, double* force, double* virial,
                 double* q_out, double* fp_out);

/* Neighbour rows under the canonical FP64 predicate: for every centre i in
 * [c1,c2) all j in [0,N), j != i, with d2(i,j) < rc*rc where d2 is the FP64
 * minimum-image displacement (x_j - x_i) squared and summed as (x*x+y*y)+z*z.
 * Requires every periodic thickness > 2*rc (one image).  Rows are written
 * sorted ascending into rows[(i-c1)*max_nb + k]; counts[i-c1] gets the row
 * length.  Returns 0, 1 if a row overflows max_nb, 2 if the box is too thin. */
int nepo_neighbor_rows(const double* h, const int* pbc, int N, const double* pos,
                       int c1, int c2, double rc, int max_nb, int* counts, int* rows);

/* Step kernels (FP64), as src/runtime.cu:392-535. */
void nepo_velocity_verlet(int first_half, int N, double dt, const double* mass,
                          double* pos, double* vel, const double* force);
void nepo_wrap_positions(const double* h, const int* pbc, int N, double* pos);
/* thermo[8] = T, sum pe, (sum virial_ab + m v_a v_b)/V for xx yy zz xy xz yz */
void nepo_thermo(int N, const double* h, const double* mass, const double* pe,
                 const double* vel, const double* virial, double* thermo);

/* Derived angular constants (for tests): C3B-style weights per (L,m) packed
 * like the moment index (L*L-1 + 2m-1 / 2m), the q_222 and q_1111 weights,
 * and the integer polynomial Z_{L,m}(z) coefficients zc[(L*(L+1)/2+m)*9+k]. */
void nepo_constant_tables(double* c3, double* c4, double* c5, double* zc);

#ifdef __cplusplus
}
#endif
#endif
