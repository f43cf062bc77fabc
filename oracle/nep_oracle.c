/*
 * nep_oracle.c -- FP64 CPU restatement of the NEP4/NEP5 (+ZBL) hot path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker that the HIP product is
 * compared against.  It is never linked into, loaded by, or called from the
 * product library.  See nep_oracle.h for the reference file:line map and for
 * how each function is pinned.
 *
 * Formulation (the oracle's own, deliberately different from the product's
 * kernels so that agreement tests the mathematics):
 *   - angular moments s[n][L,m] = sum_j g_n(r_ij) Y_{L,m}(u_ij) with the real
 *     solid harmonics Y = Z_{L,m}(u_z) * Re/Im (u_x + i u_y)^m, Z_{L,m} the
 *     primitive integer polynomial of d^m P_L / dz^m (closed form, derived in
 *     derive_constants below, not tabulated);
 *   - per-centre adjoint a[n][L,m] = dE_i/ds_i[n][L,m] formed once, then every
 *     neighbour's dE_i/dr_ij = sum a * d(g_n Y)/dr evaluated with the exact
 *     gradient of Y(r/|r|) projected on the tangent plane;
 *   - forces by Newton scatter (as nep_cpu.cpp:1133-1160), per-atom virial
 *     assigned to the neighbour, virial[j] -= r_ij (x) dE_i/dr_ij.
 * Everything is FP64; build with -ffp-contract=off so the neighbour predicate
 * is bit-reproducible against the device's FP64 predicate.
 */
#include "nep_oracle.h"

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXT 103
#define MAXL 8
#define NABC 80 /* (MAXL+1)^2 - 1 */
#define MAXN 20 /* max n_max+1 and basis+1 */
#define MAXDIM 200
#define MAXH 512

static const double PI_D = 3.14159265358979323846;
static const double K_B = 8.617343e-5;        /* common.cuh:55 */
static const double K_C_SP = 14.399645;       /* common.cuh:57 (FP64 here) */

/* Element symbols, atomic number = index+1 (nep.cu:83 lists the same symbols). */
static const char* ELEMENTS[94] = {
  "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",
  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge",
  "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd",
  "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd",
  "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg",
  "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U",  "Np", "Pu"};

/* Covalent radii used by the typewise ZBL cutoff; the same data table as
 * nep_utilities.cuh:175-186 (a data table that has to be identical). */
static const double COVALENT_RADIUS[94] = {
  0.426667, 0.613333, 1.6,     1.25333, 1.02667, 1.0,     0.946667, 0.84,    0.853333,
  0.893333, 1.86667,  1.66667, 1.50667, 1.38667, 1.46667, 1.36,     1.32,    1.28,
  2.34667,  2.05333,  1.77333, 1.62667, 1.61333, 1.46667, 1.42667,  1.38667, 1.33333,
  1.32,     1.34667,  1.45333, 1.49333, 1.45333, 1.53333, 1.46667,  1.52,    1.56,
  2.52,     2.22667,  1.96,    1.85333, 1.76,    1.65333, 1.53333,  1.50667, 1.50667,
  1.44,     1.53333,  1.64,    1.70667, 1.68,    1.68,    1.64,     1.76,    1.74667,
  2.78667,  2.34667,  2.16,    1.96,    2.10667, 2.09333, 2.08,     2.06667, 2.01333,
  2.02667,  2.01333,  2.0,     1.98667, 1.98667, 1.97333, 2.04,     1.94667, 1.82667,
  1.74667,  1.64,     1.57333, 1.54667, 1.48,    1.49333, 1.50667,  1.76,    1.73333,
  1.73333,  1.81333,  1.74667, 1.84,    1.89333, 2.68,    2.41333,  2.22667, 2.10667,
  2.02667,  2.04,     2.05333, 2.06667};

struct nepo_model {
  int version;
  int num_types;
  int Z[MAXT];
  int zbl_enabled, zbl_flexible, zbl_typewise;
  double zbl_rc_inner, zbl_rc_outer, zbl_factor;
  double zbl_para[550];
  double rcR[MAXT], rcA[MAXT];
  int MN_R, MN_A;
  int nR, nA; /* n_max + 1 */
  int kR, kA; /* basis_size + 1 */
  int L_max, has222, has1111, num_L, dim, H;
  int num_para_ann, num_c_radial, num_para_desc, num_para_total;
  double* raw;  /* parameters as parsed (double) */
  double* par;  /* parameters as used (raw or float-rounded) */
  int round_to_float;
  /* views into par */
  const double *w0[MAXT], *b0[MAXT], *w1[MAXT], *b1, *c, *q_scaler;
  /* derived constants */
  double c3[NABC];       /* weight of s^2 per moment component (2x for m>0 folded in) */
  double zc[MAXL + 1][MAXL + 1][MAXL + 1];
  double c4[5], c5;
};

/* ----------------------------------------------------------------------- */
/* Closed-form angular constants                                            */
/* ----------------------------------------------------------------------- */

static long long gcd_ll(long long a, long long b)
{
  if (a < 0) a = -a;
  if (b < 0) b = -b;
  while (b) {
    long long t = a % b;
    a = b;
    b = t;
  }
  return a;
}

static long long binom_ll(int n, int k)
{
  long long r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

/* Z_{L,m}: 2^L P_L(z) = sum_k (-1)^k C(L,k) C(2L-2k,L) z^(L-2k) has integer
 * coefficients; differentiate m times, divide by the content.  The weight of
 * the (L,m) moment in q_L follows from the addition theorem
 *   sum_{jk} g_j g_k P_L(u_j.u_k) (2L+1)/(4 pi)
 *     = sum_m c_m (2L+1)/(4pi) (L-m)!/(L+m)! alpha^2 |S_m|^2,
 * alpha = content / 2^L, c_0 = 1, c_m = 2 (Legendre addition theorem). */
static void derive_constants(nepo_model* m)
{
  memset(m->zc, 0, sizeof(m->zc));
  for (int L = 1; L <= MAXL; ++L) {
    long long base[MAXL + 1] = {0};
    for (int k = 0; 2 * k <= L; ++k) {
      long long v = binom_ll(L, k) * binom_ll(2 * L - 2 * k, L);
      base[L - 2 * k] = (k % 2) ? -v : v;
    }
    for (int mm = 0; mm <= L; ++mm) {
      long long d[MAXL + 1] = {0};
      for (int p = mm; p <= L; ++p) {
        long long f = 1;
        for (int t = 0; t < mm; ++t) f *= (p - t);
        d[p - mm] = base[p] * f;
      }
      long long g = 0;
      for (int p = 0; p <= L - mm; ++p)
        if (d[p]) g = gcd_ll(g, d[p]);
      for (int p = 0; p <= L - mm; ++p) m->zc[L][mm][p] = (double)(d[p] / g);
      double alpha = (double)g / ldexp(1.0, L);
      double fr = 1.0;
      for (int t = L - mm + 1; t <= L + mm; ++t) fr /= (double)t;
      double w = (2.0 * L + 1.0) / (4.0 * PI_D) * fr * alpha * alpha;
      int base_idx = L * L - 1;
      if (mm == 0) {
        m->c3[base_idx] = w;
      } else {
        m->c3[base_idx + 2 * mm - 1] = 2.0 * w;
        m->c3[base_idx + 2 * mm] = 2.0 * w;
      }
    }
  }
  /* q_222 = sum_{m1+m2+m3=0} (2 2 2; m1 m2 m3) A_{2m1} A_{2m2} A_{2m3} with the
   * complex moments A_{2m} written through the real ones; expanding the Wigner
   * 3j symbol gives the five monomial weights below (k * sqrt(14)/(448 pi^1.5))
   * for s3^3, s3(s4^2+s5^2), s3(s6^2+s7^2), s6(s5^2-s4^2), s4 s5 s7. */
  const double unit4 = sqrt(14.0) / (448.0 * pow(PI_D, 1.5));
  m->c4[0] = -5.0 * unit4;
  m->c4[1] = -90.0 * unit4;
  m->c4[2] = 45.0 * unit4;
  m->c4[3] = 270.0 * unit4;
  m->c4[4] = -540.0 * unit4;
  /* q_1111 = sum over the intermediate coupling L=0,1,2 of
   * (1 1 L; m1 m2 -M)(L 1 1; M m3 m4) A^4 = 21/(80 pi^2) (s0^2+s1^2+s2^2)^2. */
  m->c5 = 21.0 / (80.0 * PI_D * PI_D);
}

void nepo_constant_tables(double* c3, double* c4, double* c5, double* zc)
{
  nepo_model tmp;
  memset(&tmp, 0, sizeof(tmp));
  derive_constants(&tmp);
  if (c3) memcpy(c3, tmp.c3, sizeof(tmp.c3));
  if (c4) memcpy(c4, tmp.c4, sizeof(tmp.c4));
  if (c5) *c5 = tmp.c5;
  if (zc) {
    for (int L = 0; L <= MAXL; ++L)
      for (int mm = 0; mm <= L; ++mm)
        for (int k = 0; k <= MAXL; ++k) zc[(L * (L + 1) / 2 + mm) * 9 + k] = tmp.zc[L][mm][k];
  }
}

/* ----------------------------------------------------------------------- */
/* Parser (GPUMD nep.txt grammar as read by nep.cu:135-415)                 */
/* ----------------------------------------------------------------------- */

#define MAXTOK 64
typedef struct {
  int n;
  char tok[MAXTOK][64];
} toks_t;

static int read_tokens(FILE* f, toks_t* t)
{
  char line[4096];
  t->n = 0;
  if (!fgets(line, sizeof(line), f)) return -1;
  char* p = line;
  while (*p) {
    while (*p && isspace((unsigned char)*p)) ++p;
    if (!*p) break;
    char* q = p;
    while (*q && !isspace((unsigned char)*q)) ++q;
    if (t->n < MAXTOK) {
      int len = (int)(q - p);
      if (len > 63) len = 63;
      memcpy(t->tok[t->n], p, len);
      t->tok[t->n][len] = 0;
      t->n++;
    }
    p = q;
  }
  return t->n;
}

static int parse_int(const char* s, int* v)
{
  char* e;
  long x = strtol(s, &e, 10);
  if (e == s) return -1;
  *v = (int)x;
  return 0;
}

static int parse_double(const char* s, double* v)
{
  char* e;
  double x = strtod(s, &e);
  if (e == s || !isfinite(x)) return -1;
  *v = x;
  return 0;
}

#define FAIL(...)                              \
  do {                                         \
    if (err) snprintf(err, errlen, __VA_ARGS__); \
    goto fail;                                 \
  } while (0)

static void set_views(nepo_model* m)
{
  const double* p = m->par;
  for (int t = 0; t < m->num_types; ++t) {
    m->w0[t] = p;
    p += m->H * m->dim;
    m->b0[t] = p;
    p += m->H;
    m->w1[t] = p;
    p += m->H;
    if (m->version == 5) p += 1; /* typewise output bias stored after w1 */
  }
  m->b1 = p;
  m->c = m->par + m->num_para_ann;
  m->q_scaler = m->par + m->num_para_ann + m->num_para_desc;
}

void nepo_set_param_rounding(nepo_model* m, int round_to_float)
{
  m->round_to_float = round_to_float;
  for (int i = 0; i < m->num_para_total; ++i)
    m->par[i] = round_to_float ? (double)(float)m->raw[i] : m->raw[i];
  if (m->zbl_flexible) {
    /* flexible ZBL parameters are kept as parsed; the GPU reference stores
     * them as float (nep.cuh ZBL::para) */
  }
  set_views(m);
}

int nepo_load(const char* path, nepo_model** out, char* err, int errlen)
{
  nepo_model* m = (nepo_model*)calloc(1, sizeof(nepo_model));
  FILE* f = fopen(path, "r");
  toks_t t;
  if (!m) FAIL("out of memory");
  if (!f) FAIL("cannot open %s", path);
  derive_constants(m);
  if (read_tokens(f, &t) < 3) FAIL("first line needs at least 3 items");
  if (!strcmp(t.tok[0], "nep4")) {
    m->version = 4;
  } else if (!strcmp(t.tok[0], "nep4_zbl")) {
    m->version = 4;
    m->zbl_enabled = 1;
  } else if (!strcmp(t.tok[0], "nep5")) {
    m->version = 5;
  } else if (!strcmp(t.tok[0], "nep5_zbl")) {
    m->version = 5;
    m->zbl_enabled = 1;
  } else {
    FAIL("unsupported model keyword %s", t.tok[0]);
  }
  if (parse_int(t.tok[1], &m->num_types) || m->num_types < 1 || m->num_types > MAXT)
    FAIL("bad number of types");
  if (t.n != 2 + m->num_types) FAIL("first line should list %d symbols", m->num_types);
  for (int n = 0; n < m->num_types; ++n) {
    m->Z[n] = 0;
    for (int e = 0; e < 94; ++e)
      if (!strcmp(t.tok[2 + n], ELEMENTS[e])) {
        m->Z[n] = e + 1;
        break;
      }
  }
  if (m->zbl_enabled) {
    if (read_tokens(f, &t) != 3 && t.n != 4) FAIL("zbl line should be zbl rc_inner rc_outer [factor]");
    if (parse_double(t.tok[1], &m->zbl_rc_inner) || parse_double(t.tok[2], &m->zbl_rc_outer))
      FAIL("bad zbl cutoffs");
    if (m->zbl_rc_inner == 0 && m->zbl_rc_outer == 0) {
      m->zbl_flexible = 1;
    } else if (t.n == 4) {
      if (parse_double(t.tok[3], &m->zbl_factor)) FAIL("bad zbl factor");
      m->zbl_typewise = 1;
    }
  }
  read_tokens(f, &t);
  if (t.n != 5 && t.n != 2 * m->num_types + 3) FAIL("cutoff line has %d tokens", t.n);
  if (t.n == 5) {
    double r, a;
    if (parse_double(t.tok[1], &r) || parse_double(t.tok[2], &a)) FAIL("bad cutoff");
    for (int n = 0; n < m->num_types; ++n) {
      m->rcR[n] = r;
      m->rcA[n] = a;
    }
  } else {
    for (int n = 0; n < m->num_types; ++n)
      if (parse_double(t.tok[1 + 2 * n], &m->rcR[n]) || parse_double(t.tok[2 + 2 * n], &m->rcA[n]))
        FAIL("bad typewise cutoff");
  }
  if (parse_int(t.tok[t.n - 2], &m->MN_R) || parse_int(t.tok[t.n - 1], &m->MN_A)) FAIL("bad MN");
  if (read_tokens(f, &t) != 3 || strcmp(t.tok[0], "n_max")) FAIL("expected n_max line");
  if (parse_int(t.tok[1], &m->nR) || parse_int(t.tok[2], &m->nA)) FAIL("bad n_max");
  m->nR += 1;
  m->nA += 1;
  if (read_tokens(f, &t) != 3 || strcmp(t.tok[0], "basis_size")) FAIL("expected basis_size line");
  if (parse_int(t.tok[1], &m->kR) || parse_int(t.tok[2], &m->kA)) FAIL("bad basis_size");
  m->kR += 1;
  m->kA += 1;
  if (m->nR > MAXN || m->nA > MAXN || m->kR > MAXN || m->kA > MAXN) FAIL("n_max/basis too large");
  if (read_tokens(f, &t) < 4 || strcmp(t.tok[0], "l_max")) FAIL("expected l_max line");
  if (parse_int(t.tok[1], &m->L_max) || parse_int(t.tok[2], &m->has222) ||
      parse_int(t.tok[3], &m->has1111))
    FAIL("bad l_max");
  for (int k = 4; k < t.n; ++k) {
    int v = 0;
    if (parse_int(t.tok[k], &v)) FAIL("bad l_max flag");
    if (v) FAIL("q_112/q_123/q_233/q_134 invariants are not supported");
  }
  if (m->L_max < 1 || m->L_max > MAXL) FAIL("l_max out of range");
  if ((m->has222 || m->has1111) && m->L_max < 2) FAIL("4/5-body terms need l_max >= 2");
  m->has222 = m->has222 != 0;
  m->has1111 = m->has1111 != 0;
  m->num_L = m->L_max + m->has222 + m->has1111;
  m->dim = m->nR + m->nA * m->num_L;
  if (m->dim > MAXDIM) FAIL("descriptor too long");
  if (read_tokens(f, &t) != 3 || strcmp(t.tok[0], "ANN")) FAIL("expected ANN line");
  if (parse_int(t.tok[1], &m->H) || m->H < 1 || m->H > MAXH) FAIL("bad ANN width");
  if (m->version == 4)
    m->num_para_ann = (m->dim + 2) * m->H * m->num_types + 1;
  else
    m->num_para_ann = ((m->dim + 2) * m->H + 1) * m->num_types + 1;
  int T2 = m->num_types * m->num_types;
  m->num_c_radial = T2 * m->nR * m->kR;
  m->num_para_desc = T2 * (m->nR * m->kR + m->nA * m->kA);
  m->num_para_total = m->num_para_ann + m->num_para_desc + m->dim;
  m->raw = (double*)malloc(sizeof(double) * m->num_para_total);
  m->par = (double*)malloc(sizeof(double) * m->num_para_total);
  if (!m->raw || !m->par) FAIL("out of memory");
  for (int i = 0; i < m->num_para_total; ++i) {
    if (read_tokens(f, &t) < 1 || parse_double(t.tok[0], &m->raw[i]))
      FAIL("parameter %d missing or malformed", i);
  }
  if (m->zbl_flexible) {
    int nz = 10 * (m->num_types * (m->num_types + 1)) / 2;
    for (int i = 0; i < nz; ++i)
      if (read_tokens(f, &t) < 1 || parse_double(t.tok[0], &m->zbl_para[i]))
        FAIL("flexible zbl parameter %d missing", i);
  }
  fclose(f);
  nepo_set_param_rounding(m, 0);
  *out = m;
  return 0;
fail:
  if (f) fclose(f);
  if (m) {
    free(m->raw);
    free(m->par);
    free(m);
  }
  return 1;
}

void nepo_free(nepo_model* m)
{
  if (!m) return;
  free(m->raw);
  free(m->par);
  free(m);
}

void nepo_info(const nepo_model* m, int* ints, double* reals)
{
  ints[0] = m->version;
  ints[1] = m->num_types;
  ints[2] = m->nR - 1;
  ints[3] = m->nA - 1;
  ints[4] = m->kR - 1;
  ints[5] = m->kA - 1;
  ints[6] = m->L_max;
  ints[7] = m->num_L;
  ints[8] = m->dim;
  ints[9] = m->H;
  ints[10] = m->zbl_enabled;
  ints[11] = m->zbl_flexible;
  double rmax = 0, amax = 0;
  for (int t = 0; t < m->num_types; ++t) {
    if (m->rcR[t] > rmax) rmax = m->rcR[t];
    if (m->rcA[t] > amax) amax = m->rcA[t];
  }
  reals[0] = rmax;
  reals[1] = amax;
  reals[2] = m->zbl_rc_inner;
  reals[3] = m->zbl_rc_outer;
}

/* ----------------------------------------------------------------------- */
/* Box                                                                      */
/* ----------------------------------------------------------------------- */

typedef struct {
  double h[9], hi[9];
  int pbc[3];
  int ortho;
  double thick[3];
  double volume;
} box_t;

static void box_init(box_t* b, const double* h, const int* pbc)
{
  memcpy(b->h, h, 9 * sizeof(double));
  for (int d = 0; d < 3; ++d) b->pbc[d] = pbc[d];
  const double* H = h;
  double det = H[0] * (H[4] * H[8] - H[5] * H[7]) + H[1] * (H[5] * H[6] - H[3] * H[8]) +
               H[2] * (H[3] * H[7] - H[4] * H[6]);
  /* inverse by cofactors, box.cu:101-122 */
  b->hi[0] = (H[4] * H[8] - H[5] * H[7]) / det;
  b->hi[1] = (H[2] * H[7] - H[1] * H[8]) / det;
  b->hi[2] = (H[1] * H[5] - H[2] * H[4]) / det;
  b->hi[3] = (H[5] * H[6] - H[3] * H[8]) / det;
  b->hi[4] = (H[0] * H[8] - H[2] * H[6]) / det;
  b->hi[5] = (H[2] * H[3] - H[0] * H[5]) / det;
  b->hi[6] = (H[3] * H[7] - H[4] * H[6]) / det;
  b->hi[7] = (H[1] * H[6] - H[0] * H[7]) / det;
  b->hi[8] = (H[0] * H[4] - H[1] * H[3]) / det;
  b->volume = fabs(det);
  b->ortho = (H[1] == 0 && H[2] == 0 && H[3] == 0 && H[5] == 0 && H[6] == 0 && H[7] == 0);
  /* thickness = V / |cross of the other two cell vectors| */
  double a[3] = {H[0], H[3], H[6]}, bb[3] = {H[1], H[4], H[7]}, c[3] = {H[2], H[5], H[8]};
  const double* vv[3][2] = {{bb, c}, {c, a}, {a, bb}};
  for (int d = 0; d < 3; ++d) {
    const double* u = vv[d][0];
    const double* w = vv[d][1];
    double x = u[1] * w[2] - u[2] * w[1];
    double y = u[2] * w[0] - u[0] * w[2];
    double z = u[0] * w[1] - u[1] * w[0];
    b->thick[d] = b->volume / sqrt(x * x + y * y + z * z);
  }
}

/* Minimum image of an FP64 displacement (box.cuh:77-123). */
static void mic(const box_t* b, double* x, double* y, double* z)
{
  if (b->ortho) {
    double L[3] = {b->h[0], b->h[4], b->h[8]};
    double* r[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) {
      if (!b->pbc[d]) continue;
      if (*r[d] < -L[d] * 0.5)
        *r[d] += L[d];
      else if (*r[d] > L[d] * 0.5)
        *r[d] -= L[d];
    }
  } else {
    const double* hi = b->hi;
    const double* h = b->h;
    double sx = hi[0] * *x + hi[1] * *y + hi[2] * *z;
    double sy = hi[3] * *x + hi[4] * *y + hi[5] * *z;
    double sz = hi[6] * *x + hi[7] * *y + hi[8] * *z;
    if (b->pbc[0]) sx -= nearbyint(sx);
    if (b->pbc[1]) sy -= nearbyint(sy);
    if (b->pbc[2]) sz -= nearbyint(sz);
    *x = h[0] * sx + h[1] * sy + h[2] * sz;
    *y = h[3] * sx + h[4] * sy + h[5] * sz;
    *z = h[6] * sx + h[7] * sy + h[8] * sz;
  }
}

static double dist2(double x, double y, double z) { return (x * x + y * y) + z * z; }

/* ----------------------------------------------------------------------- */
/* Neighbour enumeration                                                    */
/* ----------------------------------------------------------------------- */

typedef struct {
  int n, cap;
  int* j;
  double* r; /* 3 per entry */
} nblist_t;

static int nb_push(nblist_t* l, int j, double x, double y, double z)
{
  if (l->n == l->cap) {
    int cap = l->cap ? 2 * l->cap : 64;
    int* nj = (int*)realloc(l->j, sizeof(int) * cap);
    double* nr = (double*)realloc(l->r, sizeof(double) * 3 * cap);
    if (!nj || !nr) return -1;
    l->j = nj;
    l->r = nr;
    l->cap = cap;
  }
  l->j[l->n] = j;
  l->r[3 * l->n + 0] = x;
  l->r[3 * l->n + 1] = y;
  l->r[3 * l->n + 2] = z;
  l->n++;
  return 0;
}

/* Collect, for each centre in [c1,c2), every (j, image) with d2 < rc2.
 * One image (minimum image) when every periodic thickness exceeds 2 rc,
 * explicit image enumeration otherwise (the small-box case, nep.cu:2002). */
static int build_neighbors(const box_t* b, int N, const double* pos, int c1, int c2, double rc,
                           nblist_t* lists)
{
  const double rc2 = rc * rc;
  const double* X = pos;
  const double* Y = pos + N;
  const double* Zp = pos + 2 * N;
  int one_image = 1;
  for (int d = 0; d < 3; ++d)
    if (b->pbc[d] && !(b->thick[d] > 2.0 * rc)) one_image = 0;
  if (one_image) {
    /* cell binning in fractional space, cells of edge >= rc */
    int nc[3];
    for (int d = 0; d < 3; ++d) {
      nc[d] = b->pbc[d] ? (int)floor(b->thick[d] / rc) : 1;
      if (nc[d] < 3) nc[d] = 1;
    }
    int ncell = nc[0] * nc[1] * nc[2];
    int* head = (int*)malloc(sizeof(int) * ncell);
    int* next = (int*)malloc(sizeof(int) * N);
    int* cell_of = (int*)malloc(sizeof(int) * N * 3);
    if (!head || !next || !cell_of) return -1;
    for (int c = 0; c < ncell; ++c) head[c] = -1;
    for (int i = N - 1; i >= 0; --i) {
      double s[3];
      s[0] = b->hi[0] * X[i] + b->hi[1] * Y[i] + b->hi[2] * Zp[i];
      s[1] = b->hi[3] * X[i] + b->hi[4] * Y[i] + b->hi[5] * Zp[i];
      s[2] = b->hi[6] * X[i] + b->hi[7] * Y[i] + b->hi[8] * Zp[i];
      int ci[3];
      for (int d = 0; d < 3; ++d) {
        double u = s[d] - floor(s[d]);
        int k = (int)floor(u * nc[d]);
        if (k < 0) k = 0;
        if (k >= nc[d]) k = nc[d] - 1;
        ci[d] = k;
        cell_of[3 * i + d] = k;
      }
      int c = ci[0] + nc[0] * (ci[1] + nc[1] * ci[2]);
      next[i] = head[c];
      head[c] = i;
    }
    for (int i = c1; i < c2; ++i) {
      nblist_t* l = &lists[i - c1];
      int ci[3] = {cell_of[3 * i], cell_of[3 * i + 1], cell_of[3 * i + 2]};
      int rng[3];
      for (int d = 0; d < 3; ++d) rng[d] = nc[d] >= 3 ? 1 : 0;
      for (int dz = -rng[2]; dz <= rng[2]; ++dz)
        for (int dy = -rng[1]; dy <= rng[1]; ++dy)
          for (int dx = -rng[0]; dx <= rng[0]; ++dx) {
            int cc[3] = {ci[0] + dx, ci[1] + dy, ci[2] + dz};
            for (int d = 0; d < 3; ++d) cc[d] = (cc[d] % nc[d] + nc[d]) % nc[d];
            int c = cc[0] + nc[0] * (cc[1] + nc[1] * cc[2]);
            for (int j = head[c]; j >= 0; j = next[j]) {
              if (j == i) continue;
              double x = X[j] - X[i], y = Y[j] - Y[i], z = Zp[j] - Zp[i];
              mic(b, &x, &y, &z);
              if (dist2(x, y, z) < rc2)
                if (nb_push(l, j, x, y, z)) return -1;
            }
          }
    }
    free(head);
    free(next);
    free(cell_of);
  } else {
    int nim[3];
    for (int d = 0; d < 3; ++d) nim[d] = b->pbc[d] ? (int)ceil(rc / b->thick[d] + 0.5) : 0;
    for (int i = c1; i < c2; ++i) {
      nblist_t* l = &lists[i - c1];
      for (int j = 0; j < N; ++j) {
        double x0 = X[j] - X[i], y0 = Y[j] - Y[i], z0 = Zp[j] - Zp[i];
        mic(b, &x0, &y0, &z0);
        for (int ia = -nim[0]; ia <= nim[0]; ++ia)
          for (int ib = -nim[1]; ib <= nim[1]; ++ib)
            for (int ic = -nim[2]; ic <= nim[2]; ++ic) {
              if (j == i && ia == 0 && ib == 0 && ic == 0) continue;
              double x = x0 + (b->h[0] * ia + b->h[1] * ib + b->h[2] * ic);
              double y = y0 + (b->h[3] * ia + b->h[4] * ib + b->h[5] * ic);
              double z = z0 + (b->h[6] * ia + b->h[7] * ib + b->h[8] * ic);
              if (dist2(x, y, z) < rc2)
                if (nb_push(l, j, x, y, z)) return -1;
            }
      }
    }
  }
  return 0;
}

static int cmp_int(const void* a, const void* b)
{
  int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

int nepo_neighbor_rows(const double* h, const int* pbc, int N, const double* pos, int c1, int c2,
                       double rc, int max_nb, int* counts, int* rows)
{
  box_t b;
  box_init(&b, h, pbc);
  for (int d = 0; d < 3; ++d)
    if (b.pbc[d] && !(b.thick[d] > 2.0 * rc)) return 2;
  nblist_t* lists = (nblist_t*)calloc(c2 - c1 > 0 ? c2 - c1 : 1, sizeof(nblist_t));
  if (!lists) return 3;
  int rcode = 0;
  if (build_neighbors(&b, N, pos, c1, c2, rc, lists)) rcode = 3;
  for (int i = 0; i < c2 - c1 && !rcode; ++i) {
    counts[i] = lists[i].n;
    if (lists[i].n > max_nb) {
      rcode = 1;
      break;
    }
    memcpy(rows + (size_t)i * max_nb, lists[i].j, sizeof(int) * lists[i].n);
    qsort(rows + (size_t)i * max_nb, lists[i].n, sizeof(int), cmp_int);
  }
  for (int i = 0; i < c2 - c1; ++i) {
    free(lists[i].j);
    free(lists[i].r);
  }
  free(lists);
  return rcode;
}

/* ----------------------------------------------------------------------- */
/* Radial functions                                                         */
/* ----------------------------------------------------------------------- */

/* cosine cutoff and derivative, nep_utilities.cuh:317-342 */
static void cutoff_fn(double rc, double d, double* fc, double* fcp)
{
  if (d < rc) {
    double x = PI_D * d / rc;
    *fc = 0.5 * cos(x) + 0.5;
    *fcp = -0.5 * PI_D / rc * sin(x);
  } else {
    *fc = 0.0;
    *fcp = 0.0;
  }
}

/* f_k(d) = (T_k(x)+1)/2 * fc(d), x = 2 (d/rc - 1)^2 - 1, and d f_k / dd
 * (nep_utilities.cuh:440-494).  T'_k = k U_{k-1}. */
static void cheb_basis(int kmax1, double rc, double d, double* f, double* fp)
{
  double fc, fcp;
  cutoff_fn(rc, d, &fc, &fcp);
  double t = d / rc - 1.0;
  double x = 2.0 * t * t - 1.0;
  double dxdd = 4.0 * t / rc;
  double T[MAXN + 1], U[MAXN + 1];
  T[0] = 1.0;
  T[1] = x;
  U[0] = 1.0;
  U[1] = 2.0 * x;
  for (int k = 2; k <= kmax1; ++k) {
    T[k] = 2.0 * x * T[k - 1] - T[k - 2];
    U[k] = 2.0 * x * U[k - 1] - U[k - 2];
  }
  for (int k = 0; k < kmax1; ++k) {
    double val = 0.5 * (T[k] + 1.0);
    double der = (k == 0) ? 0.0 : 0.5 * k * U[k - 1] * dxdd;
    f[k] = val * fc;
    fp[k] = der * fc + val * fcp;
  }
}

/* ----------------------------------------------------------------------- */
/* Real solid harmonics and gradients on the unit sphere                    */
/* ----------------------------------------------------------------------- */

/* Y[idx] for all components up to L_max, and dY/du (unprojected) */
static void harmonics(const nepo_model* m, const double* u, double* Y, double* dY /*3 per*/)
{
  double zp[MAXL + 1];
  zp[0] = 1.0;
  for (int k = 1; k <= MAXL; ++k) zp[k] = zp[k - 1] * u[2];
  /* complex powers P_m = (ux + i uy)^m */
  double pr[MAXL + 1], pi[MAXL + 1];
  pr[0] = 1.0;
  pi[0] = 0.0;
  for (int k = 1; k <= MAXL; ++k) {
    pr[k] = pr[k - 1] * u[0] - pi[k - 1] * u[1];
    pi[k] = pr[k - 1] * u[1] + pi[k - 1] * u[0];
  }
  for (int L = 1; L <= m->L_max; ++L) {
    for (int mm = 0; mm <= L; ++mm) {
      double Zv = 0.0, dZ = 0.0;
      for (int k = 0; k <= L - mm; ++k) {
        double c = m->zc[L][mm][k];
        if (c == 0.0) continue;
        Zv += c * zp[k];
        if (k > 0) dZ += c * k * zp[k - 1];
      }
      int base = L * L - 1;
      if (mm == 0) {
        Y[base] = Zv;
        dY[3 * base + 0] = 0.0;
        dY[3 * base + 1] = 0.0;
        dY[3 * base + 2] = dZ;
      } else {
        int ir = base + 2 * mm - 1, ii = base + 2 * mm;
        Y[ir] = Zv * pr[mm];
        Y[ii] = Zv * pi[mm];
        dY[3 * ir + 0] = Zv * mm * pr[mm - 1];
        dY[3 * ir + 1] = -Zv * mm * pi[mm - 1];
        dY[3 * ir + 2] = dZ * pr[mm];
        dY[3 * ii + 0] = Zv * mm * pi[mm - 1];
        dY[3 * ii + 1] = Zv * mm * pr[mm - 1];
        dY[3 * ii + 2] = dZ * pi[mm];
      }
    }
  }
}

/* ----------------------------------------------------------------------- */
/* Per-centre descriptor, ANN, adjoint and pair derivatives                 */
/* ----------------------------------------------------------------------- */

static inline double rc_pair(const double* rc, int t1, int t2) { return (rc[t1] + rc[t2]) * 0.5; }

static double c_radial(const nepo_model* m, int n, int k, int t1, int t2)
{
  int T = m->num_types;
  return m->c[(n * m->kR + k) * T * T + t1 * T + t2];
}

static double c_angular(const nepo_model* m, int n, int k, int t1, int t2)
{
  int T = m->num_types;
  return m->c[m->num_c_radial + (n * m->kA + k) * T * T + t1 * T + t2];
}

/* Computes E_i, the scaled descriptor, dE/dq, and dE_i/dr_ij for every
 * neighbour (grad, 3 per entry; zero for entries beyond both cutoffs). */
static void centre(const nepo_model* m, int ti, const nblist_t* l, const int* type, double* E_out,
                   double* q_out, double* fp_out, double* grad)
{
  const int nR = m->nR, nA = m->nA, NABCL = (m->L_max + 1) * (m->L_max + 1) - 1;
  double q[MAXDIM] = {0};
  double s[MAXN][NABC];
  memset(s, 0, sizeof(s));
  double f[MAXN], fp[MAXN], Y[NABC], dY[3 * NABC];

  for (int e = 0; e < l->n; ++e) {
    const double* r = l->r + 3 * e;
    int tj = type[l->j[e]];
    double d2 = dist2(r[0], r[1], r[2]);
    double rcr = rc_pair(m->rcR, ti, tj);
    if (d2 < rcr * rcr) {
      double d = sqrt(d2);
      cheb_basis(m->kR, rcr, d, f, fp);
      for (int n = 0; n < nR; ++n) {
        double g = 0.0;
        for (int k = 0; k < m->kR; ++k) g += c_radial(m, n, k, ti, tj) * f[k];
        q[n] += g;
      }
    }
    double rca = rc_pair(m->rcA, ti, tj);
    if (d2 < rca * rca) {
      double d = sqrt(d2);
      double u[3] = {r[0] / d, r[1] / d, r[2] / d};
      cheb_basis(m->kA, rca, d, f, fp);
      harmonics(m, u, Y, dY);
      for (int n = 0; n < nA; ++n) {
        double g = 0.0;
        for (int k = 0; k < m->kA; ++k) g += c_angular(m, n, k, ti, tj) * f[k];
        for (int a = 0; a < NABCL; ++a) s[n][a] += g * Y[a];
      }
    }
  }
  /* invariants */
  for (int n = 0; n < nA; ++n) {
    for (int L = 1; L <= m->L_max; ++L) {
      double v = 0.0;
      for (int a = L * L - 1; a < (L + 1) * (L + 1) - 1; ++a) v += m->c3[a] * s[n][a] * s[n][a];
      q[nR + (L - 1) * nA + n] = v;
    }
    int li = m->L_max;
    if (m->has222) {
      const double* x = &s[n][3];
      const double* C = m->c4;
      q[nR + li * nA + n] = C[0] * x[0] * x[0] * x[0] + C[1] * x[0] * (x[1] * x[1] + x[2] * x[2]) +
                            C[2] * x[0] * (x[3] * x[3] + x[4] * x[4]) +
                            C[3] * x[3] * (x[2] * x[2] - x[1] * x[1]) + C[4] * x[1] * x[2] * x[4];
      ++li;
    }
    if (m->has1111) {
      double r2 = s[n][0] * s[n][0] + s[n][1] * s[n][1] + s[n][2] * s[n][2];
      q[nR + li * nA + n] = m->c5 * r2 * r2;
    }
  }
  /* ANN (nep_utilities.cuh:209-310) */
  const int D = m->dim, H = m->H;
  double qs[MAXDIM], dq[MAXDIM];
  for (int d = 0; d < D; ++d) {
    qs[d] = q[d] * m->q_scaler[d];
    dq[d] = 0.0;
  }
  const double* w0 = m->w0[ti];
  const double* b0 = m->b0[ti];
  const double* w1 = m->w1[ti];
  double E = 0.0;
  for (int hh = 0; hh < H; ++hh) {
    double a = 0.0;
    for (int d = 0; d < D; ++d) a += w0[hh * D + d] * qs[d];
    double x = tanh(a - b0[hh]);
    E += w1[hh] * x;
    double g = w1[hh] * (1.0 - x * x);
    for (int d = 0; d < D; ++d) dq[d] += g * w0[hh * D + d];
  }
  E -= (m->version == 5) ? (w1[H] + m->b1[0]) : m->b1[0];
  double Fp[MAXDIM];
  for (int d = 0; d < D; ++d) Fp[d] = dq[d] * m->q_scaler[d]; /* dE/dq (unscaled q) */
  *E_out = E;
  if (q_out)
    for (int d = 0; d < D; ++d) q_out[d] = qs[d];
  if (fp_out)
    for (int d = 0; d < D; ++d) fp_out[d] = Fp[d];

  /* adjoint a[n][abc] = dE/ds[n][abc] */
  double adj[MAXN][NABC];
  for (int n = 0; n < nA; ++n) {
    for (int a = 0; a < NABCL; ++a) adj[n][a] = 0.0;
    for (int L = 1; L <= m->L_max; ++L) {
      double F = Fp[nR + (L - 1) * nA + n];
      for (int a = L * L - 1; a < (L + 1) * (L + 1) - 1; ++a)
        adj[n][a] += 2.0 * m->c3[a] * s[n][a] * F;
    }
    int li = m->L_max;
    if (m->has222) {
      double F = Fp[nR + li * nA + n];
      const double* x = &s[n][3];
      const double* C = m->c4;
      adj[n][3] += F * (3.0 * C[0] * x[0] * x[0] + C[1] * (x[1] * x[1] + x[2] * x[2]) +
                        C[2] * (x[3] * x[3] + x[4] * x[4]));
      adj[n][4] += F * (2.0 * C[1] * x[0] * x[1] - 2.0 * C[3] * x[3] * x[1] + C[4] * x[2] * x[4]);
      adj[n][5] += F * (2.0 * C[1] * x[0] * x[2] + 2.0 * C[3] * x[3] * x[2] + C[4] * x[1] * x[4]);
      adj[n][6] += F * (2.0 * C[2] * x[0] * x[3] + C[3] * (x[2] * x[2] - x[1] * x[1]));
      adj[n][7] += F * (2.0 * C[2] * x[0] * x[4] + C[4] * x[1] * x[2]);
      ++li;
    }
    if (m->has1111) {
      double F = Fp[nR + li * nA + n];
      double r2 = s[n][0] * s[n][0] + s[n][1] * s[n][1] + s[n][2] * s[n][2];
      for (int a = 0; a < 3; ++a) adj[n][a] += F * 4.0 * m->c5 * r2 * s[n][a];
    }
  }

  /* dE_i/dr_ij per neighbour */
  for (int e = 0; e < l->n; ++e) {
    const double* r = l->r + 3 * e;
    double* gr = grad + 3 * e;
    gr[0] = gr[1] = gr[2] = 0.0;
    int tj = type[l->j[e]];
    double d2 = dist2(r[0], r[1], r[2]);
    double d = sqrt(d2);
    double u[3] = {r[0] / d, r[1] / d, r[2] / d};
    double rcr = rc_pair(m->rcR, ti, tj);
    if (d2 < rcr * rcr) {
      cheb_basis(m->kR, rcr, d, f, fp);
      double dEdd = 0.0;
      for (int n = 0; n < nR; ++n) {
        double gp = 0.0;
        for (int k = 0; k < m->kR; ++k) gp += c_radial(m, n, k, ti, tj) * fp[k];
        dEdd += Fp[n] * gp;
      }
      for (int c = 0; c < 3; ++c) gr[c] += dEdd * u[c];
    }
    double rca = rc_pair(m->rcA, ti, tj);
    if (d2 < rca * rca) {
      cheb_basis(m->kA, rca, d, f, fp);
      harmonics(m, u, Y, dY);
      double radial = 0.0, G[3] = {0, 0, 0};
      for (int n = 0; n < nA; ++n) {
        double g = 0.0, gp = 0.0;
        for (int k = 0; k < m->kA; ++k) {
          double c = c_angular(m, n, k, ti, tj);
          g += c * f[k];
          gp += c * fp[k];
        }
        for (int a = 0; a < NABCL; ++a) {
          double w = adj[n][a];
          radial += w * gp * Y[a];
          G[0] += w * g * dY[3 * a + 0];
          G[1] += w * g * dY[3 * a + 1];
          G[2] += w * g * dY[3 * a + 2];
        }
      }
      double gu = G[0] * u[0] + G[1] * u[1] + G[2] * u[2];
      for (int c = 0; c < 3; ++c) gr[c] += radial * u[c] + (G[c] - gu * u[c]) / d;
    }
  }
}

/* Universal / flexible ZBL pair energy f(d) and df/dd (nep_utilities.cuh:347-433);
 * screening constants are the published ZBL universal function. */
static void zbl_pair(const nepo_model* m, int t1, int t2, double d, double* f, double* fp)
{
  int zi = m->Z[t1], zj = m->Z[t2];
  double a_inv = (pow((double)zi, 0.23) + pow((double)zj, 0.23)) * 2.134563;
  double zizj = K_C_SP * zi * zj;
  double ab[8] = {0.18175, 3.1998, 0.50986, 0.94229, 0.28022, 0.4029, 0.02817, 0.20162};
  double r1 = m->zbl_rc_inner, r2 = m->zbl_rc_outer;
  if (m->zbl_flexible) {
    int a = t1 < t2 ? t1 : t2, b = t1 < t2 ? t2 : t1;
    int idx = a * m->num_types - (a * (a - 1)) / 2 + (b - a);
    const double* p = m->zbl_para + 10 * idx;
    r1 = p[0];
    r2 = p[1];
    for (int k = 0; k < 8; ++k) ab[k] = p[2 + k];
  } else if (m->zbl_typewise) {
    double rt = (COVALENT_RADIUS[zi - 1] + COVALENT_RADIUS[zj - 1]) * m->zbl_factor;
    r2 = rt < r2 ? rt : r2;
    r1 = 0.0;
  }
  double x = d * a_inv, phi = 0.0, dphi = 0.0;
  for (int k = 0; k < 4; ++k) {
    double e = ab[2 * k] * exp(-ab[2 * k + 1] * x);
    phi += e;
    dphi -= ab[2 * k + 1] * e;
  }
  double v = zizj * phi / d;
  double vp = zizj * dphi * a_inv / d - v / d;
  double fc = 0.0, fcp = 0.0;
  if (d < r1) {
    fc = 1.0;
  } else if (d < r2) {
    double w = PI_D / (r2 - r1);
    fc = 0.5 * cos(w * (d - r1)) + 0.5;
    fcp = -0.5 * w * sin(w * (d - r1));
  }
  *f = v * fc;
  *fp = vp * fc + v * fcp;
}

int nepo_compute(const nepo_model* m, const double* h, const int* pbc, int N, int n1, int n2,
                 int nd1, int nd2, const int* type, const double* pos, double* pe, double* force,
                 double* virial, double* q_out, double* fp_out)
{
  if (N < 0 || n1 < 0 || n2 < n1 || n2 > N || nd1 < 0 || nd2 < nd1 || nd2 > N) return 1;
  for (int i = 0; i < N; ++i)
    if (type[i] < 0 || type[i] >= m->num_types) return 2;
  box_t b;
  box_init(&b, h, pbc);
  double rmax = 0.0;
  for (int t = 0; t < m->num_types; ++t) {
    if (m->rcR[t] > rmax) rmax = m->rcR[t];
    if (m->rcA[t] > rmax) rmax = m->rcA[t];
  }
  memset(pe, 0, sizeof(double) * N);
  memset(force, 0, sizeof(double) * 3 * N);
  memset(virial, 0, sizeof(double) * 9 * N);
  int nc = nd2 - nd1;
  nblist_t* lists = (nblist_t*)calloc(nc > 0 ? nc : 1, sizeof(nblist_t));
  if (!lists) return 3;
  if (build_neighbors(&b, N, pos, nd1, nd2, rmax, lists)) return 3;
  int maxn = 0;
  for (int c = 0; c < nc; ++c)
    if (lists[c].n > maxn) maxn = lists[c].n;
  double* grad = (double*)malloc(sizeof(double) * 3 * (maxn > 0 ? maxn : 1));
  for (int c = 0; c < nc; ++c) {
    int i = nd1 + c;
    const nblist_t* l = &lists[c];
    double E;
    centre(m, type[i], l, type, &E, q_out ? q_out + (size_t)i * m->dim : NULL,
           fp_out ? fp_out + (size_t)i * m->dim : NULL, grad);
    pe[i] += E;
    for (int e = 0; e < l->n; ++e) {
      int j = l->j[e];
      const double* r = l->r + 3 * e;
      double fx = grad[3 * e], fy = grad[3 * e + 1], fz = grad[3 * e + 2];
      if (m->zbl_enabled) {
        int tj = type[j];
        double d2 = dist2(r[0], r[1], r[2]);
        double rca = rc_pair(m->rcA, type[i], tj);
        if (d2 < rca * rca) {
          double d = sqrt(d2), zf, zfp;
          zbl_pair(m, type[i], tj, d, &zf, &zfp);
          fx += 0.5 * zfp * r[0] / d;
          fy += 0.5 * zfp * r[1] / d;
          fz += 0.5 * zfp * r[2] / d;
          pe[i] += 0.5 * zf;
        }
      }
      force[i] += fx;
      force[N + i] += fy;
      force[2 * N + i] += fz;
      force[j] -= fx;
      force[N + j] -= fy;
      force[2 * N + j] -= fz;
      /* virial order xx yy zz xy xz yz yx zx zy */
      virial[0 * N + j] -= r[0] * fx;
      virial[1 * N + j] -= r[1] * fy;
      virial[2 * N + j] -= r[2] * fz;
      virial[3 * N + j] -= r[0] * fy;
      virial[4 * N + j] -= r[0] * fz;
      virial[5 * N + j] -= r[1] * fz;
      virial[6 * N + j] -= r[1] * fx;
      virial[7 * N + j] -= r[2] * fx;
      virial[8 * N + j] -= r[2] * fy;
    }
  }
  free(grad);
  for (int c = 0; c < nc; ++c) {
    free(lists[c].j);
    free(lists[c].r);
  }
  free(lists);
  (void)n1;
  (void)n2;
  return 0;
}

/* ----------------------------------------------------------------------- */
/* Step kernels                                                             */
/* ----------------------------------------------------------------------- */

void nepo_velocity_verlet(int first_half, int N, double dt, const double* mass, double* pos,
                          double* vel, const double* force)
{
  const double half = dt * 0.5;
  for (int i = 0; i < N; ++i) {
    double im = 1.0 / mass[i];
    for (int c = 0; c < 3; ++c) {
      double v = vel[c * N + i] + force[c * N + i] * im * half;
      vel[c * N + i] = v;
      if (first_half) pos[c * N + i] += v * dt;
    }
  }
}

void nepo_wrap_positions(const double* h, const int* pbc, int N, double* pos)
{
  box_t b;
  box_init(&b, h, pbc);
  for (int i = 0; i < N; ++i) {
    double x = pos[i], y = pos[N + i], z = pos[2 * N + i];
    double s[3] = {b.hi[0] * x + b.hi[1] * y + b.hi[2] * z, b.hi[3] * x + b.hi[4] * y + b.hi[5] * z,
                   b.hi[6] * x + b.hi[7] * y + b.hi[8] * z};
    for (int d = 0; d < 3; ++d) {
      if (!b.pbc[d]) continue;
      if (s[d] < 0.0)
        s[d] += 1.0;
      else if (s[d] > 1.0)
        s[d] -= 1.0;
    }
    pos[i] = h[0] * s[0] + h[1] * s[1] + h[2] * s[2];
    pos[N + i] = h[3] * s[0] + h[4] * s[1] + h[5] * s[2];
    pos[2 * N + i] = h[6] * s[0] + h[7] * s[1] + h[8] * s[2];
  }
}

/* Neumaier-compensated sum */
typedef struct {
  double s, c;
} ksum_t;
static void kadd(ksum_t* k, double x)
{
  double t = k->s + x;
  if (fabs(k->s) >= fabs(x))
    k->c += (k->s - t) + x;
  else
    k->c += (x - t) + k->s;
  k->s = t;
}

void nepo_thermo(int N, const double* h, const double* mass, const double* pe, const double* vel,
                 const double* virial, double* thermo)
{
  box_t b;
  int pbc[3] = {1, 1, 1};
  box_init(&b, h, pbc);
  ksum_t acc[8];
  memset(acc, 0, sizeof(acc));
  const int va[6] = {0, 1, 2, 0, 0, 1}, vb[6] = {0, 1, 2, 1, 2, 2};
  for (int i = 0; i < N; ++i) {
    double v[3] = {vel[i], vel[N + i], vel[2 * N + i]};
    kadd(&acc[0], mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
    kadd(&acc[1], pe[i]);
    for (int c = 0; c < 6; ++c) kadd(&acc[2 + c], virial[c * N + i] + mass[i] * v[va[c]] * v[vb[c]]);
  }
  for (int q = 0; q < 8; ++q) thermo[q] = acc[q].s + acc[q].c;
  thermo[0] /= 3.0 * N * K_B;
  for (int q = 2; q < 8; ++q) thermo[q] /= b.volume;
}
