// nep_model.cpp -- nep.txt parser and device parameter image.
//
// Accepts the model keywords, cutoff forms, l_max flags and parameter counts
// of the reference loader (nep.cu:135-415): nep4, nep4_zbl, nep5, nep5_zbl;
// `cutoff rc_R rc_A MN_R MN_A` or the typewise form with 2*num_types radii;
// `l_max L has_q_222 has_q_1111 [...]`; NEP4 ANN count (dim+2)*H*T+1, NEP5
// ((dim+2)*H+1)*T+1; descriptor coefficients in the file's basis-major order
// [(n*K+k)*T^2 + t1*T + t2], radial block then angular block; then dim
// q_scaler values; then 10*T(T+1)/2 flexible-ZBL values.  Malformed input is
// reported by exception (the C ABI turns it into DMGX_EINVAL) instead of the
// reference's exit(1).
#include "nep_model.hpp"

#include <cmath>
#include <fstream>
#include <iterator>
#include <numeric>
#include <sstream>
#include <stdexcept>

namespace dmgx {

namespace {

const char* kElements[94] = {
  "H",  "He", "Li", "Be", "B",  "C",  "N",  "O",  "F",  "Ne", "Na", "Mg", "Al", "Si", "P",  "S",
  "Cl", "Ar", "K",  "Ca", "Sc", "Ti", "V",  "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge",
  "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd",
  "In", "Sn", "Sb", "Te", "I",  "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm", "Sm", "Eu", "Gd",
  "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W",  "Re", "Os", "Ir", "Pt", "Au", "Hg",
  "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U",  "Np", "Pu"};

std::vector<std::string> line_tokens(std::ifstream& in)
{
  std::string line;
  if (!std::getline(in, line)) return {};
  std::istringstream iss(line);
  return {std::istream_iterator<std::string>{iss}, std::istream_iterator<std::string>{}};
}

[[noreturn]] void bad(const std::string& what) { throw std::invalid_argument("nep.txt: " + what); }

int to_int(const std::string& s, const char* what)
{
  try {
    size_t pos = 0;
    int v = std::stoi(s, &pos);
    return v;
  } catch (...) {
    bad(std::string("expected an integer for ") + what + ", got '" + s + "'");
  }
}

double to_double(const std::string& s, const char* what)
{
  double v;
  try {
    v = std::stod(s);
  } catch (...) {
    bad(std::string("expected a number for ") + what + ", got '" + s + "'");
  }
  if (!std::isfinite(v)) bad(std::string("non-finite value for ") + what);
  return v;
}

}  // namespace

void NepModel::load(const std::string& path)
{
  std::ifstream in(path);
  if (!in) bad("cannot open " + path);
  auto t = line_tokens(in);
  if (t.size() < 3) bad("the first line should have at least 3 items");
  if (t[0] == "nep4") {
    version = 4;
  } else if (t[0] == "nep4_zbl") {
    version = 4;
    zbl_enabled = true;
  } else if (t[0] == "nep5") {
    version = 5;
  } else if (t[0] == "nep5_zbl") {
    version = 5;
    zbl_enabled = true;
  } else {
    bad(t[0] + " is not a supported model (nep4, nep4_zbl, nep5, nep5_zbl)");
  }
  num_types = to_int(t[1], "number of types");
  if (num_types < 1 || num_types > kMaxTypes) bad("number of types must be 1.." + std::to_string(kMaxTypes));
  if ((int)t.size() != 2 + num_types) bad("the first line should list " + std::to_string(num_types) + " symbols");
  symbols.assign(t.begin() + 2, t.end());
  for (int n = 0; n < num_types; ++n) {
    Z[n] = 0;
    for (int e = 0; e < 94; ++e)
      if (symbols[n] == kElements[e]) {
        Z[n] = e + 1;
        break;
      }
  }
  if (zbl_enabled) {
    t = line_tokens(in);
    if (t.size() != 3 && t.size() != 4) bad("this line should be zbl rc_inner rc_outer [zbl_factor]");
    zbl_rc_inner = to_double(t[1], "zbl rc_inner");
    zbl_rc_outer = to_double(t[2], "zbl rc_outer");
    if (zbl_rc_inner == 0 && zbl_rc_outer == 0) {
      zbl_flexible = true;
    } else if (t.size() == 4) {
      zbl_factor = to_double(t[3], "zbl factor");
      zbl_typewise = true;
    }
    for (int n = 0; n < num_types; ++n)
      if (Z[n] == 0) bad("ZBL needs a known element symbol for every type");
  }
  t = line_tokens(in);
  if (t.empty() || t[0] != "cutoff") bad("expected the cutoff line");
  if (t.size() != 5 && (int)t.size() != 2 * num_types + 3)
    bad("cutoff should have 4 or num_types * 2 + 2 parameters");
  if (t.size() == 5) {
    double r = to_double(t[1], "rc_radial"), a = to_double(t[2], "rc_angular");
    for (int n = 0; n < num_types; ++n) {
      rcR[n] = r;
      rcA[n] = a;
    }
  } else {
    for (int n = 0; n < num_types; ++n) {
      rcR[n] = to_double(t[1 + 2 * n], "rc_radial");
      rcA[n] = to_double(t[2 + 2 * n], "rc_angular");
    }
  }
  rcR_max = rcA_max = 0.0;
  for (int n = 0; n < num_types; ++n) {
    if (!(rcR[n] > 0) || !(rcA[n] > 0)) bad("cutoffs must be positive");
    if (rcA[n] > rcR[n]) bad("the angular cutoff may not exceed the radial cutoff");
    rcR_max = std::max(rcR_max, rcR[n]);
    rcA_max = std::max(rcA_max, rcA[n]);
  }
  int mnr = to_int(t[t.size() - 2], "MN_radial");
  int mna = to_int(t[t.size() - 1], "MN_angular");
  if (mnr > 819) bad("the maximum number of neighbors exceeds 819");
  if (mnr < 1 || mna < 1) bad("MN_radial and MN_angular must be positive");
  MN_R = (int)std::ceil(mnr * 1.25);
  MN_A = (int)std::ceil(mna * 1.25);

  t = line_tokens(in);
  if (t.size() != 3 || t[0] != "n_max") bad("this line should be n_max n_max_radial n_max_angular");
  nR = to_int(t[1], "n_max_radial") + 1;
  nA = to_int(t[2], "n_max_angular") + 1;
  t = line_tokens(in);
  if (t.size() != 3 || t[0] != "basis_size") bad("this line should be basis_size basis_size_radial basis_size_angular");
  kR = to_int(t[1], "basis_size_radial") + 1;
  kA = to_int(t[2], "basis_size_angular") + 1;
  if (nR < 1 || nA < 1 || kR < 1 || kA < 1) bad("n_max/basis_size must be >= 0");
  if (nR > kMaxNR || nA > kMaxNA || kR > kMaxK || kA > kMaxK) bad("n_max/basis_size larger than supported");
  t = line_tokens(in);
  if (t.size() < 4 || t[0] != "l_max") bad("this line should be l_max l_max_3body has_q_222 has_q_1111 [...]");
  L_max = to_int(t[1], "l_max");
  has222 = to_int(t[2], "l_max_4body") != 0;
  has1111 = to_int(t[3], "l_max_5body") != 0;
  for (size_t k = 4; k < t.size(); ++k)
    if (to_int(t[k], "many-body flag") != 0) bad("q_112/q_123/q_233/q_134 invariants are not supported");
  if (L_max < 1 || L_max > kMaxL) bad("l_max_3body must be 1.." + std::to_string(kMaxL));
  if (has222 && L_max < 2) bad("q_222 needs l_max_3body >= 2");
  num_L = L_max + has222 + has1111;
  dim = nR + nA * num_L;
  if (dim > kMaxDim) bad("descriptor dimension larger than supported");
  t = line_tokens(in);
  if (t.size() != 3 || t[0] != "ANN") bad("this line should be ANN num_neurons 0");
  H = to_int(t[1], "num_neurons");
  if (H < 1 || H > kMaxNeurons) bad("num_neurons out of range");
  num_para_ann = version == 4 ? (dim + 2) * H * num_types + 1 : ((dim + 2) * H + 1) * num_types + 1;
  num_para_desc = num_types * num_types * (nR * kR + nA * kA);
  const int total = num_para_ann + num_para_desc + dim;
  params.resize(total);
  for (int i = 0; i < total; ++i) {
    t = line_tokens(in);
    if (t.empty()) bad("parameter " + std::to_string(i) + " of " + std::to_string(total) + " is missing");
    params[i] = to_double(t[0], "parameter");
  }
  if (zbl_flexible) {
    const int nz = 10 * (num_types * (num_types + 1)) / 2;
    zbl_para.resize(nz);
    for (int i = 0; i < nz; ++i) {
      t = line_tokens(in);
      if (t.empty()) bad("flexible ZBL parameter missing");
      zbl_para[i] = to_double(t[0], "zbl parameter");
    }
  }
}

NepModel::Image NepModel::device_image() const
{
  Image im;
  const int T = num_types, D = dim;
  std::vector<float>& v = im.data;
  const double* p = params.data();
  // ANN: per-type w0 [H][D], b0 [H], w1 [H] (+1 NEP5 bias), shared b1
  im.off_w0 = 0;
  im.off_b0 = T * H * D;
  im.off_w1 = im.off_b0 + T * H;
  im.off_w1x = im.off_w1 + T * H;
  im.off_b1 = im.off_w1x + T;
  im.off_cR = im.off_b1 + 1;
  im.off_cA = im.off_cR + T * T * nR * kR;
  im.off_qs = im.off_cA + T * T * nA * kA;
  im.off_zbl = im.off_qs + D;
  v.assign(im.off_zbl + (int)zbl_para.size() + 1, 0.0f);
  const double* q = p;
  for (int t = 0; t < T; ++t) {
    for (int i = 0; i < H * D; ++i) v[im.off_w0 + t * H * D + i] = (float)q[i];
    q += H * D;
    for (int i = 0; i < H; ++i) v[im.off_b0 + t * H + i] = (float)q[i];
    q += H;
    for (int i = 0; i < H; ++i) v[im.off_w1 + t * H + i] = (float)q[i];
    q += H;
    if (version == 5) {
      v[im.off_w1x + t] = (float)q[0];
      q += 1;
    }
  }
  v[im.off_b1] = (float)q[0];
  // descriptor coefficients: file [(n*K+k)*T^2 + t1*T + t2] -> [(t1*T+t2)][n][k]
  const double* c = p + num_para_ann;
  for (int t1 = 0; t1 < T; ++t1)
    for (int t2 = 0; t2 < T; ++t2) {
      for (int n = 0; n < nR; ++n)
        for (int k = 0; k < kR; ++k)
          v[im.off_cR + ((t1 * T + t2) * nR + n) * kR + k] = (float)c[(n * kR + k) * T * T + t1 * T + t2];
      const double* ca = c + T * T * nR * kR;
      for (int n = 0; n < nA; ++n)
        for (int k = 0; k < kA; ++k)
          v[im.off_cA + ((t1 * T + t2) * nA + n) * kA + k] = (float)ca[(n * kA + k) * T * T + t1 * T + t2];
    }
  for (int d = 0; d < D; ++d) v[im.off_qs + d] = (float)p[num_para_ann + num_para_desc + d];
  for (size_t i = 0; i < zbl_para.size(); ++i) v[im.off_zbl + i] = (float)zbl_para[i];
  return im;
}

AngularConstants derive_angular_constants()
{
  AngularConstants a{};
  const double pi = 3.14159265358979323846;
  for (int L = 1; L <= kMaxL; ++L) {
    // 2^L P_L(z) = sum_k (-1)^k C(L,k) C(2L-2k,L) z^(L-2k)
    long long base[kMaxL + 1] = {};
    auto binom = [](int n, int k) {
      long long r = 1;
      for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
      return r;
    };
    for (int k = 0; 2 * k <= L; ++k) {
      long long val = binom(L, k) * binom(2 * L - 2 * k, L);
      base[L - 2 * k] = (k % 2) ? -val : val;
    }
    for (int m = 0; m <= L; ++m) {
      long long d[kMaxL + 1] = {};
      for (int p = m; p <= L; ++p) {
        long long f = 1;
        for (int s = 0; s < m; ++s) f *= (p - s);
        d[p - m] = base[p] * f;
      }
      long long g = 0;
      for (int p = 0; p <= L - m; ++p)
        if (d[p]) g = std::gcd(g, d[p] < 0 ? -d[p] : d[p]);
      for (int p = 0; p <= L - m; ++p) a.zc[L][m][p] = (float)(d[p] / g);
      const double alpha = (double)g / std::ldexp(1.0, L);
      double fr = 1.0;
      for (int s = L - m + 1; s <= L + m; ++s) fr /= (double)s;
      const double w = (2.0 * L + 1.0) / (4.0 * pi) * fr * alpha * alpha;
      const int b = L * L - 1;
      if (m == 0) {
        a.c3[b] = (float)w;
      } else {
        a.c3[b + 2 * m - 1] = (float)(2.0 * w);
        a.c3[b + 2 * m] = (float)(2.0 * w);
      }
    }
  }
  // q_222: Wigner-3j (2 2 2; m1 m2 m3) coupling of three l=2 moments, weights
  // k*sqrt(14)/(448 pi^1.5) for s3^3, s3(s4^2+s5^2), s3(s6^2+s7^2),
  // s6(s5^2-s4^2), s4 s5 s7.
  const double u4 = std::sqrt(14.0) / (448.0 * std::pow(pi, 1.5));
  const double k4[5] = {-5.0, -90.0, 45.0, 270.0, -540.0};
  for (int i = 0; i < 5; ++i) a.c4[i] = (float)(k4[i] * u4);
  // q_1111: sum over intermediate L of (1 1 L)(L 1 1) couplings = 21/(80 pi^2) |s_1|^4
  a.c5 = (float)(21.0 / (80.0 * pi * pi));
  return a;
}

}  // namespace dmgx
