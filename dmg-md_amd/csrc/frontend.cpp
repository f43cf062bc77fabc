// frontend.cpp -- run.in / model.xyz front end and the output writers (see
// frontend.hpp).
#include "frontend.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <numeric>
#include <random>
#include <sstream>
#include <vector>

#include "engine.hpp"

namespace dmgx {

namespace {

// ---------------------------------------------------------------------------
// tokens and numbers
// ---------------------------------------------------------------------------

std::vector<std::string> split(const std::string& line)
{
  std::vector<std::string> t;
  std::istringstream in(line);
  std::string w;
  while (in >> w) t.push_back(w);
  return t;
}

// run.in tokens end at the first token that starts with '#'
std::vector<std::string> command_tokens(const std::string& line)
{
  std::vector<std::string> t;
  for (auto& w : split(line)) {
    if (!w.empty() && w[0] == '#') break;
    t.push_back(w);
  }
  return t;
}

struct Where {
  std::string file;
  size_t line = 0;
  std::string text;
};

[[noreturn]] void fail(const Where& at, const std::string& msg)
{
  std::string m = msg;
  if (!at.file.empty()) m += " at " + at.file + (at.line ? ":" + std::to_string(at.line) : std::string());
  if (!at.text.empty()) m += " [" + at.text + "]";
  throw InputError(m);
}

int to_int(const std::string& s, const Where& at, const char* what)
{
  char* end = nullptr;
  errno = 0;
  const long v = std::strtol(s.c_str(), &end, 0);
  if (errno || end == s.c_str() || *end || v < INT32_MIN || v > INT32_MAX)
    fail(at, std::string(what) + " should be an integer");
  return (int)v;
}

double to_real(const std::string& s, const Where& at, const char* what)
{
  char* end = nullptr;
  errno = 0;
  const double v = std::strtod(s.c_str(), &end);
  if (errno || end == s.c_str() || *end || !std::isfinite(v))
    fail(at, std::string(what) + " should be a finite real number");
  return v;
}

std::string lower(std::string s)
{
  for (char& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

// ---------------------------------------------------------------------------
// run.in
// ---------------------------------------------------------------------------

struct DumpXyz {
  int interval = 0;
  std::string file;
  bool dbl = false;
  bool mass = false, charge = false, velocity = false, force = false, potential = false, virial = false,
       labels = false, unwrapped = false;
  int gmethod = -1, gid = -1;  // `group method id`: only that group's atoms
};

struct Cmd {
  enum Kind { potential, velocity, time_step, ensemble, dump_thermo, dump_xyz, dump_restart, run, correct_velocity } kind;
  Where at;
  std::string text;        // potential file
  double a = 0, b = 0, c = 0, d = 0, e = 0, f = 0;
  int i = 0;               // interval / steps / seed
  int j = -1;              // correct_velocity grouping method (-1: all atoms)
  bool has_seed = false;
  EnsembleConfig ens;
  DumpXyz xyz;
};

Cmd parse_command(const std::vector<std::string>& t, const Where& at)
{
  Cmd c;
  c.at = at;
  const std::string& k = t[0];
  auto need = [&](size_t n, const char* msg) {
    if (t.size() != n) fail(at, msg);
  };
  if (k == "potential") {
    need(2, "potential should have 1 parameter");
    c.kind = Cmd::potential;
    c.text = t[1];
  } else if (k == "velocity") {
    if (t.size() != 2 && t.size() != 4) fail(at, "velocity should have 1 or 3 parameters");
    c.kind = Cmd::velocity;
    c.a = to_real(t[1], at, "initial temperature");
    if (c.a <= 0) fail(at, "initial temperature should be > 0");
    if (t.size() == 4) {
      c.i = to_int(t[3], at, "seed");
      c.has_seed = true;
    }
  } else if (k == "time_step") {
    // time_step dt [maximum distance per step] (run_parser.cpp:184-197)
    if (t.size() != 2 && t.size() != 3) fail(at, "time_step should have 1 or 2 parameters");
    c.kind = Cmd::time_step;
    c.a = to_real(t[1], at, "time_step");
    if (t.size() == 3) {
      c.b = to_real(t[2], at, "maximum distance per step");
      if (c.b <= 0.0) fail(at, "maximum distance per step should be > 0");
    }
  } else if (k == "ensemble") {
    if (t.size() < 2) fail(at, "ensemble type is missing");
    c.kind = Cmd::ensemble;
    const std::string& e = t[1];
    if (e == "nve") {
      need(2, "ensemble nve should have 0 parameters");
      c.ens.kind = Ensemble::nve;
    } else if (e == "nvt_ber" || e == "nvt_nhc" || e == "nvt_bdp") {
      need(5, ("ensemble " + e + " should have 3 parameters").c_str());
      c.ens.kind = e == "nvt_ber" ? Ensemble::nvt_ber : e == "nvt_nhc" ? Ensemble::nvt_nhc : Ensemble::nvt_bdp;
      c.ens.t_init = to_real(t[2], at, "initial temperature");
      c.ens.t_final = to_real(t[3], at, "final temperature");
      c.ens.t_coupling = to_real(t[4], at, "temperature coupling");
      if (c.ens.t_init <= 0 || c.ens.t_final <= 0) fail(at, "initial and final temperatures should be > 0");
      if (c.ens.t_coupling < 1) fail(at, "temperature coupling should be >= 1");
    } else if (e == "npt_ber" || e == "npt_scr") {
      need(8, ("ensemble " + e + " should have 6 parameters (isotropic)").c_str());
      c.ens.kind = e == "npt_ber" ? Ensemble::npt_ber : Ensemble::npt_scr;
      c.ens.t_init = to_real(t[2], at, "initial temperature");
      c.ens.t_final = to_real(t[3], at, "final temperature");
      c.ens.t_coupling = to_real(t[4], at, "temperature coupling");
      c.ens.p_target_gpa = to_real(t[5], at, "target pressure (GPa)");
      c.ens.bulk_modulus_gpa = to_real(t[6], at, "bulk modulus (GPa)");
      c.ens.p_coupling = to_real(t[7], at, "pressure coupling time");
      if (c.ens.t_init <= 0 || c.ens.t_final <= 0) fail(at, "initial and final temperatures should be > 0");
      if (c.ens.t_coupling < 1) fail(at, "temperature coupling should be >= 1");
      if (c.ens.bulk_modulus_gpa <= 0) fail(at, "bulk modulus should be > 0");
      if (c.ens.p_coupling < 1) fail(at, "pressure coupling time should be >= 1");
    } else {
      fail(at, "unsupported ensemble '" + e + "'");
    }
  } else if (k == "dump_thermo") {
    need(2, "dump_thermo should have 1 parameter");
    c.kind = Cmd::dump_thermo;
    c.i = to_int(t[1], at, "thermo dump interval");
    if (c.i <= 0) fail(at, "thermo dump interval should be > 0");
  } else if (k == "dump_restart") {
    need(2, "dump_restart should have 1 parameter");
    c.kind = Cmd::dump_restart;
    c.i = to_int(t[1], at, "restart dump interval");
    if (c.i <= 0) fail(at, "restart dump interval should be > 0");
  } else if (k == "dump_xyz") {
    if (t.size() < 3) fail(at, "dump_xyz should have at least 2 parameters");
    c.kind = Cmd::dump_xyz;
    c.xyz.interval = to_int(t[1], at, "dump interval");
    if (c.xyz.interval <= 0) fail(at, "dump interval should be > 0");
    c.xyz.file = t[2];
    bool prec = false;
    for (size_t q = 3; q < t.size(); ++q) {
      const std::string& w = t[q];
      if (w == "precision") {
        if (prec) fail(at, "option 'precision' is specified more than once in dump_xyz");
        if (q + 1 >= t.size()) fail(at, "option 'precision' requires single or double");
        if (t[q + 1] == "double") c.xyz.dbl = true;
        else if (t[q + 1] != "single") fail(at, "invalid dump_xyz precision '" + t[q + 1] + "'");
        prec = true;
        ++q;
      } else if (w == "mass") {
        c.xyz.mass = true;
      } else if (w == "charge") {
        c.xyz.charge = true;
      } else if (w == "velocity") {
        c.xyz.velocity = true;
      } else if (w == "force") {
        c.xyz.force = true;
      } else if (w == "potential") {
        c.xyz.potential = true;
      } else if (w == "virial") {
        c.xyz.virial = true;
      } else if (w == "group_labels") {
        c.xyz.labels = true;
      } else if (w == "unwrapped_position") {
        c.xyz.unwrapped = true;
      } else if (w == "group") {
        if (c.xyz.gmethod >= 0) fail(at, "option 'group' is specified more than once in dump_xyz");
        if (q + 2 >= t.size()) fail(at, "option 'group' requires a grouping method and group ID");
        c.xyz.gmethod = to_int(t[q + 1], at, "grouping method");
        c.xyz.gid = to_int(t[q + 2], at, "group ID");
        if (c.xyz.gmethod < 0 || c.xyz.gid < 0) fail(at, "grouping method and group ID should be >= 0");
        q += 2;
      } else {
        fail(at, "unrecognized dump_xyz argument '" + w + "'");
      }
    }
  } else if (k == "run") {
    need(2, "run should have 1 parameter");
    c.kind = Cmd::run;
    c.i = to_int(t[1], at, "number of steps");
    if (c.i < 0) fail(at, "number of steps should be >= 0");
  } else if (k == "correct_velocity") {
    // correct_velocity interval [grouping method] (run_parser.cpp:256-271)
    if (t.size() != 2 && t.size() != 3) fail(at, "correct_velocity should have 1 or 2 parameters");
    c.kind = Cmd::correct_velocity;
    c.i = to_int(t[1], at, "velocity correction interval");
    if (c.i < 10) fail(at, "velocity correction interval should be >= 10");
    if (t.size() == 3) {
      c.j = to_int(t[2], at, "velocity correction grouping method");
      if (c.j < 0) fail(at, "velocity correction grouping method should be >= 0");
    }
  } else {
    fail(at, "unsupported command '" + k + "'");
  }
  return c;
}

std::vector<Cmd> parse_run(const std::string& path)
{
  std::ifstream in(path);
  if (!in) throw InputError("cannot open " + path);
  std::vector<Cmd> cmds;
  std::string line;
  size_t n = 0;
  while (std::getline(in, line)) {
    ++n;
    const auto t = command_tokens(line);
    if (t.empty()) continue;
    const Where at{path, n, line};
    if (t.size() > 32) fail(at, "the number of tokens should be <= 32");
    cmds.push_back(parse_command(t, at));
  }
  return cmds;
}

// ---------------------------------------------------------------------------
// model.xyz
// ---------------------------------------------------------------------------

// standard atomic weights (amu), H..Lr, the default masses of model files
// without a mass column
const char* kSym[103] = {
  "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar", "K", "Ca",
  "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y",
  "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd", "Ag", "Cd", "In", "Sn", "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce",
  "Pr", "Nd", "Pm", "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os", "Ir",
  "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn", "Fr", "Ra", "Ac", "Th", "Pa", "U", "Np", "Pu", "Am", "Cm",
  "Bk", "Cf", "Es", "Fm", "Md", "No", "Lr"};
const double kMass[103] = {
  1.008, 4.002602, 6.94, 9.0121831, 10.81, 12.011, 14.007, 15.999, 18.998403163, 20.1797, 22.98976928, 24.305,
  26.9815385, 28.085, 30.973761998, 32.06, 35.45, 39.948, 39.0983, 40.078, 44.955908, 47.867, 50.9415, 51.9961,
  54.938044, 55.845, 58.933194, 58.6934, 63.546, 65.38, 69.723, 72.63, 74.921595, 78.971, 79.904, 83.798, 85.4678,
  87.62, 88.90584, 91.224, 92.90637, 95.95, 98.0, 101.07, 102.9055, 106.42, 107.8682, 112.414, 114.818, 118.71,
  121.76, 127.6, 126.90447, 131.293, 132.90545196, 137.327, 138.90547, 140.116, 140.90766, 144.242, 145.0, 150.36,
  151.964, 157.25, 158.92535, 162.5, 164.93033, 167.259, 168.93422, 173.045, 174.9668, 178.49, 180.94788, 183.84,
  186.207, 190.23, 192.217, 195.084, 196.966569, 200.592, 204.38, 207.2, 208.9804, 210.0, 210.0, 222.0, 223.0,
  226.0, 227.0, 232.0377, 231.03588, 238.02891, 237.0, 244.0, 243.0, 247.0, 247.0, 251.0, 252.0, 257.0, 258.0,
  259.0, 262.0};

struct Model {
  Box box{};
  int N = 0;
  std::vector<std::string> species;
  std::vector<int> type, labels;  // labels: [group method][atom]
  int n_label = 0;
  std::vector<double> mass, charge, pos, vel;  // SoA; velocity in natural units
  bool has_vel = false;
};

// value of key=... in the comment line (key case-insensitive, optional
// blanks around '=', quoted or bare value); "" when absent
std::string assignment(const std::string& line, const std::string& key, const Where& at)
{
  const std::string l = lower(line), k = lower(key);
  size_t p = 0;
  while ((p = l.find(k, p)) != std::string::npos) {
    const bool starts = p == 0 || std::isspace((unsigned char)l[p - 1]);
    size_t q = p + k.size();
    while (q < l.size() && std::isspace((unsigned char)l[q])) ++q;
    if (starts && q < l.size() && l[q] == '=') {
      ++q;
      while (q < l.size() && std::isspace((unsigned char)l[q])) ++q;
      if (q >= l.size()) return "";
      if (line[q] == '"') {
        const size_t e = line.find('"', q + 1);
        if (e == std::string::npos) fail(at, "unterminated quoted value for " + key);
        return line.substr(q + 1, e - q - 1);
      }
      const size_t e = line.find_first_of(" \t\r\n", q);
      return line.substr(q, e == std::string::npos ? std::string::npos : e - q);
    }
    p += k.size();
  }
  return "";
}

Model parse_model(const std::string& path, const std::vector<std::string>& symbols)
{
  std::ifstream in(path);
  if (!in) throw InputError("cannot open " + path);
  std::string l1, l2;
  if (!std::getline(in, l1) || !std::getline(in, l2)) fail({path, 1, l1}, "model.xyz should contain at least two lines");
  const auto c = split(l1);
  if (c.size() != 1) fail({path, 1, l1}, "the first line should have one value");
  Model m;
  m.N = to_int(c[0], {path, 1, l1}, "number of atoms");
  if (m.N < 2) fail({path, 1, l1}, "number of atoms should be >= 2");
  const Where w2{path, 2, l2};
  m.box.pbc[0] = m.box.pbc[1] = m.box.pbc[2] = 1;
  const std::string pbc = assignment(l2, "pbc", w2);
  if (!pbc.empty()) {
    const auto f = split(pbc);
    if (f.size() != 3) fail(w2, "pbc should contain three T/F values");
    for (int d = 0; d < 3; ++d) {
      const std::string v = lower(f[d]);
      if (v != "t" && v != "f") fail(w2, "pbc should contain only T or F");
      m.box.pbc[d] = v == "t";
    }
  }
  const auto lat = split(assignment(l2, "lattice", w2));
  if (lat.size() != 9) fail(w2, "Lattice should contain nine values");
  // Lattice lists a, b, c; h holds them as columns
  const int col[9] = {0, 3, 6, 1, 4, 7, 2, 5, 8};
  for (int k = 0; k < 9; ++k) m.box.h[col[k]] = to_real(lat[k], w2, "lattice value");
  // Properties=name:type:width triples
  const std::string schema = lower(assignment(l2, "properties", w2));
  std::vector<std::string> f;
  {
    size_t s = 0;
    for (;;) {
      const size_t e = schema.find(':', s);
      f.push_back(schema.substr(s, e == std::string::npos ? std::string::npos : e - s));
      if (e == std::string::npos) break;
      s = e + 1;
    }
  }
  if (f.empty() || f.size() % 3 != 0) fail(w2, "Properties should contain name:type:width triples");
  int off = 0, o_sp = -1, o_pos = -1, o_mass = -1, o_q = -1, o_vel = -1, o_grp = -1;
  for (size_t k = 0; k < f.size(); k += 3) {
    const int width = to_int(f[k + 2], w2, "property width");
    if (width <= 0 || f[k + 1].empty()) fail(w2, "property width should be > 0");
    const std::string& name = f[k];
    if (name == "species") o_sp = off;
    else if (name == "pos") o_pos = off;
    else if (name == "mass") o_mass = off;
    else if (name == "charge") o_q = off;
    else if (name == "vel") o_vel = off;
    else if (name == "group") {
      o_grp = off;
      m.n_label = width;
    }
    off += width;
  }
  if (o_sp < 0 || o_pos < 0) fail(w2, "Properties should contain species and pos");
  const int N = m.N;
  m.has_vel = o_vel >= 0;
  m.species.resize(N);
  m.type.resize(N);
  m.mass.resize(N);
  m.charge.assign(N, 0.0);
  m.pos.resize(3 * (size_t)N);
  m.vel.assign(3 * (size_t)N, 0.0);
  m.labels.assign((size_t)m.n_label * N, 0);
  for (int i = 0; i < N; ++i) {
    std::string line;
    const Where at{path, (size_t)i + 3, ""};
    if (!std::getline(in, line)) fail(at, "model.xyz has fewer atom lines than declared");
    const Where wl{path, (size_t)i + 3, line};
    const auto t = split(line);
    if ((int)t.size() != off) fail(wl, "number of columns does not match Properties");
    m.species[i] = t[o_sp];
    const auto it = std::find(symbols.begin(), symbols.end(), m.species[i]);
    if (it == symbols.end()) fail(wl, "species '" + m.species[i] + "' is not allowed by the potential");
    m.type[i] = (int)(it - symbols.begin());
    for (int d = 0; d < 3; ++d) m.pos[(size_t)d * N + i] = to_real(t[o_pos + d], wl, "position");
    if (o_mass >= 0) {
      m.mass[i] = to_real(t[o_mass], wl, "atom mass");
    } else {
      int e = -1;
      for (int k = 0; k < 103; ++k)
        if (m.species[i] == kSym[k]) e = k;
      if (e < 0) fail(wl, "no default mass for species '" + m.species[i] + "'");
      m.mass[i] = kMass[e];
    }
    if (m.mass[i] <= 0) fail(wl, "atom mass should be > 0");
    if (o_q >= 0) m.charge[i] = (float)to_real(t[o_q], wl, "charge");
    if (o_vel >= 0)
      for (int d = 0; d < 3; ++d) m.vel[(size_t)d * N + i] = to_real(t[o_vel + d], wl, "velocity") * kTimeUnit;
    for (int g = 0; g < m.n_label; ++g) {
      const int v = to_int(t[o_grp + g], wl, "group label");
      if (v < 0 || v >= N) fail(wl, "group label should be >= 0 and < N");
      m.labels[(size_t)g * N + i] = v;
    }
  }
  return m;
}

// Maxwell-Boltzmann velocities at temperature T: Gaussian components from
// std::mt19937_64, centre-of-mass momentum removed, rescaled to exactly T
// (a fresh stream of this build; not the reference's generator)
// Removes the total linear momentum of `atoms`, then their rigid rotation
// about their centre of mass: omega = I^-1 L, v -= omega x (r - r_c); the
// rotation is kept when the inertia tensor is singular (runtime.cu:979-1076).
void correct_velocity_subset(const std::vector<double>& mass, const double* pos, double* vel, int N,
                             const std::vector<int>& atoms)
{
  if (atoms.empty()) return;
  double p[3] = {0, 0, 0}, mt = 0.0;
  for (int a : atoms) {
    mt += mass[a];
    for (int d = 0; d < 3; ++d) p[d] += mass[a] * vel[(size_t)d * N + a];
  }
  if (mt != 0.0) {
    for (double& x : p) x /= mt;
    for (int a : atoms)
      for (int d = 0; d < 3; ++d) vel[(size_t)d * N + a] -= p[d];
  }
  double c[3] = {0, 0, 0};
  mt = 0.0;
  for (int a : atoms) {
    mt += mass[a];
    for (int d = 0; d < 3; ++d) c[d] += pos[(size_t)d * N + a] * mass[a];
  }
  for (double& x : c) x /= mt;
  double L[3] = {0, 0, 0}, I[3][3] = {};
  for (int a : atoms) {
    const double m = mass[a];
    const double dx = pos[a] - c[0], dy = pos[(size_t)N + a] - c[1], dz = pos[2 * (size_t)N + a] - c[2];
    const double vx = vel[a], vy = vel[(size_t)N + a], vz = vel[2 * (size_t)N + a];
    L[0] += m * (dy * vz - dz * vy);
    L[1] += m * (dz * vx - dx * vz);
    L[2] += m * (dx * vy - dy * vx);
    I[0][0] += m * (dy * dy + dz * dz);
    I[1][1] += m * (dx * dx + dz * dz);
    I[2][2] += m * (dx * dx + dy * dy);
    I[0][1] -= m * dx * dy;
    I[1][2] -= m * dy * dz;
    I[0][2] -= m * dx * dz;
  }
  I[1][0] = I[0][1];
  I[2][1] = I[1][2];
  I[2][0] = I[0][2];
  // cofactor inverse (I is symmetric)
  double C[3][3];
  for (int r = 0; r < 3; ++r)
    for (int q = 0; q < 3; ++q) {
      const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, q1 = (q + 1) % 3, q2 = (q + 2) % 3;
      C[q][r] = I[r1][q1] * I[r2][q2] - I[r1][q2] * I[r2][q1];
    }
  const double det = I[0][0] * C[0][0] + I[0][1] * C[1][0] + I[0][2] * C[2][0];
  if (det > -1.0e-10 && det < 1.0e-10) return;
  double w[3] = {0, 0, 0};
  for (int r = 0; r < 3; ++r)
    for (int q = 0; q < 3; ++q) w[r] += C[r][q] / det * L[q];
  for (int a : atoms) {
    const double dx = pos[a] - c[0], dy = pos[(size_t)N + a] - c[1], dz = pos[2 * (size_t)N + a] - c[2];
    vel[a] -= w[1] * dz - w[2] * dy;
    vel[(size_t)N + a] -= w[2] * dx - w[0] * dz;
    vel[2 * (size_t)N + a] -= w[0] * dy - w[1] * dx;
  }
}

// The `velocity T [seed S]` initialisation (runtime.cu:1078-1109): every
// component uniform in [-1, 1] from the C library's rand() (seeded per
// component with S + 3 i + axis when a seed is given, so the draw is the
// reference's on the same C library), momentum and rotation removed, then
// scaled to T over 3N degrees of freedom.
void random_velocities(Model& m, double T, bool has_seed, int seed)
{
  const int N = m.N;
  for (int i = 0; i < N; ++i)
    for (int d = 0; d < 3; ++d) {
      if (has_seed) std::srand((unsigned)(seed + i * 3 + d));
      m.vel[(size_t)d * N + i] = -1.0 + (std::rand() * 2.0) / (double)RAND_MAX;
    }
  std::vector<int> all((size_t)N);
  for (int i = 0; i < N; ++i) all[i] = i;
  correct_velocity_subset(m.mass, m.pos.data(), m.vel.data(), N, all);
  double t = 0.0;
  for (int i = 0; i < N; ++i) {
    const double vx = m.vel[i], vy = m.vel[(size_t)N + i], vz = m.vel[2 * (size_t)N + i];
    t += m.mass[i] * (vx * vx + vy * vy + vz * vz);
  }
  t /= 3.0 * kBoltzmann * N;
  const double f = std::sqrt(T / t);
  for (double& v : m.vel) v *= f;
}

// ---------------------------------------------------------------------------
// writers
// ---------------------------------------------------------------------------

struct Snapshot {
  std::vector<double> pos, vel, force, pe, vir, unwrapped;
};

void write_tensor(FILE* f, const char* name, const char* fmt, const double* v)
{
  std::fprintf(f, " %s=\"", name);
  for (int c = 0; c < 9; ++c) std::fprintf(f, c == 0 ? fmt + 1 : fmt, v[c]);
  std::fprintf(f, "\"");
}

FILE* open_or_throw(const std::string& path, const char* mode)
{
  FILE* f = std::fopen(path.c_str(), mode);
  if (!f) throw std::runtime_error("cannot open output '" + path + "'");
  return f;
}

void write_thermo_header(const std::string& dir, int interval, int N, double dt_natural)
{
  FILE* f = open_or_throw(dir + "/thermo.out", "a");
  std::fprintf(f, "# dump_thermo %d\n# format_version 1\n# num_atoms %d\n# dt_output %.10e fs\n", interval, N,
               dt_natural * interval * kTimeUnit);
  std::fprintf(f, "# columns T KE PE sxx syy szz syz sxz sxy ax ay az bx by bz cx cy cz\n");
  std::fclose(f);
}

void write_thermo_row(const std::string& dir, const double* th, int N, const Box& box)
{
  FILE* f = open_or_throw(dir + "/thermo.out", "a");
  const double ke = 1.5 * N * kBoltzmann * th[0];
  const double P = kPressureUnit;
  std::fprintf(f, "%20.10e%20.10e%20.10e%20.10e%20.10e%20.10e%20.10e%20.10e%20.10e", th[0], ke, th[1], th[2] * P,
               th[3] * P, th[4] * P, th[7] * P, th[6] * P, th[5] * P);
  const double* h = box.h;
  std::fprintf(f, "%20.10e%20.10e%20.10e%20.10e%20.10e%20.10e%20.10e%20.10e%20.10e\n", h[0], h[3], h[6], h[1], h[4],
               h[7], h[2], h[5], h[8]);
  std::fclose(f);
}

void write_xyz(const std::string& dir, const DumpXyz& d, int steps_done, double time_natural, const Box& box,
               const Model& m, const Snapshot& s, const double* th)
{
  std::string name = d.file;
  const bool separate = !name.empty() && name.back() == '*';
  if (separate) {
    name.pop_back();
    name += std::to_string(steps_done);
  }
  FILE* f = open_or_throw(dir + "/" + name, separate ? "w" : "a");
  const char* fmt = d.dbl ? " %.17g" : " %.9g";
  const int N = m.N;
  const double* h = box.h;
  // `group method id`: only that group's atoms (the header sums stay over all
  // atoms; output_order, runtime.cu:1111-1137)
  std::vector<int> order;
  const int* lab = d.gmethod >= 0 ? m.labels.data() + (size_t)d.gmethod * N : nullptr;
  for (int i = 0; i < N; ++i)
    if (!lab || lab[i] == d.gid) order.push_back(i);
  std::fprintf(f, "%zu\nTime=%.8f", order.size(), time_natural * kTimeUnit);
  std::fprintf(f, " pbc=\"%c %c %c\"", box.pbc[0] ? 'T' : 'F', box.pbc[1] ? 'T' : 'F', box.pbc[2] ? 'T' : 'F');
  const double lat[9] = {h[0], h[3], h[6], h[1], h[4], h[7], h[2], h[5], h[8]};
  write_tensor(f, "Lattice", fmt, lat);
  std::fprintf(f, " energy=");
  std::fprintf(f, fmt + 1, th[1]);
  double vs[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < N; ++i)
    for (int c = 0; c < 6; ++c) vs[c] += s.vir[(size_t)c * N + i];
  const double vir[9] = {vs[0], vs[3], vs[4], vs[3], vs[1], vs[5], vs[4], vs[5], vs[2]};
  write_tensor(f, "virial", fmt, vir);
  const double st[9] = {th[2], th[5], th[6], th[5], th[3], th[7], th[6], th[7], th[4]};
  write_tensor(f, "stress", fmt, st);
  std::fprintf(f, " Properties=species:S:1:pos:R:3");
  if (d.mass) std::fprintf(f, ":mass:R:1");
  if (d.charge) std::fprintf(f, ":charge:R:1");
  if (d.velocity) std::fprintf(f, ":vel:R:3");
  if (d.force) std::fprintf(f, ":forces:R:3");
  if (d.potential) std::fprintf(f, ":energy_atom:R:1");
  if (d.unwrapped) std::fprintf(f, ":unwrapped_position:R:3");
  if (d.virial) std::fprintf(f, ":virial:R:9");
  if (d.labels) std::fprintf(f, ":group:I:%d", m.n_label);
  std::fprintf(f, "\n");
  // per-atom virial as a row-major 3x3 from the xx yy zz xy xz yz yx zx zy order
  const int vi[9] = {0, 3, 4, 6, 1, 5, 7, 8, 2};
  for (int i : order) {
    std::fprintf(f, "%s", m.species[i].c_str());
    for (int c = 0; c < 3; ++c) std::fprintf(f, fmt, s.pos[(size_t)c * N + i]);
    if (d.mass) std::fprintf(f, fmt, m.mass[i]);
    if (d.charge) std::fprintf(f, fmt, m.charge[i]);
    if (d.velocity)
      for (int c = 0; c < 3; ++c) std::fprintf(f, fmt, s.vel[(size_t)c * N + i] / kTimeUnit);
    if (d.force)
      for (int c = 0; c < 3; ++c) std::fprintf(f, fmt, s.force[(size_t)c * N + i]);
    if (d.potential) std::fprintf(f, fmt, s.pe[i]);
    if (d.unwrapped)
      for (int c = 0; c < 3; ++c) std::fprintf(f, fmt, s.unwrapped[(size_t)c * N + i]);
    if (d.virial)
      for (int c : vi) std::fprintf(f, fmt, s.vir[(size_t)c * N + i]);
    if (d.labels)
      for (int g = 0; g < m.n_label; ++g) std::fprintf(f, " %d", m.labels[(size_t)g * N + i]);
    std::fprintf(f, "\n");
  }
  std::fclose(f);
}

void write_restart(const std::string& dir, const Box& box, const Model& m, const Snapshot& s)
{
  FILE* f = open_or_throw(dir + "/restart.xyz", "w");
  const int N = m.N;
  const double* h = box.h;
  std::fprintf(f, "%d\n", N);
  std::fprintf(f, "pbc=\"%c %c %c\" ", box.pbc[0] ? 'T' : 'F', box.pbc[1] ? 'T' : 'F', box.pbc[2] ? 'T' : 'F');
  std::fprintf(f, "Lattice=\"%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\" ", h[0], h[3], h[6], h[1], h[4],
               h[7], h[2], h[5], h[8]);
  if (m.n_label == 0) std::fprintf(f, "Properties=species:S:1:pos:R:3:mass:R:1:vel:R:3\n");
  else std::fprintf(f, "Properties=species:S:1:pos:R:3:mass:R:1:vel:R:3:group:I:%d\n", m.n_label);
  for (int i = 0; i < N; ++i) {
    std::fprintf(f, "%s %.17g %.17g %.17g %.17g %.17g %.17g %.17g ", m.species[i].c_str(), s.pos[i],
                 s.pos[(size_t)N + i], s.pos[2 * (size_t)N + i], m.mass[i], s.vel[i] / kTimeUnit,
                 s.vel[(size_t)N + i] / kTimeUnit, s.vel[2 * (size_t)N + i] / kTimeUnit);
    for (int g = 0; g < m.n_label; ++g) std::fprintf(f, "%d ", m.labels[(size_t)g * N + i]);
    std::fprintf(f, "\n");
  }
  std::fclose(f);
}

void write_neighbor(const std::string& dir, long long call, NepEngine& nep)
{
  int mr = 0, ma = 0, mv = 0, rb = 0;
  nep.stats(&mr, &ma, &mv, &rb);
  std::ofstream out(dir + "/neighbor.out", std::ios_base::app);
  out << "Neighbor info at step " << call << ": radial(max=" << nep.model().MN_R << ",actual=" << mr
      << "), angular(max=" << nep.model().MN_A << ",actual=" << ma << ")." << std::endl;
}

// The whole program is checked before anything runs, so an unsupported
// later run cannot leave a partial trajectory (validate_runtime_ensembles,
// run_parser.cpp:344-386): correct_velocity and an adaptive time_step only
// with nve / nvt_ber, and a finite positive time_step for the thermostats
// and barostats that integrate their own state.  n_label < 0 skips the
// grouping-method range (the model is not read yet).
void validate_program(const std::vector<Cmd>& cmds, int n_label)
{
  Ensemble kind = Ensemble::nve;
  bool adaptive = false;
  int correct_group = -2;  // -2: no correct_velocity since the last run
  double dt_fs = 1.0;
  for (const Cmd& c : cmds) {
    if (c.kind == Cmd::time_step) {
      dt_fs = c.a;
      adaptive = c.b > 0.0;
    } else if (c.kind == Cmd::correct_velocity) {
      correct_group = c.j;
    } else if (c.kind == Cmd::ensemble) {
      kind = c.ens.kind;
    } else if (c.kind == Cmd::run) {
      const bool npt = kind == Ensemble::npt_ber || kind == Ensemble::npt_scr;
      const bool bdp_nhc = kind == Ensemble::nvt_bdp || kind == Ensemble::nvt_nhc;
      const std::string name = kind == Ensemble::npt_ber   ? "npt_ber"
                               : kind == Ensemble::npt_scr ? "npt_scr"
                               : kind == Ensemble::nvt_bdp ? "nvt_bdp"
                                                           : "nvt_nhc";
      const bool correct = correct_group != -2;
      if ((npt || bdp_nhc) && correct) fail(c.at, "ensemble " + name + " does not support correct_velocity");
      if ((npt || bdp_nhc) && adaptive) fail(c.at, "ensemble " + name + " does not support adaptive time_step");
      if ((npt || bdp_nhc) && !(dt_fs > 0.0 && std::isfinite(dt_fs)))
        fail(c.at, "ensemble " + name + " requires a finite positive time_step");
      if (correct && n_label >= 0 && correct_group >= n_label)
        fail(c.at, "correct_velocity grouping method is out of range");
      adaptive = false;
      correct_group = -2;
    }
  }
}

void hcheck(hipError_t e, const char* what)
{
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

Snapshot snapshot(MdState& md, int N)
{
  double *pos, *vel, *force, *pe, *vir, *mass;
  int* type;
  md.arrays(&pos, &vel, &force, &pe, &vir, &mass, &type);
  Snapshot s;
  s.pos.resize(3 * (size_t)N);
  s.vel.resize(3 * (size_t)N);
  s.force.resize(3 * (size_t)N);
  s.pe.resize(N);
  s.vir.resize(9 * (size_t)N);
  hcheck(hipDeviceSynchronize(), "snapshot");
  hcheck(hipMemcpy(s.pos.data(), pos, s.pos.size() * 8, hipMemcpyDeviceToHost), "snapshot");
  hcheck(hipMemcpy(s.vel.data(), vel, s.vel.size() * 8, hipMemcpyDeviceToHost), "snapshot");
  hcheck(hipMemcpy(s.force.data(), force, s.force.size() * 8, hipMemcpyDeviceToHost), "snapshot");
  hcheck(hipMemcpy(s.pe.data(), pe, s.pe.size() * 8, hipMemcpyDeviceToHost), "snapshot");
  hcheck(hipMemcpy(s.vir.data(), vir, s.vir.size() * 8, hipMemcpyDeviceToHost), "snapshot");
  if (md.unwrapped()) {
    s.unwrapped.resize(3 * (size_t)N);
    hcheck(hipMemcpy(s.unwrapped.data(), md.unwrapped(), s.unwrapped.size() * 8, hipMemcpyDeviceToHost), "snapshot");
  }
  return s;
}

}  // namespace

void run_directory(const std::string& dir, const std::string& run_file, const std::string& model_file)
{
  const std::vector<Cmd> cmds = parse_run(dir + "/" + run_file);
  std::string potential;
  for (const Cmd& c : cmds)
    if (c.kind == Cmd::potential) {
      if (!potential.empty()) fail(c.at, "multiple potential commands are not supported");
      potential = c.text;
    }
  if (potential.empty()) throw InputError(dir + "/" + run_file + " does not contain a potential command");
  validate_program(cmds, -1);  // before any device work
  NepEngine nep(potential[0] == '/' ? potential : dir + "/" + potential);
  Model m = parse_model(dir + "/" + model_file, nep.model().symbols);
  validate_program(cmds, m.n_label);
  // no velocities in the model: 300 K (the reference seeds this draw from the
  // clock; here it is seeded, so runs repeat)
  if (!m.has_vel) random_velocities(m, 300.0, true, 12345678);
  std::unique_ptr<MdState> md(new MdState(&nep, m.box, m.N, m.type.data(), m.mass.data(), m.pos.data(), m.vel.data()));

  double dt_fs = 1.0;
  double max_distance = 0.0;      // adaptive time_step (A per step), this run only
  int correct_every = 0, correct_group = -1;  // correct_velocity, this run only
  EnsembleConfig ens;
  int thermo_every = 0, restart_every = 0;
  std::vector<DumpXyz> xyz;
  double time_natural = 0.0;
  long long next_neighbor_call = 0;
  for (const Cmd& c : cmds) {
    switch (c.kind) {
      case Cmd::potential:
        break;
      case Cmd::velocity:
        if (!m.has_vel) {
          random_velocities(m, c.a, c.has_seed, c.i);
          double* vel;
          md->arrays(nullptr, &vel, nullptr, nullptr, nullptr, nullptr, nullptr);
          hcheck(hipMemcpy(vel, m.vel.data(), m.vel.size() * 8, hipMemcpyHostToDevice), "velocity upload");
        }
        break;
      case Cmd::time_step:
        dt_fs = c.a;
        max_distance = c.b;
        break;
      case Cmd::correct_velocity:
        correct_every = c.i;
        correct_group = c.j;
        break;
      case Cmd::ensemble:
        ens = c.ens;
        break;
      case Cmd::dump_thermo:
        if (thermo_every) fail(c.at, "multiple dump_thermo commands within one run");
        thermo_every = c.i;
        break;
      case Cmd::dump_restart:
        if (restart_every) fail(c.at, "multiple dump_restart commands within one run");
        restart_every = c.i;
        break;
      case Cmd::dump_xyz:
        if (c.xyz.labels && m.n_label == 0) fail(c.at, "cannot output group labels without model groups");
        if (c.xyz.gmethod >= 0) {
          if (c.xyz.gmethod >= m.n_label) fail(c.at, "dump_xyz grouping method is out of range");
          const int* lab = m.labels.data() + (size_t)c.xyz.gmethod * m.N;
          if (c.xyz.gid > *std::max_element(lab, lab + m.N)) fail(c.at, "dump_xyz group ID is out of range");
        }
        if (c.xyz.unwrapped) md->enable_unwrapped();
        xyz.push_back(c.xyz);
        break;
      case Cmd::run: {
        const int steps = c.i;
        const double dt_natural = dt_fs / kTimeUnit;

        if (thermo_every) write_thermo_header(dir, thermo_every, m.N, dt_natural);
        int every = thermo_every;
        for (const DumpXyz& d : xyz) every = every ? std::gcd(every, d.interval) : d.interval;
        if (restart_every) every = every ? std::gcd(every, restart_every) : restart_every;
        const double t0 = time_natural;
        int done_before = 0;
        double t_now = t0;
        const bool adaptive = max_distance > 0.0;
        auto on_output = [&](int done, const double* th) {
          if (adaptive) {
            t_now = t0 + md->elapsed();
          } else {
            for (int k = done_before; k < done; ++k) t_now += dt_natural;
          }
          done_before = done;
          if (md->force_calls() - 1 >= next_neighbor_call) {
            write_neighbor(dir, next_neighbor_call, nep);
            next_neighbor_call += 1000;
          }
          const bool want_thermo = thermo_every && done % thermo_every == 0;
          bool want_snap = restart_every && done % restart_every == 0;
          for (const DumpXyz& d : xyz) want_snap = want_snap || done % d.interval == 0;
          Snapshot s;
          if (want_snap) s = snapshot(*md, m.N);
          const Box box = md->box();
          if (want_thermo) write_thermo_row(dir, th, m.N, box);
          for (const DumpXyz& d : xyz)
            if (done % d.interval == 0) write_xyz(dir, d, done, t_now, box, m, s, th);
          if (restart_every && done % restart_every == 0) write_restart(dir, box, m, s);
        };
        if (steps > 0) {
          double th[8];
          MdState::RunOptions opt;
          opt.every = every > 0 ? every : steps;
          opt.on_output = on_output;
          opt.max_distance = max_distance;
          if (correct_every > 0) {
            // correct_device_velocity (runtime.cu:620-656): per group of the
            // chosen grouping method, or all atoms
            opt.correct_every = correct_every;
            opt.on_correct = [&](int) {
              double *pos, *vel;
              md->arrays(&pos, &vel, nullptr, nullptr, nullptr, nullptr, nullptr);
              std::vector<double> hp(3 * (size_t)m.N), hv(3 * (size_t)m.N);
              hcheck(hipMemcpy(hp.data(), pos, hp.size() * 8, hipMemcpyDeviceToHost), "position download");
              hcheck(hipMemcpy(hv.data(), vel, hv.size() * 8, hipMemcpyDeviceToHost), "velocity download");
              if (correct_group < 0) {
                std::vector<int> all((size_t)m.N);
                for (int i = 0; i < m.N; ++i) all[i] = i;
                correct_velocity_subset(m.mass, hp.data(), hv.data(), m.N, all);
              } else {
                const int* lab = m.labels.data() + (size_t)correct_group * m.N;
                const int gmax = *std::max_element(lab, lab + m.N);
                for (int g = 0; g <= gmax; ++g) {
                  std::vector<int> sub;
                  for (int i = 0; i < m.N; ++i)
                    if (lab[i] == g) sub.push_back(i);
                  correct_velocity_subset(m.mass, hp.data(), hv.data(), m.N, sub);
                }
              }
              hcheck(hipMemcpy(vel, hv.data(), hv.size() * 8, hipMemcpyHostToDevice), "velocity upload");
            };
          }
          md->run(steps, dt_fs, ens, 0u, th, opt);
        } else {
          double th[8];
          md->run(0, dt_fs, ens, 0u, th);
        }
        if (next_neighbor_call == 0 && md->force_calls() > 0) {
          write_neighbor(dir, 0, nep);
          next_neighbor_call = 1000;
        }
        if (adaptive) {
          t_now = t0 + (steps > 0 ? md->elapsed() : 0.0);
        } else {
          for (int k = done_before; k < steps; ++k) t_now += dt_natural;
        }
        time_natural = t_now;
        thermo_every = restart_every = 0;
        xyz.clear();
        // both runtimes clear these at each run boundary (runtime.cu:950-952)
        max_distance = 0.0;
        correct_every = 0;
        correct_group = -1;
        break;
      }
    }
  }
}

}  // namespace dmgx
