// ensemble.hpp -- host FP64 thermostat / barostat scalars.
//
// Restates, with the same arithmetic order so results are bit-identical to
// the reference fixtures (tests/thermostats/fixtures/{svr,nhc}.json):
//   SvrState, gasdev, gamdev, resamplekin_sumnoises, resamplekin
//     (src/gpumd_compat/thermostat_svr.cpp:29-124; Bussi, Donadio &
//      Parrinello 2007 canonical sampling by velocity rescaling with the
//      Numerical-Recipes Gaussian/Gamma deviates)
//   make_nhc_state, nhc_half_step (thermostat_nhc.cpp:30-117; Tuckerman's
//     4-chain Nose-Hoover with a 7-term Suzuki-Yoshida split and 4 RESPA steps)
//   npt_scr_isotropic_scale (npt_scr.cpp:21-34; Bernetti & Bussi 2020)
//   EnsembleController decisions (src/ensemble_control.cpp:57-304).
#pragma once

#include <array>
#include <cstdint>
#include <optional>
#include <random>

namespace dmgx {

constexpr double kBoltzmann = 8.617343e-5;          // eV/K (common.cuh:55)
constexpr double kPressureUnit = 1.602177e+2;       // natural -> GPa (common.cuh:60)
constexpr double kTimeUnit = 1.018051e+1;           // natural -> fs (common.cuh:61)
constexpr std::uint32_t kDefaultSeed = 12345678u;   // ensemble_control.hpp:12

struct Rng {
  explicit Rng(std::uint32_t seed) : engine(seed) {}
  std::mt19937 engine;
  int iset = 0;
  double gset = 0.0;
  double uniform()
  {
    std::uniform_real_distribution<double> u(0, 1);
    return u(engine);
  }
};

double gasdev(Rng& r);
double gamdev(int ia, Rng& r);
double resamplekin_sumnoises(int nn, Rng& r);
double resamplekin(double kk, double sigma, int ndeg, double taut, Rng& r);

struct NhcChain {
  std::array<double, 4> position{}, momentum{}, mass{};
};
NhcChain make_nhc_chain(int atom_count, double temperature, double damping, double dt);
double nhc_half_step(NhcChain& c, double ek2, double kT, double dN, double dt_half);

double npt_scr_scale(double p_target, double p_coupling, double p_inst, double t_target, double volume,
                     double gaussian);

enum class Ensemble { nve = 0, nvt_ber = 1, nvt_bdp = 2, nvt_nhc = 3, npt_ber = 4, npt_scr = 5 };

struct EnsembleConfig {
  Ensemble kind = Ensemble::nve;
  double t_init = 0, t_final = 0, t_coupling = 0;
  double p_target_gpa = 0, bulk_modulus_gpa = 0, p_coupling = 0;
  std::uint32_t seed = kDefaultSeed;
};

// Per-segment controller with the decisions of EnsembleController.
class Controller {
 public:
  Controller(const EnsembleConfig& cfg, int atom_count, double dt_natural);
  bool needs_pre_force_ke() const { return cfg_.kind == Ensemble::nvt_nhc; }
  bool needs_thermo() const { return cfg_.kind != Ensemble::nve; }
  bool has_barostat() const { return cfg_.kind == Ensemble::npt_ber || cfg_.kind == Ensemble::npt_scr; }
  double target_temperature(int step, int steps) const;
  // NHC pre-force half step on the kinetic energy (eV)
  std::optional<double> before_force(double dt_natural, double kinetic_energy, int step, int steps);
  // thermo[0] = T; returns the velocity factor if one applies
  std::optional<double> after_force(const double* thermo, int step, int steps);
  // isotropic box factor (thermo before velocity scaling)
  double barostat_scale(const double* thermo, double volume, int step, int steps);
  // The state-independent Bussi draws of one step, in resamplekin's order:
  // the Gaussian rr, then the sum of ndeg-1 squared Gaussians.
  void draw_bdp(double* rr, double* sumsq)
  {
    *rr = gasdev(rng_);
    *sumsq = resamplekin_sumnoises(3 * n_ - 1, rng_);
  }
  const Rng& rng() const { return rng_; }
  void set_rng(const Rng& r) { rng_ = r; }

 private:
  EnsembleConfig cfg_;
  int n_ = 0;
  double dt_ = 0.0;
  double p_target_ = 0.0, p_coupling_ = 0.0;
  Rng rng_;
  NhcChain chain_;
};

}  // namespace dmgx
