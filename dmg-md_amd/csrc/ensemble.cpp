// ensemble.cpp -- thermostat/barostat scalars (see ensemble.hpp).
//
// The random-number draws and floating-point expressions keep the order of
// the reference so that the bitwise fixtures hold; every function is written
// from the published algorithms named in ensemble.hpp.
#include "ensemble.hpp"

#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>

namespace dmgx {

// Marsaglia polar Gaussian deviate, second value cached (thermostat_svr.cpp:31-51)
double gasdev(Rng& r)
{
  if (r.iset) {
    r.iset = 0;
    return r.gset;
  }
  double a, b, rsq;
  do {
    a = 2.0 * r.uniform() - 1.0;
    b = 2.0 * r.uniform() - 1.0;
    rsq = a * a + b * b;
  } while (rsq >= 1.0 || rsq == 0.0);
  const double fac = std::sqrt(-2.0 * std::log(rsq) / rsq);
  r.gset = a * fac;
  r.iset = 1;
  return b * fac;
}

// Gamma(ia) deviate: product of uniforms below 6, rejection above
// (thermostat_svr.cpp:53-85)
double gamdev(const int ia, Rng& r)
{
  if (ia < 1) throw std::invalid_argument("gamdev requires ia >= 1");
  double x;
  if (ia < 6) {
    x = 1.0;
    for (int j = 1; j <= ia; j++) x *= r.uniform();
    return -std::log(x);
  }
  double e;
  do {
    double y, am, s;
    do {
      double v1, v2;
      do {
        v1 = r.uniform();
        v2 = 2.0 * r.uniform() - 1.0;
      } while (v1 * v1 + v2 * v2 > 1.0);
      y = v2 / v1;
      am = ia - 1;
      s = std::sqrt(2.0 * am + 1.0);
      x = s * y + am;
    } while (x <= 0.0);
    e = (1.0 + y * y) * std::exp(am * std::log(x / am) - s * y);
  } while (r.uniform() > e);
  return x;
}

// sum of nn squared unit Gaussians (thermostat_svr.cpp:87-105)
double resamplekin_sumnoises(int nn, Rng& r)
{
  if (nn == 0) return 0.0;
  if (nn == 1) {
    const double g = gasdev(r);
    return g * g;
  }
  if (nn % 2 == 0) return 2.0 * gamdev(nn / 2, r);
  const double g = gasdev(r);
  return 2.0 * gamdev((nn - 1) / 2, r) + g * g;
}

// Bussi velocity-rescaling kinetic energy update (thermostat_svr.cpp:107-124)
double resamplekin(double kk, double sigma, int ndeg, double taut, Rng& r)
{
  const double factor = taut > 0.1 ? std::exp(-1.0 / taut) : 0.0;
  const double rr = gasdev(r);
  return kk + (1.0 - factor) * (sigma * (resamplekin_sumnoises(ndeg - 1, r) + rr * rr) / ndeg - kk) +
         2.0 * rr * std::sqrt(kk * sigma / ndeg * (1.0 - factor) * factor);
}

NhcChain make_nhc_chain(int atom_count, double temperature, double damping, double dt)
{
  NhcChain c;
  c.momentum = {1.0, -1.0, 1.0, -1.0};
  const double tau = dt * damping;
  const double kT = kBoltzmann * temperature;
  for (int i = 0; i < 4; ++i) c.mass[i] = kT * tau * tau;
  c.mass[0] *= 3.0 * atom_count;
  return c;
}

// Suzuki-Yoshida 7-term factorisation weights (Tuckerman, Statistical
// Mechanics, eq. 4.11.x): w1=w7=1/(4-4^(1/3)) ... as the published sextic
// scheme; the literal values are the ones the reference integrates with.
double nhc_half_step(NhcChain& c, double ek2, double kT, double dN, double dt_half)
{
  constexpr int M = 4, nsy = 7, nrespa = 4;
  const double w[nsy] = {0.784513610477560, 0.235573213359357, -1.17767998417887, 1.31518632068391,
                         -1.17767998417887, 0.235573213359357, 0.784513610477560};
  double factor = 1.0;
  auto force_on = [&](int m) {
    return m == 0 ? ek2 - dN * kT : c.momentum[m - 1] * c.momentum[m - 1] / c.mass[m - 1] - kT;
  };
  for (int a = 0; a < nsy; a++) {
    const double dt2 = dt_half * w[a] / nrespa;
    const double dt4 = dt2 * 0.5;
    const double dt8 = dt4 * 0.5;
    for (int b = 0; b < nrespa; b++) {
      double G = c.momentum[M - 2] * c.momentum[M - 2] / c.mass[M - 2] - kT;
      c.momentum[M - 1] += dt4 * G;
      for (int m = M - 2; m >= 0; m--) {
        const double t = std::exp(-dt8 * c.momentum[m + 1] / c.mass[m + 1]);
        G = force_on(m);
        c.momentum[m] = t * (t * c.momentum[m] + dt4 * G);
      }
      for (int m = M - 1; m >= 0; m--) c.position[m] += dt2 * c.momentum[m] / c.mass[m];
      const double f = std::exp(-dt2 * c.momentum[0] / c.mass[0]);
      ek2 *= f * f;
      factor *= f;
      for (int m = 0; m < M - 1; m++) {
        const double t = std::exp(-dt8 * c.momentum[m + 1] / c.mass[m + 1]);
        G = force_on(m);
        c.momentum[m] = t * (t * c.momentum[m] + dt4 * G);
      }
      G = c.momentum[M - 2] * c.momentum[M - 2] / c.mass[M - 2] - kT;
      c.momentum[M - 1] += dt4 * G;
    }
  }
  return factor;
}

double npt_scr_scale(double p_target, double p_coupling, double p_inst, double t_target, double volume,
                     double gaussian)
{
  const double berendsen = 1.0 - p_coupling * (p_target - p_inst);
  const double stochastic =
      std::sqrt(0.666666666666667 * p_coupling * kBoltzmann * t_target / volume) * gaussian;
  return berendsen + stochastic;
}

namespace {
void positive(double v, const char* what)
{
  if (!std::isfinite(v) || v <= 0.0) throw std::runtime_error(std::string(what) + " must be finite and positive");
}
}  // namespace

Controller::Controller(const EnsembleConfig& cfg, int atom_count, double dt_natural)
    : cfg_(cfg), n_(atom_count), dt_(dt_natural), rng_(cfg.seed)
{
  switch (cfg_.kind) {
    case Ensemble::nve:
    case Ensemble::nvt_ber:
      break;
    case Ensemble::npt_ber:
    case Ensemble::npt_scr:
      positive(cfg_.t_init, "T1");
      positive(cfg_.t_final, "T2");
      if (!(cfg_.t_coupling >= 1.0) || !std::isfinite(cfg_.p_target_gpa) || !(cfg_.bulk_modulus_gpa > 0.0) ||
          !(cfg_.p_coupling >= 1.0))
        throw std::invalid_argument("NPT requires Tdamp >= 1, finite P, B > 0 and tau_p >= 1");
      p_target_ = cfg_.p_target_gpa / kPressureUnit;
      if (cfg_.bulk_modulus_gpa > 2000.0) {
        p_coupling_ = 0.0;
      } else {
        p_coupling_ = 1.0 / (3.0 * cfg_.p_coupling * cfg_.bulk_modulus_gpa);
        p_coupling_ *= kPressureUnit;
      }
      break;
    case Ensemble::nvt_nhc:
      positive(cfg_.t_init, "T1");
      positive(cfg_.t_final, "T2");
      if (!(cfg_.t_coupling >= 1.0) || atom_count <= 0 || !(dt_natural > 0.0))
        throw std::invalid_argument("NHC requires Tdamp >= 1, positive N and finite positive dt");
      chain_ = make_nhc_chain(atom_count, cfg_.t_init, cfg_.t_coupling, dt_natural);
      break;
    case Ensemble::nvt_bdp:
      positive(cfg_.t_init, "T1");
      positive(cfg_.t_final, "T2");
      if (atom_count <= 0 || atom_count > std::numeric_limits<int>::max() / 3 || !(cfg_.t_coupling >= 1.0) ||
          !(dt_natural > 0.0))
        throw std::invalid_argument("BDP requires 0 < N <= INT_MAX/3, Tdamp >= 1 and finite positive dt");
      break;
    default:
      throw std::invalid_argument("unsupported ensemble");
  }
}

double Controller::target_temperature(int step, int steps) const
{
  const double fraction = static_cast<double>(step) / static_cast<double>(steps);
  return cfg_.t_init + (cfg_.t_final - cfg_.t_init) * fraction;
}

std::optional<double> Controller::before_force(double dt_natural, double kinetic_energy, int step, int steps)
{
  if (!needs_pre_force_ke()) return std::nullopt;
  positive(kinetic_energy, "NHC global kinetic energy");
  dt_ = dt_natural;
  const double f = nhc_half_step(chain_, kinetic_energy * 2.0, kBoltzmann * target_temperature(step, steps),
                                 3.0 * n_, dt_natural * 0.5);
  positive(f, "NHC velocity factor");
  return f;
}

std::optional<double> Controller::after_force(const double* thermo, int step, int steps)
{
  if (cfg_.kind == Ensemble::nve) return std::nullopt;
  const double target = target_temperature(step, steps);
  if (cfg_.kind == Ensemble::nvt_nhc) {
    double ek2 = thermo[0];
    positive(ek2, "NHC global temperature");
    ek2 *= 3.0 * n_ * kBoltzmann;
    const double f = nhc_half_step(chain_, ek2, kBoltzmann * target, 3.0 * n_, dt_ * 0.5);
    positive(f, "NHC velocity factor");
    return f;
  }
  if (cfg_.kind == Ensemble::nvt_bdp || cfg_.kind == Ensemble::npt_scr) {
    const int ndeg = 3 * n_;
    double old_ke = thermo[0];
    positive(old_ke, "BDP global temperature");
    old_ke *= ndeg * kBoltzmann * 0.5;
    const double sigma = ndeg * kBoltzmann * target * 0.5;
    const double new_ke = resamplekin(old_ke, sigma, ndeg, cfg_.t_coupling, rng_);
    if (!std::isfinite(new_ke) || new_ke < 0.0) throw std::runtime_error("BDP sampled kinetic energy invalid");
    return std::sqrt(new_ke / old_ke);
  }
  // Berendsen (nvt_ber, npt_ber)
  const double coupling = 1.0 / cfg_.t_coupling;
  if (coupling > 1.0e-5) return std::sqrt(1.0 + coupling * (target / thermo[0] - 1.0));
  return std::nullopt;
}

double Controller::barostat_scale(const double* thermo, double volume, int step, int steps)
{
  const double p_inst = (thermo[2] + thermo[3] + thermo[4]) * 0.3333333333333333;
  const double target = target_temperature(step, steps);
  double scale;
  if (cfg_.kind == Ensemble::npt_scr) {
    const double g = gasdev(rng_);  // drawn after BDP, same stream (ensemble_control.cpp:274-286)
    scale = npt_scr_scale(p_target_, p_coupling_, p_inst, target, volume, g);
  } else {
    scale = 1.0 - p_coupling_ * (p_target_ - p_inst);
  }
  if (!std::isfinite(scale) || scale <= 0.0) throw std::runtime_error("invalid barostat box factor");
  return scale;
}

}  // namespace dmgx
