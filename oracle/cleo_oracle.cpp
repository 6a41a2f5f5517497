/*
 * cleo_oracle.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU restatement of the reference's super-droplet hot path as it runs under
 * Kokkos::Serial, written from the reference's source text (CLEO, libs/...).  Only the
 * tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker or as the timed CPU baseline -- never as the product path.
 *
 * Built twice (see oracle/Makefile):
 *   liboracle_serial.so  -- the checker: sequential loops, Serial reduction order,
 *                           stable counting sort in array order.
 *   liboracle_omp.so     -- the CPU baseline: the same code with OpenMP over gridboxes
 *                           (TeamPolicy(ngbxs) -> omp for schedule(dynamic)) and an
 *                           atomic-scatter counting sort, as the reference runs on the
 *                           Kokkos OpenMP backend (sortsupers.hpp:63-76).
 *
 * Storage mirrors the reference: an array of 72-byte Superdrop records (superdrop.hpp:36-60)
 * with per-gridbox refs (supersingbx.hpp:32-102).
 *
 * Parity status of each part (what pins it: tests/test_oracle.py, tests/oracle_pins.py,
 * tests/test_collision_kinetics.py, tests/test_observers.py, tests/test_coupling.py):
 *   pinned by closed forms / invariants / brute force: units and constants, Newton solve,
 *   Golovin/Long/Const/Low-List probabilities, Simmel terminal velocity, coalescence / breakup /
 *   coalbure enactment, Fisher-Yates with injected indices, stable counting sort and refs,
 *   predictor-corrector motion and gridbox update, Philox4x32-10 (published known-answer vectors);
 *   pinned by SI evaluations of the published equations (each constant perturbed by 1e-6 fails
 *   one, tests/test_oracle_mutations.py): saturation pressure, saturation ratio, diffusion
 *   factors, ventilation factor, Koehler factors, Rogers-Yau and Rogers-GK terminal velocities,
 *   dry-air density (through the qcond recompute) and the State update of condensation.
 *   Parity unpinned: Kokkos Random_XorShift64_Pool streams (replaced by the injected /
 *   Philox streams on both sides) and Kokkos sort_team's permutation of a gridbox after
 *   nulls appear (restated as a stable partition).
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/cleo_hip.h"

namespace orc {

/* ---------------- constants: libs/cleoconstants.hpp:32-118 ---------------- */
namespace DC {
constexpr double G = 9.80665;
constexpr double RGAS_UNIV = 8.314462618;
constexpr double MR_WATER = 0.01801528;
constexpr double MR_DRY = 0.028966216;
constexpr double RGAS_DRY = RGAS_UNIV / MR_DRY;
constexpr double RGAS_V = RGAS_UNIV / MR_WATER;
constexpr double LATENT_V = 2500930;
constexpr double CP_DRY = 1004.64;
constexpr double CP_V = 1865.01;
constexpr double C_L = 4192.664;
constexpr double RHO_L = 998.203;
constexpr double RHO_SOL = 2016.5;
constexpr double MR_SOL = 0.05844277;
constexpr int IONIC = 2;
}  // namespace DC
namespace dlc {
constexpr double W0 = 1.0;
constexpr double TIME0 = 1000.0;
constexpr double COORD0 = TIME0 * W0;
constexpr double VOL0 = COORD0 * COORD0 * COORD0;
constexpr double CP0 = DC::CP_DRY;
constexpr double MR0 = DC::MR_DRY;
constexpr double R0 = 1e-6;
constexpr double P0 = 100000.0;
constexpr double TEMP0 = 273.15;
constexpr double RHO0 = P0 / (CP0 * TEMP0);
constexpr double MASS0 = R0 * R0 * R0 * RHO0;
constexpr double MASS0grams = MASS0 * 1000;
constexpr double F0 = TIME0 / (RHO0 * R0 * R0);
constexpr double Mr_ratio = DC::MR_WATER / DC::MR_DRY;
constexpr double Cp_dry = DC::CP_DRY / CP0;
constexpr double Cp_v = DC::CP_V / CP0;
constexpr double C_l = DC::C_L / CP0;
constexpr double Latent_v = DC::LATENT_V / (TEMP0 * CP0);
constexpr double Rgas_dry = DC::RGAS_DRY / CP0;
constexpr double Rgas_v = DC::RGAS_V / CP0;
constexpr double Rho_l = DC::RHO_L / RHO0;
constexpr double Rho_sol = DC::RHO_SOL / RHO0;
constexpr double Mr_sol = DC::MR_SOL / MR0;
constexpr int IONIC = DC::IONIC;
}  // namespace dlc
constexpr double PI = 3.141592653589793238462643383279502884;
constexpr uint32_t OOB = 0xFFFFFFFFu;

/* model_step = 1e-5 dimensionless time (timesteps.hpp:29-59) */
static inline double step2dimlesstime(uint32_t s) { return static_cast<double>(s) / 100000.0; }
static inline double step2realtime(uint32_t s) { return (static_cast<double>(s) / 100000.0) * dlc::TIME0; }

/* ---------------- data model: superdrop.hpp:36-60, superdrop_attrs.hpp:61-260 ----- */
struct Attrs {
  char solute; /* empty SoluteProperties occupies one byte */
  uint64_t xi;
  double radius;
  double msol;
};
struct Superdrop {
  uint32_t sdgbxindex;
  double coord3, coord1, coord2;
  Attrs attrs;
  uint64_t sdId;
};
static_assert(sizeof(Superdrop) == 72, "reference Superdrop is 72 bytes");

struct Solute {
  double rho_sol = dlc::Rho_sol;
  double mr_sol = dlc::Mr_sol;
  double ionic = static_cast<double>(dlc::IONIC);
};

struct ErrWord {
  int code = 0;
  uint32_t gbx = 0;
  uint64_t drop = 0;
};

struct State {
  double volume, press, temp, qvap, qcond;
  double wlo, whi, ulo, uhi, vlo, vhi;
};

/* error recording: the reference aborts; the oracle keeps the first failure */
struct ErrSink {
  ErrWord* w;
  uint32_t gbx;
  uint64_t drop;
  void set(int code) const {
#ifdef _OPENMP
#pragma omp critical(orc_err)
#endif
    {
      if (w->code == 0) {
        w->code = code;
        w->gbx = gbx;
        w->drop = drop;
      }
    }
  }
};

/* superdrop_attrs.hpp:194-205 */
static inline double dryradius(const Superdrop& d, const Solute& s) {
  constexpr double vconst = 3.0 / (4.0 * PI);
  const double dryrcubed = vconst * d.attrs.msol / s.rho_sol;
  return std::pow(dryrcubed, 1.0 / 3.0);
}
static inline double rcubed(const Superdrop& d) { return d.attrs.radius * d.attrs.radius * d.attrs.radius; }
static inline double vol(const Superdrop& d) { return 4.0 / 3.0 * PI * rcubed(d); }
/* superdrop_attrs.cpp:55-62 */
static inline double mass(const Superdrop& d, const Solute& s) {
  constexpr double massconst(4.0 / 3.0 * PI * dlc::Rho_l);
  const double density_factor = 1.0 - dlc::Rho_l / s.rho_sol;
  double m = d.attrs.msol * density_factor;
  m += massconst * rcubed(d);
  return m;
}
/* superdrop_attrs.hpp:207-216 */
static inline double condensate_mass(const Superdrop& d, const Solute& s, const ErrSink& e) {
  double m_cond = mass(d, s) - d.attrs.msol;
  if (m_cond <= -0.0001 * d.attrs.msol) e.set(CLEO_ERR_NEG_CONDENSATE);
  return std::fmax(0.0, m_cond);
}
static inline void set_xi(Superdrop& d, uint64_t xi, const ErrSink& e) {
  if (xi <= 0) e.set(CLEO_ERR_XI_ZERO);
  d.attrs.xi = xi;
}
static inline void set_radius(Superdrop& d, double r, const Solute& s, const ErrSink& e) {
  if (r - dryradius(d, s) <= -1e-12 / dlc::R0) e.set(CLEO_ERR_RADIUS_BELOW_DRY);
  d.attrs.radius = r;
}
/* superdrop_attrs.cpp:35-45 */
static inline double change_radius(Superdrop& d, double newr, const Solute& s) {
  const double old = d.attrs.radius;
  const double dryr = dryradius(d, s);
  d.attrs.radius = std::fmax(newr, dryr);
  return d.attrs.radius - old;
}
static inline void set_null(Superdrop& d) {
  d.attrs.xi = 0;
  d.sdgbxindex = OOB;
}

/* ---------------- thermodynamics: thermodynamic_equations.{hpp,cpp} ---------------- */
static inline double dry_air_density(double press, double temp, double qvap) {
  const double rgas_equiv = dlc::Rgas_dry + dlc::Rgas_v * qvap;
  return press / (rgas_equiv * temp);
}
static inline double moist_specifc_heat(double qvap, double qcond) {
  return dlc::Cp_dry + dlc::Cp_v * (qvap) + dlc::C_l * (qcond);
}
static inline double supersaturation_ratio(double press, double qvap, double psat) {
  return (press * qvap) / ((dlc::Mr_ratio + qvap) * psat);
}
static inline double saturation_pressure(double temp, const ErrSink& e) {
  if (temp <= 0) e.set(CLEO_ERR_TEMP_NONPOSITIVE);
  constexpr double A = 17.4146, B = 33.639, TREF = 273.16, PREF = 611.655;
  const double T = temp * dlc::TEMP0;
  return (PREF * std::exp(A * (T - TREF) / (T - B))) / dlc::P0;
}
static inline double diffusion_factor(double press, double temp, double psat) {
  constexpr double A = 7.11756e-5;
  constexpr double B = 4.38127686e-3;
  constexpr double LATENT_RGAS_V = DC::LATENT_V / DC::RGAS_V;
  constexpr double D = 4.012182971e-5;
  const double TEMP = temp * dlc::TEMP0;
  const double PRESS = press * dlc::P0;
  const double PSAT = psat * dlc::P0;
  const double THERMK = A * std::pow(TEMP, 2.0) + TEMP * B;
  const double DIFFUSE_V = (D / PRESS * std::pow(TEMP, 1.94)) / DC::RGAS_V;
  const double fkl = (LATENT_RGAS_V / TEMP - 1.0) * DC::LATENT_V / (THERMK * dlc::F0);
  const double fdl = TEMP / (DIFFUSE_V * PSAT) / dlc::F0;
  return dlc::Rho_l * (fkl + fdl);
}
static inline double ventilation_factor(double radius) {
  constexpr double vent_factor_max = 20;
  constexpr double C1 = 6.954e+7, A = 1.963, C2 = 1.069e+3, B = 0.702;
  const double a = 1.0 / (C1 * std::pow(radius * dlc::R0, A));
  const double b = 1.0 / (C2 * std::pow(radius * dlc::R0, B));
  const double vent = 1.0 + 1.0 / (a + b);
  return std::fmin(vent, vent_factor_max);
}
static inline void kohler_factors(const Superdrop& d, double temp, const Solute& s, double& akoh,
                                  double& bkoh) {
  constexpr double akoh_constant = 3.3e-7 / (dlc::TEMP0 * dlc::R0);
  akoh = akoh_constant / temp;
  constexpr double bkoh_constant = 4.3e-6 * dlc::RHO0 / dlc::MR0;
  bkoh = bkoh_constant * d.attrs.msol * s.ionic / s.mr_sol;
}

/* ---------------- implicit Euler: impliciteuler.{hpp,cpp} ---------------- */
struct ODEConsts {
  double s_ratio, akoh, bkoh, ffactor_fv;
};
struct NewtonCfg {
  double delt, minsubdelt, rtol, atol;
  uint64_t maxniters;
};
struct NewtonCount {
  uint32_t iters = 0;
  bool substep = false;
  bool restarted = false;
};
static inline double ode_gfunc(const ODEConsts& c, double subdelt, double rprev, double rsqrd) {
  const double radius = std::sqrt(rsqrd);
  const double alpha = c.s_ratio - 1 - c.akoh / radius + c.bkoh / std::pow(radius, 3.0);
  const double beta = 2.0 * subdelt / (rsqrd * c.ffactor_fv);
  const double gamma = std::pow(rprev / radius, 2.0);
  return 1 - gamma - alpha * beta;
}
static inline double ode_gfuncderivative(const ODEConsts& c, double subdelt, double rsqrd) {
  const double radius = std::sqrt(rsqrd);
  const double alpha = c.akoh / radius - 3.0 * c.bkoh / std::pow(radius, 3.0);
  const double beta = subdelt / (rsqrd * c.ffactor_fv);
  return 1 - alpha * beta;
}
static inline bool iterate(const NewtonCfg& cfg, const ODEConsts& c, double subdelt, double rprev,
                           double& ziter, NewtonCount& cnt) {
  ++cnt.iters;
  const double numerator = ode_gfunc(c, subdelt, rprev, ziter);
  const double denominator = ode_gfuncderivative(c, subdelt, ziter);
  ziter = ziter * (1 - numerator / denominator);
  ziter = std::fmax(ziter, 1e-8);
  const double newnumerator = ode_gfunc(c, subdelt, rprev, ziter);
  const double converged = cfg.rtol * std::abs(newnumerator) + cfg.atol;
  const double current = std::abs(newnumerator - numerator);
  return current < converged;
}
/* newtonraphson_untilconverged (impliciteuler.cpp:207-227); limit exceeded -> ok=false */
static inline double until_converged(const NewtonCfg& cfg, const ODEConsts& c, uint64_t limit,
                                     double subdelt, double rprev, double ziter, NewtonCount& cnt,
                                     bool& ok) {
  bool conv = false;
  uint64_t niter = 1;
  while (!conv) {
    if (niter > limit) {
      ok = false;
      return ziter;
    }
    conv = iterate(cfg, c, subdelt, rprev, ziter, cnt);
    niter += 1;
  }
  ok = true;
  return ziter;
}
/* integrate_condensation_ode (impliciteuler.cpp:128-141) */
static inline double integrate(const NewtonCfg& cfg, const ODEConsts& c, double subdelt,
                               double rprev, double ziter, NewtonCount& cnt, const ErrSink& e) {
  double z = ziter;
  bool conv = false;
  for (int it = 0; it < 2; ++it) conv = iterate(cfg, c, subdelt, rprev, z, cnt);
  if (conv) return z;
  bool ok = true;
  const double r = until_converged(cfg, c, cfg.maxniters, subdelt, rprev, ziter, cnt, ok);
  if (!ok) e.set(CLEO_ERR_NEWTON_MAXITERS);
  return r;
}
static inline double initialguess(const ODEConsts& c, double rprev) {
  const double s_act = 1 + std::sqrt(4.0 * std::pow(c.akoh, 3.0) / 27 / c.bkoh);
  if (c.s_ratio > s_act) {
    constexpr double bigr(1e-3 / dlc::R0);
    const double maxradius = std::fmax(bigr, rprev);
    return maxradius * maxradius;
  }
  return rprev * rprev;
}
static inline bool first_unique_criteria(const ODEConsts& c, double rprev, double ziter) {
  const double rcritsqrd = 3.0 * c.bkoh / c.akoh;
  const bool is_ziter_unactivated = ziter < rcritsqrd;
  const bool is_unactivated = (rprev * rprev < rcritsqrd && is_ziter_unactivated);
  const double sqrdval = 4.0 * c.akoh * c.akoh * c.akoh / 27.0 / c.bkoh;
  const bool is_subact = (c.s_ratio <= 1.0 + std::pow(sqrdval, 0.5));
  return (is_unactivated && is_subact);
}
static inline double critical_timestep(const ODEConsts& c) {
  const double cuberoot = std::pow(5.0 * c.bkoh / c.akoh, 1.5);
  return 2.5 * c.ffactor_fv / c.akoh * cuberoot;
}
/* solve_condensation (impliciteuler.cpp:35-54); warm = NaN for the plain path.
   Learned-guess variant (no reference counterpart): the reference's Newton update and
   convergence test iterated from `warm` until converged, at most `warm_limit` iterations; if it
   has not converged by then, the reference path from the plain guess.  Where |g| at the plain
   guess is below atol, the same iteration starts from the plain guess instead. */
static inline double solve_condensation(const NewtonCfg& cfg, double s_ratio, double akoh,
                                        double bkoh, double ffactor, double rprev, double warm,
                                        uint64_t warm_limit, NewtonCount& cnt, const ErrSink& e) {
  const double ffactor_fv = ffactor / ventilation_factor(rprev);
  const ODEConsts c{s_ratio, akoh, bkoh, ffactor_fv};
  double ziter = initialguess(c, rprev);
  const bool ucrit1 = first_unique_criteria(c, rprev, ziter);
  const bool ucrit2 = (cfg.delt <= critical_timestep(c));
  double rsqrd = 0.0;
  if (ucrit1 || ucrit2) {
    /* learned path from the learned guess, or from the plain guess where that is already the
       root to atol (|g| below it there) */
    if (warm == warm && std::abs(ode_gfunc(c, cfg.delt, rprev, ziter)) < cfg.atol) warm = ziter;
    if (warm == warm) {
      bool ok = true;
      rsqrd = until_converged(cfg, c, warm_limit, cfg.delt, rprev, warm, cnt, ok);
      if (!ok) {
        cnt.restarted = true;
        rsqrd = integrate(cfg, c, cfg.delt, rprev, ziter, cnt, e);
      }
    } else {
      rsqrd = integrate(cfg, c, cfg.delt, rprev, ziter, cnt, e);
    }
  } else {
    cnt.substep = true;
    const double critdelt = critical_timestep(c);
    const double mindelt = std::fmax(critdelt, cfg.minsubdelt);
    double remdelt = cfg.delt;
    double rp = rprev;
    while (remdelt > 0.0) {
      const double subdelt = std::fmin(mindelt, remdelt);
      ziter = integrate(cfg, c, subdelt, rp, ziter, cnt, e);
      rp = std::pow(ziter, 0.5);
      remdelt -= subdelt;
    }
    rsqrd = ziter;
  }
  return std::sqrt(rsqrd);
}

/* ---------------- terminal velocity: terminalvelocity.cpp:33-146 ---------------- */
static inline double tv_simmel(double radius) {
  constexpr double r1 = 6.7215e-5 / dlc::R0;
  constexpr double r2 = 7.5582e-4 / dlc::R0;
  constexpr double r3 = 1.73892e-3 / dlc::R0;
  constexpr double VELCONST = (100.0 * dlc::W0);
  constexpr double a1 = 457950 / VELCONST;
  constexpr double a2 = 4962 / VELCONST;
  constexpr double a3 = 1732 / VELCONST;
  constexpr double a4 = 917 / VELCONST;
  if (radius >= r3) return a4;
  constexpr double massconst(4.0 / 3.0 * PI * dlc::Rho_l);
  const double MASS = (massconst * radius * radius * radius) * dlc::MASS0grams;
  if (radius >= r2) return a3 * std::pow(MASS, 1.0 / 6.0);
  if (radius >= r1) return a2 * std::pow(MASS, 1.0 / 3.0);
  return a1 * std::pow(MASS, 2.0 / 3.0);
}
static inline double tv_rogersyau(double radius) {
  constexpr double r1 = 3e-5 / dlc::R0;
  constexpr double r2 = 6e-4 / dlc::R0;
  constexpr double r3 = 2e-3 / dlc::R0;
  constexpr double k1 = 1.19e8 * dlc::R0 * dlc::R0 / dlc::W0;
  constexpr double k2 = 8000 * dlc::R0 / dlc::W0;
  constexpr double k3 = 201 / dlc::W0;
  constexpr double k4 = 9 / dlc::W0;
  if (radius < r1) return k1 * std::pow(radius, 2.0);
  if (radius < r2) return k2 * radius;
  if (radius < r3) return k3 * std::pow((radius * dlc::R0), 0.5);
  return k4;
}
static inline double tv_rogersgk(double radius) {
  constexpr double radius0 = 3.725 * 1e-4 / dlc::R0;
  constexpr double kcaps = 2.0 * 4.0 * 1000 * dlc::R0 / dlc::W0;
  constexpr double smallk = -1.0 * 2.0 * 12.0 * 1000 * dlc::R0;
  constexpr double acaps = 9.65 / dlc::W0;
  constexpr double bcaps = 10.43 / dlc::W0;
  constexpr double ccaps = -1.0 * 2.0 * 0.6 * 1000 * dlc::R0;
  if (radius < radius0) {
    const double term = 1.0 - std::exp(smallk * radius);
    return term * kcaps * radius;
  }
  return acaps - bcaps * std::exp(ccaps * radius);
}
static inline double terminalv(int kind, double radius) {
  switch (kind) {
    case CLEO_TV_SIMMEL: return tv_simmel(radius);
    case CLEO_TV_ROGERS_YAU: return tv_rogersyau(radius);
    case CLEO_TV_ROGERS_GK: return tv_rogersgk(radius);
    default: return 0.0;
  }
}

/* ---------------- pair probabilities ---------------- */
static inline double prob_golovin(const Superdrop& a, const Superdrop& b, double DELT, double VOLUME) {
  constexpr double prob_jk_const = 1.5e3 * dlc::R0 * dlc::R0 * dlc::R0;
  const double DELT_DELVOL = DELT / VOLUME;
  const double k = prob_jk_const * (vol(a) + vol(b));
  return k * DELT_DELVOL;
}
static inline double prob_long(const Superdrop& a, const Superdrop& b, double coaleff, double DELT,
                               double VOLUME) {
  constexpr double k1 = 4.5e4 * dlc::R0 * dlc::R0 * 100 * 100;
  constexpr double k2 = 3e-4 / dlc::R0 / 100;
  constexpr double rlim = 5e-5 / dlc::R0;
  constexpr double colleff_min = 0.001;
  const double rsmall = std::fmin(a.attrs.radius, b.attrs.radius);
  const double rbig = std::fmax(a.attrs.radius, b.attrs.radius);
  double colleff = 1.0;
  if (rbig < rlim) {
    colleff = k1 * rbig * rbig * (1 - (k2 / rsmall));
    colleff = std::fmax(colleff, colleff_min);
  }
  const double eff = colleff * coaleff;
  constexpr double prob_jk_const = PI * dlc::R0 * dlc::R0 * dlc::W0;
  const double DELT_DELVOL = DELT / VOLUME;
  const double sumr = a.attrs.radius + b.attrs.radius;
  const double sumrsqrd = sumr * sumr;
  const double vdiff = std::abs(tv_simmel(a.attrs.radius) - tv_simmel(b.attrs.radius));
  const double hydro_kernel = prob_jk_const * eff * sumrsqrd * vdiff;
  return hydro_kernel * DELT_DELVOL;
}

/* ---------------- collision kinetics (Low and List 1982a) ---------------- */
/* surface tension energies [J] of one drop, of two and of their coalesced drop
   (collisionkinetics.hpp:72-93, collisionkinetics.cpp:62-71; surfconst cleoconstants.hpp:102) */
constexpr double SURF_SIGMA = 7.28e-2;
constexpr double surfconst = 4.0 * SURF_SIGMA * PI * dlc::R0 * dlc::R0;
static inline double surfenergy(double r) { return surfconst * (r * r); }
static inline double total_surfenergy(double r1, double r2) { return surfconst * ((r1 * r1) + (r2 * r2)); }
static inline double coal_surfenergy(double r1, double r2) {
  return surfconst * std::pow(r1 * r1 * r1 + r2 * r2 * r2, 2.0 / 3.0);
}
/* collisionkinetics.cpp:33-46 */
static inline double coll_kinetic_energy(double r1, double r2, double v1, double v2) {
  constexpr double ckeconst = dlc::R0 * dlc::R0 * dlc::R0 * 2.0 / 3.0 * DC::RHO_L * PI * dlc::W0 * dlc::W0;
  const double r1cubed = r1 * r1 * r1;
  const double ratio = r1 / r2;
  const double rratio = r1cubed / (1 + ratio * ratio * ratio);
  const double vdiff = v1 - v2;
  return ckeconst * rratio * vdiff * vdiff;
}
static inline double cke_of(int tv, const Superdrop& a, const Superdrop& b) {
  return coll_kinetic_energy(a.attrs.radius, b.attrs.radius, terminalv(tv, a.attrs.radius),
                             terminalv(tv, b.attrs.radius));
}
/* Straub et al. 2010 eq. 5 (coalescence_efficiency.cpp:27-37) */
static inline double ecoal_straub(double r1, double r2, double cke) {
  return std::exp(-1.15 * (cke / coal_surfenergy(r1, r2)));
}
/* LowListCoalProb::coaleff, Low and List 1982a eqs. 4.5-4.6 (lowlistprob.hpp:50-91) */
static inline double lowlist_coaleff(int tv, const Superdrop& a, const Superdrop& b) {
  const double r1 = a.attrs.radius, r2 = b.attrs.radius;
  const double etot = cke_of(tv, a, b) + total_surfenergy(r1, r2) - coal_surfenergy(r1, r2);
  if (etot < 5e-6) {
    const double alpha = 1 + std::fmin(r1, r2) / std::fmax(r1, r2);
    const double expo = -2.62e6 * SURF_SIGMA * etot * etot / coal_surfenergy(r1, r2);
    return 0.778 * (1.0 / (alpha * alpha)) * std::exp(expo);
  }
  return 0.0;
}
/* NFragments: ConstNFrags, or CollisionKineticEnergyNFrags (breakup_nfrags.hpp:38-91) */
static inline double nfrags_of(const cleo_coll_params& p, const Superdrop& a, const Superdrop& b) {
  if (p.nfrags_kind != CLEO_NFRAGS_CKE) return p.nfrags;
  const double cke_cap = std::fmin(16.49789599 / 1000000, cke_of(p.nfrags_tv, a, b));
  return 1.0 / (1.5 - std::pow(cke_cap * 1e6, 0.135)) + 11.0 / 6.0;
}
/* CoalBuReFlag: 1 coalescence, 2 breakup, 0 rebound (coalbure_flag.hpp:41-198) */
static inline unsigned coalbure_flag(const cleo_coll_params& p, double phi, const Superdrop& a,
                                     const Superdrop& b) {
  if (p.flag_kind == CLEO_FLAG_CONST) return phi < p.coalbure_coaleff ? 1 : 2;
  const double r1 = a.attrs.radius, r2 = b.attrs.radius;
  const double cke = cke_of(p.flag_tv, a, b);
  switch (p.flag_kind) {
    case CLEO_FLAG_SU:
      if (cke < surfenergy(std::fmin(r1, r2))) return 0;
      if (cke < coal_surfenergy(r1, r2)) return 1;
      return 2;
    case CLEO_FLAG_TS:
      if (cke < surfenergy(std::fmin(r1, r2))) return phi < ecoal_straub(r1, r2, cke) ? 1 : 0;
      if (cke < surfenergy(std::fmax(r1, r2))) return phi < ecoal_straub(r1, r2, cke) ? 1 : 2;
      return 2;
    default: return phi < ecoal_straub(r1, r2, cke) ? 1 : 2; /* Straub */
  }
}

static inline double pair_prob(const cleo_coll_params& p, const Superdrop& a, const Superdrop& b,
                               double DELT, double VOLUME) {
  switch (p.prob_kind) {
    case CLEO_PROB_GOLOVIN: return prob_golovin(a, b, DELT, VOLUME);
    case CLEO_PROB_LONG: return prob_long(a, b, p.prob_param, DELT, VOLUME);
    case CLEO_PROB_LOWLIST_COAL: /* lowlistprob.hpp:117-123 */
      return prob_long(a, b, 1.0, DELT, VOLUME) * lowlist_coaleff(p.prob_tv, a, b);
    case CLEO_PROB_LOWLIST_BU: { /* lowlistprob.hpp:144-153 */
      const double bueff = 1.0 - lowlist_coaleff(p.prob_tv, a, b);
      return prob_long(a, b, 1.0, DELT, VOLUME) * bueff;
    }
    default: return p.prob_param * DELT / VOLUME; /* constprob.hpp:45-48 */
  }
}

/* ---------------- enactments ---------------- */
/* coalescence.cpp:68-79 (also coalbure.hpp:74-84).  A probability of 2^64 or more (or NaN)
   overflows the reference's double -> uint64 conversion, whose result is then undefined (0 on
   x86-64, saturated in a CUDA build); here gamma is the cap floor(xi1/xi2) that any such
   gamma reaches, as on the GPU. */
static inline uint64_t coal_gamma(uint64_t xi1, uint64_t xi2, double prob, double phi) {
  if (!(prob < 0x1p64)) return xi1 / xi2;
  uint64_t gamma = static_cast<uint64_t>(std::floor(prob));
  if (phi < (prob - gamma)) ++gamma;
  const uint64_t maxgamma = xi1 / xi2;
  return static_cast<uint64_t>(std::fmin(static_cast<double>(gamma), static_cast<double>(maxgamma)));
}
/* coalescence.cpp:92-196 */
static inline size_t coalesce_pair(uint64_t gamma, Superdrop& d1, Superdrop& d2, const Solute& s,
                                   const ErrSink& e) {
  const uint64_t xi1 = d1.attrs.xi, xi2 = d2.attrs.xi;
  if (xi1 - gamma * xi2 > 0) {
    const uint64_t new_xi = xi1 - gamma * xi2;
    if (new_xi >= xi1) e.set(CLEO_ERR_COAL_UNDEFINED);
    set_xi(d1, new_xi, e);
    const double new_rcubed = rcubed(d2) + gamma * rcubed(d1);
    set_radius(d2, std::pow(new_rcubed, (1.0 / 3.0)), s, e);
    d2.attrs.msol = d2.attrs.msol + gamma * d1.attrs.msol;
    return 0;
  } else if (xi1 - gamma * xi2 == 0) {
    const uint64_t old_xi = xi2;
    const uint64_t new_xi = old_xi / 2;
    if (new_xi >= old_xi) e.set(CLEO_ERR_COAL_UNDEFINED);
    const double new_rcubed = rcubed(d2) + gamma * rcubed(d1);
    const double new_r = std::pow(new_rcubed, (1.0 / 3.0));
    const double new_msol = d2.attrs.msol + gamma * d1.attrs.msol;
    if (new_xi == 0) set_null(d1);
    else set_xi(d1, new_xi, e);
    set_xi(d2, old_xi - new_xi, e);
    set_radius(d1, new_r, s, e);
    set_radius(d2, new_r, s, e);
    d1.attrs.msol = new_msol;
    d2.attrs.msol = new_msol;
    return d1.attrs.xi == 0 ? 1 : 0;
  }
  e.set(CLEO_ERR_COAL_UNDEFINED);
  return 0;
}
/* breakup.hpp:176-268 with ConstNFrags */
static inline void breakup_pair(double nfrags, Superdrop& d1, Superdrop& d2, const Solute& s,
                                const ErrSink& e) {
  if (d1.attrs.xi == d2.attrs.xi) {
    const uint64_t old_xi = d2.attrs.xi;
    const double totnfrags = nfrags * old_xi;
    if ((totnfrags / old_xi) <= 2.5) e.set(CLEO_ERR_BREAKUP_NFRAGS);
    const uint64_t new_xi1 = static_cast<uint64_t>(std::round(totnfrags / 2));
    const uint64_t new_xi2 = static_cast<uint64_t>(std::round(totnfrags - new_xi1));
    const uint64_t new_xitot = new_xi1 + new_xi2;
    if (new_xi2 <= old_xi) e.set(CLEO_ERR_BREAKUP_NFRAGS);
    if (new_xitot <= (old_xi * 2)) e.set(CLEO_ERR_BREAKUP_NFRAGS);
    const double sum_rcubed = rcubed(d1) + rcubed(d2);
    const double new_rcubed = sum_rcubed * old_xi / new_xitot;
    const double new_r = std::pow(new_rcubed, (1.0 / 3.0));
    const double sum_msol = d1.attrs.msol + d2.attrs.msol;
    const double new_msol = sum_msol * old_xi / new_xitot;
    set_xi(d1, new_xi1, e);
    set_xi(d2, new_xi2, e);
    set_radius(d1, new_r, s, e);
    set_radius(d2, new_r, s, e);
    d1.attrs.msol = new_msol;
    d2.attrs.msol = new_msol;
  } else {
    const uint64_t old_xi1 = d1.attrs.xi, old_xi2 = d2.attrs.xi;
    const uint64_t new_xi1 = old_xi1 - old_xi2;
    set_xi(d1, new_xi1, e);
    const double totnfrags = nfrags * old_xi2;
    const uint64_t new_xi2 = static_cast<uint64_t>(std::round(totnfrags));
    if ((totnfrags / old_xi2) <= 2.5) e.set(CLEO_ERR_BREAKUP_NFRAGS);
    if (new_xi2 <= old_xi2) e.set(CLEO_ERR_BREAKUP_NFRAGS);
    if ((new_xi1 + new_xi2) <= (old_xi1 + old_xi2)) e.set(CLEO_ERR_BREAKUP_NFRAGS);
    const double sum_rcubed = rcubed(d1) + rcubed(d2);
    const double new_rcubed = sum_rcubed * old_xi2 / new_xi2;
    const double new_r = std::pow(new_rcubed, (1.0 / 3.0));
    const double sum_msol = d1.attrs.msol + d2.attrs.msol;
    const double new_msol = sum_msol * old_xi2 / new_xi2;
    set_xi(d2, new_xi2, e);
    set_radius(d2, new_r, s, e);
    d2.attrs.msol = new_msol;
  }
}
/* returns nulls; `enacted` set if gamma != 0 */
static inline size_t enact(const cleo_coll_params& p, Superdrop& d1, Superdrop& d2, double prob,
                           double phi_coll, double phi_out, const Solute& s, const ErrSink& e,
                           bool& enacted) {
  enacted = false;
  if (p.enact_kind == CLEO_ENACT_COAL) {
    const uint64_t gamma = coal_gamma(d1.attrs.xi, d2.attrs.xi, prob, phi_coll);
    if (gamma != 0) {
      enacted = true;
      return coalesce_pair(gamma, d1, d2, s, e);
    }
    return 0;
  }
  if (p.enact_kind == CLEO_ENACT_BREAKUP) {
    if (phi_coll < (prob - std::floor(prob))) { /* breakup.hpp:157-166 */
      enacted = true;
      breakup_pair(nfrags_of(p, d1, d2), d1, d2, s, e);
    }
    return 0;
  }
  /* CoalBuRe (coalbure.hpp:207-256): the flag picks coalescence, breakup or rebound */
  const uint64_t gamma = coal_gamma(d1.attrs.xi, d2.attrs.xi, prob, phi_coll);
  if (gamma != 0) {
    enacted = true;
    const unsigned flag = coalbure_flag(p, phi_out, d1, d2);
    if (flag == 1) return coalesce_pair(gamma, d1, d2, s, e);
    if (flag == 2) breakup_pair(nfrags_of(p, d1, d2), d1, d2, s, e);
  }
  return 0;
}

/* ---------------- Philox4x32-10 (Salmon et al., SC'11) ---------------- */
static inline void philox4x32_10(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
    const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n1 = static_cast<uint32_t>(p1);
    const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c3 ^ k1;
    const uint32_t n3 = static_cast<uint32_t>(p0);
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
static inline void rng_block(uint64_t seed, uint64_t event, uint32_t stream, uint32_t gbx, uint32_t item,
                             uint32_t out[4]) {
  const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
  const uint32_t ctr[4] = {item, gbx, static_cast<uint32_t>(event),
                           stream + 2u * static_cast<uint32_t>(event >> 32)};
  philox4x32_10(ctr, key, out);
}
static inline double u53(uint32_t lo, uint32_t hi) {
  const uint64_t v = (static_cast<uint64_t>(hi) << 32) | lo;
  return static_cast<double>(v >> 11) * 0x1.0p-53;
}

/* ---------------- context ---------------- */
struct Ctx {
  uint32_t nz = 0, nx = 0, ny = 0, ngbx = 0;
  int per[3] = {0, 1, 1};
  std::vector<double> zh, xh, yh;
  uint64_t maxn = 0;
  std::vector<Superdrop> sup, tmp;
  uint64_t ndomain = 0;
  std::vector<uint64_t> ref0, ref1, nact;
  std::vector<State> st;
  Solute sol;
  int rng_mode = 0; /* 0 none, 1 counter, 2 injected */
  uint64_t seed = 0, event = 0;
  std::vector<uint32_t> inj_j;
  std::vector<double> inj_phi;
  std::vector<cleo_process> procs; /* microphysics procs[0] >> procs[1] >> ... */
  bool motion_on = false;
  cleo_motion_params motion{};
  bool sdm_set = false;
  ErrWord err;
  cleo_stats stats{};
  bool rec_niters = false;
  std::vector<uint32_t> niters;
  std::vector<double> guess;
  std::string msg;

  uint32_t iz(uint32_t g) const { return g % nz; }
  uint32_t ix(uint32_t g) const { return (g / nz) % nx; }
  uint32_t iy(uint32_t g) const { return g / (nz * nx); }
};

static int fail(Ctx* c, int code, const char* m) {
  if (c) c->msg = m;
  return code;
}

/* ---------------- sort + refs: sortsupers.hpp:44-226, findrefs.hpp:29-179 ---------------- */
static void set_refs(Ctx& c) {
  /* find_domainrefs: first position whose sdgbxindex > gbxindex_max */
  const uint32_t gmax = c.ngbx - 1;
  auto part = [&](uint64_t first, uint64_t len, auto pred) {
    while (len > 0) {
      const uint64_t half = len / 2, mid = first + half;
      if (pred(c.sup[mid])) {
        first = mid + 1;
        len -= (half + 1);
      } else {
        len = half;
      }
    }
    return first;
  };
  const uint64_t ntot = c.sup.size();
  c.ndomain = part(0, ntot, [&](const Superdrop& d) { return d.sdgbxindex <= gmax; });
  for (uint32_t g = 0; g < c.ngbx; ++g) {
    c.ref0[g] = part(0, c.ndomain, [&](const Superdrop& d) { return d.sdgbxindex < g; });
    c.ref1[g] = part(0, c.ndomain, [&](const Superdrop& d) { return d.sdgbxindex <= g; });
    c.nact[g] = c.ref1[g] - c.ref0[g];
  }
}
static void sort_supers(Ctx& c) {
  const uint64_t ntot = c.sup.size();
  const uint32_t gmax = c.ngbx - 1;
  std::vector<uint64_t> cuml(c.ngbx + 1, 0);
  auto pos = [&](uint32_t k) -> uint64_t { return !(k <= gmax) ? c.ngbx : k; };
  for (uint64_t kk = 0; kk < ntot; ++kk) ++cuml[pos(c.sup[kk].sdgbxindex)];
  uint64_t run = 0;
  for (uint32_t b = 0; b <= c.ngbx; ++b) {
    const uint64_t n = cuml[b];
    cuml[b] = run;
    run += n;
  }
  c.tmp.resize(ntot);
#ifdef _OPENMP
  /* reference-structured parallel scatter with atomic fetch-add (sortsupers.hpp:72-76) */
#pragma omp parallel for schedule(static)
  for (uint64_t kk = 0; kk < ntot; ++kk) {
    const uint64_t b = pos(c.sup[kk].sdgbxindex);
    uint64_t nk;
#pragma omp atomic capture
    nk = cuml[b]++;
    c.tmp[nk] = c.sup[kk];
  }
#else
  for (uint64_t kk = 0; kk < ntot; ++kk) {
    const uint64_t nk = cuml[pos(c.sup[kk].sdgbxindex)]++;
    c.tmp[nk] = c.sup[kk];
    c.sup[kk].sdgbxindex = OOB;
  }
#endif
  std::swap(c.sup, c.tmp);
  set_refs(c);
  c.stats.sorts += 1;
}

/* ---------------- motion: predcorr.hpp, predcorrmotion.hpp, cartesianmaps.cpp ---------------- */
static inline void bounds3(const Ctx& c, uint32_t g, double& lo, double& hi) {
  if (g == OOB) { lo = -1.0 * 1.7976931348623157e308; hi = 1.7976931348623157e308; return; }
  lo = c.zh[c.iz(g)]; hi = c.zh[c.iz(g) + 1];
}
static inline void bounds1(const Ctx& c, uint32_t g, double& lo, double& hi) {
  if (g == OOB) { lo = -1.0 * 1.7976931348623157e308; hi = 1.7976931348623157e308; return; }
  lo = c.xh[c.ix(g)]; hi = c.xh[c.ix(g) + 1];
}
static inline void bounds2(const Ctx& c, uint32_t g, double& lo, double& hi) {
  if (g == OOB) { lo = -1.0 * 1.7976931348623157e308; hi = 1.7976931348623157e308; return; }
  lo = c.yh[c.iy(g)]; hi = c.yh[c.iy(g) + 1];
}
static inline double interpolation(double blo, double bhi, double vlo, double vhi, double sdcoord) {
  const double coord = std::fmin(bhi, std::fmax(blo, sdcoord));
  const double alpha = (coord - blo) / (bhi - blo);
  return alpha * vhi + (1 - alpha) * vlo;
}
/* neighbour along one axis: axis 0 z (increment 1), 1 x (nz), 2 y (nz*nx)
   (domainboundaries.hpp:34-90) */
static inline uint32_t nghbr(const Ctx& c, uint32_t idx, int axis, bool forward, bool& at_edge) {
  const uint32_t inc = axis == 0 ? 1u : (axis == 1 ? c.nz : c.nz * c.nx);
  const uint32_t nd = axis == 0 ? c.nz : (axis == 1 ? c.nx : c.ny);
  if (forward) {
    const uint32_t fw = idx + inc;
    at_edge = ((fw / inc) % nd == 0);
    if (!at_edge) return fw;
    return c.per[axis] ? idx - (nd - 1) * inc : OOB;
  }
  at_edge = ((idx / inc) % nd == 0);
  if (!at_edge) return idx - inc;
  return c.per[axis] ? idx + (nd - 1) * inc : OOB;
}
static inline void axis_bounds(const Ctx& c, int axis, uint32_t g, double& lo, double& hi) {
  if (axis == 0) bounds3(c, g, lo, hi);
  else if (axis == 1) bounds1(c, g, lo, hi);
  else bounds2(c, g, lo, hi);
}
/* change_if_coordXnghbr (cartesianmaps.cpp:197-276) */
static inline uint32_t change_if_nghbr(const Ctx& c, uint32_t idx, int axis, double& coord) {
  if (idx == OOB) return idx;
  double lo, hi;
  axis_bounds(c, axis, idx, lo, hi);
  int flag = 0;
  if (coord < lo) flag = 1;
  else if (coord >= hi) flag = 2;
  if (flag == 0) return idx;
  bool at_edge = false;
  const bool fw = (flag == 2);
  const uint32_t nb = nghbr(c, idx, axis, fw, at_edge);
  if (at_edge && c.per[axis]) {
    double nlo, nhi, ilo, ihi;
    axis_bounds(c, axis, nb, nlo, nhi);
    axis_bounds(c, axis, idx, ilo, ihi);
    const double lim1 = fw ? nlo : nhi;
    const double lim2 = fw ? ihi : ilo;
    coord = coord + lim1 - lim2;
  }
  return nb;
}

static void motion_drop(const Ctx& c, uint32_t g, const State& s, double delt, int tvkind, Superdrop& d,
                        const ErrSink& e) {
  double zlo, zhi, xlo, xhi, ylo, yhi;
  bounds3(c, g, zlo, zhi);
  bounds1(c, g, xlo, xhi);
  bounds2(c, g, ylo, yhi);
  /* PredCorr::delta_coord3/1/2 (predcorr.hpp:84-145) */
  double coord3 = d.coord3;
  const double terminal = terminalv(tvkind, d.attrs.radius);
  double vel3 = interpolation(zlo, zhi, s.wlo, s.whi, coord3);
  vel3 -= terminal;
  coord3 += vel3 * delt;
  double corrvel3 = interpolation(zlo, zhi, s.wlo, s.whi, coord3);
  corrvel3 -= terminal;
  const double delta3 = (vel3 + corrvel3) * (delt / 2);

  double coord1 = d.coord1;
  const double vel1 = interpolation(xlo, xhi, s.ulo, s.uhi, coord1);
  coord1 += vel1 * delt;
  const double corrvel1 = interpolation(xlo, xhi, s.ulo, s.uhi, coord1);
  const double delta1 = (vel1 + corrvel1) * (delt / 2);

  double coord2 = d.coord2;
  const double vel2 = interpolation(ylo, yhi, s.vlo, s.vhi, coord2);
  coord2 += vel2 * delt;
  const double corrvel2 = interpolation(ylo, yhi, s.vlo, s.vhi, coord2);
  const double delta2 = (vel2 + corrvel2) * (delt / 2);

  /* cfl_criteria (cfl_criteria.hpp:40-57) */
  bool cfl = (std::abs(delta3) <= std::abs(zhi - zlo));
  cfl = (std::abs(delta1) <= std::abs(xhi - xlo)) && cfl;
  cfl = (std::abs(delta2) <= std::abs(yhi - ylo)) && cfl;
  if (!cfl) e.set(CLEO_ERR_CFL);
  d.coord3 += delta3;
  d.coord1 += delta1;
  d.coord2 += delta2;

  /* superdrop_gbx (predcorrmotion.hpp:85-106) */
  double c3 = d.coord3, c1 = d.coord1, c2 = d.coord2;
  uint32_t idx = change_if_nghbr(c, g, 0, c3);
  idx = change_if_nghbr(c, idx, 1, c1);
  idx = change_if_nghbr(c, idx, 2, c2);
  d.coord3 = c3;
  d.coord1 = c1;
  d.coord2 = c2;
  d.sdgbxindex = idx;
  if (idx >= c.ngbx) return;
  double lo, hi;
  bool bad = false;
  bounds3(c, idx, lo, hi);
  bad |= (d.coord3 < lo) | (d.coord3 >= hi);
  bounds1(c, idx, lo, hi);
  bad |= (d.coord1 < lo) | (d.coord1 >= hi);
  bounds2(c, idx, lo, hi);
  bad |= (d.coord2 < lo) | (d.coord2 >= hi);
  if (bad) e.set(CLEO_ERR_BOUNDS);
}

/* EffectOnHydrometeorStatesFunctor (movesupersindomain.hpp:95-136) */
static void hydrometeor_qcond(Ctx& c) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int64_t gi = 0; gi < static_cast<int64_t>(c.ngbx); ++gi) {
    const uint32_t g = static_cast<uint32_t>(gi);
    double tot = 0.0;
    for (uint64_t k = c.ref0[g]; k < c.ref1[g]; ++k) {
      const ErrSink e{&c.err, g, k};
      const Superdrop& d = c.sup[k];
      tot += condensate_mass(d, c.sol, e) * d.attrs.xi;
    }
    constexpr double R0cubed_VOL0 = dlc::R0 * dlc::R0 * dlc::R0 / dlc::VOL0;
    State& s = c.st[g];
    const double rho_cond = tot / s.volume * R0cubed_VOL0;
    const double rho_dry = dry_air_density(s.press, s.temp, s.qvap);
    s.qcond = rho_cond / rho_dry;
  }
}

static void do_motion(Ctx& c, const cleo_motion_params& mp) {
  const double delt = step2dimlesstime(mp.interval);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int64_t gi = 0; gi < static_cast<int64_t>(c.ngbx); ++gi) {
    const uint32_t g = static_cast<uint32_t>(gi);
    for (uint64_t k = c.ref0[g]; k < c.ref1[g]; ++k) {
      const ErrSink e{&c.err, g, k};
      motion_drop(c, g, c.st[g], delt, mp.terminalv, c.sup[k], e);
    }
  }
  c.stats.motion_drops += c.ndomain;
  sort_supers(c);
  hydrometeor_qcond(c);
}

/* ---------------- condensation per gridbox: condensation.cpp:44-137 ---------------- */
static void cond_gbx(Ctx& c, uint32_t g, const cleo_cond_params& p, uint64_t& iters, uint64_t& subs,
                     uint64_t& restarts) {
  State& s = c.st[g];
  const ErrSink eg{&c.err, g, c.ref0[g]};
  const double psat = saturation_pressure(s.temp, eg);
  const double s_ratio = supersaturation_ratio(s.press, s.qvap, psat);
  const double ffactor = diffusion_factor(s.press, s.temp, psat);
  NewtonCfg cfg{step2dimlesstime(p.interval), p.minsubdelt_s / dlc::TIME0, p.rtol, p.atol,
                static_cast<uint64_t>(p.maxniters)};
  double tot = 0.0;
  const uint64_t k0 = c.ref0[g], k1 = c.ref0[g] + c.nact[g];
  for (uint64_t k = k0; k < k1; ++k) {
    const ErrSink e{&c.err, g, k};
    Superdrop& d = c.sup[k];
    const double old_m = condensate_mass(d, c.sol, e);
    double akoh, bkoh;
    kohler_factors(d, s.temp, c.sol, akoh, bkoh);
    NewtonCount cnt;
    const double warm = (p.warm_start && !c.guess.empty()) ? c.guess[k] : NAN;
    const double newr = solve_condensation(cfg, s_ratio, akoh, bkoh, ffactor, d.attrs.radius, warm,
                                           p.warm_restart_iters, cnt, e);
    change_radius(d, newr, c.sol);
    tot += (condensate_mass(d, c.sol, e) - old_m) * d.attrs.xi;
    iters += cnt.iters;
    subs += cnt.substep ? 1 : 0;
    restarts += cnt.restarted ? 1 : 0;
    if (c.rec_niters) c.niters[k] = cnt.iters;
  }
  if (p.do_alter_thermo) {
    constexpr double R0cubed_VOL0 = dlc::R0 * dlc::R0 * dlc::R0 / dlc::VOL0;
    const double totrho = tot / s.volume * R0cubed_VOL0;
    const double rho_dry = dry_air_density(s.press, s.temp, s.qvap);
    const double dq = totrho / rho_dry;
    s.qvap -= dq;
    s.qcond += dq;
    const double dtemp = (dlc::Latent_v / moist_specifc_heat(s.qvap, s.qcond)) * dq;
    s.temp += dtemp;
  }
}

/* ---------------- collisions per gridbox: collisions.hpp:140-278 ---------------- */
struct CollCounters {
  uint64_t pairs = 0, events = 0, nulls = 0;
};
static bool coll_gbx(Ctx& c, uint32_t g, const cleo_coll_params& p, uint64_t event, uint64_t inj_j0,
                     uint64_t inj_p0, CollCounters& cc) {
  const uint64_t n = c.nact[g];
  Superdrop* sub = c.sup.data() + c.ref0[g];
  /* shuffle_supers / fisher_yates_shuffle (shuffle.cpp:38-53, shuffle.hpp:80-87) */
  if (n >= 2) {
    uint64_t t = 0;
    uint32_t r[4] = {0u, 0u, 0u, 0u};
    for (uint64_t iter = n - 1; iter > 0; --iter, ++t) {
      uint64_t j;
      if (c.rng_mode == 2) {
        j = c.inj_j[inj_j0 + t];
      } else {
        /* counter stream: word t%4 of the block for item t/4 */
        if ((t & 3u) == 0u) rng_block(c.seed, event, 0u, g, static_cast<uint32_t>(t >> 2), r);
        j = (static_cast<uint64_t>(r[t & 3u]) * (iter + 1)) >> 32;
      }
      std::swap(sub[iter], sub[j]);
    }
  }
  /* collide_supers (collisions.hpp:215-229) */
  const uint64_t npairs = n / 2;
  const double scale_p = n * (n - 1.0) / (2.0 * npairs);
  const double VOLUME = c.st[g].volume * dlc::VOL0;
  const double DELT = step2realtime(p.interval);
  size_t nulls = 0;
  for (uint64_t jj = 0; jj < npairs; ++jj) {
    const uint64_t kk = 2 * jj;
    Superdrop& A = sub[kk];
    Superdrop& B = sub[kk + 1];
    Superdrop& d1 = !(A.attrs.xi < B.attrs.xi) ? A : B;
    Superdrop& d2 = !(A.attrs.xi < B.attrs.xi) ? B : A;
    const double prob_jk = pair_prob(p, d1, d2, DELT, VOLUME);
    const double large_xi = static_cast<double>(d1.attrs.xi);
    const double prob = scale_p * large_xi * prob_jk;
    double phi_coll, phi_out;
    if (c.rng_mode == 2) {
      phi_coll = c.inj_phi[inj_p0 + 2 * jj];
      phi_out = c.inj_phi[inj_p0 + 2 * jj + 1];
    } else {
      uint32_t r[4];
      rng_block(c.seed, event, 1u, g, static_cast<uint32_t>(jj), r);
      phi_coll = u53(r[0], r[1]);
      phi_out = u53(r[2], r[3]);
    }
    const ErrSink e{&c.err, g, c.ref0[g] + kk};
    bool enacted = false;
    nulls += enact(p, d1, d2, prob, phi_coll, phi_out, c.sol, e, enacted);
    cc.events += enacted ? 1 : 0;
  }
  cc.pairs += npairs;
  cc.nulls += nulls;
  if (nulls > 0) {
    /* remove_null_supers: stands in for Kokkos sort_team (collisions.hpp:240-248) */
    std::stable_partition(sub, sub + n, [](const Superdrop& d) { return d.attrs.xi != 0; });
    c.nact[g] = n - nulls;
    return true;
  }
  return false;
}

/* injected stream offsets: per gridbox max(n-1,0) indices and 2*floor(n/2) phis */
static bool injected_offsets(const Ctx& c, std::vector<uint64_t>& oj, std::vector<uint64_t>& op) {
  oj.assign(c.ngbx + 1, 0);
  op.assign(c.ngbx + 1, 0);
  for (uint32_t g = 0; g < c.ngbx; ++g) {
    const uint64_t n = c.nact[g];
    oj[g + 1] = oj[g] + (n >= 2 ? n - 1 : 0);
    op[g + 1] = op[g] + 2 * (n / 2);
  }
  return oj[c.ngbx] == c.inj_j.size() && op[c.ngbx] == c.inj_phi.size();
}

static uint32_t interval_of(const cleo_process& q) {
  return q.kind == CLEO_PROCESS_CONDENSATION ? q.cond.interval : q.coll.interval;
}

/* one microphysics pass over [t_sdm, t_next) for every gridbox (sdmmethods.hpp:70-92,184-230) of
   the chain procs[0] >> procs[1] >> ... (CombinedMicrophysicalProcess, microphysicalprocess.hpp:
   65-120): next_step is the minimum over the chain, on each sub-step the processes on step run in
   chain order */
static int microphysics(Ctx& c, uint32_t t_sdm, uint32_t t_next, const std::vector<cleo_process>& procs) {
  auto next_step = [&](uint32_t subt) -> uint32_t {
    uint32_t t = 0xFFFFFFFFu;
    for (const cleo_process& q : procs) t = std::min(t, ((subt / interval_of(q)) + 1) * interval_of(q));
    return t;
  };
  auto on_step = [&](const cleo_process& q, uint32_t subt) {
    return (subt % interval_of(q) == 0) && interval_of(q) != 0xFFFFFFFFu;
  };
  /* (sub-step, process) pairs of this pass in order; the same for every gridbox */
  std::vector<std::pair<uint32_t, size_t>> runs;
  for (uint32_t subt = t_sdm; subt < t_next; subt = next_step(subt))
    for (size_t i = 0; i < procs.size(); ++i)
      if (on_step(procs[i], subt)) runs.emplace_back(subt, i);
  uint64_t nevents = 0;
  for (const auto& r : runs) nevents += procs[r.second].kind == CLEO_PROCESS_COLLISIONS ? 1 : 0;
  std::vector<uint64_t> oj, op;
  if (nevents > 0 && c.rng_mode == 2) {
    if (nevents != 1) return fail(&c, CLEO_E_STATE, "injected RNG covers exactly one collision event");
    if (!injected_offsets(c, oj, op)) return fail(&c, CLEO_E_ARG, "injected stream size mismatch");
  }
  if (nevents > 0 && c.rng_mode == 0) return fail(&c, CLEO_E_STATE, "no RNG set");
  uint64_t iters = 0, subs = 0, restarts = 0, ndrops = 0;
  CollCounters cc;
  bool any_change = false;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : iters, subs, restarts, ndrops) reduction(|| : any_change)
#endif
  for (int64_t gi = 0; gi < static_cast<int64_t>(c.ngbx); ++gi) {
    const uint32_t g = static_cast<uint32_t>(gi);
    uint64_t ev = 0;
    CollCounters lc;
    for (const auto& r : runs) {
      const cleo_process& q = procs[r.second];
      if (q.kind == CLEO_PROCESS_CONDENSATION) {
        ndrops += c.nact[g];
        cond_gbx(c, g, q.cond, iters, subs, restarts);
      } else {
        const uint64_t j0 = oj.empty() ? 0 : oj[g];
        const uint64_t p0 = op.empty() ? 0 : op[g];
        any_change = coll_gbx(c, g, q.coll, c.event + ev, j0, p0, lc) || any_change;
        ++ev;
      }
    }
#ifdef _OPENMP
#pragma omp critical(orc_coll)
#endif
    {
      cc.pairs += lc.pairs;
      cc.events += lc.events;
      cc.nulls += lc.nulls;
    }
  }
  c.event += nevents;
  if (c.rng_mode == 2 && nevents > 0) {
    c.inj_j.clear();
    c.inj_phi.clear();
    c.rng_mode = 0;
  }
  c.stats.cond_drops += ndrops;
  c.stats.newton_iters += iters;
  c.stats.substep_drops += subs;
  c.stats.warm_restarts += restarts;
  c.stats.coll_pairs += cc.pairs;
  c.stats.coll_events += cc.events;
  c.stats.nulls += cc.nulls;
  if (any_change) sort_supers(c);
  return c.err.code ? CLEO_E_DEVICE : CLEO_OK;
}

}  // namespace orc

using namespace orc;

/* ---------------- C API mirroring include/cleo_hip.h with the orc_ prefix ---------------- */
extern "C" {

struct orc_ctx;

int orc_create(orc_ctx** out, uint64_t maxn, const cleo_domain_desc* dom) {
  if (!out || !dom || !dom->zhalf || !dom->xhalf || !dom->yhalf) return CLEO_E_ARG;
  if (dom->ndims[0] == 0 || dom->ndims[1] == 0 || dom->ndims[2] == 0) return CLEO_E_ARG;
  Ctx* c = new Ctx();
  c->nz = dom->ndims[0];
  c->nx = dom->ndims[1];
  c->ny = dom->ndims[2];
  c->ngbx = c->nz * c->nx * c->ny;
  for (int a = 0; a < 3; ++a) c->per[a] = dom->periodic[a];
  c->zh.assign(dom->zhalf, dom->zhalf + c->nz + 1);
  c->xh.assign(dom->xhalf, dom->xhalf + c->nx + 1);
  c->yh.assign(dom->yhalf, dom->yhalf + c->ny + 1);
  c->maxn = maxn;
  c->ref0.assign(c->ngbx, 0);
  c->ref1.assign(c->ngbx, 0);
  c->nact.assign(c->ngbx, 0);
  c->st.assign(c->ngbx, State{});
  *out = reinterpret_cast<orc_ctx*>(c);
  return CLEO_OK;
}
int orc_destroy(orc_ctx* h) {
  delete reinterpret_cast<Ctx*>(h);
  return CLEO_OK;
}
const char* orc_last_error_message(const orc_ctx* h) {
  return reinterpret_cast<const Ctx*>(h)->msg.c_str();
}
int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
/* threads of the OpenMP build's later parallel regions (the 1-thread CPU baseline) */
void orc_set_num_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_upload_supers(orc_ctx* h, const cleo_supers_soa* s, uint64_t ntot) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  if (!s || ntot > c.maxn) return fail(&c, CLEO_E_ARG, "ntot beyond capacity");
  c.sup.resize(ntot);
  for (uint64_t k = 0; k < ntot; ++k) {
    Superdrop& d = c.sup[k];
    std::memset(&d, 0, sizeof(d));
    d.sdgbxindex = s->sdgbxindex[k];
    d.coord3 = s->coord3[k];
    d.coord1 = s->coord1[k];
    d.coord2 = s->coord2[k];
    d.attrs.xi = s->xi[k];
    d.attrs.radius = s->radius[k];
    d.attrs.msol = s->msol[k];
    d.sdId = s->sdid[k];
  }
  sort_supers(c);
  c.stats.sorts = 0;
  return CLEO_OK;
}
int orc_download_supers(orc_ctx* h, cleo_supers_soa* s, uint64_t* ntot, uint64_t* ndomain) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  const uint64_t n = c.sup.size();
  for (uint64_t k = 0; k < n; ++k) {
    const Superdrop& d = c.sup[k];
    s->sdgbxindex[k] = d.sdgbxindex;
    s->coord3[k] = d.coord3;
    s->coord1[k] = d.coord1;
    s->coord2[k] = d.coord2;
    s->xi[k] = d.attrs.xi;
    s->radius[k] = d.attrs.radius;
    s->msol[k] = d.attrs.msol;
    s->sdid[k] = d.sdId;
  }
  if (ntot) *ntot = n;
  if (ndomain) *ndomain = c.ndomain;
  return CLEO_OK;
}
int orc_download_refs(orc_ctx* h, uint32_t* refs) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  for (uint32_t g = 0; g < c.ngbx; ++g) refs[g] = static_cast<uint32_t>(c.ref0[g]);
  refs[c.ngbx] = static_cast<uint32_t>(c.ndomain);
  return CLEO_OK;
}
/* FromFileComms::update_gridbox_state for every gridbox (coupldyn_fromfile/fromfilecomms.cpp:51-62)
   with the face positions of fromfile_cartesian_dynamics.cpp:150-215 (wvel: lower z face
   nx (nz+1) j + (nz+1) i + k, upper + 1; uvel: lower x face (nx+1) nz j + nz i + k, upper + nz;
   vvel: lower y face at the gridbox index, upper + nx nz); a null wind array is nullwinds (0, 0) */
int orc_receive_dynamics(orc_ctx* h, const cleo_dynamics_step* d) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  if (!d || !d->press || !d->temp || !d->qvap || !d->qcond) return CLEO_E_ARG;
  const uint64_t nz = c.nz, nx = c.nx;
  for (uint32_t g = 0; g < c.ngbx; ++g) {
    const uint64_t k = c.iz(g), i = c.ix(g), j = c.iy(g);
    State& t = c.st[g];
    t.press = d->press[g];
    t.temp = d->temp[g];
    t.qvap = d->qvap[g];
    t.qcond = d->qcond[g];
    const uint64_t wl = nx * (nz + 1) * j + (nz + 1) * i + k;
    const uint64_t ul = (nx + 1) * nz * j + nz * i + k;
    t.wlo = d->wvel_zfaces ? d->wvel_zfaces[wl] : 0.0;
    t.whi = d->wvel_zfaces ? d->wvel_zfaces[wl + 1] : 0.0;
    t.ulo = d->uvel_xfaces ? d->uvel_xfaces[ul] : 0.0;
    t.uhi = d->uvel_xfaces ? d->uvel_xfaces[ul + nz] : 0.0;
    t.vlo = d->vvel_yfaces ? d->vvel_yfaces[g] : 0.0;
    t.vhi = d->vvel_yfaces ? d->vvel_yfaces[g + nx * nz] : 0.0;
  }
  return CLEO_OK;
}
/* calculate_massmoments / calculate_rainmassmoments (observers/massmoments_observer.cpp:23-73) in
   Serial order: per gridbox, over its drops in order, mom0 += xi, mom1 += (float)(xi * mass),
   mom2 += (float)(xi * mass * mass); rain: only drops with radius >= 40 micrometres */
int orc_mass_moments(orc_ctx* h, int32_t rain, uint64_t* mom0, float* mom1, float* mom2) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  if (!mom0 || !mom1 || !mom2) return CLEO_E_ARG;
  constexpr double rlim(40e-6 / dlc::R0);
  for (uint32_t g = 0; g < c.ngbx; ++g) {
    uint64_t m0 = 0;
    float m1 = 0.f, m2 = 0.f;
    for (uint64_t k = c.ref0[g]; k < c.ref1[g]; ++k) {
      const Superdrop& d = c.sup[k];
      if (d.attrs.xi >= UINT64_MAX) ErrSink{&c.err, g, k}.set(CLEO_ERR_XI_OVERFLOW);
      const bool binary = rain ? d.attrs.radius >= rlim : true;
      m0 += static_cast<uint64_t>(d.attrs.xi * binary);
      const double m = mass(d, c.sol);
      const double xi = static_cast<double>(d.attrs.xi);
      m1 += static_cast<float>(xi * m * binary);
      m2 += static_cast<float>(xi * m * m * binary);
    }
    mom0[g] = m0;
    mom1[g] = m1;
    mom2[g] = m2;
  }
  return c.err.code ? CLEO_E_DEVICE : CLEO_OK;
}
int orc_upload_gbx_state(orc_ctx* h, const cleo_gbx_state* s) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  for (uint32_t g = 0; g < c.ngbx; ++g) {
    c.st[g] = State{s->volume[g], s->press[g], s->temp[g], s->qvap[g], s->qcond[g], s->wvel_lo[g],
                    s->wvel_hi[g], s->uvel_lo[g], s->uvel_hi[g], s->vvel_lo[g], s->vvel_hi[g]};
  }
  return CLEO_OK;
}
int orc_download_gbx_state(orc_ctx* h, cleo_gbx_state* s) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  for (uint32_t g = 0; g < c.ngbx; ++g) {
    const State& t = c.st[g];
    s->volume[g] = t.volume;
    s->press[g] = t.press;
    s->temp[g] = t.temp;
    s->qvap[g] = t.qvap;
    s->qcond[g] = t.qcond;
    s->wvel_lo[g] = t.wlo;
    s->wvel_hi[g] = t.whi;
    s->uvel_lo[g] = t.ulo;
    s->uvel_hi[g] = t.uhi;
    s->vvel_lo[g] = t.vlo;
    s->vvel_hi[g] = t.vhi;
  }
  return CLEO_OK;
}
int orc_set_solute(orc_ctx* h, const cleo_solute* s) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  c.sol.rho_sol = s->rho_sol;
  c.sol.mr_sol = s->mr_sol;
  c.sol.ionic = s->ionic;
  return CLEO_OK;
}
int orc_set_rng_counter(orc_ctx* h, uint64_t seed, uint64_t first_event) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  c.rng_mode = 1;
  c.seed = seed;
  c.event = first_event;
  return CLEO_OK;
}
int orc_set_rng_injected(orc_ctx* h, const uint32_t* j, uint64_t nj, const double* phi, uint64_t np) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  c.inj_j.assign(j, j + nj);
  c.inj_phi.assign(phi, phi + np);
  c.rng_mode = 2;
  return CLEO_OK;
}
/* test hook: per-drop learned guesses (z = r^2, array position order) for warm_start */
int orc_set_cond_guess(orc_ctx* h, const double* z, uint64_t n) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  if (z) c.guess.assign(z, z + n);
  else c.guess.clear();
  return CLEO_OK;
}
int orc_set_sdm_chain(orc_ctx* h, const cleo_process* procs, uint32_t n, const cleo_motion_params* motion) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  if (n > CLEO_MAX_PROCESSES || (n && !procs)) return fail(&c, CLEO_E_ARG, "process list");
  for (uint32_t i = 0; i < n; ++i) {
    if (procs[i].kind != CLEO_PROCESS_CONDENSATION && procs[i].kind != CLEO_PROCESS_COLLISIONS)
      return fail(&c, CLEO_E_ARG, "unknown process kind");
    if (interval_of(procs[i]) == 0) return fail(&c, CLEO_E_ARG, "zero interval");
  }
  if (motion && motion->interval == 0) return fail(&c, CLEO_E_ARG, "zero interval");
  c.procs.assign(procs, procs + n);
  c.motion_on = motion != nullptr;
  if (motion) c.motion = *motion;
  c.sdm_set = true;
  return CLEO_OK;
}
int orc_set_sdm(orc_ctx* h, const cleo_sdm_params* p) {
  cleo_process q[2]{};
  uint32_t n = 0;
  if (p->enable_cond) {
    q[n].kind = CLEO_PROCESS_CONDENSATION;
    q[n++].cond = p->cond;
  }
  if (p->enable_colls) {
    q[n].kind = CLEO_PROCESS_COLLISIONS;
    q[n++].coll = p->coll;
  }
  return orc_set_sdm_chain(h, q, n, p->enable_motion ? &p->motion : nullptr);
}
int orc_sort(orc_ctx* h) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  sort_supers(c);
  return CLEO_OK;
}
int orc_motion(orc_ctx* h, const cleo_motion_params* p) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  do_motion(c, *p);
  return c.err.code ? CLEO_E_DEVICE : CLEO_OK;
}
int orc_condensation(orc_ctx* h, const cleo_cond_params* p) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  if (p->interval == 0) return fail(&c, CLEO_E_ARG, "interval 0");
  if (step2dimlesstime(p->interval) < p->minsubdelt_s / dlc::TIME0) {
    c.err.code = c.err.code ? c.err.code : CLEO_ERR_SUBSTEP;
    return CLEO_E_DEVICE;
  }
  if (c.rec_niters) c.niters.assign(c.sup.size(), 0);
  cleo_process q{};
  q.kind = CLEO_PROCESS_CONDENSATION;
  q.cond = *p;
  return microphysics(c, 0, 1, {q});
}
int orc_collisions(orc_ctx* h, const cleo_coll_params* p) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  if (p->interval == 0) return fail(&c, CLEO_E_ARG, "interval 0");
  cleo_process q{};
  q.kind = CLEO_PROCESS_COLLISIONS;
  q.coll = *p;
  return microphysics(c, 0, 1, {q});
}
int orc_run_step(orc_ctx* h, uint32_t t_mdl, uint32_t t_mdl_next) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  if (!c.sdm_set) return fail(&c, CLEO_E_STATE, "cleo_set_sdm not called");
  if (c.rec_niters) c.niters.assign(c.sup.size(), 0);
  uint32_t t_sdm = t_mdl;
  while (t_sdm < t_mdl_next) {
    uint32_t next_move = 0xFFFFFFFFu;
    if (c.motion_on) next_move = ((t_sdm / c.motion.interval) + 1) * c.motion.interval;
    const uint32_t t_sdm_next = std::min(t_mdl_next, next_move);
    if (c.motion_on && (t_sdm % c.motion.interval == 0)) do_motion(c, c.motion);
    const int rc = microphysics(c, t_sdm, t_sdm_next, c.procs);
    if (rc != CLEO_OK) return rc;
    t_sdm = t_sdm_next;
  }
  return c.err.code ? CLEO_E_DEVICE : CLEO_OK;
}
int orc_get_error(orc_ctx* h, cleo_error* e) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  e->code = c.err.code;
  e->gbxindex = c.err.gbx;
  e->drop = c.err.drop;
  return CLEO_OK;
}
int orc_clear_error(orc_ctx* h) {
  reinterpret_cast<Ctx*>(h)->err = ErrWord{};
  return CLEO_OK;
}
int orc_get_stats(orc_ctx* h, cleo_stats* s) {
  *s = reinterpret_cast<Ctx*>(h)->stats;
  return CLEO_OK;
}
int orc_reset_stats(orc_ctx* h) {
  reinterpret_cast<Ctx*>(h)->stats = cleo_stats{};
  return CLEO_OK;
}
int orc_record_niters(orc_ctx* h, int32_t en) {
  reinterpret_cast<Ctx*>(h)->rec_niters = en != 0;
  return CLEO_OK;
}
int orc_download_niters(orc_ctx* h, uint32_t* out, uint64_t n) {
  Ctx& c = *reinterpret_cast<Ctx*>(h);
  for (uint64_t k = 0; k < n; ++k) out[k] = k < c.niters.size() ? c.niters[k] : 0;
  return CLEO_OK;
}

/* ---- pieces exposed for closed-form tests (pure functions) ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox4x32_10(ctr, key, out);
}
double orc_terminalv(int kind, double radius) { return terminalv(kind, radius); }
double orc_saturation_pressure(double temp) {
  ErrWord w;
  const ErrSink e{&w, 0, 0};
  return saturation_pressure(temp, e);
}
double orc_diffusion_factor(double press, double temp, double psat) {
  return diffusion_factor(press, temp, psat);
}
double orc_ventilation_factor(double r) { return ventilation_factor(r); }
double orc_supersaturation_ratio(double press, double qvap, double psat) {
  return supersaturation_ratio(press, qvap, psat);
}
/* Koehler a and b of one drop of solute mass msol (thermodynamic_equations.hpp:92-107) */
void orc_kohler_factors(double temp, double msol, double rho_sol, double mr_sol, double ionic, double* ab) {
  Superdrop d{};
  d.attrs.msol = msol;
  Solute s;
  s.rho_sol = rho_sol;
  s.mr_sol = mr_sol;
  s.ionic = ionic;
  kohler_factors(d, temp, s, ab[0], ab[1]);
}
/* one drop through solve_condensation; returns new radius, writes iteration count */
double orc_solve_condensation(double s_ratio, double akoh, double bkoh, double ffactor, double rprev,
                              double delt, double minsubdelt, double rtol, double atol,
                              uint64_t maxniters, double warm, uint64_t warm_limit, uint32_t* niters,
                              int32_t* substep, int32_t* err) {
  ErrWord w;
  const ErrSink e{&w, 0, 0};
  NewtonCfg cfg{delt, minsubdelt, rtol, atol, maxniters};
  NewtonCount cnt;
  const double r =
      solve_condensation(cfg, s_ratio, akoh, bkoh, ffactor, rprev, warm, warm_limit, cnt, e);
  if (niters) *niters = cnt.iters;
  if (substep) *substep = cnt.substep ? 1 : 0;
  if (err) *err = w.code;
  return r;
}
double orc_pair_prob(int kind, double param, double r1, double xi1, double r2, double xi2, double DELT,
                     double VOLUME) {
  Superdrop a{}, b{};
  a.attrs.radius = r1;
  a.attrs.xi = static_cast<uint64_t>(xi1);
  b.attrs.radius = r2;
  b.attrs.xi = static_cast<uint64_t>(xi2);
  cleo_coll_params p{};
  p.prob_kind = kind;
  p.prob_param = param;
  return pair_prob(p, a, b, DELT, VOLUME);
}
/* collision kinetics of two drops of dimensionless radii r1, r2 */
double orc_cke(double r1, double r2, double v1, double v2) { return coll_kinetic_energy(r1, r2, v1, v2); }
void orc_surf_energies(double r1, double r2, double* out) {
  out[0] = surfenergy(r1);
  out[1] = total_surfenergy(r1, r2);
  out[2] = coal_surfenergy(r1, r2);
}
static inline void two_drops(double r1, double r2, Superdrop& a, Superdrop& b) {
  a = Superdrop{};
  b = Superdrop{};
  a.attrs.radius = r1;
  b.attrs.radius = r2;
  a.attrs.xi = b.attrs.xi = 1;
}
double orc_lowlist_coaleff(int tv, double r1, double r2) {
  Superdrop a, b;
  two_drops(r1, r2, a, b);
  return lowlist_coaleff(tv, a, b);
}
double orc_nfrags(const cleo_coll_params* p, double r1, double r2) {
  Superdrop a, b;
  two_drops(r1, r2, a, b);
  return nfrags_of(*p, a, b);
}
int32_t orc_coalbure_flag(const cleo_coll_params* p, double phi, double r1, double r2) {
  Superdrop a, b;
  two_drops(r1, r2, a, b);
  return static_cast<int32_t>(coalbure_flag(*p, phi, a, b));
}
double orc_pair_prob_params(const cleo_coll_params* p, double r1, double r2, double DELT, double VOLUME) {
  Superdrop a, b;
  two_drops(r1, r2, a, b);
  return pair_prob(*p, a, b, DELT, VOLUME);
}
/* enact one collision on two drops given as (xi, r, msol) arrays of length 2 (drop1, drop2) */
int32_t orc_enact_pair(const cleo_coll_params* p, uint64_t* xi, double* r, double* msol, double prob,
                       double phi_coll, double phi_out, int32_t* err) {
  ErrWord w;
  const ErrSink e{&w, 0, 0};
  Superdrop d1{}, d2{};
  d1.attrs.xi = xi[0];
  d1.attrs.radius = r[0];
  d1.attrs.msol = msol[0];
  d2.attrs.xi = xi[1];
  d2.attrs.radius = r[1];
  d2.attrs.msol = msol[1];
  Solute s;
  bool enacted = false;
  const size_t nulls = enact(*p, d1, d2, prob, phi_coll, phi_out, s, e, enacted);
  xi[0] = d1.attrs.xi;
  r[0] = d1.attrs.radius;
  msol[0] = d1.attrs.msol;
  xi[1] = d2.attrs.xi;
  r[1] = d2.attrs.radius;
  msol[1] = d2.attrs.msol;
  if (err) *err = w.code;
  return static_cast<int32_t>(nulls) | (enacted ? 2 : 0);
}
/* reference constants for unit checks */
void orc_constants(double* out) {
  out[0] = dlc::RHO0;
  out[1] = dlc::MASS0;
  out[2] = dlc::F0;
  out[3] = dlc::Latent_v;
  out[4] = dlc::Rho_l;
  out[5] = dlc::Rho_sol;
  out[6] = dlc::Mr_sol;
  out[7] = dlc::Mr_ratio;
  out[8] = step2realtime(100);
  out[9] = step2dimlesstime(100);
}

}  // extern "C"
