// physics.hpp -- device-side super-droplet physics for gfx950.
//
// Each function states the reference formula it evaluates (CLEO libs/...); the floating-point
// operation order follows the reference expression so that, compiled with -ffp-contract=off,
// every +,-,*,/,sqrt rounds as the reference's host build does.  Transcendentals (pow, exp)
// come from the device math library and may differ from glibc by an ulp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cleo_hip.h"

namespace cleo {

// ---- constants: libs/cleoconstants.hpp:32-104 -------------------------------------------
namespace dimmed {
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
constexpr double SURFSIGMA = 7.28e-2;  // surface tension of water [J/m^2]
}  // namespace dimmed
namespace dimless {
constexpr double W0 = 1.0;
constexpr double TIME0 = 1000.0;
constexpr double COORD0 = TIME0 * W0;
constexpr double VOL0 = COORD0 * COORD0 * COORD0;
constexpr double CP0 = dimmed::CP_DRY;
constexpr double MR0 = dimmed::MR_DRY;
constexpr double R0 = 1e-6;
constexpr double P0 = 100000.0;
constexpr double TEMP0 = 273.15;
constexpr double RHO0 = P0 / (CP0 * TEMP0);
constexpr double MASS0 = R0 * R0 * R0 * RHO0;
constexpr double MASS0grams = MASS0 * 1000;
constexpr double F0 = TIME0 / (RHO0 * R0 * R0);
constexpr double Mr_ratio = dimmed::MR_WATER / dimmed::MR_DRY;
constexpr double Cp_dry = dimmed::CP_DRY / CP0;
constexpr double Cp_v = dimmed::CP_V / CP0;
constexpr double C_l = dimmed::C_L / CP0;
constexpr double Latent_v = dimmed::LATENT_V / (TEMP0 * CP0);
constexpr double Rgas_dry = dimmed::RGAS_DRY / CP0;
constexpr double Rgas_v = dimmed::RGAS_V / CP0;
constexpr double Rho_l = dimmed::RHO_L / RHO0;
constexpr double R0cubed_VOL0 = R0 * R0 * R0 / VOL0;
}  // namespace dimless
constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr uint32_t kOOB = 0xFFFFFFFFu;
constexpr double kDblMax = 1.7976931348623157e308;

struct Solute {
  double rho_sol, mr_sol, ionic;
  double dens_factor;  // 1 - Rho_l / rho_sol (of drop_mass), rounded once when the solute is set
};

// the solute with its derived constant (host and device evaluate the same IEEE expression)
inline Solute make_solute(double rho_sol, double mr_sol, double ionic) {
  return Solute{rho_sol, mr_sol, ionic, 1.0 - dimless::Rho_l / rho_sol};
}

// ---- device control block ------------------------------------------------------------------
enum StatIdx {
  ST_COND_DROPS = 0,
  ST_NEWTON_ITERS,
  ST_SUBSTEP_DROPS,
  ST_WARM_RESTARTS,
  ST_COLL_PAIRS,
  ST_COLL_EVENTS,
  ST_NULLS,
  ST_SORTS,
  ST_MOTION_DROPS,
  ST_COUNT
};
struct DevCtl {
  int32_t err_code;
  uint32_t err_gbx;
  uint64_t err_drop;
  uint32_t resort_flag;  // nulls appeared in this microphysics pass
  uint32_t pad0;
  uint64_t event;        // collision events so far (counter RNG)
  unsigned long long stats[ST_COUNT];
};

__device__ __forceinline__ void record_error(DevCtl* ctl, int code, uint32_t gbx, uint64_t drop) {
  if (atomicCAS(&ctl->err_code, 0, code) == 0) {
    ctl->err_gbx = gbx;
    ctl->err_drop = drop;
  }
}

// ---- reciprocals from the hardware estimates -------------------------------------------------
// 1/x for finite nonzero normal x: v_rcp_f64 and one Newton-Raphson step, within about an ulp.
__device__ __forceinline__ double rcp_nr(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}
// n / d for the Newton update: n times the refined reciprocal of d and one correction step
// (q + r (n - d q)), within an ulp of the quotient; a d or n outside [1e-290, 1e290] in magnitude
// (or zero, inf, NaN) takes the division itself, so the special cases are those of n / d.
__device__ __forceinline__ double quotient(double n, double d) {
  const double ad = fabs(d), an = fabs(n);
  if (ad > 1e-290 && ad < 1e290 && an < 1e290) {
    const double r = rcp_nr(d);
    const double q = n * r;
    return fma(fma(-d, q, n), r, q);
  }
  return n / d;
}
// 1/sqrt(x): v_rsq_f64 and two Newton-Raphson steps y <- y + y/2 (1 - x y^2), within an ulp or
// two of 1.0 / sqrt(x) (itself two roundings) for positive finite x; 0 and +inf give the
// estimate's own inf and 0, negative or NaN x gives NaN, as 1.0 / sqrt(x) does.
__device__ __forceinline__ double rsqrt_nr(double x) {
  const double y0 = __builtin_amdgcn_rsq(x);
  double y = y0;
  for (int i = 0; i < 2; ++i) {
    const double e = fma(-(x * y), y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return (x > 0.0 && x < __builtin_huge_val()) ? y : y0;
}

// ---- superdrop attributes: superdrop_attrs.{hpp,cpp} ---------------------------------------
__device__ __forceinline__ double rcubed(double r) { return r * r * r; }
__device__ __forceinline__ double dryradius(double msol, const Solute& s) {
  constexpr double vconst = 3.0 / (4.0 * kPi);
  const double dryrcubed = vconst * msol / s.rho_sol;
  return cbrt(dryrcubed);  // pow(x, 1.0 / 3.0): the exponent is 1/3 rounded; cbrt is within 1 ulp of it
}
// fmax(r, dryradius(msol)) of change_radius (superdrop_attrs.cpp:35-45).  The cube root is taken
// only when r^3 is not clearly above the dry radius cubed (1e-9 relative, far beyond the rounding
// of either side): for every other drop the maximum is r itself.
__device__ __forceinline__ double at_least_dry(double r, double msol, const Solute& s) {
  constexpr double vconst = 3.0 / (4.0 * kPi);
  const double dry3 = vconst * msol * rcp_nr(s.rho_sol);
  if (r * r * r > dry3 * (1.0 + 1e-9)) return r;
  return fmax(r, dryradius(msol, s));
}
__device__ __forceinline__ double drop_mass(double r, double msol, const Solute& s) {
  constexpr double massconst(4.0 / 3.0 * kPi * dimless::Rho_l);
  double m = msol * s.dens_factor;
  m += massconst * rcubed(r);
  return m;
}
// condensate_mass; *bad set when below the reference's tolerance (superdrop_attrs.hpp:207-216)
__device__ __forceinline__ double condensate_mass(double r, double msol, const Solute& s, bool* bad) {
  const double m_cond = drop_mass(r, msol, s) - msol;
  if (m_cond <= -0.0001 * msol) *bad = true;
  return fmax(0.0, m_cond);
}
__device__ __forceinline__ bool radius_below_dry(double r, double msol, const Solute& s) {
  return r - dryradius(msol, s) <= -1e-12 / dimless::R0;
}

// ---- correctly rounded exp ------------------------------------------------------------------
// Double-double values hi + lo (|lo| <= ulp(hi) / 2) and their error-free sums and products.
struct DD {
  double hi, lo;
};
__device__ __forceinline__ DD dd_norm(double a, double b) {  // |a| >= |b|
  const double s = a + b;
  return DD{s, b - (s - a)};
}
__device__ __forceinline__ DD dd_add_d(DD a, double b) {  // a + b for any magnitudes
  const double s = a.hi + b;
  const double bb = s - a.hi;
  const double e = (a.hi - (s - bb)) + (b - bb);
  return dd_norm(s, e + a.lo);
}
__device__ __forceinline__ DD dd_mul(DD a, DD b) {
  const double p = a.hi * b.hi;
  const double e = fma(a.hi, b.hi, -p) + (a.hi * b.lo + a.lo * b.hi);
  return dd_norm(p, e);
}
__device__ __forceinline__ DD dd_div_n(DD a, double n) {  // a / n for a small integer n
  const double q1 = a.hi / n;
  const double rem = fma(-q1, n, a.hi);  // a.hi - q1 n exactly
  return dd_norm(q1, (rem + a.lo) / n);
}
// e^x rounded to nearest, for the arguments of saturation_pressure (|x| well below 700): x = k ln2
// + r with ln2 in two parts (k ln2_hi exact through an FMA), e^(r / 256) by its Taylor series to
// the 10th term in double-double arithmetic (|r / 256| < 1.4e-3, remainder below 2^-110), squared 8
// times (relative error about 2^-95), times 2^k.  The result is the correctly rounded e^x unless
// e^x lies within ~2^-95 relative of a midpoint between doubles.  glibc's exp errs by at most about
// 0.51 ulp, so it too returns the correctly rounded value except where e^x lies within about 0.01
// ulp of a midpoint, and the two agree bit for bit outside those arguments; the device library's
// exp is accurate to about an ulp instead.
__device__ __forceinline__ double exp_rn(double x) {
#pragma clang fp contract(off)
  constexpr double kLn2Hi = 6.93147180559945286e-01, kLn2Lo = 2.31904681384629956e-17;
  constexpr double kInvLn2 = 1.4426950408889634;
  const double k = rint(x * kInvLn2);
  const double p = k * kLn2Hi;
  const double pe = fma(k, kLn2Hi, -p);  // k ln2_hi - p exactly
  DD r = dd_add_d(DD{x - p, 0.0}, -(pe + k * kLn2Lo));  // x - p is exact (Sterbenz) for k != 0
  r.hi = ldexp(r.hi, -8);
  r.lo = ldexp(r.lo, -8);
  DD t{1.0, 0.0};
#pragma nounroll
  for (int n = 10; n >= 1; --n) t = dd_add_d(dd_div_n(dd_mul(t, r), static_cast<double>(n)), 1.0);
#pragma nounroll
  for (int i = 0; i < 8; ++i) t = dd_mul(t, t);
  return ldexp(t.hi + t.lo, static_cast<int>(k));
}

// ---- thermodynamics: thermodynamic_equations.{hpp,cpp} -------------------------------------
__device__ __forceinline__ double dry_air_density(double press, double temp, double qvap) {
  const double rgas_equiv = dimless::Rgas_dry + dimless::Rgas_v * qvap;
  return press / (rgas_equiv * temp);
}
__device__ __forceinline__ double moist_specific_heat(double qvap, double qcond) {
  return dimless::Cp_dry + dimless::Cp_v * (qvap) + dimless::C_l * (qcond);
}
__device__ __forceinline__ double supersaturation_ratio(double press, double qvap, double psat) {
  return (press * qvap) / ((dimless::Mr_ratio + qvap) * psat);
}
__device__ __forceinline__ double saturation_pressure(double temp) {
  constexpr double A = 17.4146, B = 33.639, TREF = 273.16, PREF = 611.655;
  const double T = temp * dimless::TEMP0;
  return (PREF * exp_rn(A * (T - TREF) / (T - B))) / dimless::P0;
}
__device__ __forceinline__ double diffusion_factor(double press, double temp, double psat) {
  constexpr double A = 7.11756e-5;
  constexpr double B = 4.38127686e-3;
  constexpr double LATENT_RGAS_V = dimmed::LATENT_V / dimmed::RGAS_V;
  constexpr double D = 4.012182971e-5;
  const double TEMP = temp * dimless::TEMP0;
  const double PRESS = press * dimless::P0;
  const double PSAT = psat * dimless::P0;
  const double THERMK = A * (TEMP * TEMP) + TEMP * B;  // pow(TEMP, 2.0) is exactly TEMP*TEMP rounded
  const double DIFFUSE_V = (D / PRESS * pow(TEMP, 1.94)) / dimmed::RGAS_V;
  const double fkl = (LATENT_RGAS_V / TEMP - 1.0) * dimmed::LATENT_V / (THERMK * dimless::F0);
  const double fdl = TEMP / (DIFFUSE_V * PSAT) / dimless::F0;
  return dimless::Rho_l * (fkl + fdl);
}
// ffactor / ventilation_factor(radius) (impliciteuler.cpp:38, thermodynamic_equations.cpp:132-145)
// and its reciprocal with two divisions instead of five.  The reference's factor is
// min(1 + 1/(a + b), 20) with a = 1/(C1 x^A), b = 1/(C2 x^B), x the radius in metres.  With
// u = C1 x^A and v = C2 x^B, 1 + 1/(1/u + 1/v) = 1 + uv/(u + v), so below the cap
// ffactor / vent = ffactor (u + v) / (u + v + uv).  x^A and x^B come from one logarithm:
// exp(A log x) is within ~|A log x| ulp of pow(x, A) (a few 1e-15 relative here), and both
// results are within a few ulp of the reference's quotients.
// The quotient is returned as num / den (ffv_num, ffv_den), with inv_ffv = den / num rounded once:
// the Newton evaluations use inv_ffv, and ffactor / vent itself (num / den, rounded once as
// before) is formed only where the critical timestep needs it.
__device__ __forceinline__ void ffactor_over_vent(double ffactor, double radius, double& ffv_num, double& ffv_den,
                                                  double& inv_ffv) {
  constexpr double vent_factor_max = 20;
  constexpr double C1 = 6.954e+7, A = 1.963, C2 = 1.069e+3, B = 0.702;
  const double lx = log(radius * dimless::R0);
  const double u = C1 * exp(A * lx), v = C2 * exp(B * lx);
  const double sum = u + v, prod = u * v;
  double num = ffactor * sum, den = sum + prod;
  if (prod > (vent_factor_max - 1.0) * sum) {
    num = ffactor;
    den = vent_factor_max;
  } else if (!(sum > 0.0)) {  // radius 0: 1/u and 1/v infinite, factor 1
    num = ffactor;
    den = 1.0;
  }
  ffv_num = num;
  ffv_den = den;
  inv_ffv = den / num;
}
// the Koehler a factor depends on the gridbox's temperature only (kept per gridbox with the
// saturation ratio), b on the drop's solute mass
__device__ __forceinline__ double kohler_a(double temp) {
  constexpr double akoh_constant = 3.3e-7 / (dimless::TEMP0 * dimless::R0);
  return akoh_constant / temp;
}
__device__ __forceinline__ double kohler_b(double msol, const Solute& s) {
  constexpr double bkoh_constant = 4.3e-6 * dimless::RHO0 / dimless::MR0;
  return bkoh_constant * msol * s.ionic / s.mr_sol;
}
__device__ __forceinline__ void kohler_factors(double msol, double temp, const Solute& s, double& akoh,
                                               double& bkoh) {
  akoh = kohler_a(temp);
  bkoh = kohler_b(msol, s);
}

// ---- implicit Euler: impliciteuler.{hpp,cpp} -------------------------------------------------
// x^3 with a single rounding: x*x is split exactly into hi + lo with an FMA, hi*x likewise, and
// the small terms are added last, so the result is the correctly rounded cube except when x^3
// lies within ~2^-100 relative of a rounding boundary.  The reference spells it pow(x, 3.0);
// glibc's pow is accurate to 0.52 ulp, so the two differ by at most one ulp (in about 0.1% of
// arguments).  Likewise pow(x, 0.5) is replaced by the correctly rounded sqrt(x).
__device__ __forceinline__ double cube_rn(double x) {
  const double hi = x * x;
  const double lo = fma(x, x, -hi);
  const double p = hi * x;
  const double e = fma(hi, x, -p);
  return p + (e + lo * x);
}

struct ODEConsts {
  double s_ratio, akoh, bkoh;
  double ffv_num, ffv_den;  // ffactor / ventilation factor = ffv_num / ffv_den
  double inv_ffv;           // ffv_den / ffv_num
  double inv_akoh;          // rcp_nr(akoh), kept per gridbox
};
struct NewtonCfg {
  double delt, minsubdelt, rtol, atol;
  uint32_t maxniters;
  uint32_t budget;  // iterations after which a solve gives up as deferred (0: never)
};
// g(z) (impliciteuler.cpp:270-281) and dg/dz (impliciteuler.cpp:291-301) at the same z.  The
// reference's quotients akoh/r, bkoh/r^3, subdelt/(z ffactor) and rprev/r are products with one
// reciprocal square root of z (and 1/ffactor per drop): no FP64 division or square root per
// evaluation instead of six divisions and a root, each term within a few ulp of the reference's.
__device__ __forceinline__ void ode_g_dg(const ODEConsts& c, double subdelt, double rprev, double rsqrd,
                                         double& g, double& dg) {
  // products feeding a sum may fuse into FMAs here (one rounding fewer each)
#pragma clang fp contract(fast)
  const double inv_r = rsqrt_nr(rsqrd);
  const double inv_r2 = inv_r * inv_r;
  const double a_r = c.akoh * inv_r;
  const double b_r3 = c.bkoh * (inv_r2 * inv_r);
  const double alpha = c.s_ratio - 1 - a_r + b_r3;
  const double beta_d = subdelt * (inv_r2 * c.inv_ffv);
  const double beta = 2.0 * beta_d;
  const double q = rprev * inv_r;
  const double gamma = q * q;
  g = 1 - gamma - alpha * beta;
  const double alpha_d = a_r - 3.0 * b_r3;
  dg = 1 - alpha_d * beta_d;
}
// Newton state: z with g(z) and g'(z) already evaluated
struct NewtonState {
  double z, g, dg;
};
// one Newton update + convergence test (impliciteuler.cpp:240-255, impliciteuler.hpp:214-219).
// The reference evaluates g at the new z for the test and again as the next update's numerator;
// both are the same expression of the same operands, so the value is carried over.
__device__ __forceinline__ bool newton_iterate(const NewtonCfg& cfg, const ODEConsts& c, double subdelt,
                                               double rprev, NewtonState& s, uint32_t& iters) {
  ++iters;
  double z = s.z * (1 - quotient(s.g, s.dg));
  z = fmax(z, 1e-8);
  double g1, dg1;
  ode_g_dg(c, subdelt, rprev, z, g1, dg1);
  const double converged = cfg.rtol * fabs(g1) + cfg.atol;
  const bool conv = fabs(g1 - s.g) < converged;
  s = NewtonState{z, g1, dg1};
  return conv;
}
// outcome of a run of Newton iterations: converged, `limit` exceeded, or abandoned as deferred
// (cfg.budget reached; the drop is then solved again from the start, without a budget).  Returned
// by value, not through flags passed by reference, which the compiler kept in scratch memory.
enum : int { kNewtonDeferred = -1, kNewtonLimit = 0, kNewtonConverged = 1 };
// newtonraphson_untilconverged (impliciteuler.cpp:207-227)
__device__ __forceinline__ int newton_until(const NewtonCfg& cfg, const ODEConsts& c, uint32_t limit, double subdelt,
                                            double rprev, NewtonState& s, uint32_t& iters) {
  for (uint32_t niter = 1;; ++niter) {
    if (niter > limit) return kNewtonLimit;
    if (cfg.budget && iters >= cfg.budget) return kNewtonDeferred;
    if (newton_iterate(cfg, c, subdelt, rprev, s, iters)) return kNewtonConverged;
  }
}
// integrate_condensation_ode (impliciteuler.cpp:128-141): 2 iterations, then restart from the guess
// (only the second iteration's test decides); the new z and the outcome
struct NewtonRun {
  double z;
  int status;
};
__device__ __forceinline__ NewtonRun newton_integrate(const NewtonCfg& cfg, const ODEConsts& c, double subdelt,
                                                      double rprev, double zguess, uint32_t& iters) {
  NewtonState s0;
  s0.z = zguess;
  ode_g_dg(c, subdelt, rprev, zguess, s0.g, s0.dg);
  NewtonState s = s0;
  (void)newton_iterate(cfg, c, subdelt, rprev, s, iters);
  if (newton_iterate(cfg, c, subdelt, rprev, s, iters)) return NewtonRun{s.z, kNewtonConverged};
  s = s0;
  const int st = newton_until(cfg, c, cfg.maxniters, subdelt, rprev, s, iters);
  return NewtonRun{s.z, st};
}
__device__ __forceinline__ double initialguess(const ODEConsts& c, double rprev) {
  const double s_act = 1 + sqrt(4.0 * cube_rn(c.akoh) / 27 / c.bkoh);
  if (c.s_ratio > s_act) {
    constexpr double bigr(1e-3 / dimless::R0);
    const double maxradius = fmax(bigr, rprev);
    return maxradius * maxradius;
  }
  return rprev * rprev;
}
__device__ __forceinline__ bool first_unique_criteria(const ODEConsts& c, double rprev, double ziter) {
  const double rcritsqrd = 3.0 * c.bkoh / c.akoh;
  const bool is_ziter_unactivated = ziter < rcritsqrd;
  const bool is_unactivated = (rprev * rprev < rcritsqrd && is_ziter_unactivated);
  const double sqrdval = 4.0 * c.akoh * c.akoh * c.akoh / 27.0 / c.bkoh;
  const bool is_subact = (c.s_ratio <= 1.0 + sqrt(sqrdval));  // pow(x, 0.5) rounds as sqrt(x)
  return is_unactivated && is_subact;
}
__device__ __forceinline__ double critical_timestep(const ODEConsts& c) {
  const double x = 5.0 * c.bkoh / c.akoh;
  const double cuberoot = x * sqrt(x);  // pow(x, 1.5) to within 1 ulp
  const double ffactor_fv = c.ffv_num / c.ffv_den;
  return 2.5 * ffactor_fv / c.akoh * cuberoot;
}

// Three-valued comparisons: +1 true, -1 false, 0 too close to call from approximate operands.
// and = min, or = max.
// x < t where t is known to 1e-13 relative
__device__ __forceinline__ int tri_less(double x, double t) {
  return x < t * (1.0 - 1e-12) ? 1 : (x > t * (1.0 + 1e-12) ? -1 : 0);
}
// sign of d - sqrt(y) for y >= 0 known to 1e-13 relative, decided when it also fixes the sign of
// s - (1 + sqrt(y)) rounded (d = s - 1): the margins cover the 1e-15 relative error of sqrt(y)
// and the half-ulp of 1 in the sum.
__device__ __forceinline__ int tri_sqrt_sign(double d, double y) {
  if (y != y) return 0;
  if (d <= 0.0) return -1;
  const double d2 = d * d;
  if (d2 > fma(y, 1.0 + 1e-8, 1e-20)) return 1;
  if (d2 < fma(y, 1.0 - 1e-8, -1e-20)) return -1;
  return 0;
}
// initialguess, first_unique_criteria and the critical-timestep test of solve_condensation
// (impliciteuler.cpp:35-54, 80-126).  Each comparison is decided from products with approximate
// reciprocals when its sides differ by more than 1e-8 relative (the reference's quotients and
// roots are exact to a few ulp, so the outcome is the same); a drop with any comparison inside
// that band evaluates the reference's expressions instead.
__device__ __forceinline__ void ode_setup(const ODEConsts& c, double rprev, double delt, double& ziter, bool& ucrit) {
  const double ia = c.inv_akoh, ib = rcp_nr(c.bkoh);
  const double y = (c.akoh * c.akoh * c.akoh) * ib * (4.0 / 27.0);  // (s_act - 1)^2
  const int sact = tri_sqrt_sign(c.s_ratio - 1.0, y);                // s_ratio > s_act
  constexpr double bigr(1e-3 / dimless::R0);
  const double maxradius = fmax(bigr, rprev);
  ziter = sact > 0 ? maxradius * maxradius : rprev * rprev;
  const double rcrit = 3.0 * c.bkoh * ia;  // rcritsqrd
  const int unact = min(tri_less(rprev * rprev, rcrit), tri_less(ziter, rcrit));
  const int ucrit1 = min(unact, -sact);
  // delt <= critdelt = 2.5 (num/den)/akoh x^1.5 with x = 5 bkoh/akoh, compared squared and
  // multiplied by den^2 (no division)
  const double x = 5.0 * c.bkoh * ia, k = 2.5 * c.ffv_num * ia;
  const double crit2 = (k * k) * (x * x * x);
  const double d2 = (delt * delt) * (c.ffv_den * c.ffv_den);
  const int ucrit2 = delt > 0.0 ? (d2 < crit2 * (1.0 - 1e-8) ? 1 : (d2 > crit2 * (1.0 + 1e-8) ? -1 : 0)) : 0;
  const int uc = max(ucrit1, ucrit2);
  ucrit = uc > 0;
  if (sact == 0 || uc == 0) {
    ziter = initialguess(c, rprev);
    ucrit = first_unique_criteria(c, rprev, ziter) || (delt <= critical_timestep(c));
  }
}

struct SolveOut {
  uint32_t iters;
  bool substep, restarted, maxed;
  bool deferred;  // the budget ran out: nothing of this solve is kept
};
// solve_condensation (impliciteuler.cpp:35-54). `warm` is a learned guess for z = r^2 or NaN.
// Learned path (no reference counterpart): the reference's Newton update and convergence test
// iterated from `warm` until converged, at most `warm_limit` iterations, and if it has not
// converged by then the reference path from the plain guess.  Where |g| at the plain guess is
// below atol the plain guess is already the root to the tolerance (a drop at equilibrium), and the
// same iteration starts from the plain guess instead: its test passes after one update, where the
// reference path always makes two and a learned guess (a few 1e-3 off) needs three or four.
// inv_akoh = rcp_nr(akoh).
__device__ __forceinline__ double solve_condensation(const NewtonCfg& cfg, double s_ratio, double akoh,
                                                     double bkoh, double inv_akoh, double ffactor, double rprev,
                                                     double warm, uint32_t warm_limit, SolveOut& o) {
  ODEConsts c{s_ratio, akoh, bkoh, 0.0, 0.0, 0.0, inv_akoh};
  ffactor_over_vent(ffactor, rprev, c.ffv_num, c.ffv_den, c.inv_ffv);
  double ziter;
  bool ucrit;
  ode_setup(c, rprev, cfg.delt, ziter, ucrit);
  double rsqrd;
  NewtonRun run{0.0, kNewtonConverged};
  if (ucrit) {
    bool plain = true;
    if (warm == warm) {
      // the start: the plain guess where it is already the root to the tolerance, else the learned
      NewtonState s;
      s.z = ziter;
      ode_g_dg(c, cfg.delt, rprev, ziter, s.g, s.dg);
      if (!(fabs(s.g) < cfg.atol)) {
        s.z = warm;
        ode_g_dg(c, cfg.delt, rprev, warm, s.g, s.dg);
      }
      const int st = newton_until(cfg, c, warm_limit, cfg.delt, rprev, s, o.iters);
      if (st == kNewtonDeferred) {
        o.deferred = true;
        return 0.0;
      }
      plain = st != kNewtonConverged;
      o.restarted = plain;
      run.z = s.z;
    }
    if (plain) {
      run = newton_integrate(cfg, c, cfg.delt, rprev, ziter, o.iters);
      if (run.status == kNewtonDeferred) {
        o.deferred = true;
        return 0.0;
      }
      o.maxed = run.status == kNewtonLimit;
    }
    rsqrd = run.z;
  } else {
    o.substep = true;
    const double critdelt = critical_timestep(c);
    const double mindelt = fmax(critdelt, cfg.minsubdelt);
    double remdelt = cfg.delt;
    double rp = rprev;
    bool maxed = false;
    while (remdelt > 0.0) {
      const double subdelt = fmin(mindelt, remdelt);
      run = newton_integrate(cfg, c, subdelt, rp, ziter, o.iters);
      if (run.status == kNewtonDeferred) {
        o.deferred = true;
        return 0.0;
      }
      maxed = maxed || run.status == kNewtonLimit;
      ziter = run.z;
      rp = sqrt(ziter);  // pow(ziter, 0.5)
      remdelt -= subdelt;
    }
    o.maxed = maxed;
    rsqrd = ziter;
  }
  return sqrt(rsqrd);
}

// ---- terminal velocity: terminalvelocity.cpp:33-146 ---------------------------------------
__device__ __forceinline__ double tv_simmel(double radius) {
  constexpr double r1 = 6.7215e-5 / dimless::R0;
  constexpr double r2 = 7.5582e-4 / dimless::R0;
  constexpr double r3 = 1.73892e-3 / dimless::R0;
  constexpr double VELCONST = (100.0 * dimless::W0);
  constexpr double a1 = 457950 / VELCONST;
  constexpr double a2 = 4962 / VELCONST;
  constexpr double a3 = 1732 / VELCONST;
  constexpr double a4 = 917 / VELCONST;
  if (radius >= r3) return a4;
  // MASS = massconst r^3 MASS0grams with massconst = 4/3 pi Rho_l, so MASS^(1/3) = k r for
  // k = (massconst MASS0grams)^(1/3), the correctly rounded cube root of the product of the two
  // double constants (0.000161102579187357...).  The reference's pow(MASS, 1/3), ^(2/3), ^(1/6)
  // become k r, (k r)^2 and sqrt(k r), each within a few ulp of it.
  constexpr double kMassCbrt = 0.00016110257918735754;
  const double cr = kMassCbrt * radius;
  if (radius >= r2) return a3 * sqrt(cr);
  if (radius >= r1) return a2 * cr;
  return a1 * (cr * cr);
}
__device__ __forceinline__ double tv_rogersyau(double radius) {
  constexpr double r1 = 3e-5 / dimless::R0;
  constexpr double r2 = 6e-4 / dimless::R0;
  constexpr double r3 = 2e-3 / dimless::R0;
  constexpr double k1 = 1.19e8 * dimless::R0 * dimless::R0 / dimless::W0;
  constexpr double k2 = 8000 * dimless::R0 / dimless::W0;
  constexpr double k3 = 201 / dimless::W0;
  constexpr double k4 = 9 / dimless::W0;
  if (radius < r1) return k1 * (radius * radius);
  if (radius < r2) return k2 * radius;
  if (radius < r3) return k3 * sqrt(radius * dimless::R0);  // pow(x, 0.5) rounds as sqrt(x)
  return k4;
}
__device__ __forceinline__ double tv_rogersgk(double radius) {
  constexpr double radius0 = 3.725 * 1e-4 / dimless::R0;
  constexpr double kcaps = 2.0 * 4.0 * 1000 * dimless::R0 / dimless::W0;
  constexpr double smallk = -1.0 * 2.0 * 12.0 * 1000 * dimless::R0;
  constexpr double acaps = 9.65 / dimless::W0;
  constexpr double bcaps = 10.43 / dimless::W0;
  constexpr double ccaps = -1.0 * 2.0 * 0.6 * 1000 * dimless::R0;
  if (radius < radius0) {
    const double term = 1.0 - exp(smallk * radius);
    return term * kcaps * radius;
  }
  return acaps - bcaps * exp(ccaps * radius);
}
__device__ __forceinline__ double terminal_velocity(int kind, double radius) {
  switch (kind) {
    case CLEO_TV_SIMMEL: return tv_simmel(radius);
    case CLEO_TV_ROGERS_YAU: return tv_rogersyau(radius);
    case CLEO_TV_ROGERS_GK: return tv_rogersgk(radius);
    default: return 0.0;
  }
}

// ---- pair probabilities ---------------------------------------------------------------------
struct CollCfg {
  int prob_kind;
  double prob_param;
  int enact_kind;
  double nfrags;
  double coaleff;
  double DELT;
  int nfrags_kind, flag_kind;
  int prob_tv, nfrags_tv, flag_tv;
};

// ---- collision kinetics (Low and List 1982a), energies in J from dimensionless radii --------
// surface energy 4 pi sigma R^2 of one drop, of two, and of the coalesced drop of both volumes
// (collisionkinetics.hpp:41-56, collisionkinetics.cpp:62-71)
constexpr double kSurfConst = 4.0 * dimmed::SURFSIGMA * kPi * dimless::R0 * dimless::R0;
__device__ __forceinline__ double surf_energy(double r) { return kSurfConst * (r * r); }
__device__ __forceinline__ double surf_energy_pair(double r1, double r2) { return kSurfConst * ((r1 * r1) + (r2 * r2)); }
__device__ __forceinline__ double surf_energy_coal(double r1, double r2) {
  const double sum = r1 * r1 * r1 + r2 * r2 * r2;
  return kSurfConst * pow(sum, 2.0 / 3.0);
}
// kinetic energy of the collision, (2/3) pi rho_l (r1^3 r2^3 / (r1^3 + r2^3)) (v1 - v2)^2, with
// the reduced volume written as r1^3 / (1 + (r1/r2)^3) (collisionkinetics.cpp:33-46)
__device__ __forceinline__ double collision_kinetic_energy(double r1, double r2, double v1, double v2) {
  constexpr double kCke = dimless::R0 * dimless::R0 * dimless::R0 * 2.0 / 3.0 * dimmed::RHO_L * kPi * dimless::W0 *
                          dimless::W0;
  const double q = r1 / r2;
  const double rratio = (r1 * r1 * r1) / (1 + q * q * q);
  const double vdiff = v1 - v2;
  return kCke * rratio * vdiff * vdiff;
}
// Straub et al. 2010 eq. 5: exp(-1.15 We), We = cke / coalesced surface energy
// (coalescence_efficiency.cpp:27-37)
__device__ __forceinline__ double coaleff_straub(double r1, double r2, double cke) {
  return exp(-1.15 * (cke / surf_energy_coal(r1, r2)));
}
// Low and List 1982a eqs. 4.5-4.6: 0.778 (1 + rs/rl)^-2 exp(-2.62e6 sigma E^2 / S_c) when the
// total energy E = cke + S_t - S_c is below 5e-6 J, else 0 (lowlistprob.hpp:50-91)
__device__ __forceinline__ double vol_of(double r) { return 4.0 / 3.0 * kPi * rcubed(r); }
__device__ __forceinline__ double coaleff_lowlist(int tv, double r1, double r2) {
  const double cke = collision_kinetic_energy(r1, r2, terminal_velocity(tv, r1), terminal_velocity(tv, r2));
  const double surf_t = surf_energy_pair(r1, r2);
  const double surf_c = surf_energy_coal(r1, r2);
  const double etot = cke + surf_t - surf_c;
  if (!(etot < 5e-6)) return 0.0;
  const double rsmall = fmin(r1, r2), rbig = fmax(r1, r2);
  const double alpha = 1 + rsmall / rbig;
  const double ratio = 1.0 / (alpha * alpha);
  constexpr double kExp = -2.62e6 * dimmed::SURFSIGMA;
  return 0.778 * ratio * exp(kExp * etot * etot / surf_c);
}
// Long's hydrodynamic kernel (longhydroprob.cpp:30-48, hydrodynamicprob.hpp:50-66) times coaleff
__device__ __forceinline__ double prob_long(double r1, double r2, double coaleff, double DELT, double VOLUME) {
  constexpr double k1 = 4.5e4 * dimless::R0 * dimless::R0 * 100 * 100;
  constexpr double k2 = 3e-4 / dimless::R0 / 100;
  constexpr double rlim = 5e-5 / dimless::R0;
  constexpr double colleff_min = 0.001;
  const double rsmall = fmin(r1, r2);
  const double rbig = fmax(r1, r2);
  double colleff = 1.0;
  if (rbig < rlim) {
    colleff = k1 * rbig * rbig * (1 - (k2 / rsmall));
    colleff = fmax(colleff, colleff_min);
  }
  const double eff = colleff * coaleff;
  constexpr double prob_jk_const = kPi * dimless::R0 * dimless::R0 * dimless::W0;
  const double DELT_DELVOL = DELT / VOLUME;
  const double sumr = r1 + r2;
  const double sumrsqrd = sumr * sumr;
  const double vdiff = fabs(tv_simmel(r1) - tv_simmel(r2));
  const double hydro_kernel = prob_jk_const * eff * sumrsqrd * vdiff;
  return hydro_kernel * DELT_DELVOL;
}
// golovinprob.cpp:32-43, longhydroprob.cpp:30-48 + hydrodynamicprob.hpp:50-66, constprob.hpp:45-48,
// lowlistprob.hpp:117-123,144-153
// Kinetic = false compiles out the Low-List probabilities and the kinetic-energy NFragments and
// flags (launches whose CollCfg names none of them)
template <bool Kinetic>
__device__ __forceinline__ double pair_probability(const CollCfg& k, double r1, double r2, double VOLUME) {
  if (k.prob_kind == CLEO_PROB_GOLOVIN) {
    constexpr double prob_jk_const = 1.5e3 * dimless::R0 * dimless::R0 * dimless::R0;
    const double DELT_DELVOL = k.DELT / VOLUME;
    const double kern = prob_jk_const * (vol_of(r1) + vol_of(r2));
    return kern * DELT_DELVOL;
  }
  if (k.prob_kind == CLEO_PROB_LONG) return prob_long(r1, r2, k.prob_param, k.DELT, VOLUME);
  if constexpr (Kinetic) {
    if (k.prob_kind == CLEO_PROB_LOWLIST_COAL)
      return prob_long(r1, r2, 1.0, k.DELT, VOLUME) * coaleff_lowlist(k.prob_tv, r1, r2);
    if (k.prob_kind == CLEO_PROB_LOWLIST_BU) {
      const double bueff = 1.0 - coaleff_lowlist(k.prob_tv, r1, r2);
      return prob_long(r1, r2, 1.0, k.DELT, VOLUME) * bueff;
    }
  }
  return k.prob_param * k.DELT / VOLUME;
}

// ---- enactment on (xi, r, msol) registers -----------------------------------------------------
// New radii are cbrt(x) where the reference writes pow(x, 1.0 / 3.0) (within 1 ulp of it).
struct DropAttrs {
  uint64_t xi;
  double r, msol;
};
// coalescence_gamma (coalescence.cpp:68-79).  A probability of 2^64 or more (or NaN) overflows
// the reference's double -> uint64 conversion, whose result is then undefined; gamma is then the
// cap floor(xi1/xi2) that any such gamma reaches.
__device__ __forceinline__ uint64_t coal_gamma(uint64_t xi1, uint64_t xi2, double prob, double phi) {
  if (!(prob < 0x1p64)) return xi1 / xi2;
  uint64_t gamma = static_cast<uint64_t>(floor(prob));
  if (phi < (prob - static_cast<double>(gamma))) ++gamma;
  const uint64_t maxgamma = xi1 / xi2;
  return static_cast<uint64_t>(fmin(static_cast<double>(gamma), static_cast<double>(maxgamma)));
}
// returns error code (0 ok); *null set when drop1 became null (coalescence.cpp:92-196)
__device__ __forceinline__ int coalesce(uint64_t gamma, DropAttrs& d1, DropAttrs& d2, const Solute& s,
                                        bool& null1) {
  const uint64_t xi1 = d1.xi, xi2 = d2.xi;
  int err = 0;
  if (xi1 - gamma * xi2 > 0) {
    const uint64_t new_xi = xi1 - gamma * xi2;
    if (new_xi >= xi1) err = CLEO_ERR_COAL_UNDEFINED;
    d1.xi = new_xi;
    const double new_rcubed = rcubed(d2.r) + static_cast<double>(gamma) * rcubed(d1.r);
    const double nr = cbrt(new_rcubed);
    if (radius_below_dry(nr, d2.msol, s)) err = err ? err : CLEO_ERR_RADIUS_BELOW_DRY;
    d2.r = nr;
    d2.msol = d2.msol + static_cast<double>(gamma) * d1.msol;
    return err;
  }
  if (xi1 - gamma * xi2 == 0) {
    const uint64_t old_xi = xi2;
    const uint64_t new_xi = old_xi / 2;
    if (new_xi >= old_xi) err = CLEO_ERR_COAL_UNDEFINED;
    const double new_rcubed = rcubed(d2.r) + static_cast<double>(gamma) * rcubed(d1.r);
    const double new_r = cbrt(new_rcubed);
    const double new_msol = d2.msol + static_cast<double>(gamma) * d1.msol;
    if (new_xi == 0) {
      d1.xi = 0;
      null1 = true;
    } else {
      d1.xi = new_xi;
    }
    if (old_xi - new_xi == 0) err = err ? err : CLEO_ERR_XI_ZERO;
    d2.xi = old_xi - new_xi;
    if (radius_below_dry(new_r, d1.msol, s) || radius_below_dry(new_r, d2.msol, s))
      err = err ? err : CLEO_ERR_RADIUS_BELOW_DRY;
    d1.r = new_r;
    d2.r = new_r;
    d1.msol = new_msol;
    d2.msol = new_msol;
    return err;
  }
  return CLEO_ERR_COAL_UNDEFINED;
}
// breakup with ConstNFrags (breakup.hpp:176-268)
__device__ __forceinline__ int breakup(double nfrags, DropAttrs& d1, DropAttrs& d2, const Solute& s) {
  int err = 0;
  if (d1.xi == d2.xi) {
    const uint64_t old_xi = d2.xi;
    const double totnfrags = nfrags * static_cast<double>(old_xi);
    if ((totnfrags / static_cast<double>(old_xi)) <= 2.5) err = CLEO_ERR_BREAKUP_NFRAGS;
    const uint64_t new_xi1 = static_cast<uint64_t>(round(totnfrags / 2));
    const uint64_t new_xi2 = static_cast<uint64_t>(round(totnfrags - static_cast<double>(new_xi1)));
    const uint64_t new_xitot = new_xi1 + new_xi2;
    if (new_xi2 <= old_xi) err = err ? err : CLEO_ERR_BREAKUP_NFRAGS;
    if (new_xitot <= (old_xi * 2)) err = err ? err : CLEO_ERR_BREAKUP_NFRAGS;
    const double sum_rcubed = rcubed(d1.r) + rcubed(d2.r);
    const double new_rcubed = sum_rcubed * static_cast<double>(old_xi) / static_cast<double>(new_xitot);
    const double new_r = cbrt(new_rcubed);
    const double sum_msol = d1.msol + d2.msol;
    const double new_msol = sum_msol * static_cast<double>(old_xi) / static_cast<double>(new_xitot);
    if (new_xi1 == 0 || new_xi2 == 0) err = err ? err : CLEO_ERR_XI_ZERO;
    d1.xi = new_xi1;
    d2.xi = new_xi2;
    if (radius_below_dry(new_r, d1.msol, s) || radius_below_dry(new_r, d2.msol, s))
      err = err ? err : CLEO_ERR_RADIUS_BELOW_DRY;
    d1.r = new_r;
    d2.r = new_r;
    d1.msol = new_msol;
    d2.msol = new_msol;
    return err;
  }
  const uint64_t old_xi1 = d1.xi, old_xi2 = d2.xi;
  const uint64_t new_xi1 = old_xi1 - old_xi2;
  if (new_xi1 == 0) err = CLEO_ERR_XI_ZERO;
  d1.xi = new_xi1;
  const double totnfrags = nfrags * static_cast<double>(old_xi2);
  const uint64_t new_xi2 = static_cast<uint64_t>(round(totnfrags));
  if ((totnfrags / static_cast<double>(old_xi2)) <= 2.5) err = err ? err : CLEO_ERR_BREAKUP_NFRAGS;
  if (new_xi2 <= old_xi2) err = err ? err : CLEO_ERR_BREAKUP_NFRAGS;
  if ((new_xi1 + new_xi2) <= (old_xi1 + old_xi2)) err = err ? err : CLEO_ERR_BREAKUP_NFRAGS;
  const double sum_rcubed = rcubed(d1.r) + rcubed(d2.r);
  const double new_rcubed = sum_rcubed * static_cast<double>(old_xi2) / static_cast<double>(new_xi2);
  const double new_r = cbrt(new_rcubed);
  const double sum_msol = d1.msol + d2.msol;
  const double new_msol = sum_msol * static_cast<double>(old_xi2) / static_cast<double>(new_xi2);
  if (new_xi2 == 0) err = err ? err : CLEO_ERR_XI_ZERO;
  d2.xi = new_xi2;
  if (radius_below_dry(new_r, d2.msol, s)) err = err ? err : CLEO_ERR_RADIUS_BELOW_DRY;
  d2.r = new_r;
  d2.msol = new_msol;
  return err;
}
// NFragments of the pair, from the radii before the breakup: ConstNFrags, or the collision
// kinetic energy fit of Schlottke et al. 2010 fig. 13, 1 / (1.5 - (1e6 min(cke, ckemax))^0.135)
// + 11/6 (breakup_nfrags.hpp:44-91)
__device__ __forceinline__ double breakup_nfrags(const CollCfg& k, const DropAttrs& d1, const DropAttrs& d2) {
  if (k.nfrags_kind != CLEO_NFRAGS_CKE) return k.nfrags;
  const double cke = collision_kinetic_energy(d1.r, d2.r, terminal_velocity(k.nfrags_tv, d1.r),
                                              terminal_velocity(k.nfrags_tv, d2.r));
  constexpr double ckemax = 16.49789599 / 1000000;
  const double cke_cap = fmin(ckemax, cke);
  const double g = pow(cke_cap * 1e6, 0.135);
  return 1.0 / (1.5 - g) + 11.0 / 6.0;
}
// CoalBuReFlag: 1 coalescence, 2 breakup, 0 rebound (coalbure_flag.hpp:41-198)
__device__ __forceinline__ uint32_t coalbure_flag(const CollCfg& k, double phi, const DropAttrs& d1,
                                                  const DropAttrs& d2) {
  if (k.flag_kind == CLEO_FLAG_CONST) return phi < k.coaleff ? 1u : 2u;
  const double r1 = d1.r, r2 = d2.r;
  const double cke =
      collision_kinetic_energy(r1, r2, terminal_velocity(k.flag_tv, r1), terminal_velocity(k.flag_tv, r2));
  if (k.flag_kind == CLEO_FLAG_STRAUB) return phi < coaleff_straub(r1, r2, cke) ? 1u : 2u;
  if (k.flag_kind == CLEO_FLAG_SU) {
    if (cke < surf_energy(fmin(r1, r2))) return 0u;
    return cke < surf_energy_coal(r1, r2) ? 1u : 2u;
  }
  // TS: below the small drop's surface energy rebound or coalescence, below the large drop's
  // coalescence or breakup, above it breakup
  if (cke < surf_energy(fmin(r1, r2))) return phi < coaleff_straub(r1, r2, cke) ? 1u : 0u;
  if (cke < surf_energy(fmax(r1, r2))) return phi < coaleff_straub(r1, r2, cke) ? 1u : 2u;
  return 2u;
}
// true when a CollCfg needs the Kinetic instantiations
inline bool coll_is_kinetic(const CollCfg& k) {
  return k.prob_kind == CLEO_PROB_LOWLIST_COAL || k.prob_kind == CLEO_PROB_LOWLIST_BU ||
         k.nfrags_kind != CLEO_NFRAGS_CONST || k.flag_kind != CLEO_FLAG_CONST;
}
// PairEnactX dispatch; bit0 of the return: drop1 null, bit1: enacted; error via *err
template <bool Kinetic>
__device__ __forceinline__ uint32_t enact_pair(const CollCfg& k, DropAttrs& d1, DropAttrs& d2, double prob,
                                               double phi_coll, double phi_out, const Solute& s, int& err) {
  bool null1 = false;
  if (k.enact_kind == CLEO_ENACT_BREAKUP) {
    if (phi_coll < (prob - floor(prob))) {
      err = breakup(Kinetic ? breakup_nfrags(k, d1, d2) : k.nfrags, d1, d2, s);
      return 2u;
    }
    return 0u;
  }
  const uint64_t gamma = coal_gamma(d1.xi, d2.xi, prob, phi_coll);
  if (gamma == 0) return 0u;
  uint32_t flag = 1u;
  if (k.enact_kind != CLEO_ENACT_COAL) {
    if constexpr (Kinetic) flag = coalbure_flag(k, phi_out, d1, d2);
    else flag = phi_out < k.coaleff ? 1u : 2u;
  }
  if (flag == 1u) {
    err = coalesce(gamma, d1, d2, s, null1);
  } else if (flag == 2u) {
    err = breakup(Kinetic ? breakup_nfrags(k, d1, d2) : k.nfrags, d1, d2, s);
  }
  return 2u | (null1 ? 1u : 0u);
}

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ---------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
    const uint32_t lo0 = 0xD2511F53u * c[0];
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
    const uint32_t lo1 = 0xCD9E8D57u * c[2];
    const uint32_t n0 = hi1 ^ c[1] ^ k0;
    const uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}
// stream 0: Fisher-Yates indices, four swaps per block (item = swap / 4); stream 1: (phi_coll,
// phi_out) for pair `jj`
__device__ __forceinline__ void rng_block(uint64_t seed, uint64_t event, uint32_t stream, uint32_t gbx,
                                          uint32_t item, uint32_t out[4]) {
  out[0] = item;
  out[1] = gbx;
  out[2] = static_cast<uint32_t>(event);
  out[3] = stream + 2u * static_cast<uint32_t>(event >> 32);
  philox4x32_10(out, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
}
__device__ __forceinline__ double u53(uint32_t lo, uint32_t hi) {
  const uint64_t v = (static_cast<uint64_t>(hi) << 32) | lo;
  return static_cast<double>(v >> 11) * 0x1.0p-53;
}

}  // namespace cleo
