// cleo_sdm.hpp -- C++ host layer over libcleo_hip.so mirroring CLEO's SDM vocabulary.
//
// The reference composes its hot path from compile-time "MicrophysicalProcess" objects built by
// factory functions (Condensation, CollCoal, CollBu, CoalBuRe; libs/superdrops/condensation.hpp:
// 242-253, collisions/coalescence.hpp:133-160, collisions/breakup.hpp:97-123,
// collisions/coalbure.hpp:156-185), combined with operator>> (microphysicalprocess.hpp:132-134),
// plus a Motion (cartesian_motion.hpp:60-66) and driven by SDMMethods::run_step
// (runcleo/sdmmethods.hpp:319-332).  The same names and arguments build descriptors here, and
// SDMMethods::run_step hands the whole domain to one GPU call per model step.
//
// Error behaviour: where the reference calls Kokkos::abort, the GPU records an error code and
// run_step throws cleo_gpu::Error carrying the code, gridbox and drop.
#pragma once

#include <cmath>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "cleo_hip.h"

namespace cleo_gpu {

struct Error : std::runtime_error {
  int code;
  cleo_error where;
  Error(const std::string& m, int c, cleo_error w) : std::runtime_error(m), code(c), where(w) {}
};

inline void check(cleo_ctx* ctx, int rc, const char* what) {
  if (rc == CLEO_OK) return;
  cleo_error e{};
  if (ctx) cleo_get_error(ctx, &e);
  std::string msg = std::string(what) + " failed (" + std::to_string(rc) + ")";
  if (ctx) msg += ": " + std::string(cleo_last_error_message(ctx));
  throw Error(msg, e.code ? e.code : rc, e);
}

// ---- PairProbability / NFragments / CoalBuReFlag / VelocityFormula descriptors -------------
struct GolovinProb {
  int kind() const { return CLEO_PROB_GOLOVIN; }
  double param() const { return 0.0; }
};
struct LongHydroProb {
  double coaleff = 1.0;
  LongHydroProb() = default;
  explicit LongHydroProb(double c) : coaleff(c) {}
  int kind() const { return CLEO_PROB_LONG; }
  double param() const { return coaleff; }
};
struct ConstProb {
  double kernel;
  explicit ConstProb(double k) : kernel(k) {}
  int kind() const { return CLEO_PROB_CONST; }
  double param() const { return kernel; }
};
struct NullTerminalVelocity {
  int kind() const { return CLEO_TV_NULL; }
};
struct SimmelTerminalVelocity {
  int kind() const { return CLEO_TV_SIMMEL; }
};
struct RogersYauTerminalVelocity {
  int kind() const { return CLEO_TV_ROGERS_YAU; }
};
struct RogersGKTerminalVelocity {
  int kind() const { return CLEO_TV_ROGERS_GK; }
};
// the descriptors write their kinds into the C-ABI parameters; probabilities without a terminal
// velocity leave prob_tv at CLEO_TV_NULL
template <typename Probability>
inline void apply_prob(const Probability& prob, cleo_coll_params& p) {
  p.prob_kind = prob.kind();
  p.prob_param = prob.param();
}
// LowListCoalProb<TerminalVelocity> / LowListBuProb<TerminalVelocity> (lowlistprob.hpp:40-153)
template <typename TerminalVelocity>
struct LowListCoalProb {
  TerminalVelocity terminalv;
  explicit LowListCoalProb(TerminalVelocity tv) : terminalv(tv) {}
  int kind() const { return CLEO_PROB_LOWLIST_COAL; }
  double param() const { return 1.0; }
  int tv() const { return terminalv.kind(); }
};
template <typename TerminalVelocity>
struct LowListBuProb {
  TerminalVelocity terminalv;
  explicit LowListBuProb(TerminalVelocity tv) : terminalv(tv) {}
  int kind() const { return CLEO_PROB_LOWLIST_BU; }
  double param() const { return 1.0; }
  int tv() const { return terminalv.kind(); }
};
template <typename TerminalVelocity>
inline void apply_prob(const LowListCoalProb<TerminalVelocity>& prob, cleo_coll_params& p) {
  p.prob_kind = prob.kind();
  p.prob_param = prob.param();
  p.prob_tv = prob.tv();
}
template <typename TerminalVelocity>
inline void apply_prob(const LowListBuProb<TerminalVelocity>& prob, cleo_coll_params& p) {
  p.prob_kind = prob.kind();
  p.prob_param = prob.param();
  p.prob_tv = prob.tv();
}

// NFragments (breakup_nfrags.hpp:38-91)
struct ConstNFrags {
  double nfrags;
  explicit ConstNFrags(double n) : nfrags(n) {}
  void apply(cleo_coll_params& p) const {
    p.nfrags_kind = CLEO_NFRAGS_CONST;
    p.nfrags = nfrags;
  }
};
template <typename TerminalVelocity>
struct CollisionKineticEnergyNFrags {
  TerminalVelocity terminalv;
  explicit CollisionKineticEnergyNFrags(TerminalVelocity tv) : terminalv(tv) {}
  void apply(cleo_coll_params& p) const {
    p.nfrags_kind = CLEO_NFRAGS_CKE;
    p.nfrags_tv = terminalv.kind();
  }
};

// CoalBuReFlag (coalbure_flag.hpp:41-198)
struct ConstCoalBuReFlag {
  double coaleff;
  explicit ConstCoalBuReFlag(double c) : coaleff(c) {}
  void apply(cleo_coll_params& p) const {
    p.flag_kind = CLEO_FLAG_CONST;
    p.coalbure_coaleff = coaleff;
  }
};
template <int Kind, typename TerminalVelocity>
struct KineticCoalBuReFlag {
  TerminalVelocity terminalv;
  explicit KineticCoalBuReFlag(TerminalVelocity tv) : terminalv(tv) {}
  void apply(cleo_coll_params& p) const {
    p.flag_kind = Kind;
    p.flag_tv = terminalv.kind();
  }
};
template <typename TerminalVelocity>
using SUCoalBuReFlag = KineticCoalBuReFlag<CLEO_FLAG_SU, TerminalVelocity>;
template <typename TerminalVelocity>
using TSCoalBuReFlag = KineticCoalBuReFlag<CLEO_FLAG_TS, TerminalVelocity>;
template <typename TerminalVelocity>
using StraubCoalBuReFlag = KineticCoalBuReFlag<CLEO_FLAG_STRAUB, TerminalVelocity>;

// ---- MicrophysicalProcess descriptors -------------------------------------------------------
// ConstTstepMicrophysics timing (microphysicalprocess.hpp:243-257)
struct ConstTstep {
  unsigned int interval;
  unsigned int next_step(unsigned int t) const { return ((t / interval) + 1) * interval; }
  bool on_step(unsigned int t) const { return (t % interval == 0) && (interval != 0xFFFFFFFFu); }
};

struct CondensationProcess : ConstTstep {
  cleo_cond_params params;
};
struct CollisionsProcess : ConstTstep {
  cleo_coll_params params;
};
struct NullMicrophysicalProcess {};

// the reference's conversions are arguments of its factories; the GPU evaluates the same formulas
// from the integer interval, so the given conversions must agree with them
inline void require_same(double a, double b, const char* what) {
  if (a != b) throw std::invalid_argument(std::string(what) + " does not match model_step units");
}

inline CondensationProcess Condensation(unsigned int interval,
                                        const std::function<double(unsigned int)>& step2dimlesstime,
                                        bool do_alter_thermo, size_t maxniters, double rtol, double atol,
                                        double MINSUBDELT,
                                        const std::function<double(double)>& realtime2dimless) {
  require_same(step2dimlesstime(interval), static_cast<double>(interval) / 100000.0, "step2dimlesstime");
  require_same(realtime2dimless(MINSUBDELT), MINSUBDELT / 1000.0, "realtime2dimless");
  CondensationProcess p{{interval}, {}};
  p.params = cleo_cond_params{interval, do_alter_thermo ? 1 : 0, static_cast<uint32_t>(maxniters), rtol, atol,
                              MINSUBDELT, 0, 10};
  return p;
}

inline CollisionsProcess make_colls(unsigned int interval, const std::function<double(unsigned int)>& int2realtime,
                                    int enact) {
  require_same(int2realtime(interval), (static_cast<double>(interval) / 100000.0) * 1000.0, "int2realtime");
  CollisionsProcess p{{interval}, {}};
  p.params = cleo_coll_params{};
  p.params.interval = interval;
  p.params.enact_kind = enact;
  return p;
}
template <typename Probability>
inline CollisionsProcess CollCoal(unsigned int interval, const std::function<double(unsigned int)>& int2realtime,
                                  const Probability& prob) {
  CollisionsProcess p = make_colls(interval, int2realtime, CLEO_ENACT_COAL);
  apply_prob(prob, p.params);
  return p;
}
template <typename Probability, typename NFrags>
inline CollisionsProcess CollBu(unsigned int interval, const std::function<double(unsigned int)>& int2realtime,
                                const Probability& prob, const NFrags& nfrags) {
  CollisionsProcess p = make_colls(interval, int2realtime, CLEO_ENACT_BREAKUP);
  apply_prob(prob, p.params);
  nfrags.apply(p.params);
  return p;
}
template <typename Probability, typename NFrags, typename Flag>
inline CollisionsProcess CoalBuRe(unsigned int interval, const std::function<double(unsigned int)>& int2realtime,
                                  const Probability& prob, const NFrags& nfrags, const Flag& flag) {
  CollisionsProcess p = make_colls(interval, int2realtime, CLEO_ENACT_COALBURE);
  apply_prob(prob, p.params);
  nfrags.apply(p.params);
  flag.apply(p.params);
  return p;
}

// CombinedMicrophysicalProcess (microphysicalprocess.hpp:65-120): the processes of a chain in
// order; operator>> (microphysicalprocess.hpp:132-134) joins any two processes or chains, so any
// number of processes in any order can be composed (cond >> colls, coal >> bu, cond >> coal >> bu).
// NullMicrophysicalProcess is the empty chain.
struct CombinedMicrophysicalProcess {
  std::vector<cleo_process> chain;
};
inline CombinedMicrophysicalProcess as_microphysics(const CondensationProcess& a) {
  cleo_process q{};
  q.kind = CLEO_PROCESS_CONDENSATION;
  q.cond = a.params;
  return {{q}};
}
inline CombinedMicrophysicalProcess as_microphysics(const CollisionsProcess& b) {
  cleo_process q{};
  q.kind = CLEO_PROCESS_COLLISIONS;
  q.coll = b.params;
  return {{q}};
}
inline CombinedMicrophysicalProcess as_microphysics(const CombinedMicrophysicalProcess& m) { return m; }
inline CombinedMicrophysicalProcess as_microphysics(const NullMicrophysicalProcess&) { return {}; }

template <typename T>
inline constexpr bool is_microphysical_process_v =
    std::is_same_v<T, CondensationProcess> || std::is_same_v<T, CollisionsProcess> ||
    std::is_same_v<T, CombinedMicrophysicalProcess> || std::is_same_v<T, NullMicrophysicalProcess>;

template <typename A, typename B,
          typename = std::enable_if_t<is_microphysical_process_v<A> && is_microphysical_process_v<B>>>
inline CombinedMicrophysicalProcess operator>>(const A& a, const B& b) {
  CombinedMicrophysicalProcess m = as_microphysics(a);
  const CombinedMicrophysicalProcess n = as_microphysics(b);
  m.chain.insert(m.chain.end(), n.chain.begin(), n.chain.end());
  if (m.chain.size() > CLEO_MAX_PROCESSES) throw std::invalid_argument("more processes than CLEO_MAX_PROCESSES");
  return m;
}

// ---- Motion -----------------------------------------------------------------------------------
struct PredCorrMotion : ConstTstep {
  cleo_motion_params params;
};
struct NullMotion {};
template <typename TV>
inline PredCorrMotion CartesianMotion(unsigned int motionstep, const std::function<double(unsigned int)>& int2time,
                                      const TV& terminalv) {
  require_same(int2time(motionstep), static_cast<double>(motionstep) / 100000.0, "int2time");
  PredCorrMotion m{{motionstep}, {}};
  m.params = cleo_motion_params{motionstep, terminalv.kind()};
  return m;
}

// ---- domain + driver ----------------------------------------------------------------------------
// The reference's Superdrop (superdrop.hpp:36-60): the 72-byte record cleo_upload_supers_aos reads.
struct Superdrop {
  uint32_t sdgbxindex;
  double coord3, coord1, coord2;
  struct {
    char solute;
    uint64_t xi;
    double radius;
    double msol;
  } attrs;
  uint64_t sdId;
};
static_assert(sizeof(Superdrop) == 72, "reference layout");

class GpuDomain {
 public:
  GpuDomain(uint64_t maxnsupers, const cleo_domain_desc& dom) {
    check(nullptr, cleo_create(&ctx_, maxnsupers, &dom), "cleo_create");
    ngbx_ = static_cast<uint64_t>(dom.ndims[0]) * dom.ndims[1] * dom.ndims[2];
  }
  ~GpuDomain() { cleo_destroy(ctx_); }
  GpuDomain(const GpuDomain&) = delete;
  GpuDomain& operator=(const GpuDomain&) = delete;

  cleo_ctx* ctx() const { return ctx_; }
  uint64_t ngbx() const { return ngbx_; }
  // SupersInDomain construction: sorts by sdgbxindex and sets the gridbox refs
  void upload(const std::vector<Superdrop>& supers) {
    check(ctx_, cleo_upload_supers_aos(ctx_, supers.data(), supers.size()), "upload");
    ntot_ = supers.size();
  }
  std::vector<Superdrop> download(uint64_t* ndomain = nullptr) {
    std::vector<Superdrop> out(ntot_);
    uint64_t n = 0;
    check(ctx_, cleo_download_supers_aos(ctx_, out.data(), &n, ndomain), "download");
    return out;
  }
  void set_state(const cleo_gbx_state& s) { check(ctx_, cleo_upload_gbx_state(ctx_, &s), "state upload"); }
  void get_state(cleo_gbx_state& s) { check(ctx_, cleo_download_gbx_state(ctx_, &s), "state download"); }
  void set_seed(uint64_t seed) { check(ctx_, cleo_set_rng_counter(ctx_, seed, 0), "rng"); }
  cleo_stats stats() {
    cleo_stats s{};
    check(ctx_, cleo_get_stats(ctx_, &s), "stats");
    return s;
  }
  // MassMomentsObserver's per-gridbox moments (massmoments_observer.cpp:23-73); rain: raindrops only
  struct MassMoments {
    std::vector<uint64_t> mom0;
    std::vector<float> mom1, mom2;
  };
  MassMoments mass_moments(bool rain = false) {
    MassMoments m{std::vector<uint64_t>(ngbx_), std::vector<float>(ngbx_), std::vector<float>(ngbx_)};
    check(ctx_, cleo_mass_moments(ctx_, rain ? 1 : 0, m.mom0.data(), m.mom1.data(), m.mom2.data()), "mass moments");
    return m;
  }

 private:
  cleo_ctx* ctx_ = nullptr;
  uint64_t ngbx_ = 0;
  uint64_t ntot_ = 0;
};

// SDMMethods (sdmmethods.hpp:110-333) for one GPU domain
class SDMMethods {
 public:
  template <typename Microphys, typename Motion>
  SDMMethods(unsigned int couplstep, GpuDomain& domain, const Microphys& microphys, const Motion& motion)
      : couplstep_(couplstep), domain_(domain) {
    const CombinedMicrophysicalProcess m = as_microphysics(microphys);
    set_motion(motion);
    check(domain_.ctx(),
          cleo_set_sdm_chain(domain_.ctx(), m.chain.data(), static_cast<uint32_t>(m.chain.size()),
                             motion_on_ ? &motion_ : nullptr),
          "cleo_set_sdm_chain");
  }
  unsigned int get_couplstep() const { return couplstep_; }
  unsigned int next_couplstep(unsigned int t_mdl) const { return ((t_mdl / couplstep_) + 1) * couplstep_; }
  // superdroplet motion then microphysics over [t_mdl, t_mdl_next)
  void run_step(unsigned int t_mdl, unsigned int t_mdl_next) {
    check(domain_.ctx(), cleo_run_step(domain_.ctx(), t_mdl, t_mdl_next), "run_step");
    check(domain_.ctx(), cleo_synchronize(domain_.ctx()), "run_step");
  }
  // nsteps model steps of length couplstep from t_mdl, replayed as one hipGraph
  void run_steps(unsigned int t_mdl, unsigned int nsteps) {
    check(domain_.ctx(), cleo_run_steps_graph(domain_.ctx(), t_mdl, couplstep_, nsteps), "run_steps");
    check(domain_.ctx(), cleo_synchronize(domain_.ctx()), "run_steps");
  }

 private:
  void set_motion(const PredCorrMotion& m) {
    motion_on_ = true;
    motion_ = m.params;
  }
  void set_motion(const NullMotion&) { motion_on_ = false; }
  unsigned int couplstep_;
  GpuDomain& domain_;
  bool motion_on_ = false;
  cleo_motion_params motion_{};
};

// model-step conversions (libs/initialise/timesteps.hpp:34-59)
inline unsigned int realtime2step(double TSTEP) {
  return static_cast<unsigned int>(std::nearbyint(TSTEP / 1000.0 * 100000.0));
}
inline double step2realtime(unsigned int s) { return (static_cast<double>(s) / 100000.0) * 1000.0; }
inline double step2dimlesstime(unsigned int s) { return static_cast<double>(s) / 100000.0; }
inline double realtime2dimless(double TSTEP) { return TSTEP / 1000.0; }

}  // namespace cleo_gpu
