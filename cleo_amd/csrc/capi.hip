// capi.hip -- the C ABI of include/cleo_hip.h: context, data movement and the SDM step driver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "mlp.hpp"
#include "sdm_kernels.hpp"

using namespace cleo;

#define HIP_TRY(c, expr)                                                             \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) {                                                          \
      if (c) (c)->msg = std::string(#expr) + ": " + hipGetErrorString(e_);           \
      return CLEO_E_HIP;                                                             \
    }                                                                                \
  } while (0)

namespace {

int fail(cleo_ctx* c, int code, const std::string& m) {
  if (c) c->msg = m;
  return code;
}

double step2dimlesstime(uint32_t s) { return static_cast<double>(s) / 100000.0; }
double step2realtime(uint32_t s) { return (static_cast<double>(s) / 100000.0) * dimless::TIME0; }

template <typename T>
hipError_t dalloc(T** p, uint64_t n) {
  return hipMalloc(reinterpret_cast<void**>(p), std::max<uint64_t>(n, 1) * sizeof(T));
}

hipError_t alloc_soa(SoA& s, uint64_t n) {
  hipError_t e;
  if ((e = dalloc(&s.gbx, n)) != hipSuccess) return e;
  if ((e = dalloc(&s.c3, n)) != hipSuccess) return e;
  if ((e = dalloc(&s.c1, n)) != hipSuccess) return e;
  if ((e = dalloc(&s.c2, n)) != hipSuccess) return e;
  if ((e = dalloc(&s.xi, n)) != hipSuccess) return e;
  if ((e = dalloc(&s.r, n)) != hipSuccess) return e;
  if ((e = dalloc(&s.msol, n)) != hipSuccess) return e;
  return dalloc(&s.id, n);
}
void free_soa(SoA& s) {
  hipFree(s.gbx);
  hipFree(s.c3);
  hipFree(s.c1);
  hipFree(s.c2);
  hipFree(s.xi);
  hipFree(s.r);
  hipFree(s.msol);
  hipFree(s.id);
  s = SoA{};
}

// synchronise, then report a recorded device error
int sync_check(cleo_ctx* c) {
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, ctx_copy(c, c->h_ctl, c->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost));
  if (c->h_ctl->err_code != 0) {
    c->msg = "device error " + std::to_string(c->h_ctl->err_code) + " at gridbox " +
             std::to_string(c->h_ctl->err_gbx) + " drop " + std::to_string(c->h_ctl->err_drop);
    return CLEO_E_DEVICE;
  }
  return CLEO_OK;
}

int launch_check(cleo_ctx* c) {
  HIP_TRY(c, hipGetLastError());
  return CLEO_OK;
}

int ensure_radix(cleo_ctx* c, uint64_t n) {
  if (c->rs_n >= n && c->rs_key[0]) return CLEO_OK;
  for (int i = 0; i < 2; ++i) {
    hipFree(c->rs_key[i]);
    hipFree(c->rs_idx[i]);
  }
  hipFree(c->rs_cnt);
  const uint64_t ntiles = (n + 255) / 256;
  for (int i = 0; i < 2; ++i) {
    HIP_TRY(c, dalloc(&c->rs_key[i], n));
    HIP_TRY(c, dalloc(&c->rs_idx[i], n));
  }
  HIP_TRY(c, dalloc(&c->rs_cnt, 32 * ntiles + 1));
  const size_t need = scan_tmp_words(16 * ntiles);
  if (need > c->scan_tmp_n) {
    hipFree(c->scan_tmp);
    HIP_TRY(c, dalloc(&c->scan_tmp, need));
    c->scan_tmp_n = need;
  }
  c->rs_n = n;
  return CLEO_OK;
}

uint32_t wpb_or1(uint32_t v) { return v == 0 ? 1u : std::min<uint32_t>(v, 4u); }

CondArgs cond_args(cleo_ctx* c, const cleo_cond_params& p, bool fuse_qcond) {
  CondArgs a{};
  a.s = c->buf[c->cur];
  a.off = c->off;
  a.nact = c->nact;
  a.st = c->st;
  a.sol = c->solute;
  a.cfg = NewtonCfg{step2dimlesstime(p.interval), p.minsubdelt_s / dimless::TIME0, p.rtol, p.atol, p.maxniters, 0u};
  a.ngbx = c->grid.ngbx;
  a.wpb = wpb_or1(c->tiles.cond_waves_per_block);
  a.alter = p.do_alter_thermo;
  a.fuse_qcond = fuse_qcond ? 1 : 0;
  a.ctl = c->ctl;
  a.niters = c->rec_niters ? c->niters : nullptr;
  a.warm = nullptr;
  a.warm_limit = p.warm_restart_iters;
  a.gs = c->gs;
  return a;
}

// enqueue one condensation pass (learned guesses first when requested)
int enqueue_condensation(cleo_ctx* c, const cleo_cond_params& p, bool fuse_qcond) {
  if (p.interval == 0) return fail(c, CLEO_E_ARG, "condensation interval 0");
  if (!(p.minsubdelt_s > 0.0)) return fail(c, CLEO_E_ARG, "MINSUBTSTEP must be positive");
  materialize_order(c);  // the per-gridbox sums run in drop order
  CondArgs a = cond_args(c, p, fuse_qcond);
  // the thread-per-drop sequence evaluates the learned guesses itself (after the saturation
  // ratios are known); the others read them from a separate guess pass
  const bool inline_mlp = p.warm_start && c->tiles.cond_flat == 1;
  if (p.warm_start && !inline_mlp) {
    const int rc = mlp_enqueue_guess(c, p, &a.warm);
    if (rc != CLEO_OK) return rc;
    if (c->rec_guess && c->guess)
      HIP_TRY(c, hipMemcpyAsync(c->guess, a.warm, c->ntot * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  }
  if (c->tiles.cond_flat) {
    CondFlatArgs f{a.s, a.off, a.nact, a.st, a.sol, a.cfg, a.ngbx, a.alter, a.fuse_qcond, a.ctl, a.niters,
                   a.warm, a.warm_limit, c->gs, c->dm};
    if (inline_mlp) {
      const int rc = mlp_check_ready(c);
      if (rc != CLEO_OK) return rc;
      f.mlp_inline = 1;
      f.W = mlp_weights(c);
      f.warm_buf = mlp_warm_buffer(c);
      f.guess_out = (c->rec_guess && c->guess) ? c->guess : nullptr;
    }
    launch_condensation_flat(c, f);
  } else {
    launch_condensation(c, a);
  }
  return launch_check(c);
}

// offsets of the injected stream per gridbox; validates the sizes the caller gave
int prepare_injected(cleo_ctx* c) {
  launch_inj_sizes(c, c->tot, c->oobpre);
  scan_u32(c, c->tot, c->grid.ngbx, 1u, 0u, c->inj_oj, nullptr);
  scan_u32(c, c->oobpre, c->grid.ngbx, 1u, 0u, c->inj_op, nullptr);
  uint32_t tj = 0, tp = 0;
  HIP_TRY(c, hipMemcpyAsync(&tj, c->inj_oj + c->grid.ngbx, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(&tp, c->inj_op + c->grid.ngbx, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (tj != c->inj_nj || tp != c->inj_nphi)
    return fail(c, CLEO_E_ARG, "injected stream sizes (" + std::to_string(c->inj_nj) + ", " +
                                   std::to_string(c->inj_nphi) + ") != required (" + std::to_string(tj) + ", " +
                                   std::to_string(tp) + ")");
  return CLEO_OK;
}

// one collision event over every gridbox: the drops stay in buf[cur], their shuffled order goes to
// c->order (composed with a pending order of an earlier event)
// the kinds a cleo_coll_params names must be ones the kernels know; the Low-List probabilities
// and the kinetic-energy NFrags / flags read a terminal velocity formula
bool tv_known(int32_t tv) { return tv == CLEO_TV_NULL || tv == CLEO_TV_SIMMEL || tv == CLEO_TV_ROGERS_YAU || tv == CLEO_TV_ROGERS_GK; }
const char* coll_params_error(const cleo_coll_params& p) {
  if (p.interval == 0) return "collision interval 0";
  if (p.prob_kind < CLEO_PROB_GOLOVIN || p.prob_kind > CLEO_PROB_LOWLIST_BU) return "unknown prob_kind";
  if (p.enact_kind < CLEO_ENACT_COAL || p.enact_kind > CLEO_ENACT_COALBURE) return "unknown enact_kind";
  if (p.nfrags_kind != CLEO_NFRAGS_CONST && p.nfrags_kind != CLEO_NFRAGS_CKE) return "unknown nfrags_kind";
  if (p.flag_kind < CLEO_FLAG_CONST || p.flag_kind > CLEO_FLAG_STRAUB) return "unknown flag_kind";
  if (!tv_known(p.prob_tv) || !tv_known(p.nfrags_tv) || !tv_known(p.flag_tv)) return "unknown terminal velocity";
  return nullptr;
}
CollCfg make_collcfg(const cleo_coll_params& p) {
  CollCfg k{};
  k.prob_kind = p.prob_kind;
  k.prob_param = p.prob_param;
  k.enact_kind = p.enact_kind;
  k.nfrags = p.nfrags;
  k.coaleff = p.coalbure_coaleff;
  k.DELT = step2realtime(p.interval);
  k.nfrags_kind = p.nfrags_kind;
  k.flag_kind = p.flag_kind;
  k.prob_tv = p.prob_tv;
  k.nfrags_tv = p.nfrags_tv;
  k.flag_tv = p.flag_tv;
  return k;
}

int enqueue_collision_event(cleo_ctx* c, const cleo_coll_params& p) {
  if (const char* m = coll_params_error(p)) return fail(c, CLEO_E_ARG, m);
  if (c->rng_mode == 0) return fail(c, CLEO_E_STATE, "no random-number stream set");
  const bool inj = c->rng_mode == 2;
  if (inj) {
    const int rc = prepare_injected(c);
    if (rc != CLEO_OK) return rc;
  }
  // a pending order of an earlier event is composed with this one's shuffle: the drops stay where
  // they are, and the new order goes to the other order buffer
  const bool compose = c->order_live;
  CollArgs a{};
  a.s = c->buf[c->cur];
  a.order = compose ? c->order_alt : c->order;
  a.order_in = compose ? c->order : nullptr;
  a.off = c->off;
  a.nact = c->nact;
  a.vol = c->st.vol;
  a.sol = c->solute;
  a.k = make_collcfg(p);
  a.ngbx = c->grid.ngbx;
  a.wpb = wpb_or1(c->tiles.coll_waves_per_block);
  a.ctl = c->ctl;
  a.seed = c->seed;
  a.inj_j = inj ? c->inj_j : nullptr;
  a.inj_phi = inj ? c->inj_phi : nullptr;
  a.inj_oj = c->inj_oj;
  a.inj_op = c->inj_op;
  a.big = c->big_scratch;
  launch_collisions(c, a);
  launch_advance_event(c);
  if (compose) std::swap(c->order, c->order_alt);
  c->order_live = true;
  if (inj) c->rng_mode = 0;  // consumed
  return launch_check(c);
}

// MoveSupersInDomain::run_step body: motion, stable sort + refs, (qcond unless fused later)
int enqueue_motion(cleo_ctx* c, const cleo_motion_params& p, bool qcond) {
  if (p.interval == 0) return fail(c, CLEO_E_ARG, "motion interval 0");
  MotionArgs a{};
  a.s = c->buf[c->cur];
  a.off = c->off;
  a.grid = c->grid;
  a.st = c->st;
  a.delt = step2dimlesstime(p.interval);
  a.tv = p.terminalv;
  a.wpb = wpb_or1(c->tiles.motion_waves_per_block);
  a.cnt = c->cnt;
  a.ctl = c->ctl;
  if (c->tiles.motion_flat == 2) {
    materialize_order(c);  // the fused pass ranks the drops in store order
    a.s = c->buf[c->cur];
    motion_sort_fused(c, a, c->buf[1 - c->cur], 1 - c->cur);
  } else if (c->tiles.motion_flat == 3) {
    motion_sort_gather(c, a, c->buf[1 - c->cur], 1 - c->cur);
  } else {
    if (c->tiles.motion_flat) {
      a.step = c->dstep;  // the sort reads each drop's step instead of decoding its new key
      launch_motion_flat(c, a);
    } else {
      launch_motion(c, a);
    }
    neighbour_sort(c, c->cur, 1 - c->cur, nullptr, false, a.step);
  }
  if (qcond) {
    QcondArgs q{c->buf[c->cur], c->off, c->st, c->solute, c->grid.ngbx, wpb_or1(c->tiles.cond_waves_per_block), c->ctl};
    launch_qcond(c, q);
  }
  return launch_check(c);
}

// MoveSupersInDomain::run_step followed at once by DoCondensation on the sorted drops, with the
// condensation solve in the sort's gather (thread-per-drop motion and condensation)
int enqueue_motion_condensation(cleo_ctx* c, const cleo_motion_params& mp, const cleo_cond_params& cp) {
  if (mp.interval == 0) return fail(c, CLEO_E_ARG, "motion interval 0");
  if (cp.interval == 0) return fail(c, CLEO_E_ARG, "condensation interval 0");
  if (!(cp.minsubdelt_s > 0.0)) return fail(c, CLEO_E_ARG, "MINSUBTSTEP must be positive");
  MotionArgs a{};
  a.s = c->buf[c->cur];
  a.off = c->off;
  a.grid = c->grid;
  a.st = c->st;
  a.delt = step2dimlesstime(mp.interval);
  a.tv = mp.terminalv;
  a.wpb = wpb_or1(c->tiles.motion_waves_per_block);
  a.cnt = c->cnt;
  a.ctl = c->ctl;
  const CondArgs ca = cond_args(c, cp, true);
  CondFlatArgs f{ca.s, ca.off, ca.nact, ca.st, ca.sol, ca.cfg, ca.ngbx, ca.alter, ca.fuse_qcond, ca.ctl, ca.niters,
                 nullptr, ca.warm_limit, c->gs, c->dm};
  if (cp.warm_start) {  // learned guesses, evaluated on the moved drops before the gather
    const int rc = mlp_check_ready(c);
    if (rc != CLEO_OK) return rc;
    f.mlp_inline = 1;
    f.W = mlp_weights(c);
    f.warm_buf = mlp_warm_buffer(c);
  }
  motion_sort_condense(c, a, f, c->buf[1 - c->cur], 1 - c->cur);
  return launch_check(c);
}

uint32_t process_interval(const cleo_process& q) {
  return q.kind == CLEO_PROCESS_CONDENSATION ? q.cond.interval : q.coll.interval;
}

// SDMMethods::run_step(t_mdl, t_mdl_next) (sdmmethods.hpp:319-332): host loop over SDM sub-steps;
// every gridbox follows the same schedule, so each process runs as one launch over all gridboxes.
// On each sub-step the processes that are on step run in chain order (CombinedMicrophysicalProcess,
// microphysicalprocess.hpp:65-120); the next sub-step is the earliest next step of any of them.
int enqueue_run_step(cleo_ctx* c, uint32_t t_mdl, uint32_t t_mdl_next) {
  const cleo_ctx::Chain& p = c->sdm;
  const uint32_t UMAX = 0xFFFFFFFFu;
  auto next_of = [](uint32_t t, uint32_t iv) { return ((t / iv) + 1) * iv; };
  auto on_step = [&](uint32_t i, uint32_t t) {
    const uint32_t iv = process_interval(p.proc[i]);
    return t % iv == 0 && iv != UMAX;
  };
  auto qcond_now = [&]() {
    QcondArgs q{c->buf[c->cur], c->off, c->st, c->solute, c->grid.ngbx, wpb_or1(c->tiles.cond_waves_per_block), c->ctl};
    launch_qcond(c, q);
  };
  uint32_t t_sdm = t_mdl;
  while (t_sdm < t_mdl_next) {
    const uint32_t next_move = p.motion_on ? next_of(t_sdm, p.motion.interval) : UMAX;
    const uint32_t t_sdm_next = std::min(t_mdl_next, next_move);
    bool pending_qcond = false;
    int fused = -1;  // the process whose step at t_sdm already ran inside the sort
    if (p.motion_on && t_sdm % p.motion.interval == 0) {
      // when the first process on step at once is a condensation, it is solved in the sort's gather
      // (thread-per-drop variants; not while guesses are recorded, whose positions are those of the
      // sorted store); otherwise motion and sort run alone and qcond is recomputed afterwards
      int first = -1;
      for (uint32_t i = 0; i < p.n && first < 0; ++i)
        if (on_step(i, t_sdm)) first = static_cast<int>(i);
      const bool fuse = first >= 0 && p.proc[first].kind == CLEO_PROCESS_CONDENSATION &&
                        (c->tiles.motion_flat == 1 || c->tiles.motion_flat == 3) && c->tiles.cond_flat == 1 &&
                        !(p.proc[first].cond.warm_start && c->rec_guess);
      const int rc = fuse ? enqueue_motion_condensation(c, p.motion, p.proc[first].cond)
                          : enqueue_motion(c, p.motion, false);
      if (rc != CLEO_OK) return rc;
      pending_qcond = !fuse;
      fused = fuse ? first : -1;
    }
    bool any_coll = false;
    for (uint32_t subt = t_sdm; subt < t_sdm_next;) {
      for (uint32_t i = 0; i < p.n; ++i) {
        if (!on_step(i, subt) || (static_cast<int>(i) == fused && subt == t_sdm)) continue;
        const cleo_process& q = p.proc[i];
        int rc;
        if (q.kind == CLEO_PROCESS_CONDENSATION) {
          rc = enqueue_condensation(c, q.cond, pending_qcond);  // with the qcond recompute first if due
        } else {
          if (pending_qcond) qcond_now();  // qcond after the sort comes before any drop changes
          rc = enqueue_collision_event(c, q.coll);
          any_coll = true;
        }
        if (rc != CLEO_OK) return rc;
        pending_qcond = false;
      }
      if (pending_qcond) {
        qcond_now();
        pending_qcond = false;
      }
      uint32_t nxt = UMAX;
      for (uint32_t i = 0; i < p.n; ++i) nxt = std::min(nxt, next_of(subt, process_interval(p.proc[i])));
      subt = nxt;
    }
    if (pending_qcond) qcond_now();
    if (any_coll) neighbour_sort(c, c->cur, 1 - c->cur, &c->ctl->resort_flag, true);
    t_sdm = t_sdm_next;
  }
  return launch_check(c);
}

}  // namespace

extern "C" {

int cleo_create(cleo_ctx** out, uint64_t maxn, const cleo_domain_desc* dom) {
  if (!out || !dom || !dom->zhalf || !dom->xhalf || !dom->yhalf) return CLEO_E_ARG;
  const uint64_t nz = dom->ndims[0], nx = dom->ndims[1], ny = dom->ndims[2];
  if (nz == 0 || nx == 0 || ny == 0) return CLEO_E_ARG;
  const uint64_t ngbx = nz * nx * ny;
  if (ngbx >= 0xFFFFFFF0ull || maxn >= 0xFFFFFFF0ull) return CLEO_E_ARG;
  auto increasing = [](const double* h, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i)
      if (!(h[i] < h[i + 1])) return false;
    return true;
  };
  if (!increasing(dom->zhalf, nz) || !increasing(dom->xhalf, nx) || !increasing(dom->yhalf, ny)) return CLEO_E_ARG;
  cleo_ctx* c = new cleo_ctx();
  HIP_TRY(c, hipGetDevice(&c->device));
  HIP_TRY(c, hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, c->device));
  HIP_TRY(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->grid.nz = static_cast<uint32_t>(nz);
  c->grid.nx = static_cast<uint32_t>(nx);
  c->grid.ny = static_cast<uint32_t>(ny);
  c->grid.ngbx = static_cast<uint32_t>(ngbx);
  for (int i = 0; i < 3; ++i) c->grid.per[i] = dom->periodic[i] ? 1 : 0;
  const uint64_t nh = nz + nx + ny + 3;
  HIP_TRY(c, dalloc(&c->d_halves, nh));
  HIP_TRY(c, ctx_copy(c, c->d_halves, dom->zhalf, (nz + 1) * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, ctx_copy(c, c->d_halves + nz + 1, dom->xhalf, (nx + 1) * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, ctx_copy(c, c->d_halves + nz + nx + 2, dom->yhalf, (ny + 1) * 8, hipMemcpyHostToDevice));
  c->grid.zh = c->d_halves;
  c->grid.xh = c->d_halves + nz + 1;
  c->grid.yh = c->d_halves + nz + nx + 2;
  c->maxn = maxn;
  HIP_TRY(c, alloc_soa(c->buf[0], maxn));
  HIP_TRY(c, alloc_soa(c->buf[1], maxn));
  HIP_TRY(c, dalloc(&c->off, ngbx + 1));
  HIP_TRY(c, dalloc(&c->off_alt, ngbx + 1));
  HIP_TRY(c, dalloc(&c->nact, ngbx));
  HIP_TRY(c, dalloc(&c->d_state, 11 * ngbx));
  double* s = c->d_state;
  c->st = GState{s, s + ngbx, s + 2 * ngbx, s + 3 * ngbx, s + 4 * ngbx, s + 5 * ngbx,
                 s + 6 * ngbx, s + 7 * ngbx, s + 8 * ngbx, s + 9 * ngbx, s + 10 * ngbx};
  HIP_TRY(c, dalloc(&c->cnt, 56 * ngbx));  // counts, then the run offsets of the neighbour sort
  HIP_TRY(c, dalloc(&c->tot, ngbx + 1));
  HIP_TRY(c, dalloc(&c->oobpre, ngbx + 1));
  HIP_TRY(c, dalloc(&c->oobcnt, ngbx));
  HIP_TRY(c, dalloc(&c->inj_oj, ngbx + 1));
  HIP_TRY(c, dalloc(&c->inj_op, ngbx + 1));
  c->scan_tmp_n = scan_tmp_words(28 * ngbx + 1);
  HIP_TRY(c, dalloc(&c->scan_tmp, c->scan_tmp_n));
  HIP_TRY(c, dalloc(&c->big_scratch, 2 * maxn));
  HIP_TRY(c, dalloc(&c->gs, kGsStride * ngbx));
  HIP_TRY(c, dalloc(&c->dm, maxn));
  // a gridbox launch has at most ngbx + 3 waves, a drop launch 4 per 256 drops
  c->sp_slots = std::max<uint64_t>(ngbx + 4, 4 * ((maxn + 255) / 256) + 4);
  HIP_TRY(c, dalloc(&c->sp, kStatSlotsMax * c->sp_slots));
  // deferred drops of the condensation: per-wave segments (waves of a drop launch times a quarter
  // of their drops + 16), the shared list they move to and its counter.  The list has room for
  // every drop (a launch defers each drop at most once), so no deferred drop is ever solved in
  // the launch that deferred it.
  HIP_TRY(c, dalloc(&c->dcount, 1));
  c->dlist_cap = static_cast<uint32_t>(std::min<uint64_t>(maxn + 64, 0xFFFFFFFFull));
  HIP_TRY(c, dalloc(&c->dlist, c->dlist_cap));
  HIP_TRY(c, dalloc(&c->dseg, std::min<uint64_t>(maxn / 4 + 20 * c->sp_slots + 64, 0xFFFFFFFFull)));
  HIP_TRY(c, dalloc(&c->perm, std::max<uint64_t>(maxn, 1)));
  HIP_TRY(c, dalloc(&c->dstep, std::max<uint64_t>(maxn, 1)));
  HIP_TRY(c, dalloc(&c->order, std::max<uint64_t>(maxn, 1)));
  HIP_TRY(c, dalloc(&c->order_alt, std::max<uint64_t>(maxn, 1)));
  HIP_TRY(c, dalloc(&c->qm, std::max<uint64_t>(maxn, 1)));
  c->big_scratch_n = 2 * maxn;
  HIP_TRY(c, dalloc(&c->ctl, 1));
  HIP_TRY(c, hipHostMalloc(reinterpret_cast<void**>(&c->h_ctl), sizeof(DevCtl), hipHostMallocDefault));
  HIP_TRY(c, ctx_fill(c, c->ctl, 0, sizeof(DevCtl)));
  HIP_TRY(c, ctx_fill(c, c->off, 0, (ngbx + 1) * 4));
  HIP_TRY(c, ctx_fill(c, c->off_alt, 0, (ngbx + 1) * 4));
  HIP_TRY(c, ctx_fill(c, c->nact, 0, ngbx * 4));
  HIP_TRY(c, ctx_fill(c, c->d_state, 0, 11 * ngbx * 8));
  c->solute = make_solute(2016.5 / dimless::RHO0, 0.05844277 / dimless::MR0, 2.0);
  HIP_TRY(c, mlp_create(c));
  *out = c;
  return CLEO_OK;
}

int cleo_destroy(cleo_ctx* c) {
  if (!c) return CLEO_OK;
  hipStreamSynchronize(c->stream);
  mlp_destroy(c);
  free_soa(c->buf[0]);
  free_soa(c->buf[1]);
  for (void* p : {static_cast<void*>(c->d_halves), static_cast<void*>(c->off), static_cast<void*>(c->off_alt),
                  static_cast<void*>(c->nact), static_cast<void*>(c->d_state), static_cast<void*>(c->cnt),
                  static_cast<void*>(c->tot), static_cast<void*>(c->oobpre), static_cast<void*>(c->oobcnt), static_cast<void*>(c->inj_oj),
                  static_cast<void*>(c->inj_op), static_cast<void*>(c->scan_tmp), static_cast<void*>(c->big_scratch),
                  static_cast<void*>(c->ctl), static_cast<void*>(c->inj_j), static_cast<void*>(c->inj_phi),
                  static_cast<void*>(c->niters), static_cast<void*>(c->guess), static_cast<void*>(c->rs_key[0]),
                  static_cast<void*>(c->rs_key[1]), static_cast<void*>(c->rs_idx[0]), static_cast<void*>(c->rs_idx[1]),
                  static_cast<void*>(c->rs_cnt), static_cast<void*>(c->gs), static_cast<void*>(c->dm),
                  static_cast<void*>(c->sp), static_cast<void*>(c->dlist), static_cast<void*>(c->dcount),
                  static_cast<void*>(c->dseg), c->mom, static_cast<void*>(c->perm), static_cast<void*>(c->dstep), static_cast<void*>(c->order), static_cast<void*>(c->order_alt),
                  static_cast<void*>(c->qm)})
    if (p) hipFree(p);
  if (c->h_ctl) hipHostFree(c->h_ctl);
  for (hipEvent_t& e : c->ev)
    if (e) hipEventDestroy(e);
  for (hipEvent_t& e : c->kev)
    if (e) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return CLEO_OK;
}

const char* cleo_last_error_message(const cleo_ctx* c) { return c ? c->msg.c_str() : "null context"; }

int cleo_device_info(char* buf, size_t len) {
  hipDeviceProp_t p;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return CLEO_E_HIP;
  std::snprintf(buf, len, "%s %s CUs=%d", p.name, p.gcnArchName, p.multiProcessorCount);
  return CLEO_OK;
}

int cleo_upload_supers(cleo_ctx* c, const cleo_supers_soa* h, uint64_t ntot) {
  if (!c || !h) return CLEO_E_ARG;
  if (ntot > c->maxn) return fail(c, CLEO_E_ARG, "ntot exceeds the context capacity");
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->order_live = false;  // the uploaded drops replace the store and its pending order
  SoA& d = c->buf[c->cur];
  HIP_TRY(c, ctx_copy(c, d.gbx, h->sdgbxindex, ntot * 4, hipMemcpyHostToDevice));
  HIP_TRY(c, ctx_copy(c, d.c3, h->coord3, ntot * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, ctx_copy(c, d.c1, h->coord1, ntot * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, ctx_copy(c, d.c2, h->coord2, ntot * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, ctx_copy(c, d.xi, h->xi, ntot * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, ctx_copy(c, d.r, h->radius, ntot * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, ctx_copy(c, d.msol, h->msol, ntot * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, ctx_copy(c, d.id, h->sdid, ntot * 8, hipMemcpyHostToDevice));
  c->ntot = ntot;
  const int rc = ensure_radix(c, ntot);
  if (rc != CLEO_OK) return rc;
  generic_sort(c);
  HIP_TRY(c, hipGetLastError());
  const int rs = sync_check(c);
  // the upload's own sort is the reference's constructor sort, not a step sort
  HIP_TRY(c, ctx_fill(c, &c->ctl->stats[ST_SORTS], 0, 8));
  return rs;
}

int cleo_upload_supers_device(cleo_ctx* c, const cleo_supers_soa* h, uint64_t ntot) {
  if (!c || !h) return CLEO_E_ARG;
  if (ntot > c->maxn) return fail(c, CLEO_E_ARG, "ntot exceeds the context capacity");
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->order_live = false;  // the uploaded drops replace the store and its pending order
  SoA& d = c->buf[c->cur];
  HIP_TRY(c, ctx_copy(c, d.gbx, h->sdgbxindex, ntot * 4, hipMemcpyDeviceToDevice));
  HIP_TRY(c, ctx_copy(c, d.c3, h->coord3, ntot * 8, hipMemcpyDeviceToDevice));
  HIP_TRY(c, ctx_copy(c, d.c1, h->coord1, ntot * 8, hipMemcpyDeviceToDevice));
  HIP_TRY(c, ctx_copy(c, d.c2, h->coord2, ntot * 8, hipMemcpyDeviceToDevice));
  HIP_TRY(c, ctx_copy(c, d.xi, h->xi, ntot * 8, hipMemcpyDeviceToDevice));
  HIP_TRY(c, ctx_copy(c, d.r, h->radius, ntot * 8, hipMemcpyDeviceToDevice));
  HIP_TRY(c, ctx_copy(c, d.msol, h->msol, ntot * 8, hipMemcpyDeviceToDevice));
  HIP_TRY(c, ctx_copy(c, d.id, h->sdid, ntot * 8, hipMemcpyDeviceToDevice));
  c->ntot = ntot;
  const int rc = ensure_radix(c, ntot);
  if (rc != CLEO_OK) return rc;
  generic_sort(c);
  HIP_TRY(c, hipGetLastError());
  const int rs = sync_check(c);
  HIP_TRY(c, ctx_fill(c, &c->ctl->stats[ST_SORTS], 0, 8));
  return rs;
}

// reference Superdrop layout (superdrop.hpp:36-60): u32 gbx @0, f64 coord3 @8, coord1 @16,
// coord2 @24, attrs {solute @32, u64 xi @40, f64 radius @48, f64 msol @56}, u64 id @64
int cleo_upload_supers_aos(cleo_ctx* c, const void* aos, uint64_t ntot) {
  if (!c || !aos) return CLEO_E_ARG;
  std::vector<uint32_t> g(ntot);
  std::vector<double> c3(ntot), c1(ntot), c2(ntot), r(ntot), m(ntot);
  std::vector<uint64_t> xi(ntot), id(ntot);
  const unsigned char* p = static_cast<const unsigned char*>(aos);
  for (uint64_t k = 0; k < ntot; ++k, p += 72) {
    std::memcpy(&g[k], p, 4);
    std::memcpy(&c3[k], p + 8, 8);
    std::memcpy(&c1[k], p + 16, 8);
    std::memcpy(&c2[k], p + 24, 8);
    std::memcpy(&xi[k], p + 40, 8);
    std::memcpy(&r[k], p + 48, 8);
    std::memcpy(&m[k], p + 56, 8);
    std::memcpy(&id[k], p + 64, 8);
  }
  cleo_supers_soa s{g.data(), c3.data(), c1.data(), c2.data(), xi.data(), r.data(), m.data(), id.data()};
  return cleo_upload_supers(c, &s, ntot);
}

int cleo_download_supers(cleo_ctx* c, cleo_supers_soa* h, uint64_t* ntot, uint64_t* ndomain) {
  if (!c || !h) return CLEO_E_ARG;
  materialize_order(c);  // the drops in their order
  const int rc = sync_check(c);
  const uint64_t n = c->ntot;
  const SoA& d = c->buf[c->cur];
  HIP_TRY(c, ctx_copy(c, h->sdgbxindex, d.gbx, n * 4, hipMemcpyDeviceToHost));
  HIP_TRY(c, ctx_copy(c, h->coord3, d.c3, n * 8, hipMemcpyDeviceToHost));
  HIP_TRY(c, ctx_copy(c, h->coord1, d.c1, n * 8, hipMemcpyDeviceToHost));
  HIP_TRY(c, ctx_copy(c, h->coord2, d.c2, n * 8, hipMemcpyDeviceToHost));
  HIP_TRY(c, ctx_copy(c, h->xi, d.xi, n * 8, hipMemcpyDeviceToHost));
  HIP_TRY(c, ctx_copy(c, h->radius, d.r, n * 8, hipMemcpyDeviceToHost));
  HIP_TRY(c, ctx_copy(c, h->msol, d.msol, n * 8, hipMemcpyDeviceToHost));
  HIP_TRY(c, ctx_copy(c, h->sdid, d.id, n * 8, hipMemcpyDeviceToHost));
  uint32_t nd = 0;
  HIP_TRY(c, ctx_copy(c, &nd, c->off + c->grid.ngbx, 4, hipMemcpyDeviceToHost));
  if (ntot) *ntot = n;
  if (ndomain) *ndomain = nd;
  return rc;
}

int cleo_download_supers_aos(cleo_ctx* c, void* aos, uint64_t* ntot, uint64_t* ndomain) {
  if (!c || !aos) return CLEO_E_ARG;
  const uint64_t n = c->ntot;
  std::vector<uint32_t> g(n);
  std::vector<double> c3(n), c1(n), c2(n), r(n), m(n);
  std::vector<uint64_t> xi(n), id(n);
  cleo_supers_soa s{g.data(), c3.data(), c1.data(), c2.data(), xi.data(), r.data(), m.data(), id.data()};
  const int rc = cleo_download_supers(c, &s, ntot, ndomain);
  unsigned char* p = static_cast<unsigned char*>(aos);
  for (uint64_t k = 0; k < n; ++k, p += 72) {
    std::memset(p, 0, 72);
    std::memcpy(p, &g[k], 4);
    std::memcpy(p + 8, &c3[k], 8);
    std::memcpy(p + 16, &c1[k], 8);
    std::memcpy(p + 24, &c2[k], 8);
    std::memcpy(p + 40, &xi[k], 8);
    std::memcpy(p + 48, &r[k], 8);
    std::memcpy(p + 56, &m[k], 8);
    std::memcpy(p + 64, &id[k], 8);
  }
  return rc;
}

int cleo_download_refs(cleo_ctx* c, uint32_t* refs) {
  if (!c || !refs) return CLEO_E_ARG;
  const int rc = sync_check(c);
  HIP_TRY(c, ctx_copy(c, refs, c->off, (c->grid.ngbx + 1ull) * 4, hipMemcpyDeviceToHost));
  return rc;
}

// FromFileComms::receive_dynamics from device arrays: no host round trip of the State
int cleo_receive_dynamics(cleo_ctx* c, const cleo_dynamics_step* dyn) {
  if (!c || !dyn || !dyn->press || !dyn->temp || !dyn->qvap || !dyn->qcond) return CLEO_E_ARG;
  launch_receive_dynamics(c, *dyn);
  return launch_check(c);
}

// MassMomentsObserver's reductions (observers/massmoments_observer.cpp:23-73) on the device,
// copied to the host arrays
int cleo_mass_moments(cleo_ctx* c, int32_t rain, uint64_t* mom0, float* mom1, float* mom2) {
  if (!c || !mom0 || !mom1 || !mom2) return CLEO_E_ARG;
  const uint64_t ngbx = c->grid.ngbx;
  if (!c->mom) {
    uint8_t* p = nullptr;
    HIP_TRY(c, dalloc(&p, 16 * ngbx));
    c->mom = p;
  }
  uint64_t* d0 = static_cast<uint64_t*>(c->mom);
  float* d1 = reinterpret_cast<float*>(d0 + ngbx);
  float* d2 = d1 + ngbx;
  launch_mass_moments(c, rain != 0, d0, d1, d2);
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, ctx_copy(c, mom0, d0, ngbx * 8, hipMemcpyDeviceToHost));
  HIP_TRY(c, ctx_copy(c, mom1, d1, ngbx * 4, hipMemcpyDeviceToHost));
  HIP_TRY(c, ctx_copy(c, mom2, d2, ngbx * 4, hipMemcpyDeviceToHost));
  return sync_check(c);
}

static double* const* state_fields(cleo_ctx* c, double* (&out)[11]) {
  out[0] = c->st.vol; out[1] = c->st.press; out[2] = c->st.temp; out[3] = c->st.qv; out[4] = c->st.qc;
  out[5] = c->st.wlo; out[6] = c->st.whi; out[7] = c->st.ulo; out[8] = c->st.uhi; out[9] = c->st.vlo;
  out[10] = c->st.vhi;
  return out;
}

int cleo_upload_gbx_state(cleo_ctx* c, const cleo_gbx_state* h) {
  if (!c || !h) return CLEO_E_ARG;
  double* f[11];
  state_fields(c, f);
  const double* src[11] = {h->volume, h->press, h->temp, h->qvap, h->qcond, h->wvel_lo,
                           h->wvel_hi, h->uvel_lo, h->uvel_hi, h->vvel_lo, h->vvel_hi};
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < 11; ++i) {
    if (!src[i]) return fail(c, CLEO_E_ARG, "null state field");
    HIP_TRY(c, ctx_copy(c, f[i], src[i], c->grid.ngbx * 8ull, hipMemcpyHostToDevice));
  }
  return CLEO_OK;
}

int cleo_download_gbx_state(cleo_ctx* c, cleo_gbx_state* h) {
  if (!c || !h) return CLEO_E_ARG;
  const int rc = sync_check(c);
  double* f[11];
  state_fields(c, f);
  double* dst[11] = {h->volume, h->press, h->temp, h->qvap, h->qcond, h->wvel_lo,
                     h->wvel_hi, h->uvel_lo, h->uvel_hi, h->vvel_lo, h->vvel_hi};
  for (int i = 0; i < 11; ++i)
    if (dst[i]) HIP_TRY(c, ctx_copy(c, dst[i], f[i], c->grid.ngbx * 8ull, hipMemcpyDeviceToHost));
  return rc;
}

int cleo_set_solute(cleo_ctx* c, const cleo_solute* s) {
  if (!c || !s || !(s->rho_sol > 0) || !(s->mr_sol > 0)) return CLEO_E_ARG;
  c->solute = make_solute(s->rho_sol, s->mr_sol, s->ionic);
  return CLEO_OK;
}

int cleo_set_rng_counter(cleo_ctx* c, uint64_t seed, uint64_t first_event) {
  if (!c) return CLEO_E_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->rng_mode = 1;
  c->seed = seed;
  HIP_TRY(c, ctx_copy(c, &c->ctl->event, &first_event, 8, hipMemcpyHostToDevice));
  return CLEO_OK;
}

int cleo_set_rng_injected(cleo_ctx* c, const uint32_t* j, uint64_t nj, const double* phi, uint64_t nphi) {
  if (!c || (nj && !j) || (nphi && !phi)) return CLEO_E_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (nj > c->inj_cap_j) {
    hipFree(c->inj_j);
    HIP_TRY(c, dalloc(&c->inj_j, nj));
    c->inj_cap_j = nj;
  }
  if (nphi > c->inj_cap_phi) {
    hipFree(c->inj_phi);
    HIP_TRY(c, dalloc(&c->inj_phi, nphi));
    c->inj_cap_phi = nphi;
  }
  if (!c->inj_j) HIP_TRY(c, dalloc(&c->inj_j, 1));
  if (!c->inj_phi) HIP_TRY(c, dalloc(&c->inj_phi, 1));
  if (nj) HIP_TRY(c, ctx_copy(c, c->inj_j, j, nj * 4, hipMemcpyHostToDevice));
  if (nphi) HIP_TRY(c, ctx_copy(c, c->inj_phi, phi, nphi * 8, hipMemcpyHostToDevice));
  c->inj_nj = nj;
  c->inj_nphi = nphi;
  c->rng_mode = 2;
  return CLEO_OK;
}

int cleo_set_sdm_chain(cleo_ctx* c, const cleo_process* procs, uint32_t nprocs, const cleo_motion_params* motion) {
  if (!c || (nprocs && !procs)) return CLEO_E_ARG;
  if (nprocs > CLEO_MAX_PROCESSES) return fail(c, CLEO_E_ARG, "more processes than CLEO_MAX_PROCESSES");
  for (uint32_t i = 0; i < nprocs; ++i) {
    const cleo_process& q = procs[i];
    if (q.kind == CLEO_PROCESS_CONDENSATION) {
      if (q.cond.interval == 0) return fail(c, CLEO_E_ARG, "zero interval");
    } else if (q.kind == CLEO_PROCESS_COLLISIONS) {
      if (const char* m = coll_params_error(q.coll)) return fail(c, CLEO_E_ARG, m);
    } else {
      return fail(c, CLEO_E_ARG, "unknown process kind");
    }
  }
  if (motion && motion->interval == 0) return fail(c, CLEO_E_ARG, "zero interval");
  if (motion && !tv_known(motion->terminalv)) return fail(c, CLEO_E_ARG, "unknown terminal velocity");
  c->sdm = cleo_ctx::Chain{};
  for (uint32_t i = 0; i < nprocs; ++i) c->sdm.proc[i] = procs[i];
  c->sdm.n = nprocs;
  c->sdm.motion_on = motion != nullptr;
  if (motion) c->sdm.motion = *motion;
  c->sdm_set = true;
  return CLEO_OK;
}

int cleo_set_sdm(cleo_ctx* c, const cleo_sdm_params* p) {
  if (!c || !p) return CLEO_E_ARG;
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
  return cleo_set_sdm_chain(c, q, n, p->enable_motion ? &p->motion : nullptr);
}

int cleo_sort(cleo_ctx* c) {
  if (!c) return CLEO_E_ARG;
  materialize_order(c);  // the generic sort is stable in the drops' order
  const int rc = ensure_radix(c, c->ntot);
  if (rc != CLEO_OK) return rc;
  generic_sort(c);
  return launch_check(c);
}

int cleo_motion(cleo_ctx* c, const cleo_motion_params* p) {
  if (!c || !p) return CLEO_E_ARG;
  if (!tv_known(p->terminalv)) return fail(c, CLEO_E_ARG, "unknown terminal velocity");
  return enqueue_motion(c, *p, true);
}

static int set_substep_error(cleo_ctx* c) {
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, ctx_copy(c, c->h_ctl, c->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost));
  if (c->h_ctl->err_code == 0) {
    int32_t code = CLEO_ERR_SUBSTEP;
    HIP_TRY(c, ctx_copy(c, &c->ctl->err_code, &code, 4, hipMemcpyHostToDevice));
  }
  c->msg = "timestep must be at least as large as the minimum sub-timestep";
  return CLEO_E_DEVICE;
}

int cleo_condensation(cleo_ctx* c, const cleo_cond_params* p) {
  if (!c || !p) return CLEO_E_ARG;
  if (p->interval == 0) return fail(c, CLEO_E_ARG, "interval 0");
  if (step2dimlesstime(p->interval) < p->minsubdelt_s / dimless::TIME0) return set_substep_error(c);
  return enqueue_condensation(c, *p, false);
}

int cleo_collisions(cleo_ctx* c, const cleo_coll_params* p) {
  if (!c || !p) return CLEO_E_ARG;
  const int rc = enqueue_collision_event(c, *p);
  if (rc != CLEO_OK) return rc;
  neighbour_sort(c, c->cur, 1 - c->cur, &c->ctl->resort_flag, true);
  return launch_check(c);
}

int cleo_run_step(cleo_ctx* c, uint32_t t_mdl, uint32_t t_mdl_next) {
  if (!c) return CLEO_E_ARG;
  if (!c->sdm_set) return fail(c, CLEO_E_STATE, "cleo_set_sdm not called");
  for (uint32_t i = 0; i < c->sdm.n; ++i) {
    const cleo_process& q = c->sdm.proc[i];
    if (q.kind == CLEO_PROCESS_CONDENSATION &&
        step2dimlesstime(q.cond.interval) < q.cond.minsubdelt_s / dimless::TIME0)
      return set_substep_error(c);
  }
  return enqueue_run_step(c, t_mdl, t_mdl_next);
}

int cleo_run_steps_graph(cleo_ctx* c, uint32_t t_mdl, uint32_t span, uint32_t nsteps) {
  if (!c || span == 0) return CLEO_E_ARG;
  if (!c->sdm_set) return fail(c, CLEO_E_STATE, "cleo_set_sdm not called");
  const cleo_ctx::Chain& p = c->sdm;
  if (c->rng_mode != 1 && p.first_coll()) return fail(c, CLEO_E_STATE, "graph replay needs the counter RNG");
  auto aligned = [&](bool on, uint32_t iv) { return !on || (span % iv == 0 && t_mdl % iv == 0); };
  bool ok = aligned(p.motion_on, p.motion.interval);
  for (uint32_t i = 0; i < p.n; ++i) ok = ok && aligned(true, process_interval(p.proc[i]));
  if (!ok) return fail(c, CLEO_E_ARG, "graph span must be a multiple of every process interval");
  for (uint32_t i = 0; i < p.n; ++i) {
    if (p.proc[i].kind == CLEO_PROCESS_CONDENSATION && p.proc[i].cond.warm_start) {
      const int rc = mlp_check_ready(c);
      if (rc != CLEO_OK) return rc;
    }
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  // The launches a step enqueues depend on the store's state when it starts (which buffer is
  // current, whether a collision order is pending), so every replay must start in the state the
  // previous one ended in.  One step is captured, or two when one step changes the state; if two
  // still do, the start state is set to the end state outside the graph and the capture repeated.
  hipGraph_t graph = nullptr;
  uint32_t per_graph = 1;
  int cur0 = c->cur;
  bool live0 = c->order_live;
  uint32_t* ord0 = c->order;  // which of the two order buffers holds the pending order
  auto set_order = [&](uint32_t* ord) {
    if (c->order != ord) std::swap(c->order, c->order_alt);
  };
  int rc = CLEO_OK;
  for (int attempt = 0; attempt < 3; ++attempt) {
    cur0 = c->cur;
    live0 = c->order_live;
    ord0 = c->order;
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_run_step(c, t_mdl, t_mdl + span);
    per_graph = 1;
    if (rc == CLEO_OK && (c->cur != cur0 || c->order_live != live0 || c->order != ord0)) {
      rc = enqueue_run_step(c, t_mdl + span, t_mdl + 2 * span);
      per_graph = 2;
    }
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    const int cur1 = c->cur;
    const bool live1 = c->order_live;
    uint32_t* const ord1 = c->order;
    c->cur = cur0;
    c->order_live = live0;
    set_order(ord0);
    if (rc != CLEO_OK || e != hipSuccess) {
      if (graph) hipGraphDestroy(graph);
      if (rc != CLEO_OK) return rc;
      return fail(c, CLEO_E_HIP, std::string("capture: ") + hipGetErrorString(e));
    }
    if (cur1 == cur0 && live1 == live0 && ord1 == ord0) break;
    hipGraphDestroy(graph);
    graph = nullptr;
    if (live1 && !live0) {
      pending_identity_order(c);  // the same drops in the same order, as a pending order
    } else if (!live1 && live0) {
      materialize_order(c);
    } else {
      return fail(c, CLEO_E_STATE, "the step leaves the store in a state no replay can start from");
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
  }
  if (!graph) return fail(c, CLEO_E_STATE, "no consistent step graph");
  hipGraphExec_t exec = nullptr;
  hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(graph);
    return fail(c, CLEO_E_HIP, std::string("instantiate: ") + hipGetErrorString(e));
  }
  c->cur = cur0;  // every replay starts and ends in the same state
  c->order_live = live0;
  set_order(ord0);
  const uint32_t nreplay = nsteps / per_graph;
  for (uint32_t i = 0; i < nreplay; ++i) {
    e = hipGraphLaunch(exec, c->stream);
    if (e != hipSuccess) break;
  }
  hipGraphExecDestroy(exec);
  hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(c, CLEO_E_HIP, std::string("graph launch: ") + hipGetErrorString(e));
  if (nsteps % per_graph) rc = enqueue_run_step(c, t_mdl + nreplay * per_graph * span,
                                               t_mdl + (nreplay * per_graph + 1) * span);
  return rc;
}

int cleo_synchronize(cleo_ctx* c) {
  if (!c) return CLEO_E_ARG;
  return sync_check(c);
}

int cleo_event_record(cleo_ctx* c, int32_t slot) {
  if (!c || slot < 0 || slot >= 16) return CLEO_E_ARG;
  if (!c->ev[slot]) HIP_TRY(c, hipEventCreate(&c->ev[slot]));
  HIP_TRY(c, hipEventRecord(c->ev[slot], c->stream));
  return CLEO_OK;
}

int cleo_event_elapsed(cleo_ctx* c, int32_t a, int32_t b, float* ms) {
  if (!c || !ms || a < 0 || b < 0 || a >= 16 || b >= 16 || !c->ev[a] || !c->ev[b]) return CLEO_E_ARG;
  HIP_TRY(c, hipEventSynchronize(c->ev[b]));
  HIP_TRY(c, hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
  return CLEO_OK;
}

int cleo_ndomain(cleo_ctx* c, uint64_t* nd) {
  if (!c || !nd) return CLEO_E_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  uint32_t v = 0;
  HIP_TRY(c, ctx_copy(c, &v, c->off + c->grid.ngbx, 4, hipMemcpyDeviceToHost));
  *nd = v;
  return CLEO_OK;
}

int cleo_get_error(cleo_ctx* c, cleo_error* err) {
  if (!c || !err) return CLEO_E_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, ctx_copy(c, c->h_ctl, c->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost));
  err->code = c->h_ctl->err_code;
  err->gbxindex = c->h_ctl->err_gbx;
  err->drop = c->h_ctl->err_drop;
  return CLEO_OK;
}

int cleo_clear_error(cleo_ctx* c) {
  if (!c) return CLEO_E_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, ctx_fill(c, c->ctl, 0, 16));
  return CLEO_OK;
}

int cleo_get_stats(cleo_ctx* c, cleo_stats* s) {
  if (!c || !s) return CLEO_E_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, ctx_copy(c, c->h_ctl, c->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost));
  const unsigned long long* v = c->h_ctl->stats;
  s->cond_drops = v[ST_COND_DROPS];
  s->newton_iters = v[ST_NEWTON_ITERS];
  s->substep_drops = v[ST_SUBSTEP_DROPS];
  s->warm_restarts = v[ST_WARM_RESTARTS];
  s->coll_pairs = v[ST_COLL_PAIRS];
  s->coll_events = v[ST_COLL_EVENTS];
  s->nulls = v[ST_NULLS];
  s->sorts = v[ST_SORTS];
  s->motion_drops = v[ST_MOTION_DROPS];
  return CLEO_OK;
}

int cleo_reset_stats(cleo_ctx* c) {
  if (!c) return CLEO_E_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, ctx_fill(c, c->ctl->stats, 0, sizeof(c->ctl->stats)));
  return CLEO_OK;
}

int cleo_record_niters(cleo_ctx* c, int32_t enable) {
  if (!c) return CLEO_E_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->rec_niters = enable != 0;
  if (c->rec_niters && !c->niters) {
    HIP_TRY(c, dalloc(&c->niters, c->maxn));
  }
  if (c->rec_niters) HIP_TRY(c, ctx_fill(c, c->niters, 0, c->maxn * 4));
  return CLEO_OK;
}

int cleo_download_niters(cleo_ctx* c, uint32_t* out, uint64_t n) {
  if (!c || !out) return CLEO_E_ARG;
  if (!c->niters) return fail(c, CLEO_E_STATE, "niters recording not enabled");
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, ctx_copy(c, out, c->niters, std::min<uint64_t>(n, c->maxn) * 4, hipMemcpyDeviceToHost));
  return CLEO_OK;
}

int cleo_record_guess(cleo_ctx* c, int32_t enable) {
  if (!c) return CLEO_E_ARG;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->rec_guess = enable != 0;
  if (c->rec_guess && !c->guess) HIP_TRY(c, dalloc(&c->guess, c->maxn));
  return CLEO_OK;
}

int cleo_download_guess(cleo_ctx* c, double* out, uint64_t n) {
  if (!c || !out) return CLEO_E_ARG;
  if (!c->guess) return fail(c, CLEO_E_STATE, "guess recording not enabled");
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, ctx_copy(c, out, c->guess, std::min<uint64_t>(n, c->maxn) * 8, hipMemcpyDeviceToHost));
  return CLEO_OK;
}

int cleo_thermo_eval(cleo_ctx* c, uint64_t n, const double* temp, const double* press, const double* qvap,
                     double* psat, double* s_ratio, double* ffactor) {
  if (!c || (n && (!temp || !press || !qvap || !psat || !s_ratio || !ffactor))) return CLEO_E_ARG;
  if (n == 0) return CLEO_OK;
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  double* d = nullptr;
  HIP_TRY(c, dalloc(&d, 6 * n));
  const size_t b = n * 8;
  hipError_t e = ctx_copy(c, d, temp, b, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = ctx_copy(c, d + n, press, b, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = ctx_copy(c, d + 2 * n, qvap, b, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch_thermo_eval(c, n, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = ctx_copy(c, psat, d + 3 * n, b, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = ctx_copy(c, s_ratio, d + 4 * n, b, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = ctx_copy(c, ffactor, d + 5 * n, b, hipMemcpyDeviceToHost);
  hipFree(d);
  HIP_TRY(c, e);
  return CLEO_OK;
}

int cleo_set_tiles(cleo_ctx* c, const cleo_tiles* t) {
  if (!c || !t) return CLEO_E_ARG;
  for (uint32_t v : {t->cond_waves_per_block, t->coll_waves_per_block, t->motion_waves_per_block})
    if (v > 4) return fail(c, CLEO_E_ARG, "waves per block must be 0..4");
  if (t->cond_flat > 2 || t->motion_flat > 3 || t->coll_flat > 1)
    return fail(c, CLEO_E_ARG, "variant flags: cond_flat 0..2, motion_flat 0..3, coll_flat 0 or 1");
  if (t->cond_defer == 1) return fail(c, CLEO_E_ARG, "cond_defer must be 0 or at least 2");
  c->tiles = *t;
  return CLEO_OK;
}

// Runs one stage `reps` times with tiling `t` on a copy of the resident drops in the spare buffer
// (a pending collision order is applied first; the collision order array is scratch here); State
// fields that condensation writes, the active counts and the control block are restored
// afterwards.  Reports the fastest and the mean duration, of the whole stage or (kernel_only) of
// its main kernel alone.
static int time_stage(cleo_ctx* c, int kern, const cleo_tiles& t, int reps, bool kernel_only, float* tmin,
                      float* tmean) {
  const uint32_t ngbx = c->grid.ngbx;
  materialize_order(c);
  const int cur = c->cur, spare = 1 - c->cur;
  for (int s = 0; s < 2; ++s) {
    if (!c->ev[14 + s]) HIP_TRY(c, hipEventCreate(&c->ev[14 + s]));
    if (!c->kev[s]) HIP_TRY(c, hipEventCreate(&c->kev[s]));
  }
  double* saved = nullptr;
  uint32_t* saved_nact = nullptr;
  HIP_TRY(c, dalloc(&saved, 3ull * ngbx));
  HIP_TRY(c, dalloc(&saved_nact, ngbx));
  HIP_TRY(c, hipMemcpyAsync(saved, c->st.temp, 3ull * ngbx * 8, hipMemcpyDeviceToDevice, c->stream));
  HIP_TRY(c, hipMemcpyAsync(saved_nact, c->nact, ngbx * 4ull, hipMemcpyDeviceToDevice, c->stream));
  DevCtl saved_ctl;
  HIP_TRY(c, ctx_copy(c, &saved_ctl, c->ctl, sizeof(DevCtl), hipMemcpyDeviceToHost));
  const cleo_cond_params cp = c->sdm.first_cond() ? *c->sdm.first_cond()
                                                   : cleo_cond_params{100, 1, 50, 1e-6, 1e-8, 0.001, 0, 10};
  const cleo_coll_params kp = c->sdm.first_coll()
                                  ? *c->sdm.first_coll()
                                  : cleo_coll_params{100, CLEO_PROB_LONG, 1.0, CLEO_ENACT_COALBURE, 5.0, 0.85};
  const cleo_motion_params mp = c->sdm.motion_on ? c->sdm.motion : cleo_motion_params{100, CLEO_TV_SIMMEL};
  const cleo_tiles tiles0 = c->tiles;
  c->tiles = t;
  int rc = CLEO_OK;
  float best = 1e30f, total = 0.f;
  // motion is timed with its sort, into a scratch store (the current one stays untouched)
  SoA sorted{};
  if (kern >= 2 && alloc_soa(sorted, c->ntot) != hipSuccess) {
    free_soa(sorted);
    c->tiles = tiles0;
    return fail(c, CLEO_E_HIP, "stage timing buffer");
  }
  for (int rep = 0; rep < reps && rc == CLEO_OK; ++rep) {
    const SoA& a0 = c->buf[cur];
    const SoA& b0 = c->buf[spare];
    const uint64_t n = c->ntot;
    hipError_t e = hipMemcpyAsync(b0.gbx, a0.gbx, n * 4, hipMemcpyDeviceToDevice, c->stream);
    void* dst8[7] = {b0.c3, b0.c1, b0.c2, b0.xi, b0.r, b0.msol, b0.id};
    const void* src8[7] = {a0.c3, a0.c1, a0.c2, a0.xi, a0.r, a0.msol, a0.id};
    for (int i = 0; i < 7 && e == hipSuccess; ++i)
      e = hipMemcpyAsync(dst8[i], src8[i], n * 8, hipMemcpyDeviceToDevice, c->stream);
    if (e != hipSuccess) {
      rc = fail(c, CLEO_E_HIP, "stage timing copy");
      break;
    }
    c->kev_on = kernel_only;
    hipEventRecord(c->ev[14], c->stream);
    if (kern == 0) {
      CondArgs a = cond_args(c, cp, false);
      a.s = c->buf[spare];
      a.niters = nullptr;
      if (t.cond_flat) {
        CondFlatArgs f{a.s, a.off, a.nact, a.st, a.sol, a.cfg, a.ngbx, a.alter, a.fuse_qcond, a.ctl, nullptr,
                       nullptr, a.warm_limit, c->gs, c->dm};
        launch_condensation_flat(c, f);
      } else {
        launch_condensation(c, a);
      }
    } else if (kern == 1) {
      CollArgs a{};
      a.s = c->buf[spare];
      a.order = c->order;
      a.off = c->off;
      a.nact = c->nact;
      a.vol = c->st.vol;
      a.sol = c->solute;
      a.k = make_collcfg(kp);
      a.ngbx = ngbx;
      a.wpb = t.coll_waves_per_block ? t.coll_waves_per_block : 1;
      a.ctl = c->ctl;
      a.seed = c->seed;
      a.inj_oj = c->inj_oj;
      a.inj_op = c->inj_op;
      a.big = c->big_scratch;
      launch_collisions(c, a);
    } else if (kern == 3) {  // motion, sort and condensation with the solves in the sort's gather
      MotionArgs a{};
      a.s = c->buf[spare];
      a.off = c->off;
      a.grid = c->grid;
      a.st = c->st;
      a.delt = step2dimlesstime(mp.interval);
      a.tv = mp.terminalv;
      a.wpb = t.motion_waves_per_block ? t.motion_waves_per_block : 1;
      a.cnt = c->cnt;
      a.ctl = c->ctl;
      const CondArgs ca = cond_args(c, cp, true);
      CondFlatArgs f{ca.s, ca.off, ca.nact, ca.st, ca.sol, ca.cfg, ca.ngbx, ca.alter, ca.fuse_qcond, ca.ctl, nullptr,
                     nullptr, ca.warm_limit, c->gs, c->dm};
      motion_sort_condense(c, a, f, sorted, -1);
    } else {
      MotionArgs a{};
      a.s = c->buf[spare];
      a.off = c->off;
      a.grid = c->grid;
      a.st = c->st;
      a.delt = step2dimlesstime(mp.interval);
      a.tv = mp.terminalv;
      a.wpb = t.motion_waves_per_block ? t.motion_waves_per_block : 1;
      a.cnt = c->cnt;
      a.ctl = c->ctl;
      if (t.motion_flat == 2) {
        motion_sort_fused(c, a, sorted, -1);
      } else if (t.motion_flat == 3) {
        motion_sort_gather(c, a, sorted, -1);
      } else {
        if (t.motion_flat) {
          a.step = c->dstep;
          launch_motion_flat(c, a);
        } else {
          launch_motion(c, a);
        }
        const bool on = c->kev_on;
        c->kev_on = false;  // the main kernel of the unfused stage is the motion kernel
        neighbour_sort_to(c, c->buf[spare], sorted, -1, a.step);
        c->kev_on = on;
      }
    }
    hipEventRecord(c->ev[15], c->stream);
    c->kev_on = false;
    hipEventSynchronize(c->ev[15]);
    float ms = 0.f;
    if (kernel_only)
      hipEventElapsedTime(&ms, c->kev[0], c->kev[1]);
    else
      hipEventElapsedTime(&ms, c->ev[14], c->ev[15]);
    best = std::min(best, ms);
    total += ms;
    if (hipMemcpyAsync(c->st.temp, saved, 3ull * ngbx * 8, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(c->nact, saved_nact, ngbx * 4ull, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
      rc = fail(c, CLEO_E_HIP, "stage timing restore");
  }
  c->tiles = tiles0;
  hipStreamSynchronize(c->stream);
  free_soa(sorted);
  hipFree(saved);
  hipFree(saved_nact);
  HIP_TRY(c, ctx_copy(c, c->ctl, &saved_ctl, sizeof(DevCtl), hipMemcpyHostToDevice));
  if (rc != CLEO_OK) return rc;
  if (tmin) *tmin = best;
  if (tmean) *tmean = total / static_cast<float>(reps);
  return CLEO_OK;
}

// Times each stage at 1, 2 and 4 waves (gridboxes) per workgroup, the thread-per-drop, tiled and
// task-ring variants, and grids of one wave per gridbox or of 8192 / 32768 looping waves, on the
// resident population, and keeps the fastest.  Drops and State are left unchanged.
int cleo_autotune(cleo_ctx* c, cleo_tiles* chosen) {
  if (!c) return CLEO_E_ARG;
  if (c->ntot == 0) return fail(c, CLEO_E_STATE, "no superdroplets resident");
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  struct Cand {
    uint32_t wpb, flat;
  };
  constexpr int kNC = 7;
  const Cand cands[kNC] = {{1, 0}, {2, 0}, {4, 0}, {1, 1}, {4, 1}, {1, 2}, {1, 3}};  // {waves per block, variant}
  const uint32_t grid_waves[3] = {0, 8192, 32768};
  cleo_tiles best_t = c->tiles;
  for (int kern = 0; kern < 3; ++kern) {
    float best = 1e30f;
    for (int gi = 0; gi < 3 * kNC; ++gi) {
      const Cand& cd = cands[gi % kNC];
      const uint32_t gw = grid_waves[gi / kNC];
      if (kern != 1 && cd.flat && cd.wpb != 1) continue;  // thread-per-drop kernels ignore wpb
      if (kern == 1 && cd.flat >= 2) continue;              // collisions have two variants
      if (kern == 0 && cd.flat == 3) continue;              // condensation has three
      cleo_tiles t = best_t;
      if (kern == 0) {
        t.cond_waves_per_block = cd.wpb;
        t.cond_flat = cd.flat;
        t.cond_grid_waves = gw;
      } else if (kern == 1) {
        t.coll_waves_per_block = cd.wpb;
        t.coll_flat = cd.flat;
        t.coll_grid_waves = gw;
      } else {
        t.motion_waves_per_block = cd.wpb;
        t.motion_flat = cd.flat;
        t.motion_grid_waves = gw;
      }
      float tmin = 0.f;
      const int rc = time_stage(c, kern, t, 2, false, &tmin, nullptr);
      if (rc != CLEO_OK) return rc;
      if (tmin < best) {
        best = tmin;
        best_t = t;
      }
    }
  }
  // thread-per-drop condensation: the iteration budget after which slow drops are set aside
  if (best_t.cond_flat == 1) {
    float best = 1e30f;
    const cleo_tiles base_t = best_t;
    for (uint32_t budget : {0u, 6u, 8u, 12u}) {
      cleo_tiles t = base_t;
      t.cond_defer = budget;
      float tmin = 0.f;
      const int rc = time_stage(c, 0, t, 2, false, &tmin, nullptr);
      if (rc != CLEO_OK) return rc;
      if (tmin < best) {
        best = tmin;
        best_t = t;
      }
    }
  }
  // thread-per-drop motion and condensation: a condensation that follows motion runs inside the
  // sort's gather, so the motion variant (drops moved before the sort, or inside the gather), the
  // gather's grid and the deferral budget are chosen on that fused pass
  if (best_t.cond_flat == 1 && (best_t.motion_flat == 1 || best_t.motion_flat == 3)) {
    float best = 1e30f;
    const cleo_tiles base_t = best_t;
    for (uint32_t mv : {1u, 3u}) {
      for (uint32_t gw : {32768u, 65536u, 131072u}) {
        for (uint32_t budget : {6u, 8u, 12u}) {
          cleo_tiles t = base_t;
          t.motion_flat = mv;
          t.cond_grid_waves = gw;
          t.cond_defer = budget;
          float tmin = 0.f;
          const int rc = time_stage(c, 3, t, 4, false, &tmin, nullptr);  // 4 reps: with 2 the pick varied between runs
          if (rc != CLEO_OK) return rc;
          if (tmin < best) {
            best = tmin;
            best_t = t;
          }
        }
      }
    }
  }
  c->tiles = best_t;
  if (chosen) *chosen = c->tiles;
  return CLEO_OK;
}

int cleo_time_kernel(cleo_ctx* c, int32_t stage, uint32_t reps, float* mean_ms) {
  if (!c || !mean_ms || stage < 0 || stage > 4 || reps == 0) return CLEO_E_ARG;
  if (c->ntot == 0) return fail(c, CLEO_E_STATE, "no superdroplets resident");
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (stage == 4) {  // the mass-moment reduction (reads only: the drops stay as they are)
    const uint64_t ngbx = c->grid.ngbx;
    if (!c->mom) {
      uint8_t* p = nullptr;
      HIP_TRY(c, dalloc(&p, 16 * ngbx));
      c->mom = p;
    }
    uint64_t* d0 = static_cast<uint64_t*>(c->mom);
    float* d1 = reinterpret_cast<float*>(d0 + ngbx);
    materialize_order(c);
    for (int s = 0; s < 2; ++s)
      if (!c->kev[s]) HIP_TRY(c, hipEventCreate(&c->kev[s]));
    launch_mass_moments(c, false, d0, d1, d1 + ngbx);  // warm-up
    HIP_TRY(c, hipEventRecord(c->kev[0], c->stream));
    for (uint32_t r = 0; r < reps; ++r) launch_mass_moments(c, false, d0, d1, d1 + ngbx);
    HIP_TRY(c, hipEventRecord(c->kev[1], c->stream));
    HIP_TRY(c, hipEventSynchronize(c->kev[1]));
    float ms = 0.f;
    HIP_TRY(c, hipEventElapsedTime(&ms, c->kev[0], c->kev[1]));
    *mean_ms = ms / static_cast<float>(reps);
    return sync_check(c);
  }
  return time_stage(c, stage, c->tiles, static_cast<int>(reps), true, nullptr, mean_ms);
}

int cleo_get_tiles(cleo_ctx* c, cleo_tiles* t) {
  if (!c || !t) return CLEO_E_ARG;
  *t = c->tiles;
  return CLEO_OK;
}

}  // extern "C"
