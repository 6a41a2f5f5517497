// sdm_kernels.hip -- gfx950 kernels of the super-droplet step.
//
// Launch shape: one wave64 per gridbox, `wpb` waves per workgroup; with a capped grid each wave
// loops over gridboxes g, g + (waves in the grid), ...  Drops of a gridbox are contiguous in the
// SoA store (sorted by sdgbxindex), so every per-drop load and store of a wave is a coalesced run.  No kernel uses atomics on drop data: sums keep the
// reference's sequential order and the sort is stable by construction.
#include <algorithm>

#include "kernels.hpp"
#include "sdm_kernels.hpp"

namespace cleo {

// ============================ per-gridbox sums ================================================
// Drop-order sums of v(k) over [lo, lo + len) for the gridboxes of a tile of 64 (lane i passes its
// own gridbox's range and gets its sum).  The tile's drops [t_lo, t_hi) are staged through an LDS
// window of kSumWin values with coalesced loads (the next window's loads in flight while the current
// one is added); each lane then adds its gridbox's values left to right, exactly as a serial loop
// does (the reference's sums run in drop order under Serial).
template <class V>
__device__ __forceinline__ double tile_sum(double* win, uint32_t t_lo, uint32_t t_hi, uint32_t lo, uint32_t len, V v,
                                           int lane) {
  constexpr uint32_t kPer = kSumWin / kWave;
  double acc = 0.0;
  if (t_lo >= t_hi) return acc;
  // the next window's values are loaded into registers while the lanes add the current one's
  // (every lane loads, with indices clamped into the tile, so the loads stay in flight)
  double r[kPer];
  auto load = [&](uint32_t w) {
    for (uint32_t i = 0; i < kPer; ++i) r[i] = v(min(w + static_cast<uint32_t>(lane) + i * kWave, t_hi - 1u));
  };
  load(t_lo);
  for (uint32_t w = t_lo; w < t_hi; w += kSumWin) {
    const uint32_t we = min(w + kSumWin, t_hi);
    for (uint32_t i = 0; i < kPer; ++i) win[static_cast<uint32_t>(lane) + i * kWave] = r[i];
    wave_sync();
    if (we < t_hi) load(we);
    const uint32_t a0 = max(lo, w), a1 = min(lo + len, we);
    for (uint32_t k = a0; k < a1; ++k) acc += win[k - w];
    wave_sync();
  }
  return acc;
}
// The same for two value sequences at once (each summed left to right on its own): half the window
// each, one pass over the tile.
template <class V1, class V2>
__device__ __forceinline__ void tile_sum2(double* win, uint32_t t_lo, uint32_t t_hi, uint32_t lo, uint32_t len, V1 v1,
                                          V2 v2, int lane, double& s1, double& s2) {
  constexpr uint32_t kHalf = kSumWin / 2, kPer = kHalf / kWave;
  double a1 = 0.0, a2 = 0.0;
  if (t_lo >= t_hi) {
    s1 = s2 = 0.0;
    return;
  }
  // the next window's values are loaded into registers while the lanes add the current one's
  // (every lane loads, with indices clamped into the tile, so the loads stay in flight)
  double r1[kPer], r2[kPer];
  auto load = [&](uint32_t w) {
    for (uint32_t i = 0; i < kPer; ++i) {
      const uint32_t k = min(w + static_cast<uint32_t>(lane) + i * kWave, t_hi - 1u);
      r1[i] = v1(k);
      r2[i] = v2(k);
    }
  };
  load(t_lo);
  for (uint32_t w = t_lo; w < t_hi; w += kHalf) {
    const uint32_t we = min(w + kHalf, t_hi);
    for (uint32_t i = 0; i < kPer; ++i) {
      const uint32_t j = static_cast<uint32_t>(lane) + i * kWave;
      win[j] = r1[i];
      win[kHalf + j] = r2[i];
    }
    wave_sync();
    if (we < t_hi) load(we);
    const uint32_t b0 = max(lo, w), b1 = min(lo + len, we);
    for (uint32_t k = b0; k < b1; ++k) {
      a1 += win[k - w];
      a2 += win[kHalf + k - w];
    }
    wave_sync();
  }
  s1 = a1;
  s2 = a2;
}
// condensed water of drop k (condensate_mass * xi, superdrop_attrs.hpp:207-216)
__device__ __forceinline__ double qcond_term(const SoA& s, const Solute& sol, DevCtl* ctl, uint32_t k) {
  bool bad = false;
  const double m = condensate_mass(s.r[k], s.msol[k], sol, &bad) * static_cast<double>(s.xi[k]);
  if (bad) set_error(ctl, CLEO_ERR_NEG_CONDENSATE, s.gbx[k], k);
  return m;
}

// ============================ condensation ====================================================
// DoCondensation::do_condensation for one gridbox per wave (condensation.hpp:141-150,
// condensation.cpp:44-137).  With `fuse_qcond`, the gridbox qcond of
// EffectOnHydrometeorStatesFunctor (movesupersindomain.hpp:95-136) is recomputed first, as
// happens between motion and condensation in the reference step.
template <bool kFuseQcond>
__device__ __forceinline__ void cond_gbx(const CondArgs& a, uint32_t g, int lane, uint32_t (&acc)[4]) {
  const uint32_t base = a.off[g];
  const uint32_t nfull = a.off[g + 1] - base;
  const uint32_t n = a.nact[g];
  const double temp = a.st.temp[g];
  const double press = a.st.press[g];
  double qv = a.st.qv[g];
  double qc = a.st.qc[g];
  const double volume = a.st.vol[g];
  bool bad_cm = false;

  if (kFuseQcond) {
    double tot = 0.0;
    for (uint32_t c = 0; c < nfull; c += kWave) {
      const uint32_t k = base + c + lane;
      double v = 0.0;
      if (c + lane < nfull) {
        v = condensate_mass(a.s.r[k], a.s.msol[k], a.sol, &bad_cm) * static_cast<double>(a.s.xi[k]);
      }
      const int m = static_cast<int>(min(static_cast<uint32_t>(kWave), nfull - c));
      tot = ordered_add(tot, v, m);
    }
    const double rho_cond = tot / volume * dimless::R0cubed_VOL0;
    const double rho_dry = dry_air_density(press, temp, qv);
    qc = rho_cond / rho_dry;
  }
  if (temp <= 0 && lane == 0) set_error(a.ctl, CLEO_ERR_TEMP_NONPOSITIVE, g, base);
  const double s_ratio = a.gs[kGsStride * g], ffactor = a.gs[kGsStride * g + 1];  // from k_gbx_thermo
  const double akoh_g = a.gs[kGsStride * g + 2], inv_akoh_g = a.gs[kGsStride * g + 3];

  double tot = 0.0;
  uint32_t iters = 0, subs = 0, restarts = 0;
  for (uint32_t c = 0; c < n; c += kWave) {
    const uint32_t k = base + c + lane;
    double dm = 0.0;
    if (c + lane < n) {
      const double r = a.s.r[k];
      const double msol = a.s.msol[k];
      const uint64_t xi = a.s.xi[k];
      const double old_m = condensate_mass(r, msol, a.sol, &bad_cm);
      const double akoh = akoh_g, bkoh = kohler_b(msol, a.sol);
      const double warm = a.warm ? a.warm[k] : __longlong_as_double(0x7FF8000000000000ll);
      SolveOut o{0, false, false, false, false};
      const double newr = solve_condensation(a.cfg, s_ratio, akoh, bkoh, inv_akoh_g, ffactor, r, warm, a.warm_limit, o);
      const double rnew = at_least_dry(newr, msol, a.sol);
      a.s.r[k] = rnew;
      dm = (condensate_mass(rnew, msol, a.sol, &bad_cm) - old_m) * static_cast<double>(xi);
      if (o.maxed) set_error(a.ctl, CLEO_ERR_NEWTON_MAXITERS, g, k);
      if (a.niters) a.niters[k] = o.iters;
      iters += o.iters;
      subs += o.substep ? 1u : 0u;
      restarts += o.restarted ? 1u : 0u;
    }
    const int m = static_cast<int>(min(static_cast<uint32_t>(kWave), n - c));
    tot = ordered_add(tot, dm, m);
  }
  if (bad_cm) set_error(a.ctl, CLEO_ERR_NEG_CONDENSATE, g, base);

  // EffectOnThermodynamicStateFunctor::state_change (condensation.cpp:126-137)
  double tnew = temp;
  if (a.alter) {
    const double totrho = tot / volume * dimless::R0cubed_VOL0;
    const double rho_dry = dry_air_density(press, temp, qv);
    const double dq = totrho / rho_dry;
    qv -= dq;
    qc += dq;
    const double dtemp = (dimless::Latent_v / moist_specific_heat(qv, qc)) * dq;
    tnew += dtemp;
  }
  iters = wave_sum_u32(iters);
  subs = wave_sum_u32(subs);
  restarts = wave_sum_u32(restarts);
  if (lane == 0) {
    if (a.alter) {
      a.st.temp[g] = tnew;
      a.st.qv[g] = qv;
    }
    if (a.alter || kFuseQcond) a.st.qc[g] = qc;
  }
  acc[0] += n;
  acc[1] += iters;
  acc[2] += subs;
  acc[3] += restarts;
}
template <bool kFuseQcond>
__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(3) k_condensation(CondArgs a) {
  const int lane = lane_id();
  uint32_t acc[4] = {0, 0, 0, 0};
  for (uint32_t g = wave_index(); g < a.ngbx; g += wave_count()) cond_gbx<kFuseQcond>(a, g, lane, acc);
  if (lane == 0)
    for (int j = 0; j < 4; ++j) put_stat(a.stat, j, wave_index(), acc[j]);
}

// EffectOnHydrometeorStatesFunctor (movesupersindomain.hpp:95-136), standalone: 64 gridboxes per
// wave, drop-order sums of condensed mass, then one lane per gridbox for its qcond
__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(4) k_qcond(QcondArgs a) {
  __shared__ double win_all[4][kSumWin];
  double* win = win_all[threadIdx.x >> 6];
  const int lane = lane_id();
  const uint32_t ntiles = (a.ngbx + kWave - 1) / kWave;
  for (uint32_t tile = wave_index(); tile < ntiles; tile += wave_count()) {
    const uint32_t g0 = tile * kWave, g = g0 + static_cast<uint32_t>(lane);
    const bool valid = g < a.ngbx;
    const uint32_t t_lo = a.off[g0], t_hi = a.off[min(g0 + kWave, a.ngbx)];
    const uint32_t lo = valid ? a.off[g] : t_hi, hi = valid ? a.off[g + 1] : t_hi;
    const double tot = tile_sum(win, t_lo, t_hi, lo, hi - lo, [&](uint32_t k) { return qcond_term(a.s, a.sol, a.ctl, k); },
                                lane);
    if (valid) {
      const double rho_cond = tot / a.st.vol[g] * dimless::R0cubed_VOL0;
      const double rho_dry = dry_air_density(a.st.press[g], a.st.temp[g], a.st.qv[g]);
      a.st.qc[g] = rho_cond / rho_dry;
    }
  }
}

// ============================ collisions ======================================================
// DoCollisions::do_collisions (collisions.hpp:140-278): Fisher-Yates shuffle of a gridbox's drops,
// pairs (2jj, 2jj+1) of the shuffled order, enactment, nulls moved behind the active drops
// (stable partition).  The drops stay where they are: the shuffled order is recorded in
// CollArgs::order (slot -> position) and only the attributes of enacted pairs are written, in
// place.  The reference moves every drop to its shuffled place; here the next sort moves each drop
// once, to its sorted place in that order.
// word t%4 of a block (selected without indexing the array by a lane-varying value)
__device__ __forceinline__ uint32_t rng_word(const uint32_t r4[4], uint32_t t) {
  const uint32_t w = t & 3u;
  return w == 0u ? r4[0] : (w == 1u ? r4[1] : (w == 2u ? r4[2] : r4[3]));
}
// uniform u32 -> index in [0, iter]
__device__ __forceinline__ uint32_t fy_scale(uint32_t u, uint32_t iter) {
  return static_cast<uint32_t>((static_cast<uint64_t>(u) * (iter + 1ull)) >> 32);
}
// j_t of Fisher-Yates swap t < n-1 of gridbox g: uniform in [0, iter], iter = n-1-t
// (shuffle.hpp:80-87), from the counter stream (word t%4 of the block for item t/4, so one block
// serves four consecutive swaps) or the injected indices
__device__ __forceinline__ uint32_t fy_index(const CollArgs& a, uint32_t g, uint32_t base, uint32_t n, uint32_t t,
                                             uint64_t event) {
  const uint32_t iter = n - 1 - t;
  uint32_t j;
  if (a.inj_j) {
    j = a.inj_j[a.inj_oj[g] + t];
    if (j > iter) {  // an injected index outside [0, iter] is an argument error
      set_error(a.ctl, CLEO_ERR_INJECTED_INDEX, g, base + iter);
      j = iter;
    }
  } else {
    uint32_t r4[4];
    rng_block(a.seed, event, 0u, g, t >> 2, r4);
    j = fy_scale(rng_word(r4, t), iter);
  }
  return j;
}

// position of the drop at slot `slot` of gridbox `base`'s logical order: through a pending order
// of an earlier event when there is one (the drops then stay where they are and the orders
// compose), else the slot itself
__device__ __forceinline__ uint32_t coll_pos(const CollArgs& a, uint32_t base, uint32_t slot) {
  return a.order_in ? a.order_in[base + slot] : base + slot;
}

// One collision pair (2jj, 2jj+1) of gridbox g (n active drops) from the drops at ka, kb
// (collisions.hpp:180-230): probability and enactment on (xi, r, msol) of both drops.
struct PairOut {
  DropAttrs A, B;  // the drops at ka and kb after the enactment
  bool nullA, nullB;
  bool enacted;
  int err;
};
template <bool Kinetic>
__device__ __forceinline__ PairOut coll_pair_eval(const CollArgs& a, uint32_t g, uint32_t n, uint32_t jj, uint32_t ka,
                                                  uint32_t kb, uint64_t event) {
  const uint32_t npairs = n / 2;
  const double scale_p = n * (n - 1.0) / (2.0 * npairs);
  const double VOLUME = a.vol[g] * dimless::VOL0;
  PairOut o;
  const DropAttrs A{a.s.xi[ka], a.s.r[ka], a.s.msol[ka]};
  const DropAttrs B{a.s.xi[kb], a.s.r[kb], a.s.msol[kb]};
  const bool a_first = !(A.xi < B.xi);  // assign_drops (collisions.hpp:94-101)
  // d1 is the drop with the larger multiplicity; selected field by field (a reference picked by a
  // lane-varying flag would put both drops in scratch memory)
  DropAttrs d1{a_first ? A.xi : B.xi, a_first ? A.r : B.r, a_first ? A.msol : B.msol};
  DropAttrs d2{a_first ? B.xi : A.xi, a_first ? B.r : A.r, a_first ? B.msol : A.msol};
  const double prob_jk = pair_probability<Kinetic>(a.k, d1.r, d2.r, VOLUME);
  const double prob = scale_p * static_cast<double>(d1.xi) * prob_jk;
  double phi_coll, phi_out;
  if (a.inj_phi) {
    const uint64_t q = a.inj_op[g] + 2ull * jj;
    phi_coll = a.inj_phi[q];
    phi_out = a.inj_phi[q + 1];
  } else {
    uint32_t r4[4];
    rng_block(a.seed, event, 1u, g, jj, r4);
    phi_coll = u53(r4[0], r4[1]);
    phi_out = u53(r4[2], r4[3]);
  }
  o.err = 0;
  const uint32_t res = enact_pair<Kinetic>(a.k, d1, d2, prob, phi_coll, phi_out, a.sol, o.err);
  o.enacted = (res >> 1) & 1u;
  o.nullA = (res & 1u) && a_first;
  o.nullB = (res & 1u) && !a_first;
  o.A = DropAttrs{a_first ? d1.xi : d2.xi, a_first ? d1.r : d2.r, a_first ? d1.msol : d2.msol};
  o.B = DropAttrs{a_first ? d2.xi : d1.xi, a_first ? d2.r : d1.r, a_first ? d2.msol : d1.msol};
  return o;
}
// the attributes of drop k after an enacted pair, in place; a null drop gets the out-of-domain key
__device__ __forceinline__ void coll_store(const CollArgs& a, uint32_t k, const DropAttrs& D, bool isnull) {
  a.s.xi[k] = D.xi;
  a.s.r[k] = D.r;
  a.s.msol[k] = D.msol;
  if (isnull) a.s.gbx[k] = kOOB;
}
// pair jj of the drops at ka, kb: the slots 2jj, 2jj+1 of the logical order, and the new attributes
// when it was enacted; true when one of the two became null; `events` counts enactments
template <bool Kinetic>
__device__ __forceinline__ bool coll_pair(const CollArgs& a, uint32_t g, uint32_t base, uint32_t n, uint32_t jj,
                                          uint32_t ka, uint32_t kb, uint64_t event, uint32_t& events) {
  const PairOut o = coll_pair_eval<Kinetic>(a, g, n, jj, ka, kb, event);
  if (o.err) set_error(a.ctl, o.err, g, base + 2 * jj);
  a.order[base + 2 * jj] = ka;
  a.order[base + 2 * jj + 1] = kb;
  if (o.enacted) {
    events += 1u;
    coll_store(a, ka, o.A, o.nullA);
    coll_store(a, kb, o.B, o.nullB);
  }
  return o.nullA || o.nullB;
}

// remove_null_supers (collisions.hpp:240-248) for gridbox g after its pairs: the reference's
// sort_team is restated as a stable partition of the gridbox's logical order (active drops first,
// nulls behind, each in order).  The partitioned order is staged in the gridbox's HBM scratch
// (2 u32 per drop from big + 2 base, free once the shuffle is done) and copied back.  Only this
// wave reads what it stored, so a workgroup-scope fence orders it (an agent-scope fence would
// write back the XCD's whole L2).
__device__ __forceinline__ void coll_order_nulls(const CollArgs& a, uint32_t g, uint32_t base, uint32_t n,
                                                 uint32_t nulls, int lane) {
  __threadfence_block();
  wave_sync();
  uint32_t* tmp = a.big + 2ull * base;
  const uint32_t nn = n - nulls;
  uint32_t run_nn = 0, run_null = 0;
  for (uint32_t c = 0; c < n; c += kWave) {
    const uint32_t k = c + lane;
    const bool act = k < n;
    const uint32_t pk = act ? a.order[base + k] : 0u;
    const bool isnull = act && a.s.xi[pk] == 0;
    const uint64_t mnull = __ballot(isnull);
    const uint64_t mnn = __ballot(act && !isnull);
    if (act) {
      const uint32_t pos = isnull ? nn + run_null + __popcll(mnull & lanemask_lt())
                                  : run_nn + __popcll(mnn & lanemask_lt());
      tmp[pos] = pk;
    }
    run_null += __popcll(mnull);
    run_nn += __popcll(mnn);
  }
  __threadfence_block();
  wave_sync();
  for (uint32_t k = lane; k < n; k += kWave) a.order[base + k] = tmp[k];
  if (lane == 0) {
    a.nact[g] = nn;
    a.ctl->resort_flag = 1u;
  }
}

// The permutation `perm` and the drawn indices `jb` live in LDS (kLds) or, for gridboxes above
// kCollLdsCap drops, in HBM scratch; the two instantiations keep LDS accesses as ds_* instructions.
template <bool kLds, bool Kinetic>
__device__ __forceinline__ void coll_gbx(const CollArgs& a, int lane, uint32_t g, uint32_t base, uint32_t nfull,
                                         uint32_t n, uint32_t* perm, uint32_t* jb, uint32_t (&acc)[3]) {
  wave_sync();  // the previous gridbox of this wave is done with perm / jb
  const uint64_t event = a.ctl->event;
  auto draw = [&](uint32_t t) -> uint32_t { return fy_index(a, g, base, n, t, event); };
  if (kLds && n <= 2 * kWave) {
    // Fisher-Yates in registers: element e of the permutation is lane e % 64 of P[e / 64]; the
    // n-1 swaps are wave-uniform readlane / writelane pairs instead of a dependent LDS chain.
    uint32_t P0 = static_cast<uint32_t>(lane), P1 = static_cast<uint32_t>(lane) + kWave;
    const uint32_t J0 = (n >= 2 && static_cast<uint32_t>(lane) < n - 1) ? draw(lane) : 0u;
    const uint32_t J1 = (n >= 2 && static_cast<uint32_t>(lane) + kWave < n - 1) ? draw(lane + kWave) : 0u;
    for (uint32_t t = 0; t + 1 < n; ++t) {
      const uint32_t iter = n - 1 - t;
      const uint32_t j = t < kWave ? readlane_u32(J0, t) : readlane_u32(J1, t - kWave);
      const uint32_t vi = iter < kWave ? readlane_u32(P0, iter) : readlane_u32(P1, iter - kWave);
      const uint32_t vj = j < kWave ? readlane_u32(P0, j) : readlane_u32(P1, j - kWave);
      if (iter < kWave)
        P0 = writelane_u32(P0, vj, iter);
      else
        P1 = writelane_u32(P1, vj, iter - kWave);
      if (j < kWave)
        P0 = writelane_u32(P0, vi, j);
      else
        P1 = writelane_u32(P1, vi, j - kWave);
    }
    perm[lane] = P0;
    perm[lane + kWave] = P1;
    wave_sync();
  } else {
    for (uint32_t k = lane; k < n; k += kWave) perm[k] = k;
    if (n >= 2) {
      for (uint32_t t = lane; t < n - 1; t += kWave) jb[t] = draw(t);
      if (!kLds) __threadfence_block();
      wave_sync();
      if (lane == 0) {  // fisher_yates_shuffle (shuffle.hpp:80-87)
        for (uint32_t t = 0; t + 1 < n; ++t) {
          const uint32_t iter = n - 1 - t;
          const uint32_t j = jb[t];
          const uint32_t pi = perm[iter];
          perm[iter] = perm[j];
          perm[j] = pi;
        }
      }
    }
    if (!kLds) __threadfence_block();
    wave_sync();
  }

  const uint32_t npairs = n / 2;
  uint32_t nulls = 0, events = 0;
  for (uint32_t c = 0; c < npairs; c += kWave) {
    const uint32_t jj = c + lane;
    bool nul = false;
    if (jj < npairs)
      nul = coll_pair<Kinetic>(a, g, base, n, jj, coll_pos(a, base, perm[2 * jj]), coll_pos(a, base, perm[2 * jj + 1]),
                               event, events);
    nulls += static_cast<uint32_t>(__popcll(__ballot(nul)));
  }
  // unpaired last drop and the inactive tail (nulls of earlier sub-steps) keep their places
  for (uint32_t k = 2 * npairs + lane; k < nfull; k += kWave) a.order[base + k] = coll_pos(a, base, k < n ? perm[k] : k);
  if (nulls > 0) coll_order_nulls(a, g, base, n, nulls, lane);
  acc[0] += npairs;
  acc[1] += wave_sum_u32(events);
  acc[2] += nulls;
}

template <bool Kinetic>
__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(3) k_collisions(CollArgs a) {
  extern __shared__ uint32_t lds_dyn[];
  const int lane = lane_id();
  const int w = static_cast<int>(threadIdx.x >> 6);
  uint32_t acc[3] = {0, 0, 0};
  for (uint32_t g = wave_index(); g < a.ngbx; g += wave_count()) {
    const uint32_t base = a.off[g];
    const uint32_t nfull = a.off[g + 1] - base;
    const uint32_t n = a.nact[g];
    if (n <= kCollLdsCap) {
      uint32_t* perm = lds_dyn + static_cast<uint32_t>(w) * 2u * kCollLdsCap;
      coll_gbx<true, Kinetic>(a, lane, g, base, nfull, n, perm, perm + kCollLdsCap, acc);
    } else {
      uint32_t* perm = a.big + 2ull * base;
      coll_gbx<false, Kinetic>(a, lane, g, base, nfull, n, perm, perm + n, acc);
    }
  }
  if (lane == 0)
    for (int j = 0; j < 3; ++j) put_stat(a.stat, j, wave_index(), acc[j]);
}

// ---- tiled collisions: 64 gridboxes per wave ---------------------------------------------------
// Lane i shuffles gridbox 64*tile + i: its permutation (one byte per drop) lives in the wave's LDS
// tile and the swaps of all 64 gridboxes run side by side, with the same index sequence as the
// per-gridbox path.  The tile's pairs are then flattened over the lanes (a lane finds its gridbox
// by binary search over the prefix of per-gridbox pair counts), so every lane evaluates a pair
// whatever the gridbox sizes.  Gridboxes above kTileMaxN drops, or beyond the tile's LDS, take the
// per-gridbox path through HBM scratch.  The constant-kinetics instantiation runs at four waves per
// SIMD (20 registers spilled, their scratch accesses mostly outside the tile loop: 1.18 -> 1.07 ms on
// the coupled workload); the kinetic-energy one, which spills more, at three.
struct CollTileLds {
  uint32_t perm32[kTilePermBytes / 4];  // byte permutations, gridbox i from byte loff[i]
  uint32_t ppre[kWave + 1];             // first pair of gridbox i among the tile's pairs
  uint32_t base[kWave];                 // first drop of gridbox i
  uint32_t loff[kWave];                 // kBigGbx: the gridbox takes the per-gridbox path
  uint32_t n[kWave];
  uint32_t nulls[kWave];
};
static_assert(sizeof(CollTileLds) == kCollTileLdsBytes, "tile LDS size");

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, int lane, uint32_t& total) {
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  total = readlane_u32(x, kWave - 1);
  return x - v;
}
// the i with pre[i] <= f < pre[i+1], given pre[0] <= f < pre[kWave]
__device__ __forceinline__ uint32_t tile_owner(const uint32_t* pre, uint32_t f) {
  uint32_t lo = 0, hi = kWave;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pre[mid] <= f)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <bool Kinetic>
__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(Kinetic ? 3 : 4) k_coll_tiles(CollArgs a) {
  extern __shared__ uint32_t lds_dyn[];
  const int lane = lane_id();
  CollTileLds& L = reinterpret_cast<CollTileLds*>(lds_dyn)[threadIdx.x >> 6];
  uint8_t* pb = reinterpret_cast<uint8_t*>(L.perm32);
  const uint64_t event = a.ctl->event;
  uint32_t acc[3] = {0, 0, 0};
  uint32_t pairs = 0, events = 0, nulls = 0;
  const uint32_t ntiles = (a.ngbx + kWave - 1) / kWave;
  for (uint32_t tile = wave_index(); tile < ntiles; tile += wave_count()) {
    wave_sync();  // the previous tile is done with L
    const uint32_t g = tile * kWave + static_cast<uint32_t>(lane);
    const bool valid = g < a.ngbx;
    uint32_t base = 0, nfull = 0, n = 0;
    if (valid) {
      base = a.off[g];
      nfull = a.off[g + 1] - base;
      n = a.nact[g];
    }
    const bool fits = valid && n <= kTileMaxN;
    uint32_t ltot;
    const uint32_t loff = wave_excl_scan(fits ? (n + 3u) & ~3u : 0u, lane, ltot);
    const bool small = fits && loff + n <= kTilePermBytes;
    // fisher_yates_shuffle (shuffle.hpp:80-87), one gridbox per lane
    if (small) {
      for (uint32_t k = 0; k < n; k += 4) L.perm32[(loff + k) >> 2] = k | ((k + 1) << 8) | ((k + 2) << 16) | ((k + 3) << 24);
      uint8_t* p = pb + loff;
      uint32_t r4[4] = {0u, 0u, 0u, 0u};
      for (uint32_t t = 0; t + 1 < n; ++t) {
        const uint32_t iter = n - 1 - t;
        uint32_t j;
        if (a.inj_j) {
          j = fy_index(a, g, base, n, t, event);
        } else {
          // t is the same for every lane still shuffling, so the block is drawn once per 4 swaps
          if ((t & 3u) == 0u) rng_block(a.seed, event, 0u, g, t >> 2, r4);
          j = fy_scale(rng_word(r4, t), iter);
        }
        const uint8_t pi = p[iter];
        p[iter] = p[j];
        p[j] = pi;
      }
    }
    const uint32_t np = small ? n / 2 : 0u;
    uint32_t ptot;
    const uint32_t pfirst = wave_excl_scan(np, lane, ptot);
    L.ppre[lane] = pfirst;
    if (lane == kWave - 1) L.ppre[kWave] = ptot;
    L.base[lane] = base;
    L.loff[lane] = small ? loff : kBigGbx;
    L.n[lane] = n;
    L.nulls[lane] = 0u;
    wave_sync();
    // the tile's pairs flattened over the lanes (pair f belongs to the gridbox i with
    // ppre[i] <= f < ppre[i+1]): each pair is evaluated once, by one lane, which records its two
    // slots of the logical order and stores the drops' new attributes when it was enacted
    for (uint32_t f = lane; f < ptot; f += kWave) {
      const uint32_t i = tile_owner(L.ppre, f);
      const uint32_t jj = f - L.ppre[i], bi = L.base[i];
      const uint8_t* pi = pb + L.loff[i];
      if (coll_pair<Kinetic>(a, tile * kWave + i, bi, L.n[i], jj, coll_pos(a, bi, pi[2 * jj]),
                             coll_pos(a, bi, pi[2 * jj + 1]), event, events))
        atomicAdd(&L.nulls[i], 1u);
    }
    // the unpaired last drop and the inactive tail of the lane's own gridbox keep their places
    if (small) {
      const uint8_t* p = pb + loff;
      for (uint32_t k = 2u * np; k < nfull; ++k)
        a.order[base + k] = coll_pos(a, base, k < n ? static_cast<uint32_t>(p[k]) : k);
    }
    wave_sync();
    const uint32_t mynulls = L.nulls[lane];
    pairs += np;
    nulls += mynulls;
    uint64_t m = __ballot(mynulls > 0u);
    while (m) {
      const int i = __ffsll(static_cast<unsigned long long>(m)) - 1;
      m &= m - 1;
      coll_order_nulls(a, tile * kWave + static_cast<uint32_t>(i), readlane_u32(base, i), readlane_u32(n, i),
                       readlane_u32(mynulls, i), lane);
    }
    uint64_t mb = __ballot(valid && !small);
    while (mb) {
      const int i = __ffsll(static_cast<unsigned long long>(mb)) - 1;
      mb &= mb - 1;
      const uint32_t bb = readlane_u32(base, i), nb = readlane_u32(n, i);
      coll_gbx<false, Kinetic>(a, lane, tile * kWave + static_cast<uint32_t>(i), bb, readlane_u32(nfull, i), nb,
                      a.big + 2ull * bb, a.big + 2ull * bb + nb, acc);
    }
  }
  acc[0] += wave_sum_u32(pairs);
  acc[1] += wave_sum_u32(events);
  acc[2] += wave_sum_u32(nulls);
  if (lane == 0)
    for (int j = 0; j < 3; ++j) put_stat(a.stat, j, wave_index(), acc[j]);
}

// per-gridbox sizes of the injected stream: n-1 indices and 2*floor(n/2) phis
__global__ void k_inj_sizes(const uint32_t* nact, uint32_t ngbx, uint32_t* nj, uint32_t* np) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngbx) return;
  const uint32_t n = nact[g];
  nj[g] = n >= 2 ? n - 1 : 0;
  np[g] = 2 * (n / 2);
}

__global__ void k_advance_event(DevCtl* ctl) { ctl->event += 1; }

// DevCtl::stats[first + blockIdx.y] += sum of the nslot slots of counter blockIdx.y; each block
// sums a strided share and adds it with one atomic (a few hundred atomics per launch).
__global__ void __launch_bounds__(256) k_stats_reduce(const uint32_t* sp, uint32_t nslot, int first, DevCtl* ctl) {
  __shared__ unsigned long long part[4];
  const uint32_t* col = sp + static_cast<uint64_t>(blockIdx.y) * nslot;
  unsigned long long v = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nslot; i += gridDim.x * blockDim.x) v += col[i];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (lane_id() == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = part[0] + part[1] + part[2] + part[3];
    if (t) atomicAdd(&ctl->stats[first + static_cast<int>(blockIdx.y)], t);
  }
}

// ============================ motion ==========================================================
// Source gridbox that reaches `d` by the neighbour step `s` (s = (dz+1) + 3(dx+1) + 9(dy+1)),
// or kOOB when the step would come from outside a finite axis.
__device__ __forceinline__ uint32_t source_of(const Grid& gr, uint32_t d, int s) {
  const int dz = s % 3 - 1, dx = (s / 3) % 3 - 1, dy = s / 9 - 1;
  int jz = static_cast<int>(d % gr.nz) - dz;
  int jx = static_cast<int>((d / gr.nz) % gr.nx) - dx;
  int jy = static_cast<int>(d / (gr.nz * gr.nx)) - dy;
  const int nz = static_cast<int>(gr.nz), nx = static_cast<int>(gr.nx), ny = static_cast<int>(gr.ny);
  if (jz < 0 || jz >= nz) {
    if (!gr.per[0]) return kOOB;
    jz = (jz + nz) % nz;
  }
  if (jx < 0 || jx >= nx) {
    if (!gr.per[1]) return kOOB;
    jx = (jx + nx) % nx;
  }
  if (jy < 0 || jy >= ny) {
    if (!gr.per[2]) return kOOB;
    jy = (jy + ny) % ny;
  }
  return static_cast<uint32_t>(jz + nz * (jx + nx * jy));
}

// gridbox reached from g by the step s (kOOB off a finite axis)
__device__ __forceinline__ uint32_t neighbour_of(const Grid& gr, uint32_t g, int s) { return source_of(gr, g, 26 - s); }


// wind at coordinate sd between the faces blo, bhi (predcorr.cpp:31-42).  inv_span = 1/(bhi - blo)
// of the gridbox, so the reference's alpha = (coord - blo) / (bhi - blo) is a product, within an
// ulp or two of the quotient.
__device__ __forceinline__ double interpolation(double blo, double bhi, double inv_span, double vlo, double vhi,
                                                double sd) {
  const double coord = fmin(bhi, fmax(blo, sd));
  const double alpha = (coord - blo) * inv_span;
  return alpha * vhi + (1 - alpha) * vlo;
}

// One axis of CartesianMaps::get_local_bounding_gridbox_index (cartesianmaps.cpp:68-286,
// domainboundaries.hpp:34-109): step to the backward/forward neighbour when the coordinate left
// [lower, upper) and wrap the coordinate at a periodic domain edge.  i is the drop's index along
// the axis, updated to the neighbour's.  Returns the step taken.
__device__ __forceinline__ int axis_step(uint32_t& idx, double& coord, uint32_t& i, uint32_t nd, uint32_t inc,
                                         bool periodic, const double* h) {
  const double lo = h[i], hi = h[i + 1];
  if (coord < lo) {
    if (i == 0) {
      if (periodic) {
        const uint32_t nb = idx + (nd - 1) * inc;
        coord = coord + h[nd] - lo;  // lim1 = upper bound of the backward neighbour
        idx = nb;
        i = nd - 1;
      } else {
        idx = kOOB;
      }
    } else {
      idx = idx - inc;
      i -= 1;
    }
    return -1;
  }
  if (coord >= hi) {
    if (i == nd - 1) {
      if (periodic) {
        const uint32_t nb = idx - (nd - 1) * inc;
        coord = coord + h[0] - hi;  // lim1 = lower bound of the forward neighbour
        idx = nb;
        i = 0;
      } else {
        idx = kOOB;
      }
    } else {
      idx = idx + inc;
      i += 1;
    }
    return 1;
  }
  return 0;
}

// bounds, reciprocal widths and face winds of gridbox g
struct GbxGeom {
  uint32_t iz, ix, iy;
  double zlo, zhi, xlo, xhi, ylo, yhi, wlo, whi, ulo, uhi, vlo, vhi;
  double inv_z, inv_x, inv_y;
};
__device__ __forceinline__ GbxGeom gbx_geom(const MotionArgs& a, uint32_t g) {
  const uint32_t nz = a.grid.nz, nx = a.grid.nx;
  GbxGeom G;
  G.iz = g % nz;
  G.ix = (g / nz) % nx;
  G.iy = g / (nz * nx);
  G.zlo = a.grid.zh[G.iz];
  G.zhi = a.grid.zh[G.iz + 1];
  G.xlo = a.grid.xh[G.ix];
  G.xhi = a.grid.xh[G.ix + 1];
  G.ylo = a.grid.yh[G.iy];
  G.yhi = a.grid.yh[G.iy + 1];
  G.wlo = a.st.wlo[g];
  G.whi = a.st.whi[g];
  G.ulo = a.st.ulo[g];
  G.uhi = a.st.uhi[g];
  G.vlo = a.st.vlo[g];
  G.vhi = a.st.vhi[g];
  G.inv_z = rcp_nr(G.zhi - G.zlo);
  G.inv_x = rcp_nr(G.xhi - G.xlo);
  G.inv_y = rcp_nr(G.yhi - G.ylo);
  return G;
}
// drop k of gridbox g after one predictor-corrector step: new coordinates, new gridbox (kOOB
// when it left the domain) and the step taken, dir = (dz+1) + 3(dx+1) + 9(dy+1), or 27 out of the
// domain.  PredCorr::delta_coord3/1/2 (predcorr.hpp:84-145), cfl_criteria (cfl_criteria.hpp:40-57),
// superdrop_gbx z, then x, then y (predcorrmotion.hpp:85-106) and CartesianCheckBounds
// (cartesian_motion.hpp:38-52); `report` records the CFL and bounds errors.
struct Moved {
  double c3, c1, c2;
  uint32_t idx;
  int dir;
};
// what the motion step reads of drop k
struct DropPos {
  double c3, c1, c2, r;
};
__device__ __forceinline__ DropPos drop_pos(const MotionArgs& a, uint64_t k) {
  return DropPos{a.s.c3[k], a.s.c1[k], a.s.c2[k], a.s.r[k]};
}
__device__ __forceinline__ DropPos drop_pos_nt(const MotionArgs& a, uint64_t k) {
  return DropPos{__builtin_nontemporal_load(&a.s.c3[k]), __builtin_nontemporal_load(&a.s.c1[k]),
                 __builtin_nontemporal_load(&a.s.c2[k]), __builtin_nontemporal_load(&a.s.r[k])};
}
__device__ __forceinline__ Moved move_drop(const MotionArgs& a, uint32_t g, const GbxGeom& G, const DropPos& P,
                                           uint64_t k, bool report) {
  const uint32_t nz = a.grid.nz, nx = a.grid.nx, ny = a.grid.ny;
  const double delt = a.delt;
  const double c3_0 = P.c3, c1_0 = P.c1, c2_0 = P.c2;
  double coord3 = c3_0;
  const double terminal = terminal_velocity(a.tv, P.r);
  double vel3 = interpolation(G.zlo, G.zhi, G.inv_z, G.wlo, G.whi, coord3);
  vel3 -= terminal;
  coord3 += vel3 * delt;
  double corrvel3 = interpolation(G.zlo, G.zhi, G.inv_z, G.wlo, G.whi, coord3);
  corrvel3 -= terminal;
  const double delta3 = (vel3 + corrvel3) * (delt / 2);
  double coord1 = c1_0;
  const double vel1 = interpolation(G.xlo, G.xhi, G.inv_x, G.ulo, G.uhi, coord1);
  coord1 += vel1 * delt;
  const double corrvel1 = interpolation(G.xlo, G.xhi, G.inv_x, G.ulo, G.uhi, coord1);
  const double delta1 = (vel1 + corrvel1) * (delt / 2);
  double coord2 = c2_0;
  const double vel2 = interpolation(G.ylo, G.yhi, G.inv_y, G.vlo, G.vhi, coord2);
  coord2 += vel2 * delt;
  const double corrvel2 = interpolation(G.ylo, G.yhi, G.inv_y, G.vlo, G.vhi, coord2);
  const double delta2 = (vel2 + corrvel2) * (delt / 2);
  bool cfl = (fabs(delta3) <= fabs(G.zhi - G.zlo));
  cfl = (fabs(delta1) <= fabs(G.xhi - G.xlo)) && cfl;
  cfl = (fabs(delta2) <= fabs(G.yhi - G.ylo)) && cfl;
  if (report && !cfl) set_error(a.ctl, CLEO_ERR_CFL, g, k);
  Moved M;
  M.c3 = c3_0 + delta3;
  M.c1 = c1_0 + delta1;
  M.c2 = c2_0 + delta2;
  uint32_t idx = g, jz = G.iz, jx = G.ix, jy = G.iy;
  const int dz = axis_step(idx, M.c3, jz, nz, 1u, a.grid.per[0] != 0, a.grid.zh);
  int dx = 0, dy = 0;
  if (idx != kOOB) dx = axis_step(idx, M.c1, jx, nx, nz, a.grid.per[1] != 0, a.grid.xh);
  if (idx != kOOB) dy = axis_step(idx, M.c2, jy, ny, nz * nx, a.grid.per[2] != 0, a.grid.yh);
  M.idx = idx;
  M.dir = 27;
  if (idx != kOOB) {
    const bool bad = (M.c3 < a.grid.zh[jz]) | (M.c3 >= a.grid.zh[jz + 1]) | (M.c1 < a.grid.xh[jx]) |
                     (M.c1 >= a.grid.xh[jx + 1]) | (M.c2 < a.grid.yh[jy]) | (M.c2 >= a.grid.yh[jy + 1]);
    if (report && bad) set_error(a.ctl, CLEO_ERR_BOUNDS, g, k);
    M.dir = (dz + 1) + 3 * (dx + 1) + 9 * (dy + 1);
  }
  return M;
}
__device__ __forceinline__ Moved move_drop(const MotionArgs& a, uint32_t g, const GbxGeom& G, uint64_t k, bool report) {
  return move_drop(a, g, G, drop_pos(a, k), k, report);
}

// MoveSupersInGridboxesFunctor (movesupersindomain.hpp:47-93) with PredCorrMotion over
// CartesianMaps; also counts, per gridbox, how many drops go to each of the 27 neighbour
// directions (and out of the domain) for the stable neighbour sort.
__device__ __forceinline__ uint32_t motion_gbx(const MotionArgs& a, uint32_t g, int lane) {
  const uint32_t base = a.off[g];
  const uint32_t nfull = a.off[g + 1] - base;
  const GbxGeom G = gbx_geom(a, g);
  uint32_t mycount = 0;
  for (uint32_t c = 0; c < nfull; c += kWave) {
    const uint32_t k = base + c + lane;
    const bool act = c + lane < nfull;
    int dir = 0;
    if (act) {
      const Moved M = move_drop(a, g, G, k, true);
      a.s.c3[k] = M.c3;
      a.s.c1[k] = M.c1;
      a.s.c2[k] = M.c2;
      a.s.gbx[k] = M.idx;
      dir = M.dir;
    }
    uint64_t mask = __ballot(act);
    while (mask) {
      const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
      const int v = static_cast<int>(readlane_u32(static_cast<uint32_t>(dir), leader));
      const uint64_t m = __ballot(act && dir == v);
      if (lane == v) mycount += static_cast<uint32_t>(__popcll(m));
      mask &= ~m;
    }
  }
  // destination-major counts: row d column s = drops arriving at d by the step s; column 27 of the
  // source's own row = drops leaving the domain (the table is zero-filled before motion)
  if (lane == 27)
    a.cnt[28ull * g + 27] = mycount;
  else if (lane < 27 && mycount > 0)
    a.cnt[28ull * neighbour_of(a.grid, g, lane) + lane] = mycount;
  return nfull;
}
__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(3) k_motion(MotionArgs a) {
  const int lane = lane_id();
  uint32_t moved = 0;
  for (uint32_t g = wave_index(); g < a.grid.ngbx; g += wave_count()) moved += motion_gbx(a, g, lane);
  if (lane == 0) put_stat(a.stat, 0, wave_index(), moved);
}

// ============================ neighbour sort ==================================================
__device__ __forceinline__ bool guard_off(const uint32_t* flag) { return flag && *flag == 0u; }

// Per destination d (thread per d): the drops arriving in d, and for each step s the drops that
// reach d from source gridboxes of lower index than the source of s (the stable order puts them
// first).  The counts are destination-major (row d, column s: arrivals at d by the step s) and
// read as one row; the run offsets are written source-major into `runs` (row g, column s: the
// offset of g's drops taking the step s), so the rank pass reads its gridbox's 27 offsets as one
// contiguous row.  Every entry of `runs` belongs to one destination (g + step s): no two threads
// write the same entry.  Only the entries of steps that some drop takes are written (a source off
// a finite axis has none); the others keep older values that no pass reads.
__global__ void k_nsort_tot(NsortArgs a) {
  if (guard_off(a.flag)) return;
  const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= a.grid.ngbx) return;
  const uint32_t* row = a.cnt + 28ull * d;
  a.oobcnt[d] = row[27];  // the out-of-domain count, compact for its scan
  uint32_t cnt[27];
  uint32_t t = 0;
#pragma unroll
  for (int s = 0; s < 27; ++s) {
    cnt[s] = row[s];
    t += cnt[s];
  }
  const Grid& gr = a.grid;
  const int nz = static_cast<int>(gr.nz), nx = static_cast<int>(gr.nx), ny = static_cast<int>(gr.ny);
  const int gz = static_cast<int>(d % gr.nz), gx = static_cast<int>((d / gr.nz) % gr.nx),
            gy = static_cast<int>(d / (gr.nz * gr.nx));
  auto axis = [](int i, int dd, int n, bool per) {
    int j = i - dd;
    if (j < 0) j = per ? j + n : -1;
    if (j >= n) j = per ? j - n : -1;
    return j;
  };
  uint32_t src[27];  // the source of each step (kOOB: off a finite axis, no drops)
#pragma unroll
  for (int s = 0; s < 27; ++s) {
    const int jz = axis(gz, s % 3 - 1, nz, gr.per[0] != 0), jx = axis(gx, (s / 3) % 3 - 1, nx, gr.per[1] != 0),
              jy = axis(gy, s / 9 - 1, ny, gr.per[2] != 0);
    src[s] = (jz >= 0 && jx >= 0 && jy >= 0) ? static_cast<uint32_t>(jz + nz * (jx + nx * jy)) : kOOB;
  }
  const bool wraps = (gr.per[0] && (gz == 0 || gz + 1 == nz)) || (gr.per[1] && (gx == 0 || gx + 1 == nx)) ||
                     (gr.per[2] && (gy == 0 || gy + 1 == ny));
  if (!wraps) {
    // no step wraps: the source of step s is d - (dz + nz dx + nz nx dy), so the sources in index
    // order are the steps in decreasing order
    uint32_t before = 0;
#pragma unroll
    for (int s = 26; s >= 0; --s) {
      if (cnt[s] > 0u && src[s] != kOOB) a.runs[28ull * src[s] + static_cast<uint32_t>(s)] = before;
      before += cnt[s];
    }
  } else {
#pragma unroll
    for (int s = 0; s < 27; ++s) {
      uint32_t before = 0;
#pragma unroll
      for (int q = 0; q < 27; ++q) before += src[q] < src[s] ? cnt[q] : 0u;
      if (cnt[s] > 0u && src[s] != kOOB) a.runs[28ull * src[s] + static_cast<uint32_t>(s)] = before;
    }
  }
  a.tot[d] = t;
}

// Stable scatter (sortsupers.hpp:63-76,167-193 under Serial): drop p of source gridbox g with
// key d lands at  base[d] + (drops reaching d from gridboxes < g) + (earlier drops of g with key d);
// the middle term is entry (g, step g -> d) of the `runs` table k_nsort_tot writes, so the 27
// run offsets of g are its own contiguous row.  The 27 run starts of g (one per step, with base[d])
// and the start of its out-of-domain run are loaded by lanes 0..27 together with the keys (one
// round trip; loading only the steps that occur costs a second one), so the ranking below touches
// no memory.  The destination of each drop is
// recorded as perm[dest] = p; k_nsort_gather then moves the records, so that each cache line of the
// sorted store is written whole by one wave (runs scattered from many source gridboxes into one
// line cost HBM read-modify-writes).
// The ranking of a gridbox is split in four stages so that a wave works on four gridboxes at once,
// each stage reading what the previous one loaded an iteration earlier: the head (range and run
// starts), the positions of the first two chunks of slots, the keys of those drops, then the
// ranking itself.  Every load of a stage is issued by all lanes, with indices clamped into the
// arrays, so the number of loads in flight is the same on every path and the waits of the ranking
// leave the younger stages' loads outstanding.
//
// head of gridbox g: its range of positions and, in lane s < 27, the next place of its drops that
// take step s (lane 27: of those leaving).  The neighbour reached by step s = (dz+1) + 3(dx+1) +
// 9(dy+1) is g + noff away from the edges, else wrapped on periodic axes (the same gridbox as
// neighbour_of, without a division per lane).  A g past the last gridbox gives an empty head.
struct RankHead {
  uint32_t g, sbase, nfull, trun;
  int gz, gx, gy;
};
__device__ __forceinline__ RankHead rank_head(const NsortArgs& a, uint64_t g, int gz, int gx, int gy, int32_t noff,
                                              int lane) {
  const uint32_t ngbx = a.grid.ngbx;
  const int nz = static_cast<int>(a.grid.nz), nx = static_cast<int>(a.grid.nx), ny = static_cast<int>(a.grid.ny);
  const bool valid = g < ngbx;
  const uint32_t gc = valid ? static_cast<uint32_t>(g) : ngbx - 1u;
  RankHead h{gc, 0u, 0u, 0u, gz, gx, gy};
  const uint32_t lo = a.off[gc], hi = a.off[gc + 1];
  h.sbase = lo;
  h.nfull = valid ? hi - lo : 0u;
  uint32_t d = ngbx;  // lanes 27..63: the out-of-domain run (base[ngbx] + oobpre[g])
  bool ok = valid && lane == 27;
  if (lane < 27 && !valid) {
    d = 0u;  // no gridbox: a harmless load, discarded
  } else if (lane < 27) {
    const bool interior = gz > 0 && gz + 1 < nz && gx > 0 && gx + 1 < nx && gy > 0 && gy + 1 < ny;
    d = gc + static_cast<uint32_t>(noff);
    ok = true;
    if (!interior) {
      auto axis = [](int i, int dd, int n, bool per) {
        int j = i + dd;
        if (j < 0) j = per ? j + n : -1;
        if (j >= n) j = per ? j - n : -1;
        return j;
      };
      const int jz = axis(gz, lane % 3 - 1, nz, a.grid.per[0] != 0);
      const int jx = axis(gx, (lane / 3) % 3 - 1, nx, a.grid.per[1] != 0);
      const int jy = axis(gy, lane / 9 - 1, ny, a.grid.per[2] != 0);
      ok = jz >= 0 && jx >= 0 && jy >= 0;
      d = ok ? static_cast<uint32_t>(jz + nz * (jx + nx * jy)) : 0u;
    }
  }
  const uint32_t b = a.base[d];
  const uint32_t extra = lane < 27 ? a.runs[28ull * gc + static_cast<uint32_t>(lane)] : a.oobpre[gc];
  h.trun = ok ? b + extra : 0u;
  return h;
}
// positions held by slots lane and lane + 64 of the gridbox's logical order (the order collisions
// left, or the store's own order); kOOB past the gridbox
struct RankSlots {
  uint32_t p0, p1;
};
template <bool kOrder>
__device__ __forceinline__ RankSlots rank_slots(const NsortArgs& a, const RankHead& h, uint32_t nsrc, int lane) {
  const uint32_t i0 = static_cast<uint32_t>(lane), i1 = i0 + kWave;
  const uint32_t q0 = min(h.sbase + i0, nsrc - 1u), q1 = min(h.sbase + i1, nsrc - 1u);
  uint32_t p0 = q0, p1 = q1;
  if constexpr (kOrder) {
    p0 = a.order[q0];
    p1 = a.order[q1];
  }
  return RankSlots{i0 < h.nfull ? p0 : kOOB, i1 < h.nfull ? p1 : kOOB};
}
// the step recorded for each of those drops (kStep), or its gridbox key; 0 for no drop
struct RankKeys {
  uint32_t k0, k1;
};
template <bool kStep>
__device__ __forceinline__ RankKeys rank_keys(const NsortArgs& a, const RankSlots& s, uint32_t nsrc) {
  const uint32_t c0 = s.p0 < nsrc ? s.p0 : 0u, c1 = s.p1 < nsrc ? s.p1 : 0u;
  uint32_t k0, k1;
  if constexpr (kStep) {
    k0 = a.step[c0];
    k1 = a.step[c1];
  } else {
    k0 = a.src.gbx[c0];
    k1 = a.src.gbx[c1];
  }
  return RankKeys{k0, k1};
}
// ranks the drops of the head's gridbox and records perm[dest] = p (stable scatter above); the
// first two chunks use the preloaded positions and keys, later chunks load their own
template <bool kOrder, bool kStep>
__device__ __forceinline__ void rank_gbx(const NsortArgs& a, const RankHead& h, const RankSlots& s, const RankKeys& k,
                                         uint32_t nsrc, int lane) {
  const uint32_t g = h.g, sbase = h.sbase, nfull = h.nfull, ngbx = a.grid.ngbx;
  const int nz = static_cast<int>(a.grid.nz), nx = static_cast<int>(a.grid.nx), ny = static_cast<int>(a.grid.ny);
  const int gz = h.gz, gx = h.gx, gy = h.gy;
  uint32_t trun = h.trun;
  auto near = [](int q, int r, int nd, bool per) {
    const int dd = q - r;
    if (dd >= -1 && dd <= 1) return true;
    return per && (dd == nd - 1 || dd == 1 - nd);
  };
  auto step = [](int q, int r) {
    const int dd = q - r;
    return dd > 1 ? -1 : (dd < -1 ? 1 : dd);
  };
  // chunk c of 64 slots: the drop of slot sbase + c + lane is at position p with key kv
  auto chunk = [&](uint32_t c, uint32_t p, uint32_t kv) {
    const uint32_t q = sbase + c + lane;
    const bool act = c + lane < nfull;
    // the step g -> key (27: out of the domain): recorded by motion, or found from the key; a key
    // that is no neighbour of g, or a position outside the domain, is an error
    int sv = 27;
    bool ok = act;
    if (act && p >= nsrc) {
      set_error(a.ctl, CLEO_ERR_SORT_NONLOCAL, g, q);
      ok = false;
    }
    if (ok && kStep) {
      sv = static_cast<int>(kv);
      if (sv > 27) {
        set_error(a.ctl, CLEO_ERR_SORT_NONLOCAL, g, p);
        ok = false;
      }
    } else if (ok && kv == g) {
      sv = 13;
    } else if (ok && kv < ngbx) {
      const int vz = static_cast<int>(kv % a.grid.nz), vx = static_cast<int>((kv / a.grid.nz) % a.grid.nx),
                vy = static_cast<int>(kv / (a.grid.nz * a.grid.nx));
      if (near(vz, gz, nz, a.grid.per[0]) && near(vx, gx, nx, a.grid.per[1]) && near(vy, gy, ny, a.grid.per[2])) {
        sv = (step(vz, gz) + 1) + 3 * (step(vx, gx) + 1) + 9 * (step(vy, gy) + 1);
      } else {
        set_error(a.ctl, CLEO_ERR_SORT_NONLOCAL, g, p);
        ok = false;
      }
    }
    // one round per distinct step in the chunk: its lanes take consecutive places from the run's
    // next place (held by lane v), in lane order
    uint32_t dest = 0;
    const uint64_t okm = ballot(ok);
    uint64_t mask = okm;
    while (mask) {
      const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
      const int v = static_cast<int>(readlane_u32(static_cast<uint32_t>(sv), leader));
      const uint64_t m = ballot(sv == v) & okm;
      const uint32_t start = readlane_u32(trun, v);
      const uint32_t here = start + lanes_below(m);
      dest = (ok && sv == v) ? here : dest;
      trun = lane == v ? start + static_cast<uint32_t>(__popcll(m)) : trun;
      mask &= ~m;
    }
    if (ok && dest >= a.ntot) {  // counts inconsistent with keys: never write out of range
      set_error(a.ctl, CLEO_ERR_SORT_NONLOCAL, g, p);
    } else if (ok) {
      a.perm[dest] = p;
    }
  };
  // the first two chunks from the preloaded positions and keys, straight-line; larger gridboxes
  // (rare at these sizes) load the rest as they go
  if (nfull > 0) chunk(0u, s.p0, k.k0);
  if (nfull > kWave) chunk(kWave, s.p1, k.k1);
  for (uint32_t c = 2 * kWave; c < nfull; c += kWave) {
    const bool act = c + lane < nfull;
    const uint32_t q = sbase + c + lane;
    const uint32_t p = act ? (kOrder ? a.order[q] : q) : kOOB;
    const uint32_t kv = (act && p < nsrc) ? (kStep ? a.step[p] : a.src.gbx[p]) : 0u;
    chunk(c, p, kv);
  }
}
// first pass of the neighbour sort: the destination of every drop of the domain, as perm[dest] = p.
// Gridboxes go to waves through xcd_wave_index(), so each XCD sweeps one contiguous range and the
// count-table rows that neighbouring gridboxes share are fetched into one L2.  The gridbox's
// (z, x, y) advance by the stride's with carries (no division per gridbox).
template <bool kOrder, bool kStep>
__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(kStep ? 6 : 4) k_nsort_perm(NsortArgs a) {
  if (guard_off(a.flag)) return;
  const int lane = lane_id();
  const uint32_t nz = a.grid.nz, nx = a.grid.nx;
  const uint64_t ngbx = a.grid.ngbx;
  const uint32_t g0 = xcd_wave_index(), stride = wave_count();
  const uint32_t nsrc = a.off[ngbx];
  if (g0 >= ngbx || nsrc == 0) return;
  uint32_t gz = g0 % nz, gx = (g0 / nz) % nx, gy = g0 / (nz * nx);
  const uint32_t sz = stride % nz, sx = (stride / nz) % nx, sy = stride / (nz * nx);
  const int32_t noff = (lane % 3 - 1) + static_cast<int32_t>(nz) * ((lane / 3) % 3 - 1) +
                       static_cast<int32_t>(nz * nx) * (lane / 9 - 1);
  uint64_t gh = g0;  // next gridbox whose head is loaded
  auto next_head = [&]() {
    const RankHead h = rank_head(a, gh, static_cast<int>(gz), static_cast<int>(gx), static_cast<int>(gy), noff, lane);
    gh += stride;
    gz += sz;
    uint32_t carry = gz >= nz ? 1u : 0u;
    gz -= carry * nz;
    gx += sx + carry;
    carry = gx >= nx ? 1u : 0u;
    gx -= carry * nx;
    gy += sy + carry;
    return h;
  };
  // rings of four per stage, rotated by name over a four-fold unrolled loop: no register moves
  // between iterations, which would wait for the loads in flight
  RankHead H0 = next_head(), H1 = next_head(), H2, H3;
  RankSlots S0 = rank_slots<kOrder>(a, H0, nsrc, lane), S1, S2, S3;
  H2 = next_head();
  S1 = rank_slots<kOrder>(a, H1, nsrc, lane);
  RankKeys K0 = rank_keys<kStep>(a, S0, nsrc), K1, K2, K3;
  // gridbox i: loads the head of i + 3, the slots of i + 2, the keys of i + 1, then ranks i
  auto stage = [&](const RankHead& hu, RankHead& hl, const RankHead& hs, const RankSlots& su, RankSlots& sl,
                   const RankSlots& sk, const RankKeys& ku, RankKeys& kl) {
    hl = next_head();
    sl = rank_slots<kOrder>(a, hs, nsrc, lane);
    kl = rank_keys<kStep>(a, sk, nsrc);
    rank_gbx<kOrder, kStep>(a, hu, su, ku, nsrc, lane);
  };
  for (uint64_t g = g0;;) {
    stage(H0, H3, H2, S0, S2, S1, K0, K1);
    if ((g += stride) >= ngbx) break;
    stage(H1, H0, H3, S1, S3, S2, K1, K2);
    if ((g += stride) >= ngbx) break;
    stage(H2, H1, H0, S2, S0, S3, K2, K3);
    if ((g += stride) >= ngbx) break;
    stage(H3, H2, H1, S3, S1, S0, K3, K0);
    if ((g += stride) >= ngbx) break;
  }
}

// Second pass: destination position q takes the record at source position perm[q].  Consecutive
// threads write consecutive positions, so each cache line of the sorted store is written whole by
// one wave; the reads follow the runs of the source order.  Two positions per thread, all loads of
// both issued before the stores.  A position whose perm entry is not a source position of the
// domain (counts inconsistent with the keys) is reported and skipped, so no read leaves the store.
__global__ void __launch_bounds__(256) k_nsort_gather(NsortArgs a) {
  if (guard_off(a.flag)) return;
  const uint64_t nsrc = a.off[a.grid.ngbx];
  for (uint64_t q0 = 2ull * blockIdx.x * blockDim.x; q0 < nsrc; q0 += 2ull * gridDim.x * blockDim.x) {
    const uint64_t qa = q0 + threadIdx.x, qb = qa + blockDim.x;
    const uint32_t pa = qa < nsrc ? a.perm[qa] : 0u, pb = qb < nsrc ? a.perm[qb] : 0u;
    const bool oka = qa < nsrc && pa < nsrc, okb = qb < nsrc && pb < nsrc;
    if (qa < nsrc && !oka) set_error(a.ctl, CLEO_ERR_SORT_NONLOCAL, kOOB, qa);
    if (qb < nsrc && !okb) set_error(a.ctl, CLEO_ERR_SORT_NONLOCAL, kOOB, qb);
    const uint32_t sa = oka ? pa : 0u, sb = okb ? pb : 0u;
    const uint32_t ga = a.src.gbx[sa], gb = a.src.gbx[sb];
    const double c3a = a.src.c3[sa], c3b = a.src.c3[sb];
    const double c1a = a.src.c1[sa], c1b = a.src.c1[sb];
    const double c2a = a.src.c2[sa], c2b = a.src.c2[sb];
    const uint64_t xia = a.src.xi[sa], xib = a.src.xi[sb];
    const double ra = a.src.r[sa], rb = a.src.r[sb];
    const double ma = a.src.msol[sa], mb = a.src.msol[sb];
    const uint64_t ida = a.src.id[sa], idb = a.src.id[sb];
    if (oka) {
      a.dst.gbx[qa] = ga;
      a.dst.c3[qa] = c3a;
      a.dst.c1[qa] = c1a;
      a.dst.c2[qa] = c2a;
      a.dst.xi[qa] = xia;
      a.dst.r[qa] = ra;
      a.dst.msol[qa] = ma;
      a.dst.id[qa] = ida;
    }
    if (okb) {
      a.dst.gbx[qb] = gb;
      a.dst.c3[qb] = c3b;
      a.dst.c1[qb] = c1b;
      a.dst.c2[qb] = c2b;
      a.dst.xi[qb] = xib;
      a.dst.r[qb] = rb;
      a.dst.msol[qb] = mb;
      a.dst.id[qb] = idb;
    }
  }
}

// drops already out of the domain stay behind the newly out-of-domain ones, in order
__global__ void k_nsort_oob(NsortArgs a, uint64_t ntot) {
  if (guard_off(a.flag)) return;
  const uint64_t old_end = a.off[a.grid.ngbx];
  const uint64_t dst0 = static_cast<uint64_t>(a.base[a.grid.ngbx]) + a.oobpre[a.grid.ngbx];
  for (uint64_t p = old_end + blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p < ntot;
       p += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t d = dst0 + (p - old_end);
    if (d >= ntot) {
      set_error(a.ctl, CLEO_ERR_SORT_NONLOCAL, kOOB, p);
      continue;
    }
    a.dst.gbx[d] = a.src.gbx[p];
    a.dst.c3[d] = a.src.c3[p];
    a.dst.c1[d] = a.src.c1[p];
    a.dst.c2[d] = a.src.c2[p];
    a.dst.xi[d] = a.src.xi[p];
    a.dst.r[d] = a.src.r[p];
    a.dst.msol[d] = a.src.msol[p];
    a.dst.id[d] = a.src.id[p];
  }
}

// Second pass of the fused motion + sort (the first is k_motion_drops<false>): the drops of
// gridbox g move again, with the arithmetic of the counting pass, and go with their new coordinates
// and gridbox straight to their place in the sorted store.  The place is the start of the
// (destination, step) run left in the count table by k_nsort_tot plus the drop's rank among this
// gridbox's drops that took the same step -- the stable order of the scatter above.
__device__ __forceinline__ void motion_scatter_gbx(const MotionArgs& a, const NsortArgs& n, uint32_t g, int lane) {
  const uint32_t base = a.off[g];
  const uint32_t nfull = a.off[g + 1] - base;
  const GbxGeom G = gbx_geom(a, g);
  uint32_t run = 0;   // lane L < 28: next position of the drops of step L
  uint32_t have = 0;  // bit L: the run of step L has its start
  for (uint32_t c = 0; c < nfull; c += kWave) {
    const uint32_t k = base + c + lane;
    const bool act = c + lane < nfull;
    Moved M{0.0, 0.0, 0.0, 0u, 0};
    if (act) M = move_drop(a, g, G, k, false);
    const int dir = act ? M.dir : -1;
    // the steps present in this chunk; the runs they open get their starts in one round of loads
    uint32_t present = 0;
    uint64_t mask = __ballot(act);
    while (mask) {
      const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
      const int v = static_cast<int>(readlane_u32(static_cast<uint32_t>(dir), leader));
      present |= 1u << v;
      mask &= ~__ballot(act && dir == v);
    }
    const uint32_t need = present & ~have;
    if (lane < 28 && ((need >> lane) & 1u)) {
      if (lane == 27) {
        run = n.base[a.grid.ngbx] + n.oobpre[g];
      } else {
        const uint32_t d = neighbour_of(a.grid, g, lane);
        run = n.base[d] + n.runs[28ull * g + lane];
      }
    }
    have |= need;
    uint32_t dest = 0;
    mask = __ballot(act);
    while (mask) {
      const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
      const int v = static_cast<int>(readlane_u32(static_cast<uint32_t>(dir), leader));
      const uint64_t m = __ballot(act && dir == v);
      const uint32_t start = readlane_u32(run, v);
      if ((m >> lane) & 1ull) dest = start + static_cast<uint32_t>(__popcll(m & lanemask_lt()));
      if (lane == v) run += static_cast<uint32_t>(__popcll(m));
      mask &= ~m;
    }
    if (act && dest >= n.ntot) {  // counts inconsistent with the moves: never write out of range
      set_error(n.ctl, CLEO_ERR_SORT_NONLOCAL, g, k);
    } else if (act) {
      n.dst.gbx[dest] = M.idx;
      n.dst.c3[dest] = M.c3;
      n.dst.c1[dest] = M.c1;
      n.dst.c2[dest] = M.c2;
      n.dst.xi[dest] = a.s.xi[k];
      n.dst.r[dest] = a.s.r[k];
      n.dst.msol[dest] = a.s.msol[k];
      n.dst.id[dest] = a.s.id[k];
    }
  }
}
__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(4) k_motion_scatter(MotionArgs a, NsortArgs n) {
  const int lane = lane_id();
  for (uint32_t g = wave_index(); g < a.grid.ngbx; g += wave_count()) motion_scatter_gbx(a, n, g, lane);
}

// counts of the re-sort after collisions: active drops stay, nulls leave the domain
__global__ void k_cnt_from_nact(const uint32_t* off, const uint32_t* nact, uint32_t ngbx, uint32_t* cnt,
                                const uint32_t* flag) {
  if (guard_off(flag)) return;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < 28ull * ngbx;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t g = static_cast<uint32_t>(i / 28), s = static_cast<uint32_t>(i % 28);
    const uint32_t n = nact[g], nfull = off[g + 1] - off[g];
    cnt[i] = s == 13 ? n : (s == 27 ? nfull - n : 0u);
  }
}

__global__ void k_set_nact(const uint32_t* off, uint32_t ngbx, uint32_t* nact, const uint32_t* flag) {
  if (guard_off(flag)) return;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < ngbx) nact[g] = off[g + 1] - off[g];
}

__global__ void k_copy_u32(const uint32_t* src, uint32_t* dst, uint64_t n, const uint32_t* flag) {
  if (guard_off(flag)) return;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_copy_soa(SoA src, SoA dst, uint64_t lo, uint64_t hi, const uint32_t* flag) {
  if (guard_off(flag)) return;
  for (uint64_t p = lo + blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p < hi;
       p += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    dst.gbx[p] = src.gbx[p];
    dst.c3[p] = src.c3[p];
    dst.c1[p] = src.c1[p];
    dst.c2[p] = src.c2[p];
    dst.xi[p] = src.xi[p];
    dst.r[p] = src.r[p];
    dst.msol[p] = src.msol[p];
    dst.id[p] = src.id[p];
  }
}

// copies [off_old[ngbx], ntot) from src to dst (the out-of-domain region after collisions)
__global__ void k_copy_oob_region(SoA src, SoA dst, const uint32_t* off, uint32_t ngbx, uint64_t ntot) {
  const uint64_t lo = off[ngbx];
  for (uint64_t p = lo + blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p < ntot;
       p += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    dst.gbx[p] = src.gbx[p];
    dst.c3[p] = src.c3[p];
    dst.c1[p] = src.c1[p];
    dst.c2[p] = src.c2[p];
    dst.xi[p] = src.xi[p];
    dst.r[p] = src.r[p];
    dst.msol[p] = src.msol[p];
    dst.id[p] = src.id[p];
  }
}

__global__ void k_clear_flag(DevCtl* ctl) { ctl->resort_flag = 0u; }
__global__ void k_sorts_stat(DevCtl* ctl, const uint32_t* flag) {
  if (guard_off(flag)) return;
  ctl->stats[ST_SORTS] += 1;
}

// ============================ generic stable radix sort =======================================
__global__ void k_rs_init(const uint32_t* gbx, uint64_t n, uint32_t ngbx, uint32_t* key, uint32_t* idx) {
  for (uint64_t p = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p < n;
       p += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t k = gbx[p];
    key[p] = k < ngbx ? k : ngbx;  // _get_count_position (sortsupers.hpp:44-48)
    idx[p] = static_cast<uint32_t>(p);
  }
}

// per-tile (256 keys) histogram of one 4-bit digit: cnt[digit * ntiles + tile]
__global__ void __launch_bounds__(256) k_rs_count(const uint32_t* key, uint64_t n, int shift, uint32_t ntiles,
                                                  uint32_t* cnt) {
  __shared__ uint32_t wc[4][16];
  const uint32_t t = blockIdx.x;
  const uint64_t p = static_cast<uint64_t>(t) * 256 + threadIdx.x;
  const bool act = p < n;
  const uint32_t dg = act ? (key[p] >> shift) & 15u : 0u;
  const int w = threadIdx.x >> 6, lane = lane_id();
  for (uint32_t v = 0; v < 16; ++v) {
    const uint64_t m = __ballot(act && dg == v);
    if (lane == 0) wc[w][v] = static_cast<uint32_t>(__popcll(m));
  }
  __syncthreads();
  if (threadIdx.x < 16) cnt[threadIdx.x * ntiles + t] = wc[0][threadIdx.x] + wc[1][threadIdx.x] + wc[2][threadIdx.x] + wc[3][threadIdx.x];
}

__global__ void __launch_bounds__(256) k_rs_scatter(const uint32_t* key, const uint32_t* idx, uint64_t n, int shift,
                                                    uint32_t ntiles, const uint32_t* offs, uint32_t* key2,
                                                    uint32_t* idx2) {
  __shared__ uint32_t wc[4][16];
  const uint32_t t = blockIdx.x;
  const uint64_t p = static_cast<uint64_t>(t) * 256 + threadIdx.x;
  const bool act = p < n;
  const uint32_t k = act ? key[p] : 0u;
  const uint32_t dg = (k >> shift) & 15u;
  const int w = threadIdx.x >> 6, lane = lane_id();
  uint32_t wrank = 0;
  for (uint32_t v = 0; v < 16; ++v) {
    const uint64_t m = __ballot(act && dg == v);
    if (act && dg == v) wrank = static_cast<uint32_t>(__popcll(m & lanemask_lt()));
    if (lane == 0) wc[w][v] = static_cast<uint32_t>(__popcll(m));
  }
  __syncthreads();
  if (act) {
    uint32_t pre = 0;
    for (int ww = 0; ww < w; ++ww) pre += wc[ww][dg];
    const uint32_t dst = offs[dg * ntiles + t] + pre + wrank;
    if (dst >= n) return;
    key2[dst] = k;
    idx2[dst] = idx[p];
  }
}

__global__ void k_gather_soa(SoA src, SoA dst, const uint32_t* idx, uint64_t n) {
  for (uint64_t p = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p < n;
       p += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t s = idx[p];
    dst.gbx[p] = src.gbx[s];
    dst.c3[p] = src.c3[s];
    dst.c1[p] = src.c1[s];
    dst.c2[p] = src.c2[s];
    dst.xi[p] = src.xi[s];
    dst.r[p] = src.r[s];
    dst.msol[p] = src.msol[s];
    dst.id[p] = src.id[s];
  }
}

// off[b] = first position with key >= b, b = 0..ngbx (find_refs / find_domainrefs, findrefs.hpp)
__global__ void k_bucket_bounds(const uint32_t* key, uint64_t n, uint32_t ngbx, uint32_t* off) {
  for (uint64_t p = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; p <= n;
       p += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const int64_t prev = p == 0 ? -1 : static_cast<int64_t>(key[p - 1]);
    const int64_t cur = p == n ? static_cast<int64_t>(ngbx) + 1 : static_cast<int64_t>(key[p]);
    for (int64_t b = prev + 1; b <= cur && b <= static_cast<int64_t>(ngbx); ++b) off[b] = static_cast<uint32_t>(p);
  }
}

// ============================ scan (exclusive, u32) ===========================================
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t load_in(const uint32_t* in, uint64_t i, uint32_t stride, uint32_t offset,
                                            uint64_t len) {
  return i < len ? in[i * stride + offset] : 0u;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const uint32_t* in, uint64_t len, uint32_t stride,
                                                              uint32_t offset, uint32_t* partial,
                                                              const uint32_t* flag) {
  if (guard_off(flag)) return;
  __shared__ uint32_t ws[4];
  const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kScanTile;
  uint32_t s = 0;
  for (int i = 0; i < kScanItems; ++i) s += load_in(in, t0 + i * kScanThreads + threadIdx.x, stride, offset, len);
  s = wave_sum_u32(s);
  if (lane_id() == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ void __launch_bounds__(kScanThreads) k_scan_apply(const uint32_t* in, uint64_t len, uint32_t stride,
                                                             uint32_t offset, const uint32_t* block_off,
                                                             uint32_t* out, const uint32_t* flag) {
  if (guard_off(flag)) return;
  __shared__ uint32_t ws[4];
  const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kScanTile;
  // each thread owns kScanItems consecutive items
  uint32_t v[kScanItems];
  uint32_t s = 0;
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = load_in(in, t0 + static_cast<uint64_t>(threadIdx.x) * kScanItems + i, stride, offset, len);
    s += v[i];
  }
  // inclusive wave scan of per-thread sums
  const int lane = lane_id();
  uint32_t incl = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  if (lane == 63) ws[threadIdx.x >> 6] = incl;
  __syncthreads();
  uint32_t pre = block_off ? block_off[blockIdx.x] : 0u;
  for (int ww = 0; ww < static_cast<int>(threadIdx.x >> 6); ++ww) pre += ws[ww];
  uint32_t run = pre + incl - s;
  for (int i = 0; i < kScanItems; ++i) {
    const uint64_t idx = t0 + static_cast<uint64_t>(threadIdx.x) * kScanItems + i;
    if (idx < len) out[idx] = run;
    run += v[i];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) out[len] = run;
}

}  // namespace cleo

// ============================ host-side launchers =============================================
namespace cleo {

static inline unsigned nblocks(uint64_t n, uint64_t per) {
  const uint64_t b = (n + per - 1) / per;
  return static_cast<unsigned>(b == 0 ? 1 : b);
}
__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v);
__global__ void k_iota_u32(uint32_t* p, uint64_t n, const uint32_t* flag);
template <bool kWrite>
static void launch_motion_drops(cleo_ctx* c, const MotionArgs& a, unsigned nb);
static void launch_gather_move(cleo_ctx* c, const NsortArgs& n, const MotionArgs& m);
// blocks of the motion drop loop (two drops per thread), capped at `grid_waves` waves when non-zero
static inline unsigned motion_blocks(uint64_t n, uint32_t grid_waves) {
  const unsigned b = nblocks(n, 512);
  return grid_waves ? std::min<unsigned>(b, std::max<unsigned>(1u, grid_waves / 4)) : b;
}

// events around the main kernel of a stage (stage timing), recorded only while kev_on is set
static inline void kev_mark(cleo_ctx* c, int i) {
  if (c->kev_on) (void)hipEventRecord(c->kev[i], c->stream);
}
// blocks of a gridbox loop: one wave per gridbox, capped at `grid_waves` waves when non-zero
static inline unsigned gbx_blocks(uint64_t ngbx, uint32_t wpb, uint32_t grid_waves) {
  const unsigned b = nblocks(ngbx, wpb);
  return grid_waves ? std::min<unsigned>(b, std::max<unsigned>(1u, grid_waves / wpb)) : b;
}
// blocks of 256 threads of a drop loop: one thread per drop, capped at `grid_waves` waves
static inline unsigned drop_blocks(uint64_t n, uint32_t grid_waves) {
  const unsigned b = nblocks(n, 256);
  return grid_waves ? std::min<unsigned>(b, std::max<unsigned>(1u, grid_waves / 4)) : b;
}
// launches k_nsort_perm for the order and key source of `a`
static void launch_nsort_perm(cleo_ctx* c, const NsortArgs& a, uint32_t ngbx) {
  const dim3 grid(gbx_blocks(ngbx, a.wpb, c->tiles.motion_grid_waves)), block(64 * a.wpb);
  if (a.order && a.step)
    hipLaunchKernelGGL((k_nsort_perm<true, true>), grid, block, 0, c->stream, a);
  else if (a.order)
    hipLaunchKernelGGL((k_nsort_perm<true, false>), grid, block, 0, c->stream, a);
  else if (a.step)
    hipLaunchKernelGGL((k_nsort_perm<false, true>), grid, block, 0, c->stream, a);
  else
    hipLaunchKernelGGL((k_nsort_perm<false, false>), grid, block, 0, c->stream, a);
}
static inline unsigned stride_blocks(uint64_t n) {
  const uint64_t b = (n + 255) / 256;
  return static_cast<unsigned>(b == 0 ? 1 : (b > 65536 ? 65536 : b));
}
// blocks of a grid-stride loop that usually has little or nothing to do: a launch guarded by a
// device flag that is mostly clear (the re-sort after collisions), or the out-of-domain tail.
// 2048 blocks of 256 threads still fill every CU when the work is there; a launch of many
// thousands of blocks that exit at once costs tens of microseconds.
static inline unsigned sparse_blocks(uint64_t n) { return std::min<unsigned>(stride_blocks(n), 2048u); }

__global__ void k_set_u32(uint32_t* out, uint32_t v, const uint32_t* flag) {
  if (guard_off(flag)) return;
  *out = v;
}

// supersaturation ratio and diffusion factor of every gridbox (thread per gridbox) for
// k_condensation: the correctly rounded exp of the saturation pressure, evaluated here, would
// otherwise cost the per-gridbox solve kernel its registers
__global__ void k_gbx_thermo(GState st, uint32_t ngbx, double* gs) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < ngbx; g += gridDim.x * blockDim.x) {
    const double temp = st.temp[g], press = st.press[g];
    const double psat = saturation_pressure(temp);
    gs[kGsStride * g] = supersaturation_ratio(press, st.qv[g], psat);
    gs[kGsStride * g + 1] = diffusion_factor(press, temp, psat);
    const double akoh = kohler_a(temp);
    gs[kGsStride * g + 2] = akoh;
    gs[kGsStride * g + 3] = rcp_nr(akoh);
  }
}
void launch_condensation(cleo_ctx* c, const CondArgs& args) {
  if (args.ngbx == 0) return;
  CondArgs a = args;
  hipLaunchKernelGGL(k_gbx_thermo, dim3(stride_blocks(a.ngbx)), dim3(256), 0, c->stream, a.st, a.ngbx, a.gs);
  const unsigned nb = gbx_blocks(a.ngbx, a.wpb, c->tiles.cond_grid_waves);
  a.stat = stat_slots(c, static_cast<uint64_t>(nb) * a.wpb);
  kev_mark(c, 0);
  if (a.fuse_qcond)
    hipLaunchKernelGGL(k_condensation<true>, dim3(nb), dim3(64 * a.wpb), 0, c->stream, a);
  else
    hipLaunchKernelGGL(k_condensation<false>, dim3(nb), dim3(64 * a.wpb), 0, c->stream, a);
  kev_mark(c, 1);
  launch_stats_reduce(c, nb * a.wpb, ST_COND_DROPS, 4);
}
void launch_qcond(cleo_ctx* c, const QcondArgs& a) {
  if (a.ngbx == 0) return;
  hipLaunchKernelGGL(k_qcond, dim3(gbx_blocks((a.ngbx + kWave - 1) / kWave, 4, c->tiles.cond_grid_waves)), dim3(256), 0,
                     c->stream, a);
}
void launch_collisions(cleo_ctx* c, const CollArgs& args) {
  if (args.ngbx == 0) return;
  CollArgs a = args;
  if (c->tiles.coll_flat) {
    const unsigned nb = gbx_blocks((a.ngbx + kWave - 1) / kWave, a.wpb, c->tiles.coll_grid_waves);
    a.stat = stat_slots(c, static_cast<uint64_t>(nb) * a.wpb);
    kev_mark(c, 0);
    if (coll_is_kinetic(a.k))
      hipLaunchKernelGGL(k_coll_tiles<true>, dim3(nb), dim3(64 * a.wpb), a.wpb * kCollTileLdsBytes, c->stream, a);
    else
      hipLaunchKernelGGL(k_coll_tiles<false>, dim3(nb), dim3(64 * a.wpb), a.wpb * kCollTileLdsBytes, c->stream, a);
    kev_mark(c, 1);
    launch_stats_reduce(c, nb * a.wpb, ST_COLL_PAIRS, 3);
    return;
  }
  const unsigned nb = gbx_blocks(a.ngbx, a.wpb, c->tiles.coll_grid_waves);
  a.stat = stat_slots(c, static_cast<uint64_t>(nb) * a.wpb);
  const size_t lds = static_cast<size_t>(a.wpb) * 2u * kCollLdsCap * sizeof(uint32_t);
  kev_mark(c, 0);
  if (coll_is_kinetic(a.k))
    hipLaunchKernelGGL(k_collisions<true>, dim3(nb), dim3(64 * a.wpb), lds, c->stream, a);
  else
    hipLaunchKernelGGL(k_collisions<false>, dim3(nb), dim3(64 * a.wpb), lds, c->stream, a);
  kev_mark(c, 1);
  launch_stats_reduce(c, nb * a.wpb, ST_COLL_PAIRS, 3);
}
void launch_motion(cleo_ctx* c, const MotionArgs& args) {
  if (args.grid.ngbx == 0) return;
  MotionArgs a = args;
  const unsigned nb = gbx_blocks(a.grid.ngbx, a.wpb, c->tiles.motion_grid_waves);
  a.stat = stat_slots(c, static_cast<uint64_t>(nb) * a.wpb);
  hipLaunchKernelGGL(k_fill_u32, dim3(stride_blocks(28ull * a.grid.ngbx)), dim3(256), 0, c->stream, a.cnt,
                     28ull * a.grid.ngbx, 0u);
  kev_mark(c, 0);
  hipLaunchKernelGGL(k_motion, dim3(nb), dim3(64 * a.wpb), 0, c->stream, a);
  kev_mark(c, 1);
  launch_stats_reduce(c, nb * a.wpb, ST_MOTION_DROPS, 1);
}
void launch_stats_reduce(cleo_ctx* c, uint32_t nslot, int first, int ncount) {
  if (nslot == 0) return;
  const unsigned nb = static_cast<unsigned>(std::min<uint64_t>(128, (nslot + 1023) / 1024));
  hipLaunchKernelGGL(k_stats_reduce, dim3(nb, static_cast<unsigned>(ncount)), dim3(256), 0, c->stream,
                     static_cast<const uint32_t*>(c->sp), nslot, first, c->ctl);
}
StatSlots stat_slots(cleo_ctx* c, uint64_t nslot) { return StatSlots{c->sp, static_cast<uint32_t>(nslot)}; }
void launch_advance_event(cleo_ctx* c) { hipLaunchKernelGGL(k_advance_event, dim3(1), dim3(1), 0, c->stream, c->ctl); }
void launch_inj_sizes(cleo_ctx* c, uint32_t* nj, uint32_t* np) {
  hipLaunchKernelGGL(k_inj_sizes, dim3(nblocks(c->grid.ngbx, 256)), dim3(256), 0, c->stream, c->nact, c->grid.ngbx, nj, np);
}
void launch_copy_oob_region(cleo_ctx* c, int src, int dst) {
  hipLaunchKernelGGL(k_copy_oob_region, dim3(sparse_blocks(c->ntot)), dim3(256), 0, c->stream, c->buf[src],
                     c->buf[dst], c->off, c->grid.ngbx, c->ntot);
}

static size_t scan_words_rec(uint64_t len) {
  const uint64_t nb = (len + kScanTile - 1) / kScanTile;
  if (nb <= 1) return 0;
  return nb + (nb + 1) + scan_words_rec(nb);
}
size_t scan_tmp_words(uint64_t len) { return scan_words_rec(len) + 16; }

static void scan_rec(hipStream_t stream, const uint32_t* in, uint64_t len, uint32_t stride, uint32_t offset,
                     uint32_t* out, const uint32_t* flag, uint32_t* tmp) {
  if (len == 0) {
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, stream, out, 0u, flag);
    return;
  }
  const uint64_t nb = (len + kScanTile - 1) / kScanTile;
  if (nb == 1) {
    hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(kScanThreads), 0, stream, in, len, stride, offset,
                       static_cast<const uint32_t*>(nullptr), out, flag);
    return;
  }
  uint32_t* partial = tmp;
  uint32_t* pscan = tmp + nb;
  hipLaunchKernelGGL(k_scan_reduce, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, stream, in, len, stride,
                     offset, partial, flag);
  scan_rec(stream, partial, nb, 1u, 0u, pscan, flag, tmp + nb + nb + 1);
  hipLaunchKernelGGL(k_scan_apply, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, stream, in, len, stride,
                     offset, static_cast<const uint32_t*>(pscan), out, flag);
}
void scan_u32(cleo_ctx* c, const uint32_t* in, uint64_t len, uint32_t stride, uint32_t offset, uint32_t* out,
              const uint32_t* flag) {
  scan_rec(c->stream, in, len, stride, offset, out, flag, c->scan_tmp);
}

// argument block of the neighbour-sort kernels reading `src` and writing `dst`
static NsortArgs nsort_args(cleo_ctx* c, const SoA& src, const SoA& dst, const uint32_t* flag) {
  NsortArgs a{};
  a.src = src;
  a.dst = dst;
  a.off = c->off;
  a.base = c->off_alt;
  a.cnt = c->cnt;
  a.runs = c->cnt + 28ull * c->grid.ngbx;
  a.oobpre = c->oobpre;
  a.tot = c->tot;
  a.oobcnt = c->oobcnt;
  a.grid = c->grid;
  a.wpb = c->tiles.motion_waves_per_block ? c->tiles.motion_waves_per_block : 1;
  a.ctl = c->ctl;
  a.flag = flag;
  a.ntot = c->ntot;
  a.perm = c->perm;
  a.order = c->order_live ? c->order : nullptr;
  return a;
}
// the drops of the domain to their sorted places: destinations, then the records
static void nsort_move(cleo_ctx* c, const NsortArgs& a) {
  const uint32_t ngbx = c->grid.ngbx;
  launch_nsort_perm(c, a, ngbx);
  // streaming copy: one pair of positions per thread (measured faster than a capped grid), except
  // under a device flag that is mostly clear
  const unsigned nb = a.flag ? sparse_blocks(c->ntot) : motion_blocks(c->ntot, 0);
  hipLaunchKernelGGL(k_nsort_gather, dim3(nb), dim3(256), 0, c->stream, a);
}
// totals per destination, start offsets in the count table, new refs and the out-of-domain prefix
static void nsort_tables(cleo_ctx* c, const NsortArgs& a, const uint32_t* flag) {
  const uint32_t ngbx = c->grid.ngbx;
  hipLaunchKernelGGL(k_nsort_tot, dim3(nblocks(ngbx, 256)), dim3(256), 0, c->stream, a);
  scan_u32(c, c->tot, ngbx, 1u, 0u, c->off_alt, flag);
  scan_u32(c, c->oobcnt, ngbx, 1u, 0u, c->oobpre, flag);
}
// the sorted store in buffer `dst` becomes current: refs and active counts follow
static void nsort_commit(cleo_ctx* c, int dst) {
  const uint32_t ngbx = c->grid.ngbx;
  c->cur = dst;
  c->order_live = false;  // the sorted store is in its logical order
  hipLaunchKernelGGL(k_copy_u32, dim3(stride_blocks(ngbx + 1ull)), dim3(256), 0, c->stream, c->off_alt, c->off,
                     ngbx + 1ull, static_cast<const uint32_t*>(nullptr));
  hipLaunchKernelGGL(k_set_nact, dim3(nblocks(ngbx, 256)), dim3(256), 0, c->stream, c->off, ngbx, c->nact,
                     static_cast<const uint32_t*>(nullptr));
}

void neighbour_sort_to(cleo_ctx* c, const SoA& src, const SoA& dst, int commit_dst, const uint8_t* step) {
  const uint32_t ngbx = c->grid.ngbx;
  NsortArgs a = nsort_args(c, src, dst, nullptr);
  a.step = step;
  nsort_tables(c, a, nullptr);
  nsort_move(c, a);
  hipLaunchKernelGGL(k_nsort_oob, dim3(sparse_blocks(c->ntot)), dim3(256), 0, c->stream, a, c->ntot);
  hipLaunchKernelGGL(k_sorts_stat, dim3(1), dim3(1), 0, c->stream, c->ctl, static_cast<const uint32_t*>(nullptr));
  if (commit_dst >= 0) nsort_commit(c, commit_dst);
}

void neighbour_sort(cleo_ctx* c, int src, int dst, const uint32_t* flag, bool from_nact, const uint8_t* step) {
  if (flag == nullptr && !from_nact) {
    neighbour_sort_to(c, c->buf[src], c->buf[dst], dst, step);
    return;
  }
  const uint32_t ngbx = c->grid.ngbx;
  NsortArgs a = nsort_args(c, c->buf[src], c->buf[dst], flag);
  if (from_nact) {
    const unsigned nb = flag ? sparse_blocks(28ull * ngbx) : nblocks(28ull * ngbx, 256);
    hipLaunchKernelGGL(k_cnt_from_nact, dim3(nb), dim3(256), 0, c->stream, c->off, c->nact, ngbx, c->cnt, flag);
  }
  nsort_tables(c, a, flag);
  nsort_move(c, a);
  hipLaunchKernelGGL(k_nsort_oob, dim3(sparse_blocks(c->ntot)), dim3(256), 0, c->stream, a, c->ntot);
  hipLaunchKernelGGL(k_sorts_stat, dim3(1), dim3(1), 0, c->stream, c->ctl, flag);
  if (flag == nullptr) {
    nsort_commit(c, dst);
  } else {
    hipLaunchKernelGGL(k_copy_soa, dim3(sparse_blocks(c->ntot)), dim3(256), 0, c->stream, c->buf[dst], c->buf[src],
                       0ull, c->ntot, flag);
    // the copied-back store is in its logical order: if the sort ran, the order becomes the identity
    if (c->order_live)
      hipLaunchKernelGGL(k_iota_u32, dim3(sparse_blocks(c->ntot)), dim3(256), 0, c->stream, c->order, c->ntot, flag);
    hipLaunchKernelGGL(k_copy_u32, dim3(stride_blocks(ngbx + 1ull)), dim3(256), 0, c->stream, c->off_alt, c->off,
                       ngbx + 1ull, flag);
    hipLaunchKernelGGL(k_set_nact, dim3(nblocks(ngbx, 256)), dim3(256), 0, c->stream, c->off, ngbx, c->nact, flag);
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, c->stream, c->ctl);
  }
}

// A pending collision order applied: the domain's drops are gathered into the other buffer in
// their logical order (the out-of-domain drops are copied as they are) and that buffer becomes
// current.  Nothing happens when the store is already in order.
// the store as it is, with the identity as its pending order
void pending_identity_order(cleo_ctx* c) {
  hipLaunchKernelGGL(k_iota_u32, dim3(stride_blocks(c->ntot)), dim3(256), 0, c->stream, c->order, c->ntot,
                     static_cast<const uint32_t*>(nullptr));
  c->order_live = true;
}
void materialize_order(cleo_ctx* c) {
  if (!c->order_live) return;
  const int src = c->cur, dst = 1 - c->cur;
  NsortArgs a = nsort_args(c, c->buf[src], c->buf[dst], nullptr);
  a.perm = c->order;
  hipLaunchKernelGGL(k_nsort_gather, dim3(motion_blocks(c->ntot, 0)), dim3(256), 0, c->stream, a);
  launch_copy_oob_region(c, src, dst);
  c->cur = dst;
  c->order_live = false;
}

// Motion + neighbour sort in two passes over the drops: count the moves, build the tables, then
// move again and write each drop to its sorted place in `dst` (src is left unchanged).
// commit_dst >= 0: that buffer becomes current.
void motion_sort_fused(cleo_ctx* c, const MotionArgs& args, const SoA& dst, int commit_dst) {
  const uint32_t ngbx = c->grid.ngbx;
  MotionArgs a = args;
  hipLaunchKernelGGL(k_fill_u32, dim3(stride_blocks(28ull * ngbx)), dim3(256), 0, c->stream, a.cnt, 28ull * ngbx, 0u);
  const unsigned nb = motion_blocks(c->ntot, c->tiles.motion_grid_waves);
  a.stat = stat_slots(c, 4ull * nb);
  launch_motion_drops<false>(c, a, nb);
  launch_stats_reduce(c, 4u * nb, ST_MOTION_DROPS, 1);
  NsortArgs n = nsort_args(c, a.s, dst, nullptr);
  nsort_tables(c, n, nullptr);
  kev_mark(c, 0);
  hipLaunchKernelGGL(k_motion_scatter, dim3(gbx_blocks(ngbx, n.wpb, c->tiles.motion_grid_waves)), dim3(64 * n.wpb), 0,
                     c->stream, a, n);
  kev_mark(c, 1);
  hipLaunchKernelGGL(k_nsort_oob, dim3(sparse_blocks(c->ntot)), dim3(256), 0, c->stream, n, c->ntot);
  hipLaunchKernelGGL(k_sorts_stat, dim3(1), dim3(1), 0, c->stream, c->ctl, static_cast<const uint32_t*>(nullptr));
  if (commit_dst >= 0) nsort_commit(c, commit_dst);
}

// Motion + neighbour sort with the motion arithmetic in the gather (motion_flat 3): a counting
// pass records each drop's step (nothing of the store changes), the tables and the rank pass
// follow, and the gather moves each drop again as it writes it to its sorted place in `dst`.
// commit_dst >= 0: that buffer becomes current.
void motion_sort_gather(cleo_ctx* c, const MotionArgs& args, const SoA& dst, int commit_dst) {
  const uint32_t ngbx = c->grid.ngbx;
  if (ngbx == 0) return;
  MotionArgs a = args;
  a.step = c->dstep;
  hipLaunchKernelGGL(k_fill_u32, dim3(stride_blocks(28ull * ngbx)), dim3(256), 0, c->stream, a.cnt, 28ull * ngbx, 0u);
  const unsigned nb = motion_blocks(c->ntot, c->tiles.motion_grid_waves);
  a.stat = stat_slots(c, 4ull * nb);
  kev_mark(c, 0);
  launch_motion_drops<false>(c, a, nb);
  kev_mark(c, 1);
  launch_stats_reduce(c, 4u * nb, ST_MOTION_DROPS, 1);
  NsortArgs n = nsort_args(c, a.s, dst, nullptr);
  n.step = c->dstep;
  nsort_tables(c, n, nullptr);
  launch_nsort_perm(c, n, ngbx);
  launch_gather_move(c, n, a);
  hipLaunchKernelGGL(k_nsort_oob, dim3(sparse_blocks(c->ntot)), dim3(256), 0, c->stream, n, c->ntot);
  hipLaunchKernelGGL(k_sorts_stat, dim3(1), dim3(1), 0, c->stream, c->ctl, static_cast<const uint32_t*>(nullptr));
  if (commit_dst >= 0) nsort_commit(c, commit_dst);
}

uint64_t radix_scratch_words(uint64_t n) {
  const uint64_t ntiles = (n + 255) / 256;
  return 4 * n + (32 * ntiles + 1) + scan_tmp_words(16 * ntiles);
}
RadixScratch radix_scratch_at(uint32_t* base, uint64_t n) {
  const uint64_t ntiles = (n + 255) / 256;
  RadixScratch sc{};
  sc.key[0] = base;
  sc.key[1] = base + n;
  sc.idx[0] = base + 2 * n;
  sc.idx[1] = base + 3 * n;
  sc.cnt = base + 4 * n;
  sc.scan_tmp = sc.cnt + 32 * ntiles + 1;
  return sc;
}

// stable LSD radix sort of the drops of `src` by sdgbxindex (keys >= ngbx in one bucket after every
// gridbox, as _get_count_position, sortsupers.hpp:44-48), gathered into `dst`, with off[0..ngbx]
// the first position of each bucket (find_refs, findrefs.hpp)
void radix_sort_soa(hipStream_t stream, const SoA& src, const SoA& dst, uint64_t n, uint32_t ngbx, uint32_t* off,
                    const RadixScratch& sc) {
  const uint32_t ntiles = static_cast<uint32_t>((n + 255) / 256);
  if (n == 0) {
    hipLaunchKernelGGL(k_bucket_bounds, dim3(1), dim3(256), 0, stream, static_cast<const uint32_t*>(sc.key[0]), 0ull,
                       ngbx, off);
    return;
  }
  hipLaunchKernelGGL(k_rs_init, dim3(stride_blocks(n)), dim3(256), 0, stream, src.gbx, n, ngbx, sc.key[0], sc.idx[0]);
  int passes = 1;
  while (passes < 8 && (static_cast<uint64_t>(ngbx) >> (4 * passes)) != 0) ++passes;
  int cur = 0;
  uint32_t* counts = sc.cnt;
  uint32_t* offs = sc.cnt + 16ull * ntiles;
  for (int p = 0; p < passes; ++p) {
    const int shift = 4 * p;
    hipLaunchKernelGGL(k_rs_count, dim3(ntiles), dim3(256), 0, stream, sc.key[cur], n, shift, ntiles, counts);
    scan_rec(stream, counts, 16ull * ntiles, 1u, 0u, offs, nullptr, sc.scan_tmp);
    hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(256), 0, stream, sc.key[cur], sc.idx[cur], n, shift, ntiles,
                       static_cast<const uint32_t*>(offs), sc.key[1 - cur], sc.idx[1 - cur]);
    cur = 1 - cur;
  }
  hipLaunchKernelGGL(k_gather_soa, dim3(stride_blocks(n)), dim3(256), 0, stream, src, dst,
                     static_cast<const uint32_t*>(sc.idx[cur]), n);
  hipLaunchKernelGGL(k_bucket_bounds, dim3(stride_blocks(n + 1)), dim3(256), 0, stream,
                     static_cast<const uint32_t*>(sc.key[cur]), n, ngbx, off);
}

void generic_sort(cleo_ctx* c) {
  const uint64_t n = c->ntot;
  const uint32_t ngbx = c->grid.ngbx;
  const int src = c->cur, dst = 1 - c->cur;
  RadixScratch sc{};
  for (int i = 0; i < 2; ++i) {
    sc.key[i] = c->rs_key[i];
    sc.idx[i] = c->rs_idx[i];
  }
  sc.cnt = c->rs_cnt;
  sc.scan_tmp = c->scan_tmp;
  radix_sort_soa(c->stream, c->buf[src], c->buf[dst], n, ngbx, c->off, sc);
  if (n > 0) c->cur = dst;
  hipLaunchKernelGGL(k_set_nact, dim3(nblocks(ngbx, 256)), dim3(256), 0, c->stream, c->off, ngbx, c->nact,
                     static_cast<const uint32_t*>(nullptr));
  hipLaunchKernelGGL(k_sorts_stat, dim3(1), dim3(1), 0, c->stream, c->ctl, static_cast<const uint32_t*>(nullptr));
}

}  // namespace cleo

// ============================ flat (thread-per-drop) variants =================================
namespace cleo {

// per gridbox: optional qcond recompute (ordered, before any radius changes), then the
// saturation ratio and diffusion factor condensation uses (condensation.cpp:44-56)
__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(4) k_cond_prep(CondFlatArgs a) {
  if (a.dcount && blockIdx.x == 0 && threadIdx.x == 0) *a.dcount = 0u;  // the deferred list starts empty
  __shared__ double win_all[4][kSumWin];
  double* win = win_all[threadIdx.x >> 6];
  const int lane = lane_id();
  const uint32_t ntiles = (a.ngbx + kWave - 1) / kWave;
  for (uint32_t tile = wave_index(); tile < ntiles; tile += wave_count()) {
    const uint32_t g0 = tile * kWave, g = g0 + static_cast<uint32_t>(lane);
    const bool valid = g < a.ngbx;
    double tot = 0.0;
    const bool qcond_here = a.fuse_qcond && !a.qm;  // else k_cond_state recomputes it from qm
    if (qcond_here) {
      const uint32_t t_lo = a.off[g0], t_hi = a.off[min(g0 + kWave, a.ngbx)];
      const uint32_t lo = valid ? a.off[g] : t_hi, hi = valid ? a.off[g + 1] : t_hi;
      tot = tile_sum(win, t_lo, t_hi, lo, hi - lo, [&](uint32_t k) { return qcond_term(a.s, a.sol, a.ctl, k); }, lane);
    }
    if (valid) {
      const double temp = a.st.temp[g], press = a.st.press[g], qv = a.st.qv[g];
      if (qcond_here) {
        const double rho_cond = tot / a.st.vol[g] * dimless::R0cubed_VOL0;
        a.st.qc[g] = rho_cond / dry_air_density(press, temp, qv);
      }
      if (temp <= 0) set_error(a.ctl, CLEO_ERR_TEMP_NONPOSITIVE, g, a.off[g]);
      const double psat = saturation_pressure(temp);
      a.gs[kGsStride * g] = supersaturation_ratio(press, qv, psat);
      a.gs[kGsStride * g + 1] = diffusion_factor(press, temp, psat);
      const double akoh = kohler_a(temp);
      a.gs[kGsStride * g + 2] = akoh;
      a.gs[kGsStride * g + 3] = rcp_nr(akoh);
    }
  }
}

// One active drop k of gridbox g: Koehler factors, Newton solve, radius update, mass change * xi
// into dm[k] (condensation.cpp:44-125).  With cfg.budget the solve may give up (returns false,
// nothing written) so that the drop is solved again, without a budget, by k_cond_deferred.
struct CondStats {
  uint32_t solved, iters, subs, restarts;
};
// The solve of one active drop of gridbox g from its radius r and solute mass msol: Koehler
// factors, Newton, new radius rnew, condensed-mass change * xi (dm) and condensed mass * xi before
// the solve (qm, the drop's term of the qcond recompute).  Errors and the iteration count are
// recorded under position k; xi() gives the multiplicity (read after the solve).  Returns false,
// with nothing recorded, when cfg.budget ran out: the drop is solved again later from the same r.
template <class Xi>
__device__ __forceinline__ bool cond_solve(const CondFlatArgs& a, const NewtonCfg& cfg, uint32_t g, uint64_t k, double r,
                                           double msol, Xi xi_of, double warm, CondStats& st, double& rnew, double& dm,
                                           double& qm) {
  const double s_ratio = a.gs[kGsStride * g], ffactor = a.gs[kGsStride * g + 1];
  const double akoh = a.gs[kGsStride * g + 2], bkoh = kohler_b(msol, a.sol);
  SolveOut o{0, false, false, false, false};
  const double newr =
      solve_condensation(cfg, s_ratio, akoh, bkoh, a.gs[kGsStride * g + 3], ffactor, r, warm, a.warm_limit, o);
  if (o.deferred) return false;
  bool bad = false;
  const double old_m = condensate_mass(r, msol, a.sol, &bad);
  const double xi = static_cast<double>(xi_of());
  qm = old_m * xi;
  rnew = at_least_dry(newr, msol, a.sol);
  dm = (condensate_mass(rnew, msol, a.sol, &bad) - old_m) * xi;
  if (bad) set_error(a.ctl, CLEO_ERR_NEG_CONDENSATE, g, k);
  if (o.maxed) set_error(a.ctl, CLEO_ERR_NEWTON_MAXITERS, g, k);
  if (a.niters) a.niters[k] = o.iters;
  st.iters += o.iters;
  st.subs += o.substep ? 1u : 0u;
  st.restarts += o.restarted ? 1u : 0u;
  st.solved += 1;
  return true;
}
// drop k of the store: its solve, the new radius in place, dm[k] and qm[k]
__device__ __forceinline__ bool cond_drop(const CondFlatArgs& a, const NewtonCfg& cfg, uint32_t g, uint64_t k,
                                          double warm, CondStats& st) {
  double rnew, dm, qm;
  if (!cond_solve(a, cfg, g, k, a.s.r[k], a.s.msol[k], [&] { return a.s.xi[k]; }, warm, st, rnew, dm, qm))
    return false;
  a.s.r[k] = rnew;
  a.dm[k] = dm;
  if (a.qm) a.qm[k] = qm;
  return true;
}
__device__ __forceinline__ void put_cond_stats(const StatSlots& sl, CondStats st) {
  st.solved = wave_sum_u32(st.solved);
  st.iters = wave_sum_u32(st.iters);
  st.subs = wave_sum_u32(st.subs);
  st.restarts = wave_sum_u32(st.restarts);
  if (lane_id() == 0) {
    const uint32_t slot = wave_index();
    put_stat(sl, 0, slot, st.solved);
    put_stat(sl, 1, slot, st.iters);
    put_stat(sl, 2, slot, st.subs);
    put_stat(sl, 3, slot, st.restarts);
  }
}

// the learned guess z = r^2 exp(2 y) of every in-domain drop into a.warm_buf (NaN for inactive
// drops; also into a.guess_out when recording), 64 drops per network pass; it runs after
// k_cond_prep, whose saturation ratio it reads
__global__ void __launch_bounds__(256) k_cond_guess(CondFlatArgs a) {
  __shared__ InferConst K;
  infer_const_load(K, a.W.p);
  __syncthreads();
  const InferFrags F = infer_frags(a.W);
  const uint32_t ndom = a.off[a.ngbx];
  // wave-uniform loop: every lane of a wave runs the network on each chunk of 64 drops
  for (uint64_t k0 = static_cast<uint64_t>(blockIdx.x) * blockDim.x + (threadIdx.x & ~63u); k0 < ndom;
       k0 += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t k = k0 + lane_id();
    bool valid = false;
    uint32_t g = 0;
    float f[kIn] = {0, 0, 0, 0, 0, 0};
    double r = 0.0;
    if (k < ndom) {
      g = a.s.gbx[k];
      valid = g < a.ngbx && (a.unsorted || k < a.off[g] + a.nact[g]);
      if (valid) {
        r = a.s.r[k];
        drop_features(r, a.s.msol[k], a.gs[kGsStride * g], a.st.temp[g], a.st.press[g], a.cfg.delt, f);
      }
    }
    float xn[kIn];
    infer_normalise(K, f, valid, xn);
    const float yn = mlp_infer(F, K, xn);
    if (k < ndom) {
      const double z = valid ? learned_z(K, r, yn) : __longlong_as_double(0x7FF8000000000000ll);
      a.warm_buf[k] = z;
      if (a.guess_out && valid) a.guess_out[k] = z;
    }
  }
}

// Deferred drops of a wave are first listed in the wave's own segment of a.dseg (`dsegcap` each,
// no contention); defer_drop appends those of the wave's deferring lanes (position q), and a
// deferring lane whose drop finds the segment full appends it to the shared list at once, with
// an atomic of its own.  The shared list holds dcap >= ntot entries and a launch defers each drop
// at most once, so every deferred drop is listed (an overflow would be recorded as a bounds
// error).  Every lane of the wave calls it.
__device__ __forceinline__ void defer_drop(const CondFlatArgs& a, uint32_t* seg, uint32_t& ndef, bool deferred,
                                           uint32_t q) {
  const uint64_t m = __ballot(deferred);
  if (!m) return;
  const uint32_t idx = ndef + static_cast<uint32_t>(__popcll(m & lanemask_lt()));
  ndef = min(a.dsegcap, ndef + static_cast<uint32_t>(__popcll(m)));
  if (!deferred) return;
  if (idx < a.dsegcap) {
    seg[idx] = q;
    return;
  }
  const uint32_t j = atomicAdd(a.dcount, 1u);
  if (j < a.dcap)
    a.dlist[j] = q;
  else
    set_error(a.ctl, CLEO_ERR_BOUNDS, kOOB, q);
}
// At the end of the kernel the wave moves its segment to the shared list with one atomic, so that
// k_cond_deferred finds the deferred drops of all waves packed 64 to a wave.  The shared list has
// room for every segment, so nothing is lost.
__device__ __forceinline__ void defer_flush(const CondFlatArgs& a, const uint32_t* seg, uint32_t ndef, int lane) {
  if (ndef == 0) return;
  wave_sync();  // the segment's entries were written by other lanes of this wave
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(a.dcount, ndef);
  base = readlane_u32(base, 0);
  for (uint32_t i = static_cast<uint32_t>(lane); i < ndef; i += kWave) {
    const uint64_t j = static_cast<uint64_t>(base) + i;
    if (j < a.dcap) a.dlist[j] = seg[i];
  }
}

// per active drop, one thread each.  Drops whose solve outlives cfg.budget iterations (a few per
// cent; their lanes would hold the other 63 of the wave) go to the deferred list, up to `dcap`
// of them, and are solved by k_cond_deferred, 64 to a wave.
__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(4) k_cond_drops(CondFlatArgs a) {
  const uint32_t ndom = a.off[a.ngbx];
  const int lane = lane_id();
  CondStats st{0, 0, 0, 0};
  uint32_t ndef = 0;
  uint32_t* seg = a.dseg + static_cast<uint64_t>(wave_index()) * a.dsegcap;
  for (uint64_t k0 = static_cast<uint64_t>(blockIdx.x) * blockDim.x; k0 < ndom;
       k0 += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t k = k0 + threadIdx.x;
    bool active = false;
    uint32_t g = 0;
    if (k < ndom) {
      g = a.s.gbx[k];
      active = g < a.ngbx && k < a.off[g] + a.nact[g];
    }
    double warm = __longlong_as_double(0x7FF8000000000000ll);
    if (active && a.warm) warm = a.warm[k];
    bool deferred = false;
    if (active) deferred = !cond_drop(a, a.cfg, g, k, warm, st);
    defer_drop(a, seg, ndef, deferred, static_cast<uint32_t>(k));
  }
  if (a.dcount) defer_flush(a, seg, ndef, lane);
  put_cond_stats(a.stat, st);
}

// the deferred drops, thread i taking entry i of the list (grid-stride), solved without a budget;
// every wave writes its statistics slot
__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(4) k_cond_deferred(CondFlatArgs a) {
  CondStats st{0, 0, 0, 0};
  NewtonCfg nobudget = a.cfg;
  nobudget.budget = 0;
  const uint64_t n = min(static_cast<uint64_t>(*a.dcount), static_cast<uint64_t>(a.dcap));
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t k = a.dlist[i];
    const double warm =
        a.warm ? a.warm[a.warm_perm ? a.warm_perm[k] : k] : __longlong_as_double(0x7FF8000000000000ll);
    cond_drop(a, nobudget, a.s.gbx[k], k, warm, st);
  }
  put_cond_stats(a.stat, st);
}

// Condensation in the sort's gather, for a condensation that follows motion at once: destination
// q takes the drop at position perm[q] of the unsorted store and, when it stays in the domain,
// solves it before writing its record (with the new radius) and dm[q], qm[q] to the sorted store.
// The solve does not depend on the order of the drops; the per-gridbox sums run afterwards on the
// sorted store (k_cond_state).  Drops that outlive cfg.budget iterations are written unsolved and
// listed by their sorted position for k_cond_deferred, as in k_cond_drops.  The memory traffic of
// the gather and the FP64 arithmetic of the solves overlap in one pass.
// kMove: the unsorted store holds the drops before motion (the counting pass k_motion_drops<false>
// only recorded their steps); each drop moves here, with the same arithmetic, before its solve in
// the gridbox it reaches, so motion writes nothing of its own.  kLds: the coordinate bounds are read
// from an LDS copy (dynamic shared memory of nz + nx + ny + 3 doubles).
template <bool kMove, bool kLds>
__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(kMove ? 4 : 5) k_sort_gather_cond(NsortArgs n, CondFlatArgs a, MotionArgs m) {
  extern __shared__ double bounds_lds[];
  MotionArgs b = m;
  if constexpr (kMove && kLds) {
    const uint32_t nzh = m.grid.nz + 1, nxh = m.grid.nx + 1, nyh = m.grid.ny + 1;
    for (uint32_t i = threadIdx.x; i < nzh + nxh + nyh; i += blockDim.x)
      bounds_lds[i] = i < nzh ? m.grid.zh[i] : (i < nzh + nxh ? m.grid.xh[i - nzh] : m.grid.yh[i - nzh - nxh]);
    __syncthreads();
    b.grid.zh = bounds_lds;
    b.grid.xh = bounds_lds + nzh;
    b.grid.yh = bounds_lds + nzh + nxh;
  }
  const uint64_t nsrc = n.off[n.grid.ngbx];
  const int lane = lane_id();
  CondStats st{0, 0, 0, 0};
  uint32_t ndef = 0;
  uint32_t* seg = a.dseg + static_cast<uint64_t>(wave_index()) * a.dsegcap;
  const double nowarm = __longlong_as_double(0x7FF8000000000000ll);
  // Each block takes one contiguous range of destinations, in block-uniform chunks (every wave
  // of the block stays in the loop for the ballots below).  A drop that changed gridbox is read
  // from lines that the neighbouring destinations' drops also read; with contiguous ranges those
  // readers run on the same CU at about the same time, and with the ranges in xcd_block() order
  // the neighbouring ranges run on the same XCD, behind one L2 (measured 4.82 -> 4.61 -> 4.49 ms
  // against chunks strided over the grid).
  const uint64_t per = ((nsrc + gridDim.x - 1) / gridDim.x + blockDim.x - 1) / blockDim.x * blockDim.x;
  const uint64_t c0 = static_cast<uint64_t>(xcd_block()) * per, c1 = min(c0 + per, nsrc);
  for (uint64_t q0 = c0; q0 < c1; q0 += blockDim.x) {
    const uint64_t q = q0 + threadIdx.x;
    const bool inq = q < nsrc;
    const uint32_t p = inq ? n.perm[q] : 0u;
    const bool okp = inq && p < nsrc;
    if (inq && !okp) set_error(n.ctl, CLEO_ERR_SORT_NONLOCAL, kOOB, q);
    const uint32_t sp = okp ? p : 0u;
    uint32_t g = n.src.gbx[sp];
    double c3 = n.src.c3[sp], c1 = n.src.c1[sp], c2 = n.src.c2[sp];
    const uint64_t xi = n.src.xi[sp], id = n.src.id[sp];
    const double r = n.src.r[sp], msol = n.src.msol[sp];
    if constexpr (kMove) {
      // an in-domain drop's gridbox is below ngbx; the guard keeps a corrupted one from indexing
      const uint32_t gg = g < n.grid.ngbx ? g : 0u;
      const GbxGeom G = gbx_geom(b, gg);
      const Moved M = move_drop(b, gg, G, DropPos{c3, c1, c2, r}, sp, false);
      c3 = M.c3;
      c1 = M.c1;
      c2 = M.c2;
      g = M.idx;
    }
    const bool active = okp && g < a.ngbx;
    const double warm = a.warm ? a.warm[sp] : nowarm;  // learned guesses by source position
    // non-temporal stores: the sorted store and dm, qm are read back only after the whole pass,
    // from HBM either way, and streaming them past the L2 leaves it to the source lines that
    // neighbouring destinations still read (5.31 -> 5.11 ms).  Every field but the radius is
    // final before the solve and is written first, so its registers are free during the solve.
    if (okp) {
      __builtin_nontemporal_store(g, &n.dst.gbx[q]);
      __builtin_nontemporal_store(c3, &n.dst.c3[q]);
      __builtin_nontemporal_store(c1, &n.dst.c1[q]);
      __builtin_nontemporal_store(c2, &n.dst.c2[q]);
      __builtin_nontemporal_store(xi, &n.dst.xi[q]);
      __builtin_nontemporal_store(msol, &n.dst.msol[q]);
      __builtin_nontemporal_store(id, &n.dst.id[q]);
    }
    double rnew = r, dm = 0.0, qm = 0.0;
    bool solved = false;
    if (active) solved = cond_solve(a, a.cfg, g, q, r, msol, [&] { return xi; }, warm, st, rnew, dm, qm);
    defer_drop(a, seg, ndef, active && !solved, static_cast<uint32_t>(q));
    if (okp) __builtin_nontemporal_store(solved ? rnew : r, &n.dst.r[q]);
    if (solved) {
      __builtin_nontemporal_store(dm, &a.dm[q]);
      if (a.qm) __builtin_nontemporal_store(qm, &a.qm[q]);
    }
  }
  if (a.dcount) defer_flush(a, seg, ndef, lane);
  put_cond_stats(a.stat, st);
}

// ---- condensation with a per-wave task ring ------------------------------------------------------
// Newton iteration counts differ from drop to drop (2 for about a quarter, 5-6 for most, tens for
// a few), so 64 drops solved in lockstep leave most lanes idle.  Here each lane runs the solve of
// its drop as a state machine, one g-evaluation per tick, and a lane whose drop is done takes the
// next prepared task from a ring in LDS; a setup pass prepares 64 drops at a time whenever fewer
// than 64 tasks wait.  Every drop goes through exactly the operations of solve_condensation
// (initial evaluation, two iterations, restart until converged, sub-steps, learned guess), so the
// result does not depend on the scheduling.
struct CondRing {
  double s_ratio[kRing], akoh[kRing], bkoh[kRing], inv_ffv[kRing], r0[kRing], zplain[kRing], zwarm[kRing],
      mindelt[kRing], dryr[kRing], oldm[kRing], msol[kRing];
  uint64_t xi[kRing];
  uint32_t k[kRing], unique[kRing];
};
static_assert(sizeof(CondRing) == kCondRingBytes, "ring size");

enum : uint32_t { PH_EVAL = 0, PH_FIRST = 1, PH_SECOND = 2, PH_UNTIL = 3 };
enum : uint32_t { MODE_PLAIN = 0, MODE_WARM = 1, MODE_SUB = 2 };

__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(3) k_cond_queue(CondFlatArgs a) {
  extern __shared__ uint32_t lds_dyn[];
  CondRing& R = reinterpret_cast<CondRing*>(lds_dyn)[threadIdx.x >> 6];
  const int lane = lane_id();
  const NewtonCfg cfg = a.cfg;
  const uint64_t ndom = a.off[a.ngbx];
  const uint64_t nw = wave_count(), w = wave_index();
  const uint64_t k_hi = ndom * (w + 1) / nw;
  uint64_t next = ndom * w / nw;  // first drop not yet prepared (wave-uniform)
  uint32_t head = 0, tail = 0;    // ring positions handed out / prepared (wave-uniform)
  // lane state
  bool busy = false, restarted = false, maxed = false;
  uint32_t k = 0, phase = PH_EVAL, mode = MODE_PLAIN, niter = 0, limit = 0, iters = 0;
  ODEConsts c{0, 0, 0, 0, 0, 0, 0};
  double rprev = 0, r0 = 0, subdelt = 0, remdelt = 0, mindelt = 0, zplain = 0;
  double z = 0, gz = 0, dgz = 0, z0 = 0, g0 = 0, dg0 = 0, dryr = 0, oldm = 0, msol = 0;
  uint64_t xi = 0;
  uint32_t st_solved = 0, st_iters = 0, st_subs = 0, st_restarts = 0;
  while (true) {
    if (tail - head < static_cast<uint32_t>(kWave) && next < k_hi) {
      // setup pass: Koehler factors, ventilation, initial guess and uniqueness criteria of the next
      // 64 drops (condensation.cpp:44-56, impliciteuler.cpp:35-54)
      const uint64_t kk = next + static_cast<uint64_t>(lane);
      bool act = false;
      uint32_t g = 0;
      if (kk < k_hi) {
        g = a.s.gbx[kk];
        act = g < a.ngbx && kk < a.off[g] + a.nact[g];
      }
      const uint64_t m = __ballot(act);
      if (act) {
        const uint32_t slot = (tail + static_cast<uint32_t>(__popcll(m & lanemask_lt()))) % kRing;
        const double s_ratio = a.gs[kGsStride * g], ffactor = a.gs[kGsStride * g + 1];
        const double r = a.s.r[kk], ms = a.s.msol[kk];
        bool bad = false;
        const double om = condensate_mass(r, ms, a.sol, &bad);
        if (bad) set_error(a.ctl, CLEO_ERR_NEG_CONDENSATE, g, kk);
        const double akoh = a.gs[kGsStride * g + 2], bkoh = kohler_b(ms, a.sol);
        ODEConsts cc{s_ratio, akoh, bkoh, 0.0, 0.0, 0.0, a.gs[kGsStride * g + 3]};
        ffactor_over_vent(ffactor, r, cc.ffv_num, cc.ffv_den, cc.inv_ffv);
        double zg;
        bool uniq;
        ode_setup(cc, r, cfg.delt, zg, uniq);
        const double critdelt = uniq ? 0.0 : critical_timestep(cc);
        R.s_ratio[slot] = s_ratio;
        R.akoh[slot] = akoh;
        R.bkoh[slot] = bkoh;
        R.inv_ffv[slot] = cc.inv_ffv;
        R.r0[slot] = r;
        R.zplain[slot] = zg;
        double zw = a.warm ? a.warm[kk] : __longlong_as_double(0x7FF8000000000000ll);
        if (uniq && zw == zw) {  // the learned path starts from the plain guess where that is the root
          double gp, dg;       // to the tolerance (solve_condensation)
          ode_g_dg(cc, cfg.delt, r, zg, gp, dg);
          if (fabs(gp) < cfg.atol) zw = zg;
        }
        R.zwarm[slot] = zw;
        R.mindelt[slot] = uniq ? 0.0 : fmax(critdelt, cfg.minsubdelt);
        R.dryr[slot] = dryradius(ms, a.sol);
        R.oldm[slot] = om;
        R.msol[slot] = ms;
        R.xi[slot] = a.s.xi[kk];
        R.k[slot] = static_cast<uint32_t>(kk);
        R.unique[slot] = uniq ? 1u : 0u;
      }
      tail += static_cast<uint32_t>(__popcll(m));
      next += kWave;
      wave_sync();
    }
    // hand the waiting tasks to idle lanes
    const uint64_t midle = __ballot(!busy);
    const uint32_t avail = tail - head;
    if (!busy) {
      const uint32_t rank = static_cast<uint32_t>(__popcll(midle & lanemask_lt()));
      if (rank < avail) {
        const uint32_t slot = (head + rank) % kRing;
        c = ODEConsts{R.s_ratio[slot], R.akoh[slot], R.bkoh[slot], 0.0, 0.0, R.inv_ffv[slot], 0.0};
        r0 = R.r0[slot];
        zplain = R.zplain[slot];
        mindelt = R.mindelt[slot];
        dryr = R.dryr[slot];
        oldm = R.oldm[slot];
        msol = R.msol[slot];
        xi = R.xi[slot];
        k = R.k[slot];
        const double zw = R.zwarm[slot];
        rprev = r0;
        iters = 0;
        restarted = maxed = false;
        if (R.unique[slot]) {
          subdelt = cfg.delt;
          if (zw == zw) {
            mode = MODE_WARM;
            limit = a.warm_limit;
            z = zw;
          } else {
            mode = MODE_PLAIN;
            limit = cfg.maxniters;
            z = zplain;
          }
        } else {
          mode = MODE_SUB;
          limit = cfg.maxniters;
          remdelt = cfg.delt;
          subdelt = fmin(mindelt, remdelt);
          z = zplain;
        }
        phase = PH_EVAL;
        busy = true;
      }
    }
    head += min(static_cast<uint32_t>(__popcll(midle)), avail);
    wave_sync();
    if (!__any(busy)) {
      if (next >= k_hi && head == tail) break;
      continue;
    }
    if (busy) {
      // one evaluation of g: the initial state of a segment, or a Newton iteration
      // (newton_iterate, impliciteuler.cpp:240-255)
      double zn = z;
      if (phase != PH_EVAL) zn = fmax(z * (1 - quotient(gz, dgz)), 1e-8);  // as newton_iterate
      double g1, dg1;
      ode_g_dg(c, subdelt, rprev, zn, g1, dg1);
      bool segok = false, segfail = false;
      if (phase == PH_EVAL) {
        z0 = z = zn;
        g0 = gz = g1;
        dg0 = dgz = dg1;
        if (mode == MODE_WARM) {  // from the learned guess: iterate until converged at once
          phase = PH_UNTIL;
          niter = 1;
          segfail = niter > limit;
        } else {
          phase = PH_FIRST;
        }
      } else {
        ++iters;
        const bool conv = fabs(g1 - gz) < cfg.rtol * fabs(g1) + cfg.atol;
        z = zn;
        gz = g1;
        dgz = dg1;
        if (phase == PH_FIRST) {
          phase = PH_SECOND;
        } else if (phase == PH_SECOND) {  // integrate_condensation_ode: restart unless converged
          if (conv) {
            segok = true;
          } else {
            z = z0;
            gz = g0;
            dgz = dg0;
            phase = PH_UNTIL;
            niter = 1;
            segfail = niter > limit;
          }
        } else {  // newtonraphson_untilconverged (impliciteuler.cpp:207-227)
          ++niter;
          if (conv)
            segok = true;
          else
            segfail = niter > limit;
        }
      }
      if (segfail) {
        if (mode == MODE_WARM) {  // the learned guess failed: the reference path from its guess
          restarted = true;
          mode = MODE_PLAIN;
          limit = cfg.maxniters;
          subdelt = cfg.delt;
          rprev = r0;
          z = zplain;
          phase = PH_EVAL;
        } else {
          maxed = true;
          segok = true;
        }
      }
      if (segok) {
        bool done = true;
        if (mode == MODE_SUB) {  // substepping loop of solve_condensation
          rprev = sqrt(z);
          remdelt -= subdelt;
          if (remdelt > 0.0) {
            subdelt = fmin(mindelt, remdelt);
            phase = PH_EVAL;
            done = false;
            if (!(subdelt > 0.0)) {  // no progress possible: stop with the reference's substep error
              set_error(a.ctl, CLEO_ERR_SUBSTEP, a.s.gbx[k], k);
              done = true;
            }
          }
        }
        if (done) {
          const double rnew = fmax(sqrt(z), dryr);  // change_radius (superdrop_attrs.cpp:35-45)
          a.s.r[k] = rnew;
          bool bad = false;
          a.dm[k] = (condensate_mass(rnew, msol, a.sol, &bad) - oldm) * static_cast<double>(xi);
          if (bad) set_error(a.ctl, CLEO_ERR_NEG_CONDENSATE, a.s.gbx[k], k);
          if (maxed) set_error(a.ctl, CLEO_ERR_NEWTON_MAXITERS, a.s.gbx[k], k);
          if (a.niters) a.niters[k] = iters;
          st_solved += 1;
          st_iters += iters;
          st_subs += mode == MODE_SUB ? 1u : 0u;
          st_restarts += restarted ? 1u : 0u;
          busy = false;
        }
      }
    }
  }
  st_solved = wave_sum_u32(st_solved);
  st_iters = wave_sum_u32(st_iters);
  st_subs = wave_sum_u32(st_subs);
  st_restarts = wave_sum_u32(st_restarts);
  if (lane == 0) {
    put_stat(a.stat, 0, wave_index(), st_solved);
    put_stat(a.stat, 1, wave_index(), st_iters);
    put_stat(a.stat, 2, wave_index(), st_subs);
    put_stat(a.stat, 3, wave_index(), st_restarts);
  }
}

// per gridbox: Σ dm in drop order and the thermodynamic feedback (condensation.cpp:126-137).
// With a.qm, first the qcond recompute after motion from the drops' terms before the solve (the
// same ordered sum and formula as k_cond_prep's; State is unchanged until here).
__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(4) k_cond_state(CondFlatArgs a) {
  __shared__ double win_all[4][kSumWin];
  double* win = win_all[threadIdx.x >> 6];
  const int lane = lane_id();
  const uint32_t ntiles = (a.ngbx + kWave - 1) / kWave;
  for (uint32_t tile = wave_index(); tile < ntiles; tile += wave_count()) {
    const uint32_t g0 = tile * kWave, g = g0 + static_cast<uint32_t>(lane);
    const bool valid = g < a.ngbx;
    const uint32_t t_lo = a.off[g0], t_hi = a.off[min(g0 + kWave, a.ngbx)];
    const uint32_t lo = valid ? a.off[g] : t_hi, n = valid ? a.nact[g] : 0u;
    double qc_new = 0.0, tot = 0.0;
    if (a.qm) {
      double totq = 0.0;
      tile_sum2(win, t_lo, t_hi, lo, n, [&](uint32_t k) { return a.qm[k]; }, [&](uint32_t k) { return a.dm[k]; },
                lane, totq, tot);
      if (valid) {
        const double rho_cond = totq / a.st.vol[g] * dimless::R0cubed_VOL0;
        qc_new = rho_cond / dry_air_density(a.st.press[g], a.st.temp[g], a.st.qv[g]);
        if (!a.alter) a.st.qc[g] = qc_new;
      }
    } else {
      tot = tile_sum(win, t_lo, t_hi, lo, n, [&](uint32_t k) { return a.dm[k]; }, lane);
    }
    if (valid && a.alter) {
      const double temp = a.st.temp[g], press = a.st.press[g];
      double qv = a.st.qv[g], qc = a.qm ? qc_new : a.st.qc[g];
      const double totrho = tot / a.st.vol[g] * dimless::R0cubed_VOL0;
      const double dq = totrho / dry_air_density(press, temp, qv);
      qv -= dq;
      qc += dq;
      a.st.temp[g] = temp + (dimless::Latent_v / moist_specific_heat(qv, qc)) * dq;
      a.st.qv[g] = qv;
      a.st.qc[g] = qc;
    }
  }
}

// ============================ coupling ========================================================
// FromFileComms::update_gridbox_state for every gridbox (fromfilecomms.cpp:51-62): thermodynamics
// at the gridbox, winds at its lower and upper faces (fromfile_cartesian_dynamics.cpp:150-215)
__global__ void k_receive_dynamics(GState st, Grid gr, cleo_dynamics_step dyn) {
  const uint32_t nz = gr.nz, nx = gr.nx, ngbx = gr.ngbx;
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < ngbx; g += gridDim.x * blockDim.x) {
    const uint32_t k = g % nz, i = (g / nz) % nx, j = g / (nz * nx);
    st.press[g] = dyn.press[g];
    st.temp[g] = dyn.temp[g];
    st.qv[g] = dyn.qvap[g];
    st.qc[g] = dyn.qcond[g];
    double wl = 0.0, wh = 0.0, ul = 0.0, uh = 0.0, vl = 0.0, vh = 0.0;
    if (dyn.wvel_zfaces) {
      const uint64_t l = static_cast<uint64_t>(nx) * (nz + 1) * j + static_cast<uint64_t>(nz + 1) * i + k;
      wl = dyn.wvel_zfaces[l];
      wh = dyn.wvel_zfaces[l + 1];
    }
    if (dyn.uvel_xfaces) {
      const uint64_t l = static_cast<uint64_t>(nx + 1) * nz * j + static_cast<uint64_t>(nz) * i + k;
      ul = dyn.uvel_xfaces[l];
      uh = dyn.uvel_xfaces[l + nz];
    }
    if (dyn.vvel_yfaces) {
      vl = dyn.vvel_yfaces[g];
      vh = dyn.vvel_yfaces[static_cast<uint64_t>(g) + static_cast<uint64_t>(nx) * nz];
    }
    st.wlo[g] = wl;
    st.whi[g] = wh;
    st.ulo[g] = ul;
    st.uhi[g] = uh;
    st.vlo[g] = vl;
    st.vhi[g] = vh;
  }
}
void launch_receive_dynamics(cleo_ctx* c, const cleo_dynamics_step& dyn) {
  if (c->grid.ngbx == 0) return;
  hipLaunchKernelGGL(k_receive_dynamics, dim3(stride_blocks(c->grid.ngbx)), dim3(256), 0, c->stream, c->st, c->grid,
                     dyn);
}

// ============================ observers =======================================================
// Mass moments per gridbox (observers/massmoments_observer.cpp:23-73): one lane per gridbox of a
// 64-gridbox tile.  The tile's drops pass through an LDS window with coalesced loads (the next
// window's loads in flight while the current one is added); the loading lane turns each drop into
// its three terms (xi; (float)(xi m); (float)(xi m m), m the drop mass,
// zero for a drop below the raindrop radius when kRain), and each lane then adds its gridbox's
// terms left to right, so the float sums are the reference's serial ones.
constexpr uint32_t kMomWin = 512;
template <bool kRain>
__global__ void __launch_bounds__(256) k_mass_moments(SoA s, const uint32_t* off, uint32_t ngbx, Solute sol,
                                                      DevCtl* ctl, uint64_t* mom0, float* mom1, float* mom2) {
  __shared__ uint64_t w0_all[4][kMomWin];
  __shared__ float w1_all[4][kMomWin], w2_all[4][kMomWin];
  const uint32_t wv = threadIdx.x >> 6;
  uint64_t* w0 = w0_all[wv];
  float* w1 = w1_all[wv];
  float* w2 = w2_all[wv];
  const int lane = lane_id();
  constexpr double rlim(40e-6 / dimless::R0);  // calculate_rainmassmoments: raindrop radius
  const uint32_t ntiles = (ngbx + kWave - 1) / kWave;
  for (uint32_t tile = wave_index(); tile < ntiles; tile += wave_count()) {
    const uint32_t g0 = tile * kWave, g = g0 + static_cast<uint32_t>(lane);
    const bool valid = g < ngbx;
    const uint32_t t_lo = off[g0], t_hi = off[min(g0 + kWave, ngbx)];
    const uint32_t lo = valid ? off[g] : t_hi, hi = valid ? off[g + 1] : t_hi;
    uint64_t m0 = 0;
    float m1 = 0.f, m2 = 0.f;
    if (t_lo >= t_hi) {
      if (valid) {
        mom0[g] = 0;
        mom1[g] = 0.f;
        mom2[g] = 0.f;
      }
      continue;
    }
    // the next window's drops are loaded into registers while the lanes add the current window
    // (every lane loads, with indices clamped into the tile, so the loads stay in flight)
    constexpr uint32_t kPer = kMomWin / kWave;
    uint64_t rx[kPer];
    double rr[kPer], rm[kPer];
    auto load = [&](uint32_t w) {
      for (uint32_t i = 0; i < kPer; ++i) {
        const uint32_t k = min(w + static_cast<uint32_t>(lane) + i * kWave, t_hi - 1u);
        rx[i] = __builtin_nontemporal_load(&s.xi[k]);  // read once (0.63 -> 0.60 ms)
        rr[i] = __builtin_nontemporal_load(&s.r[k]);
        rm[i] = __builtin_nontemporal_load(&s.msol[k]);
      }
    };
    load(t_lo);
    for (uint32_t w = t_lo; w < t_hi; w += kMomWin) {
      const uint32_t we = min(w + kMomWin, t_hi);
      for (uint32_t i = 0; i < kPer; ++i) {
        const uint32_t j = static_cast<uint32_t>(lane) + i * kWave;
        const uint64_t xi = rx[i];
        const double r = rr[i];
        if (j < we - w && xi == ~0ull) set_error(ctl, CLEO_ERR_XI_OVERFLOW, s.gbx[w + j], w + j);
        const bool in = !kRain || r >= rlim;
        const double m = drop_mass(r, rm[i], sol);
        const double xm = static_cast<double>(xi) * m;
        w0[j] = in ? xi : 0ull;
        w1[j] = in ? static_cast<float>(xm) : 0.f;
        w2[j] = in ? static_cast<float>(xm * m) : 0.f;
      }
      wave_sync();
      if (we < t_hi) load(we);
      const uint32_t b0 = max(lo, w), b1 = min(hi, we);
      for (uint32_t k = b0; k < b1; ++k) {
        m0 += w0[k - w];
        m1 += w1[k - w];
        m2 += w2[k - w];
      }
      wave_sync();
    }
    if (valid) {
      mom0[g] = m0;
      mom1[g] = m1;
      mom2[g] = m2;
    }
  }
}
void launch_mass_moments(cleo_ctx* c, bool rain, uint64_t* mom0, float* mom1, float* mom2) {
  const uint32_t ngbx = c->grid.ngbx;
  if (ngbx == 0) return;
  materialize_order(c);  // the sums follow the drops' logical order
  const uint64_t ntiles = (ngbx + kWave - 1) / kWave;
  const unsigned nb = gbx_blocks(ntiles, 4, 0);
  if (rain)
    hipLaunchKernelGGL(k_mass_moments<true>, dim3(nb), dim3(256), 0, c->stream, c->buf[c->cur], c->off, ngbx,
                       c->solute, c->ctl, mom0, mom1, mom2);
  else
    hipLaunchKernelGGL(k_mass_moments<false>, dim3(nb), dim3(256), 0, c->stream, c->buf[c->cur], c->off, ngbx,
                       c->solute, c->ctl, mom0, mom1, mom2);
}

// the thermodynamic functions of k_cond_prep and k_condensation at n states, thread per state
__global__ void k_thermo_eval(uint64_t n, const double* temp, const double* press, const double* qvap,
                              double* psat, double* s_ratio, double* ffactor) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const double ps = saturation_pressure(temp[i]);
    psat[i] = ps;
    s_ratio[i] = supersaturation_ratio(press[i], qvap[i], ps);
    ffactor[i] = diffusion_factor(press[i], temp[i], ps);
  }
}
void launch_thermo_eval(cleo_ctx* c, uint64_t n, const double* temp, const double* press, const double* qvap,
                        double* psat, double* s_ratio, double* ffactor) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_thermo_eval, dim3(stride_blocks(n)), dim3(256), 0, c->stream, n, temp, press, qvap, psat,
                     s_ratio, ffactor);
}

// p[i] = i for i < n, skipped unless *flag when flag is non-null
__global__ void k_iota_u32(uint32_t* p, uint64_t n, const uint32_t* flag) {
  if (flag && *flag == 0u) return;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = static_cast<uint32_t>(i);
}

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

// count-table key of a moved drop: 28*d + step for destination d, 28*g + 27 when it left the domain
__device__ __forceinline__ uint32_t motion_key(const Moved& M, uint32_t g) {
  return M.idx != kOOB ? 28u * M.idx + static_cast<uint32_t>(M.dir) : 28u * g + 27u;
}
// adds the wave's drops per distinct key to the count table (every lane of the wave calls it)
__device__ __forceinline__ void count_keys(uint32_t* cnt, uint32_t key, bool act, int lane) {
  uint64_t mask = __ballot(act);
  while (mask) {
    const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    const uint32_t v = readlane_u32(key, leader);
    const uint64_t m = __ballot(act && key == v);
    if (lane == leader) atomicAdd(&cnt[v], static_cast<uint32_t>(__popcll(m)));
    mask &= ~m;
  }
}
// largest coordinate-bound table (zh, xh, yh together) that the motion kernels stage in LDS
constexpr uint32_t kMotionLdsBounds = 2048;

// MoveSupersInGridboxesFunctor per drop, two drops per thread and iteration with every load of
// both issued before the arithmetic; the destination-major counts are aggregated per wave, then
// added with atomics, and the step of each drop goes to a.step when that is set.  kLds: the
// coordinate bounds are read from an LDS copy (dynamic shared memory of (nz + nx + ny + 3)
// doubles).  Without kWrite this is the counting pass of a fused motion + sort: the drops' new
// coordinates and gridboxes are not stored (the second pass computes them again).
template <bool kWrite, bool kLds>
__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(3) k_motion_drops(MotionArgs a) {
  extern __shared__ double bounds_lds[];
  const int lane = lane_id();
  MotionArgs b = a;
  if constexpr (kLds) {
    const uint32_t nzh = a.grid.nz + 1, nxh = a.grid.nx + 1, nyh = a.grid.ny + 1;
    for (uint32_t i = threadIdx.x; i < nzh + nxh + nyh; i += blockDim.x)
      bounds_lds[i] = i < nzh ? a.grid.zh[i] : (i < nzh + nxh ? a.grid.xh[i - nzh] : a.grid.yh[i - nzh - nxh]);
    __syncthreads();
    b.grid.zh = bounds_lds;
    b.grid.xh = bounds_lds + nzh;
    b.grid.yh = bounds_lds + nzh + nxh;
  }
  const uint32_t ndom = a.off[a.grid.ngbx];
  uint32_t nmoved = 0;
  // block-uniform chunks of 512 drops keep every wave in the loop together for the ballots
  for (uint64_t k0 = 2ull * blockIdx.x * blockDim.x; k0 < ndom; k0 += 2ull * gridDim.x * blockDim.x) {
    const uint64_t ka = k0 + threadIdx.x, kb = ka + blockDim.x;
    const bool acta = ka < ndom, actb = kb < ndom;
    // past the end a lane reads drop k0 (in the domain) and discards the result
    const uint64_t la = acta ? ka : k0, lb = actb ? kb : k0;
    // non-temporal loads: each drop's fields are read once here (2.21 -> 1.94 ms)
    const uint32_t ga = __builtin_nontemporal_load(&a.s.gbx[la]), gb = __builtin_nontemporal_load(&a.s.gbx[lb]);
    const DropPos pa = drop_pos_nt(a, la), pb = drop_pos_nt(a, lb);
    const GbxGeom Ga = gbx_geom(b, ga), Gb = gbx_geom(b, gb);
    const Moved Ma = move_drop(b, ga, Ga, pa, la, acta);
    const Moved Mb = move_drop(b, gb, Gb, pb, lb, actb);
    // non-temporal stores: the moved coordinates and steps are read back by the sort's passes,
    // after the whole launch (2.27 -> 2.23 ms)
    if (kWrite && acta) {
      __builtin_nontemporal_store(Ma.c3, &a.s.c3[ka]);
      __builtin_nontemporal_store(Ma.c1, &a.s.c1[ka]);
      __builtin_nontemporal_store(Ma.c2, &a.s.c2[ka]);
      __builtin_nontemporal_store(Ma.idx, &a.s.gbx[ka]);
    }
    if (kWrite && actb) {
      __builtin_nontemporal_store(Mb.c3, &a.s.c3[kb]);
      __builtin_nontemporal_store(Mb.c1, &a.s.c1[kb]);
      __builtin_nontemporal_store(Mb.c2, &a.s.c2[kb]);
      __builtin_nontemporal_store(Mb.idx, &a.s.gbx[kb]);
    }
    if (a.step && acta) __builtin_nontemporal_store(static_cast<uint8_t>(Ma.dir), &a.step[ka]);
    if (a.step && actb) __builtin_nontemporal_store(static_cast<uint8_t>(Mb.dir), &a.step[kb]);
    nmoved += static_cast<uint32_t>(__popcll(__ballot(acta))) + static_cast<uint32_t>(__popcll(__ballot(actb)));
    count_keys(a.cnt, motion_key(Ma, ga), acta, lane);
    count_keys(a.cnt, motion_key(Mb, gb), actb, lane);
  }
  if (lane == 0) put_stat(a.stat, 0, wave_index(), nmoved);
}

// launches k_motion_drops with the bounds in LDS when they fit
template <bool kWrite>
static void launch_motion_drops(cleo_ctx* c, const MotionArgs& a, unsigned nb) {
  const uint32_t nb_ = a.grid.nz + a.grid.nx + a.grid.ny + 3;
  if (nb_ <= kMotionLdsBounds)
    hipLaunchKernelGGL((k_motion_drops<kWrite, true>), dim3(nb), dim3(256), nb_ * sizeof(double), c->stream, a);
  else
    hipLaunchKernelGGL((k_motion_drops<kWrite, false>), dim3(nb), dim3(256), 0, c->stream, a);
}

// Motion fused into the sort's gather (motion_flat 3): destination q takes the drop at position
// perm[q] of the unmoved store, moved by the same predictor-corrector step the counting pass
// computed for it (the arithmetic is deterministic, so the new gridbox is the key the drop was
// ranked by); the new coordinates and gridbox go straight to q.  One position per thread and
// iteration (two would spill); bounds from LDS when kLds.  A perm entry that is not a position of
// the domain is reported and skipped.
template <bool kLds>
__global__ void __launch_bounds__(256) CLEO_OCCUPANCY(4) k_nsort_gather_move(NsortArgs n, MotionArgs m) {
  extern __shared__ double bounds_lds[];
  MotionArgs b = m;
  if constexpr (kLds) {
    const uint32_t nzh = m.grid.nz + 1, nxh = m.grid.nx + 1, nyh = m.grid.ny + 1;
    for (uint32_t i = threadIdx.x; i < nzh + nxh + nyh; i += blockDim.x)
      bounds_lds[i] = i < nzh ? m.grid.zh[i] : (i < nzh + nxh ? m.grid.xh[i - nzh] : m.grid.yh[i - nzh - nxh]);
    __syncthreads();
    b.grid.zh = bounds_lds;
    b.grid.xh = bounds_lds + nzh;
    b.grid.yh = bounds_lds + nzh + nxh;
  }
  const uint64_t nsrc = n.off[n.grid.ngbx];
  for (uint64_t q = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < nsrc;
       q += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t p = n.perm[q];
    if (p >= nsrc) {
      set_error(n.ctl, CLEO_ERR_SORT_NONLOCAL, kOOB, q);
      continue;
    }
    const uint32_t g = n.src.gbx[p];
    const DropPos P = drop_pos(b, p);
    // an in-domain drop's gridbox is below ngbx; the guard keeps a corrupted one from indexing
    const uint32_t gg = g < n.grid.ngbx ? g : 0u;
    const GbxGeom G = gbx_geom(b, gg);
    const Moved M = move_drop(b, gg, G, P, p, false);
    n.dst.gbx[q] = M.idx;
    n.dst.c3[q] = M.c3;
    n.dst.c1[q] = M.c1;
    n.dst.c2[q] = M.c2;
    n.dst.xi[q] = n.src.xi[p];
    n.dst.r[q] = P.r;
    n.dst.msol[q] = n.src.msol[p];
    n.dst.id[q] = n.src.id[p];
  }
}
static void launch_gather_move(cleo_ctx* c, const NsortArgs& n, const MotionArgs& m) {
  const uint32_t nb_ = m.grid.nz + m.grid.nx + m.grid.ny + 3;
  const unsigned nb = drop_blocks(c->ntot, 0);
  if (nb_ <= kMotionLdsBounds)
    hipLaunchKernelGGL(k_nsort_gather_move<true>, dim3(nb), dim3(256), nb_ * sizeof(double), c->stream, n, m);
  else
    hipLaunchKernelGGL(k_nsort_gather_move<false>, dim3(nb), dim3(256), 0, c->stream, n, m);
}

// the gather with the solves; with m, the drops also move in it (k_sort_gather_cond<true, ...>)
static void launch_gather_cond(cleo_ctx* c, const NsortArgs& n, const CondFlatArgs& a, const MotionArgs* m,
                               unsigned nb) {
  if (!m) {
    hipLaunchKernelGGL((k_sort_gather_cond<false, false>), dim3(nb), dim3(256), 0, c->stream, n, a, MotionArgs{});
    return;
  }
  const uint32_t nb_ = m->grid.nz + m->grid.nx + m->grid.ny + 3;
  if (nb_ <= kMotionLdsBounds)
    hipLaunchKernelGGL((k_sort_gather_cond<true, true>), dim3(nb), dim3(256), nb_ * sizeof(double), c->stream, n, a,
                       *m);
  else
    hipLaunchKernelGGL((k_sort_gather_cond<true, false>), dim3(nb), dim3(256), 0, c->stream, n, a, *m);
}

// per-wave segments of the deferred list for a drop launch of nb blocks: a quarter of a wave's
// drops, at least 16
static void set_defer_segments(const cleo_ctx* c, CondFlatArgs& a, unsigned nb) {
  const uint64_t per_wave = (c->ntot + 4ull * nb - 1) / (4ull * nb);
  a.dseg = c->dseg;
  a.dsegcap = a.dcount ? static_cast<uint32_t>(per_wave / 4 + 16) : 0u;
}

void launch_condensation_flat(cleo_ctx* c, const CondFlatArgs& args) {
  if (args.ngbx == 0) return;
  CondFlatArgs a = args;
  // the thread-per-drop solve records each drop's qcond term, so the qcond recompute after motion
  // joins k_cond_state (the task ring keeps it in k_cond_prep)
  a.qm = (a.fuse_qcond && c->tiles.cond_flat == 1) ? c->qm : nullptr;
  const uint32_t gw = c->tiles.cond_grid_waves;
  const uint64_t ntiles = (a.ngbx + kWave - 1) / kWave;
  const uint32_t budget = c->tiles.cond_flat == 2 ? 0u : c->tiles.cond_defer;
  a.cfg.budget = budget;
  a.dcap = budget ? c->dlist_cap : 0u;
  a.dlist = c->dlist;
  a.dcount = budget ? c->dcount : nullptr;  // zeroed by k_cond_prep
  hipLaunchKernelGGL(k_cond_prep, dim3(gbx_blocks(ntiles, 4, gw)), dim3(256), 0, c->stream, a);
  if (c->tiles.cond_flat == 2) {
    // task-ring kernel: waves of 4 per block; by default as many waves as fit on the device at once
    const uint64_t waves = gw ? gw : static_cast<uint64_t>(c->num_cu) * kCondQueueWavesPerCu;
    const unsigned nb = static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(waves / 4, (c->ntot + 255) / 256)));
    a.stat = stat_slots(c, 4ull * nb);
    kev_mark(c, 0);
    hipLaunchKernelGGL(k_cond_queue, dim3(nb), dim3(256), 4 * kCondRingBytes, c->stream, a);
    kev_mark(c, 1);
    launch_stats_reduce(c, 4u * nb, ST_COND_DROPS, 4);
  } else {
    const unsigned nb = drop_blocks(c->ntot, gw);
    a.stat = stat_slots(c, 4ull * nb);  // every wave of k_cond_drops writes its slot
    set_defer_segments(c, a, nb);
    if (a.mlp_inline) {  // the learned guesses, from the saturation ratios k_cond_prep left
      hipLaunchKernelGGL(k_cond_guess, dim3(drop_blocks(c->ntot, gw ? gw : 32768)), dim3(256), 0, c->stream, a);
      a.warm = a.warm_buf;
    }
    kev_mark(c, 0);
    hipLaunchKernelGGL(k_cond_drops, dim3(nb), dim3(256), 0, c->stream, a);
    if (budget) {
      launch_stats_reduce(c, 4u * nb, ST_COND_DROPS, 4);
      hipLaunchKernelGGL(k_cond_deferred, dim3(nb), dim3(256), 0, c->stream, a);
    }
    kev_mark(c, 1);
    launch_stats_reduce(c, 4u * nb, ST_COND_DROPS, 4);
  }
  hipLaunchKernelGGL(k_cond_state, dim3(gbx_blocks(ntiles, 4, gw)), dim3(256), 0, c->stream, a);
}

// Motion, the neighbour sort and a condensation that follows at once, with the condensation solve
// in the sort's gather (k_sort_gather_cond): motion (thread per drop), the sort tables and ranks,
// the saturation ratios (State only), the gather with the solves, the commit of the sorted store,
// the deferred drops, then the ordered sums and the State update (with the qcond recompute).
// margs.s is the store motion reads, `dst` the sorted store; commit_dst >= 0: that buffer becomes
// current (else, for timing, the refs stay those of the unsorted store).
void motion_sort_condense(cleo_ctx* c, const MotionArgs& margs, const CondFlatArgs& cond, const SoA& dst,
                          int commit_dst) {
  const uint32_t ngbx = c->grid.ngbx;
  if (ngbx == 0) return;
  MotionArgs m = margs;
  m.step = c->dstep;
  // motion variant 3: the drops move inside the gather (the counting pass writes only the steps);
  // not with learned guesses, which are evaluated on the moved store before the gather
  const bool move_in_gather = c->tiles.motion_flat == 3 && !cond.mlp_inline && !cond.warm;
  const bool on = c->kev_on;
  c->kev_on = false;  // the main kernel of this sequence is the gather with the solves
  if (move_in_gather) {
    hipLaunchKernelGGL(k_fill_u32, dim3(stride_blocks(28ull * ngbx)), dim3(256), 0, c->stream, m.cnt, 28ull * ngbx,
                       0u);
    const unsigned mb = motion_blocks(c->ntot, c->tiles.motion_grid_waves);
    m.stat = stat_slots(c, 4ull * mb);
    launch_motion_drops<false>(c, m, mb);
    launch_stats_reduce(c, 4u * mb, ST_MOTION_DROPS, 1);
  } else {
    launch_motion_flat(c, m);
  }
  c->kev_on = on;
  NsortArgs n = nsort_args(c, m.s, dst, nullptr);
  n.step = c->dstep;
  nsort_tables(c, n, nullptr);
  launch_nsort_perm(c, n, ngbx);
  CondFlatArgs a = cond;
  a.s = dst;
  a.qm = c->qm;  // the qcond recompute joins k_cond_state
  const uint32_t gw = c->tiles.cond_grid_waves;
  const uint64_t ntiles = (ngbx + kWave - 1) / kWave;
  // a deferred drop is solved again by its sorted position k; its learned guess is that of source
  // position perm[k]
  const uint32_t budget = c->tiles.cond_defer;
  if (a.mlp_inline) a.warm_perm = n.perm;
  a.cfg.budget = budget;
  a.dcap = budget ? c->dlist_cap : 0u;
  a.dlist = c->dlist;
  a.dcount = budget ? c->dcount : nullptr;  // zeroed by k_cond_prep
  hipLaunchKernelGGL(k_cond_prep, dim3(gbx_blocks(ntiles, 4, gw)), dim3(256), 0, c->stream, a);
  if (a.mlp_inline) {  // learned guesses of the drops where motion left them
    CondFlatArgs ga = a;
    ga.s = m.s;
    ga.unsorted = 1;
    ga.guess_out = nullptr;
    hipLaunchKernelGGL(k_cond_guess, dim3(drop_blocks(c->ntot, gw ? gw : 32768)), dim3(256), 0, c->stream, ga);
    a.warm = a.warm_buf;
  }
  const unsigned nb = drop_blocks(c->ntot, gw);
  a.stat = stat_slots(c, 4ull * nb);
  set_defer_segments(c, a, nb);
  kev_mark(c, 0);
  launch_gather_cond(c, n, a, move_in_gather ? &m : nullptr, nb);
  kev_mark(c, 1);
  if (budget) {
    launch_stats_reduce(c, 4u * nb, ST_COND_DROPS, 4);
    hipLaunchKernelGGL(k_cond_deferred, dim3(nb), dim3(256), 0, c->stream, a);
  }
  launch_stats_reduce(c, 4u * nb, ST_COND_DROPS, 4);
  hipLaunchKernelGGL(k_nsort_oob, dim3(sparse_blocks(c->ntot)), dim3(256), 0, c->stream, n, c->ntot);
  hipLaunchKernelGGL(k_sorts_stat, dim3(1), dim3(1), 0, c->stream, c->ctl, static_cast<const uint32_t*>(nullptr));
  if (commit_dst >= 0) nsort_commit(c, commit_dst);
  hipLaunchKernelGGL(k_cond_state, dim3(gbx_blocks(ntiles, 4, gw)), dim3(256), 0, c->stream, a);
}

void launch_motion_flat(cleo_ctx* c, const MotionArgs& args) {
  if (args.grid.ngbx == 0) return;
  MotionArgs a = args;
  const unsigned nb = motion_blocks(c->ntot, c->tiles.motion_grid_waves);
  a.stat = stat_slots(c, 4ull * nb);  // every wave of k_motion_drops writes its slot
  hipLaunchKernelGGL(k_fill_u32, dim3(stride_blocks(28ull * a.grid.ngbx)), dim3(256), 0, c->stream, a.cnt,
                     28ull * a.grid.ngbx, 0u);
  kev_mark(c, 0);
  launch_motion_drops<true>(c, a, nb);
  kev_mark(c, 1);
  launch_stats_reduce(c, 4u * nb, ST_MOTION_DROPS, 1);
}

}  // namespace cleo
