// sdm_kernels.hpp -- kernel argument blocks and host-side launchers of sdm_kernels.hip.
#pragma once
#include "ctx.hpp"
#include "mlp_device.hpp"

namespace cleo {

constexpr int kCollMaxWpb = 4;
constexpr uint32_t kCollLdsCap = 512;  // drops per gridbox held in LDS; larger gridboxes use HBM scratch
constexpr int kStatSlotsMax = 4;       // counters per wave of one kernel
// tiled collisions: byte permutations of up to kTileMaxN drops, kTilePermBytes of them per wave
constexpr uint32_t kTileMaxN = 255;
constexpr uint32_t kTilePermBytes = 8192;
constexpr uint32_t kCollTileLdsBytes = kTilePermBytes + 4u * (65u + 4u * 64u);
constexpr uint32_t kBigGbx = 0xFFFFFFFFu;  // tile LDS marker of a gridbox on the per-gridbox path
// condensation task ring: prepared drops per wave, and waves per CU of the default grid
constexpr uint32_t kRing = 128;
constexpr uint32_t kCondRingBytes = kRing * (11u * 8u + 8u + 2u * 4u);
constexpr uint32_t kCondQueueWavesPerCu = 12;
// per-gridbox condensation inputs in gs: saturation ratio, diffusion factor, Koehler a and its
// refined reciprocal (one 32-byte line each)
constexpr uint32_t kGsStride = 4;
// LDS window (values) of the per-gridbox drop-order sums
constexpr uint32_t kSumWin = 1024;

struct CondArgs {
  SoA s;
  const uint32_t* off;
  const uint32_t* nact;
  GState st;
  Solute sol;
  NewtonCfg cfg;
  uint32_t ngbx;
  uint32_t wpb;
  int32_t alter;
  int32_t fuse_qcond;
  DevCtl* ctl;
  uint32_t* niters;
  const double* warm;
  uint32_t warm_limit;
  StatSlots stat;  // ST_COND_DROPS .. ST_WARM_RESTARTS
  double* gs;      // per gridbox (kGsStride): supersaturation ratio, diffusion factor, Koehler a (k_gbx_thermo)
};

struct QcondArgs {
  SoA s;
  const uint32_t* off;
  GState st;
  Solute sol;
  uint32_t ngbx;
  uint32_t wpb;
  DevCtl* ctl;
};

struct CollArgs {
  SoA s;            // the store: drops keep their places, changed attributes are written in place
  uint32_t* order;  // out: position of the drop at each slot of the shuffled (logical) order
  // in: a pending order of an earlier event (position of the drop at each slot), composed with
  // this event's shuffle; null when the store is in its logical order
  const uint32_t* order_in;
  const uint32_t* off;
  uint32_t* nact;
  const double* vol;
  Solute sol;
  CollCfg k;
  uint32_t ngbx;
  uint32_t wpb;
  DevCtl* ctl;
  uint64_t seed;
  const uint32_t* inj_j;
  const double* inj_phi;
  const uint32_t* inj_oj;
  const uint32_t* inj_op;
  uint32_t* big;
  StatSlots stat;   // ST_COLL_PAIRS .. ST_NULLS
};

struct MotionArgs {
  SoA s;
  const uint32_t* off;
  Grid grid;
  GState st;
  double delt;
  int32_t tv;
  uint32_t wpb;
  uint32_t* cnt;
  DevCtl* ctl;
  StatSlots stat;  // ST_MOTION_DROPS
  uint8_t* step;   // non-null: the step each drop takes (27: out of the domain), position order
};

struct NsortArgs {
  SoA src, dst;
  const uint32_t* off;     // current refs (ngbx+1)
  const uint32_t* base;    // new refs (ngbx+1)
  uint32_t* cnt;           // ngbx*28: arrivals per (destination, step)
  uint32_t* runs;          // ngbx*28: per (source, step) the run's offset among its destination's arrivals
  const uint32_t* oobpre;  // ngbx+1
  uint32_t* tot;
  uint32_t* oobcnt;        // ngbx: drops of each gridbox that leave the domain (row d's last count)
  Grid grid;
  uint32_t wpb;
  DevCtl* ctl;
  const uint32_t* flag;    // non-null: skip unless *flag
  uint64_t ntot;
  uint32_t* perm;          // source position of each destination position of the domain's drops
  const uint8_t* step;     // non-null: each drop's step from motion, read instead of its key
  const uint32_t* order;   // non-null: position of the drop at each slot of the logical order
};

// thread-per-drop condensation: prep (per gridbox) -> drops -> state (per gridbox)
struct CondFlatArgs {
  SoA s;
  const uint32_t* off;
  const uint32_t* nact;
  GState st;
  Solute sol;
  NewtonCfg cfg;
  uint32_t ngbx;
  int32_t alter;
  int32_t fuse_qcond;
  DevCtl* ctl;
  uint32_t* niters;
  const double* warm;
  uint32_t warm_limit;
  double* gs;  // kGsStride per gridbox: saturation ratio, diffusion factor, Koehler a
  double* dm;  // per drop: condensed mass * xi
  // non-null (with fuse_qcond): per drop the condensed mass * xi before the solve; k_cond_state
  // recomputes the gridbox qcond from it instead of a separate pass in k_cond_prep
  double* qm;
  StatSlots stat;  // ST_COND_DROPS .. ST_WARM_RESTARTS
  uint32_t* dlist;   // deferred drops (positions) of all waves, room for `dcap`
  uint32_t* dcount;  // one counter: entries claimed in dlist, zeroed by k_cond_prep (null: no deferral)
  uint32_t dcap;
  uint32_t* dseg;    // per-wave segments where a wave lists its deferred drops first, `dsegcap` each
  uint32_t dsegcap;
  // mlp_inline: the learned guesses are evaluated by k_cond_guess into warm_buf inside the launch
  // sequence (after the saturation ratios exist) and read from there
  int32_t mlp_inline;
  MLPWeights W;
  double* warm_buf;
  // warm (learned guesses) indexed by the drop's source position perm[k] rather than by k: the
  // guesses of the sort's gather with the solves, read again by k_cond_deferred
  const uint32_t* warm_perm;
  double* guess_out;  // non-null: the guesses are recorded here too (array position order)
  // k_cond_guess on the store as motion left it (before the sort): a drop is active when its new
  // gridbox is in the domain (the refs still describe the unsorted store)
  int32_t unsorted;
};

// launchers (sdm_kernels.hip); all enqueue on ctx->stream
void launch_condensation_flat(cleo_ctx* c, const CondFlatArgs& a);
void launch_motion_flat(cleo_ctx* c, const MotionArgs& a);
void launch_condensation(cleo_ctx* c, const CondArgs& a);
void launch_qcond(cleo_ctx* c, const QcondArgs& a);
void launch_collisions(cleo_ctx* c, const CollArgs& a);
void launch_motion(cleo_ctx* c, const MotionArgs& a);
void launch_advance_event(cleo_ctx* c);
// DevCtl::stats[first + j] += sum over the nslot slots of counter j, j < ncount
void launch_stats_reduce(cleo_ctx* c, uint32_t nslot, int first, int ncount);
// slot block of the context for a launch with `nslot` waves
StatSlots stat_slots(cleo_ctx* c, uint64_t nslot);
void launch_inj_sizes(cleo_ctx* c, uint32_t* nj, uint32_t* np);
// exclusive scan of in[i*stride+offset], i < len, into out[0..len] (out[len] = total)
void scan_u32(cleo_ctx* c, const uint32_t* in, uint64_t len, uint32_t stride, uint32_t offset, uint32_t* out,
              const uint32_t* flag);
size_t scan_tmp_words(uint64_t len);
// stable neighbour sort of buf[src] into buf[dst] after motion (flag == nullptr) or after
// collisions (flag = &ctl->resort_flag, data ends back in buf[src] and refs in c->off)
void neighbour_sort(cleo_ctx* c, int src, int dst, const uint32_t* flag, bool from_nact,
                    const uint8_t* step = nullptr);
// the neighbour sort of `src` (after motion) into `dst`; commit_dst >= 0: that buffer becomes current
void neighbour_sort_to(cleo_ctx* c, const SoA& src, const SoA& dst, int commit_dst, const uint8_t* step = nullptr);
// motion + neighbour sort + a condensation that follows at once, the solve in the sort's gather
// (thread-per-drop motion and condensation): margs.s is read, the moved, sorted and
// condensed drops go to dst; commit_dst >= 0: that buffer becomes current
void motion_sort_condense(cleo_ctx* c, const MotionArgs& margs, const CondFlatArgs& cond, const SoA& dst,
                          int commit_dst);
// motion + neighbour sort with the motion arithmetic in the sort's gather (motion_flat 3): the
// store args.s is left unchanged, the moved and sorted drops go to dst
void motion_sort_gather(cleo_ctx* c, const MotionArgs& args, const SoA& dst, int commit_dst);
// motion + neighbour sort fused into two passes (motion_flat 2): src = args.s is left unchanged
void motion_sort_fused(cleo_ctx* c, const MotionArgs& args, const SoA& dst, int commit_dst);
// generic stable sort of buf[cur] into buf[1-cur]; swaps cur; sets off and nact
void generic_sort(cleo_ctx* c);
// the generic sort without a context: scratch of radix_scratch_words(n) u32 words, laid out by
// radix_scratch_at (two key and two index arrays of n words, the digit counts, the scan's partials)
struct RadixScratch {
  uint32_t* key[2];
  uint32_t* idx[2];
  uint32_t* cnt;
  uint32_t* scan_tmp;
};
uint64_t radix_scratch_words(uint64_t n);
RadixScratch radix_scratch_at(uint32_t* base, uint64_t n);
void radix_sort_soa(hipStream_t stream, const SoA& src, const SoA& dst, uint64_t n, uint32_t ngbx, uint32_t* off,
                    const RadixScratch& sc);
void launch_copy_oob_region(cleo_ctx* c, int src, int dst);
// applies a pending collision order to the store (buf[cur] then holds the drops in order)
void materialize_order(cleo_ctx* c);
// 0th, 1st and 2nd mass moments of every gridbox into device arrays of ngbx entries (raindrops
// only when rain)
void launch_mass_moments(cleo_ctx* c, bool rain, uint64_t* mom0, float* mom1, float* mom2);
// saturation pressure, supersaturation ratio and diffusion factor of n states (device arrays)
void launch_thermo_eval(cleo_ctx* c, uint64_t n, const double* temp, const double* press, const double* qvap,
                        double* psat, double* s_ratio, double* ffactor);
// one coupling step of prescribed dynamics (device arrays) into every gridbox's State
void launch_receive_dynamics(cleo_ctx* c, const cleo_dynamics_step& dyn);
// the identity as the pending order (the store unchanged)
void pending_identity_order(cleo_ctx* c);

}  // namespace cleo
