// ctx.hpp -- the device context behind the opaque cleo_ctx of include/cleo_hip.h.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "physics.hpp"

namespace cleo {

// Structure-of-arrays superdroplet store in HBM: 60 bytes per drop.
struct SoA {
  uint32_t* gbx;
  double* c3;
  double* c1;
  double* c2;
  uint64_t* xi;
  double* r;
  double* msol;
  uint64_t* id;
};

// Rectilinear grid; gbxindex = iz + nz*(ix + nx*iy).
struct Grid {
  uint32_t nz, nx, ny, ngbx;
  int32_t per[3];
  const double* zh;
  const double* xh;
  const double* yh;
};

// Gridbox State, one array per field.
struct GState {
  double* vol;
  double* press;
  double* temp;
  double* qv;
  double* qc;
  double* wlo;
  double* whi;
  double* ulo;
  double* uhi;
  double* vlo;
  double* vhi;
};

// Per-wave statistics slots: counter j of the wave in slot i is sp[j * nslot + i].
struct StatSlots {
  uint32_t* sp;
  uint32_t nslot;
};

struct MLPDev;  // mlp.hip

}  // namespace cleo

struct cleo_ctx {
  int device = 0;
  int num_cu = 256;  // compute units of the device
  hipStream_t stream = nullptr;
  cleo::Grid grid{};
  double* d_halves = nullptr;
  uint64_t maxn = 0;
  uint64_t ntot = 0;
  cleo::SoA buf[2]{};
  int cur = 0;
  uint32_t* off = nullptr;      // ngbx+1: gridbox refs, off[ngbx] = in-domain count
  uint32_t* off_alt = nullptr;  // ngbx+1: next refs while sorting
  uint32_t* nact = nullptr;     // ngbx: active drops per gridbox (subview extent)
  cleo::GState st{};
  double* d_state = nullptr;
  cleo::Solute solute{};
  // neighbour-sort scratch
  uint32_t* cnt = nullptr;      // ngbx*56: sort counts (ngbx*28), then run offsets (ngbx*28)
  uint32_t* tot = nullptr;      // ngbx+1
  uint32_t* oobpre = nullptr;   // ngbx+1
  uint32_t* oobcnt = nullptr;   // ngbx: out-of-domain count per gridbox (compact copy for its scan)
  uint32_t* scan_tmp = nullptr;
  size_t scan_tmp_n = 0;
  // generic radix sort scratch (lazily allocated)
  uint32_t* rs_key[2] = {nullptr, nullptr};
  uint32_t* rs_idx[2] = {nullptr, nullptr};
  uint32_t* rs_cnt = nullptr;
  size_t rs_n = 0;
  // control
  cleo::DevCtl* ctl = nullptr;
  cleo::DevCtl* h_ctl = nullptr;  // pinned mirror
  // RNG
  int rng_mode = 0;  // 0 none, 1 counter, 2 injected
  uint64_t seed = 0;
  uint32_t* inj_j = nullptr;
  double* inj_phi = nullptr;
  uint64_t inj_cap_j = 0, inj_cap_phi = 0, inj_nj = 0, inj_nphi = 0;
  uint32_t* inj_oj = nullptr;  // ngbx+1
  uint32_t* inj_op = nullptr;  // ngbx+1
  // large-gridbox collision scratch (permutation + indices for gridboxes above the LDS capacity)
  uint32_t* big_scratch = nullptr;
  uint64_t big_scratch_n = 0;
  // configuration
  // the SDMMethods configuration: microphysics proc[0] >> ... >> proc[n-1], then motion
  struct Chain {
    uint32_t n = 0;
    cleo_process proc[CLEO_MAX_PROCESSES]{};
    bool motion_on = false;
    cleo_motion_params motion{};
    const cleo_cond_params* first_cond() const {
      for (uint32_t i = 0; i < n; ++i)
        if (proc[i].kind == CLEO_PROCESS_CONDENSATION) return &proc[i].cond;
      return nullptr;
    }
    const cleo_coll_params* first_coll() const {
      for (uint32_t i = 0; i < n; ++i)
        if (proc[i].kind == CLEO_PROCESS_COLLISIONS) return &proc[i].coll;
      return nullptr;
    }
  } sdm;
  bool sdm_set = false;
  cleo_tiles tiles{1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
  double* gs = nullptr;  // kGsStride*ngbx: per gridbox saturation ratio, diffusion factor, Koehler a, 1/a
  double* dm = nullptr;  // maxn: per-drop condensed mass (flat condensation)
  // per-wave statistics: up to kStatSlotsMax counters per wave, reduced into DevCtl::stats
  uint32_t* sp = nullptr;
  // deferred-drop list of the thread-per-drop condensation (segments per wave) and their counts
  uint32_t* dlist = nullptr;
  uint32_t* dcount = nullptr;
  uint32_t dlist_cap = 0;
  uint32_t* dseg = nullptr;
  // neighbour sort: source position of every destination position (maxn), and the step each drop
  // took in the last thread-per-drop motion (maxn bytes, 27 = left the domain)
  uint32_t* perm = nullptr;
  uint8_t* dstep = nullptr;
  // Collisions leave the drops in place and record the shuffled order: order[q] is the position
  // of the drop at slot q of the domain's logical order.  While order_live is set the store's
  // physical order may differ from its logical order; the next sort applies the order,
  // materialize_order() applies it for callers that need the drops in order, and a further
  // collision event composes its shuffle with it.
  uint32_t* order = nullptr;
  uint32_t* order_alt = nullptr;  // the other buffer: a collision event composes order into it
  bool order_live = false;
  // per drop: condensed mass * xi before a condensation solve (the qcond recompute after motion)
  double* qm = nullptr;
  // mass-moment observer results (ngbx uint64 + 2 * ngbx float, allocated at first use)
  void* mom = nullptr;
  uint64_t sp_slots = 0;  // slots per counter: max(ngbx, waves of a thread-per-drop launch)
  // per-drop Newton iteration record
  bool rec_niters = false;
  uint32_t* niters = nullptr;
  // learned guess record (z) for tests
  bool rec_guess = false;
  double* guess = nullptr;
  // MLP
  cleo::MLPDev* mlp = nullptr;
  // timing events on `stream`
  hipEvent_t ev[16] = {};
  // events around the main kernel of a stage, recorded by the launchers while kev_on is set
  hipEvent_t kev[2] = {};
  bool kev_on = false;
  std::string msg;
};
// Copies and fills of a context run in order on its stream and have completed when they return.
// The stream is created non-blocking, so null-stream work (a plain hipMemcpy or hipMemset) would
// not be ordered with the context's kernels, and a device-to-device hipMemcpy or a device
// hipMemset may return before it has completed.
inline hipError_t ctx_copy(cleo_ctx* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
  return e != hipSuccess ? e : hipStreamSynchronize(c->stream);
}
inline hipError_t ctx_fill(cleo_ctx* c, void* dst, int value, size_t bytes) {
  const hipError_t e = hipMemsetAsync(dst, value, bytes, c->stream);
  return e != hipSuccess ? e : hipStreamSynchronize(c->stream);
}

