// noalloc.hip -- the entry point of libcleo_hip.so that allocates no device memory:
// cleo_sort_supers_noalloc, the generic superdroplet sort of cleo_upload_supers_device on
// caller-owned device arrays.  Its scratch (radix keys, indices, digit counts and the scan's
// partials) lives in `scratch`, a fixed device array of this file; every GPU the library runs on
// holds one copy until the process ends.  Calls that overlap on one GPU share that copy and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "sdm_kernels.hpp"

namespace {

constexpr uint64_t kMaxSupers = CLEO_NOALLOC_MAX_SUPERS;
// radix_scratch_words(kMaxSupers) = 4 n + 32 ceil(n / 256) + 1 + the scan's partials (a few hundred
// words); the margin covers the last
constexpr uint64_t kScratchWords = 4 * kMaxSupers + 32 * ((kMaxSupers + 255) / 256) + 1 + 4096;

__device__ __align__(256) uint32_t scratch[kScratchWords];

// device address of `scratch` on each GPU, looked up once per device with that device current
constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t* g_scratch[kMaxDevices] = {};

int scratch_on_current_device(uint32_t** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return CLEO_E_HIP;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CLEO_E_HIP;
    g_scratch[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch[dev];
  return CLEO_OK;
}

bool complete(const cleo_supers_soa* s) {
  return s->sdgbxindex && s->coord3 && s->coord1 && s->coord2 && s->xi && s->radius && s->msol && s->sdid;
}

cleo::SoA as_soa(const cleo_supers_soa* s) {
  return cleo::SoA{s->sdgbxindex, s->coord3, s->coord1, s->coord2, s->xi, s->radius, s->msol, s->sdid};
}

}  // namespace

extern "C" int cleo_sort_supers_noalloc(const cleo_supers_soa* in, cleo_supers_soa* out, uint32_t* refs, uint64_t n,
                                        uint32_t ngbx, void* stream) {
  if (!in || !out || !refs || ngbx == 0 || ngbx >= 0xFFFFFFF0u) return CLEO_E_ARG;
  if (n > kMaxSupers) return CLEO_E_ARG;
  if (n > 0 && (!complete(in) || !complete(out))) return CLEO_E_ARG;
  if (cleo::radix_scratch_words(n) > kScratchWords) return CLEO_E_ARG;
  uint32_t* base = nullptr;
  const int rc = scratch_on_current_device(&base);
  if (rc != CLEO_OK) return rc;
  cleo::radix_sort_soa(static_cast<hipStream_t>(stream), as_soa(in), as_soa(out), n, ngbx, refs,
                       cleo::radix_scratch_at(base, n));
  return hipGetLastError() == hipSuccess ? CLEO_OK : CLEO_E_HIP;
}
