// kernels.hpp -- wave-level helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "ctx.hpp"

namespace cleo {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() { return static_cast<int>(threadIdx.x & (kWave - 1)); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }
// waves of a launch: gridbox kernels visit g = wave_index(), + wave_count(), ... so that a grid
// smaller than one wave per gridbox amortises the dispatch of short waves
__device__ __forceinline__ uint32_t wave_index() { return blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); }
__device__ __forceinline__ uint32_t wave_count() { return gridDim.x * (blockDim.x >> 6); }
// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b to XCD b % 8).  This bijection
// of [0, gridDim.x) gives each XCD's workgroups one contiguous range of logical indices, in
// dispatch order, so that work on neighbouring indices shares that XCD's L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t per = nb / kXcds, rem = nb % kXcds, x = b % kXcds;
  return x * per + min(x, rem) + b / kXcds;
}
__device__ __forceinline__ uint32_t xcd_wave_index() { return xcd_block() * (blockDim.x >> 6) + (threadIdx.x >> 6); }

// LDS / memory ordering between lanes of one wavefront (no workgroup barrier needed)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int l) {
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), l));
}
// the wave's lanes where c holds, as a bit mask
__device__ __forceinline__ uint64_t ballot(bool c) { return __builtin_amdgcn_ballot_w64(c); }
// lanes below this one whose bit is set in m (v_mbcnt_lo/hi: two instructions)
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}
// occupancy floor of a kernel: at least W waves per SIMD, so at most 512/W VGPRs per lane
#define CLEO_OCCUPANCY(W) __attribute__((amdgpu_waves_per_eu(W)))

// v with lane l replaced by the wave-uniform value x
__device__ __forceinline__ uint32_t writelane_u32(uint32_t v, uint32_t x, uint32_t l) {
  return static_cast<uint32_t>(lane_id()) == l ? x : v;
}
__device__ __forceinline__ double readlane_f64(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// Adds the per-lane values v of lanes 0..m-1 to `acc` strictly in lane order, so a
// gridbox sum keeps the sequential order of the reference's Serial reduction.
__device__ __forceinline__ double ordered_add(double acc, double v, int m) {
  for (int i = 0; i < m; ++i) acc += readlane_f64(v, i);
  return acc;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ __forceinline__ uint32_t wave_or_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o);
  return v;
}

// first error wins; later errors are dropped
__device__ __forceinline__ void set_error(DevCtl* ctl, int code, uint32_t gbx, uint64_t drop) {
  if (code) record_error(ctl, code, gbx, drop);
}

// Statistics of a wave go to a slot of their own, sp[j * nslot + slot] for the kernel's j-th
// counter; k_stats_reduce adds the slots into DevCtl::stats after the launch.  Same-address
// device-scope atomics from every wave would serialise across the XCDs.
__device__ __forceinline__ void put_stat(const StatSlots& s, int j, uint32_t slot, uint32_t v) {
  s.sp[static_cast<uint64_t>(j) * s.nslot + slot] = v;
}

}  // namespace cleo
