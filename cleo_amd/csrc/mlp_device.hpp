// mlp_device.hpp -- device side of the warm-start MLP (x[6] -> 64 -> 64 -> 1, SiLU, bf16 weights,
// fp32 accumulation on v_mfma_f32_32x32x8_bf16): parameter layout, the wave-wide forward pass over
// 64 samples and the per-drop input features.  Used by the training and guess kernels (mlp.hip)
// and by the condensation kernel that evaluates the guess in place (sdm_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace cleo {

constexpr int kIn = 6;
constexpr int kInPad = 8;
constexpr int kHid = 64;
// parameter offsets in the fp32 blob (include/cleo_hip.h)
constexpr int P_W1 = 0;
constexpr int P_B1 = P_W1 + kHid * kIn;
constexpr int P_W2 = P_B1 + kHid;
constexpr int P_B2 = P_W2 + kHid * kHid;
constexpr int P_W3 = P_B2 + kHid;
constexpr int P_B3 = P_W3 + kHid;
constexpr int NPARAM = P_B3 + 1;
static_assert(NPARAM == CLEO_MLP_NPARAMS, "parameter count");
constexpr int P_XMEAN = NPARAM;
constexpr int P_XSTD = P_XMEAN + kIn;
constexpr int P_YMEAN = P_XSTD + kIn;
constexpr int P_YSTD = P_YMEAN + 1;
constexpr int NBLOB = P_YSTD + 1;
static_assert(NBLOB == CLEO_MLP_NBLOB, "blob size");
constexpr int kH1Stride = 72;  // bf16 elements per LDS row (padding avoids bank conflicts)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// device weights in MFMA-friendly bf16 layouts + fp32 copies of the small vectors
struct MLPWeights {
  const uint16_t* w1p;   // [64][8]  W1 padded with two zero columns
  const uint16_t* w2;    // [64][64] W2[out][in]
  const uint16_t* w2t;   // [64][64] W2[in][out]
  const float* p;        // fp32 blob (biases, W3, normalisation)
};

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<uint16_t>((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigm(z);
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ int drow(int v, int lane) { return 8 * (v / 4) + 4 * (lane / 32) + (v % 4); }

__device__ __forceinline__ f32x16 mfma(bf16x4 a, bf16x4 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
}



// Forward pass of the 64 samples held in L.xb; returns z1, z2 in accumulator layout (tile
// [mi][ni], element v: row drow(v,lane) + 32 mi, column lane%32 + 32 ni) and y for sample `lane`
// (lane < 64).  Each output row depends on its own input row only.  `Lds` holds xb, h1 and ys.
template <class Lds>
__device__ __forceinline__ float mlp_forward_wave(const MLPWeights& W, Lds& L, f32x16 z1[2][2], f32x16 z2[2][2]) {
  const int lane = lane_id();
  const int col = lane % 32, kq = 4 * (lane / 32);
  // layer 1: [64 x 8] . [8 x 64]
  for (int ni = 0; ni < 2; ++ni) {
    const uint16_t* wr = W.w1p + (ni * 32 + col) * kInPad + kq;
    const bf16x4 b = {static_cast<short>(wr[0]), static_cast<short>(wr[1]), static_cast<short>(wr[2]),
                      static_cast<short>(wr[3])};
    const float bias = W.p[P_B1 + ni * 32 + col];
    for (int mi = 0; mi < 2; ++mi) {
      const uint16_t* xr = &L.xb[mi * 32 + col][kq];
      const bf16x4 a = {static_cast<short>(xr[0]), static_cast<short>(xr[1]), static_cast<short>(xr[2]),
                        static_cast<short>(xr[3])};
      f32x16 acc = {};
      acc = mfma(a, b, acc);
      for (int v = 0; v < 16; ++v) {
        acc[v] += bias;
        L.h1[mi * 32 + drow(v, lane)][ni * 32 + col] = f2bf(silu(acc[v]));
      }
      z1[mi][ni] = acc;
    }
  }
  wave_sync();
  // layer 2: [64 x 64] . [64 x 64]
  for (int mi = 0; mi < 2; ++mi) {
    for (int ni = 0; ni < 2; ++ni) {
      f32x16 acc = {};
      for (int kk = 0; kk < 8; ++kk) {
        const uint16_t* ar = &L.h1[mi * 32 + col][kk * 8 + kq];
        const bf16x4 a = {static_cast<short>(ar[0]), static_cast<short>(ar[1]), static_cast<short>(ar[2]),
                          static_cast<short>(ar[3])};
        const uint16_t* br = W.w2 + (ni * 32 + col) * kHid + kk * 8 + kq;
        const bf16x4 b = {static_cast<short>(br[0]), static_cast<short>(br[1]), static_cast<short>(br[2]),
                          static_cast<short>(br[3])};
        acc = mfma(a, b, acc);
      }
      const float bias = W.p[P_B2 + ni * 32 + col];
      for (int v = 0; v < 16; ++v) acc[v] += bias;
      z2[mi][ni] = acc;
    }
  }
  // layer 3: y[s] = sum_n SiLU(z2[s][n]) w3[n] + b3, reduced over the 32 lanes of a half-wave
  const float w3a = W.p[P_W3 + col], w3b = W.p[P_W3 + 32 + col];
  for (int mi = 0; mi < 2; ++mi) {
    for (int v = 0; v < 16; ++v) {
      float t = silu(z2[mi][0][v]) * w3a + silu(z2[mi][1][v]) * w3b;
      for (int o = 16; o > 0; o >>= 1) t += __shfl_xor(t, o);
      if (col == 0) L.ys[mi * 32 + drow(v, lane)] = t;
    }
  }
  wave_sync();
  return L.ys[lane] + W.p[P_B3];
}

// ---- inference pass with register-resident weights ---------------------------------------------
// The guess kernels evaluate the network on every drop of every step, so the weights stay in
// registers for the life of a wave and no activation goes through LDS.  The pass works on the
// transposed products Z1^T = W1 X^T and Z2^T = W2 H1^T (rows: hidden units, columns: the wave's
// 64 samples).  Lane l then holds hidden units 32 m + 8 q + 4 (l/32) + r (accumulator 4 q + r) of
// sample 32 n + l%32, and layer 2 (v_mfma_f32_32x32x16_bf16, B[k][j] from lane j + 32 (k/8),
// element k%8) takes its B operand straight from those accumulators: step t contracts the 16
// hidden units 32 (t/2) + 16 (t%2) + {8 (e/4) + 4 (l/32) + e%4}, e = 0..7, the same order on the
// A (W2) side.  Layer 3 sums the lane's 32 hidden units of layer 2 and adds the partner lane l^32.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x16 mfma16(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
// SiLU with the hardware reciprocal (1 ulp): the inference pass has no gradient to match
__device__ __forceinline__ float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
// hidden unit contracted by element e of lane half hh in layer-2 step t
__device__ __forceinline__ int l2_hidden(int t, int hh, int e) {
  return 32 * (t / 2) + 16 * (t % 2) + 8 * (e / 4) + 4 * hh + e % 4;
}

// biases and output weights of the pass (fp32), one copy per block in LDS
struct alignas(16) InferConst {
  float b1[kHid];
  float b2[kHid];
  float w3[kHid];
  float xmean[kIn], xstd[kIn];
  float b3, ymean, ystd;
};
// every thread of the block calls it; the caller synchronises the block afterwards
__device__ __forceinline__ void infer_const_load(InferConst& K, const float* p) {
  for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
    K.b1[i] = p[P_B1 + i];
    K.b2[i] = p[P_B2 + i];
    K.w3[i] = p[P_W3 + i];
  }
  if (threadIdx.x < kIn) {
    K.xmean[threadIdx.x] = p[P_XMEAN + threadIdx.x];
    K.xstd[threadIdx.x] = p[P_XSTD + threadIdx.x];
  }
  if (threadIdx.x == 0) {
    K.b3 = p[P_B3];
    K.ymean = p[P_YMEAN];
    K.ystd = p[P_YSTD];
  }
}

// the lane's weight fragments: layer 1 A = W1[32 m + l%32][4 (l/32) + r], layer 2 A =
// W2[32 m + l%32][l2_hidden(t, l/32, e)]
struct InferFrags {
  bf16x4 w1[2];
  bf16x8 w2[2][4];
};
__device__ __forceinline__ InferFrags infer_frags(const MLPWeights& W) {
  const int lane = lane_id(), col = lane % 32, hh = lane / 32;
  InferFrags F;
  for (int m = 0; m < 2; ++m) {
    const uint16_t* w1 = W.w1p + (32 * m + col) * kInPad + 4 * hh;
    F.w1[m] = bf16x4{static_cast<short>(w1[0]), static_cast<short>(w1[1]), static_cast<short>(w1[2]),
                     static_cast<short>(w1[3])};
    const uint16_t* w2 = W.w2 + (32 * m + col) * kHid;
    for (int t = 0; t < 4; ++t)
      for (int e = 0; e < 8; ++e) F.w2[m][t][e] = __builtin_bit_cast(__bf16, w2[l2_hidden(t, hh, e)]);
  }
  return F;
}

// the 16 entries q[8 (v/4) + v%4], v = 0..15, of an LDS row vector in accumulator order (q 16-byte
// aligned)
__device__ __forceinline__ f32x16 lds_rows(const float* q) {
  f32x16 o;
  for (int j = 0; j < 4; ++j) {
    const float4 t = *reinterpret_cast<const float4*>(q + 8 * j);
    o[4 * j] = t.x;
    o[4 * j + 1] = t.y;
    o[4 * j + 2] = t.z;
    o[4 * j + 3] = t.w;
  }
  return o;
}

__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
  return static_cast<uint32_t>(f2bf(a)) | (static_cast<uint32_t>(f2bf(b)) << 16);
}

// normalised network output of the wave's 64 samples: lane l passes sample l's normalised inputs
// and gets sample l's output.  Every lane of the wave must call it.
__device__ __forceinline__ float mlp_infer(const InferFrags& F, const InferConst& K, const float xn[kIn]) {
  const int lane = lane_id(), hh = lane / 32;
  // B of layer 1 for column block n: X[32 n + l%32][4 (l/32) + r]; the lane has its own sample's
  // inputs and fetches the other half from lane l^32
  const uint32_t lo0 = pack_bf16(xn[0], xn[1]), lo1 = pack_bf16(xn[2], xn[3]);
  const uint32_t hi0 = pack_bf16(xn[4], xn[5]), hi1 = 0u;
  const uint32_t r0 = __shfl_xor(hh ? lo0 : hi0, 32), r1 = __shfl_xor(hh ? lo1 : hi1, 32);
  uint32_t bx[2][2];
  bx[0][0] = hh ? r0 : lo0;
  bx[0][1] = hh ? r1 : lo1;
  bx[1][0] = hh ? hi0 : r0;
  bx[1][1] = hh ? hi1 : r1;
  float y[2];
  for (int n = 0; n < 2; ++n) {
    // row offset of the lane's bias and output-weight entries; opaque to the compiler so that
    // those LDS reads are repeated per column block rather than 96 values held in registers
    int hb = 4 * hh;
    asm volatile("" : "+v"(hb));
    const bf16x4 b = {static_cast<short>(bx[n][0] & 0xffffu), static_cast<short>(bx[n][0] >> 16),
                      static_cast<short>(bx[n][1] & 0xffffu), static_cast<short>(bx[n][1] >> 16)};
    bf16x8 h[4];
    for (int m = 0; m < 2; ++m) {
      f32x16 acc = lds_rows(K.b1 + 32 * m + hb);
      acc = mfma(F.w1[m], b, acc);
      for (int t2 = 0; t2 < 2; ++t2)
        for (int e = 0; e < 8; ++e) h[2 * m + t2][e] = static_cast<__bf16>(silu_fast(acc[8 * t2 + e]));
    }
    float part = 0.0f;
    for (int m = 0; m < 2; ++m) {
      f32x16 acc = lds_rows(K.b2 + 32 * m + hb);
      for (int t = 0; t < 4; ++t) acc = mfma16(F.w2[m][t], h[t], acc);
      const f32x16 w3 = lds_rows(K.w3 + 32 * m + hb);
      for (int v = 0; v < 16; ++v) part += silu_fast(acc[v]) * w3[v];
    }
    y[n] = part + __shfl_xor(part, 32);
  }
  return (hh ? y[1] : y[0]) + K.b3;
}

// the guess z = r_new^2 = r^2 exp(2 y) from the normalised output
__device__ __forceinline__ double learned_z(const InferConst& K, double r, float yn) {
  return r * r * static_cast<double>(expf(2.0f * (yn * K.ystd + K.ymean)));
}

// the lane's normalised inputs (zero for a lane without a sample)
__device__ __forceinline__ void infer_normalise(const InferConst& K, const float f[kIn], bool valid, float xn[kIn]) {
  for (int k = 0; k < kIn; ++k) xn[k] = valid ? (f[k] - K.xmean[k]) / K.xstd[k] : 0.0f;
}

// features of one drop: (log r, log msol, s_ratio, T, p, delt), the logarithms taken in single
// precision (the features are single precision; msol is floored at 1e-30 kg)
__device__ __forceinline__ void drop_features(double r, double msol, double s_ratio, double temp, double press,
                                              double delt, float f[kIn]) {
  f[0] = logf(static_cast<float>(r));
  f[1] = logf(static_cast<float>(fmax(msol, 1e-30)));
  f[2] = static_cast<float>(s_ratio);
  f[3] = static_cast<float>(temp);
  f[4] = static_cast<float>(press);
  f[5] = static_cast<float>(delt);
}

// the lane's normalised inputs into L.xb (bf16) and, when the LDS block has one, L.xs (fp32)
template <class Lds>
__device__ __forceinline__ void load_normalised(Lds& L, const float* p, const float f[kIn], bool valid) {
  const int lane = lane_id();
  for (int k = 0; k < kInPad; ++k) {
    float v = 0.0f;
    if (k < kIn && valid) v = (f[k] - p[P_XMEAN + k]) / p[P_XSTD + k];
    if constexpr (Lds::kHasXs) L.xs[lane][k] = v;
    L.xb[lane][k] = f2bf(v);
  }
  wave_sync();
}



}  // namespace cleo
