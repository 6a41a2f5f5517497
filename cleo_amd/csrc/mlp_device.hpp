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

// The forward pass for inference only: layers 2 and 3 row block by row block, so that 32 rather
// than 64 accumulators per lane are live at once.  Every output goes through the same operations
// in the same order as in mlp_forward_wave, so y is identical.
template <class Lds>
__device__ __forceinline__ float mlp_infer_wave(const MLPWeights& W, Lds& L) {
  const int lane = lane_id();
  const int col = lane % 32, kq = 4 * (lane / 32);
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
    }
  }
  wave_sync();
  const float w3a = W.p[P_W3 + col], w3b = W.p[P_W3 + 32 + col];
  for (int mi = 0; mi < 2; ++mi) {
    f32x16 z2[2];
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
      z2[ni] = acc;
    }
    for (int v = 0; v < 16; ++v) {
      float t = silu(z2[0][v]) * w3a + silu(z2[1][v]) * w3b;
      for (int o = 16; o > 0; o >>= 1) t += __shfl_xor(t, o);
      if (col == 0) L.ys[mi * 32 + drow(v, lane)] = t;
    }
  }
  wave_sync();
  return L.ys[lane] + W.p[P_B3];
}

// features of one drop: (log r, log msol, s_ratio, T, p, delt)
__device__ __forceinline__ void drop_features(double r, double msol, double s_ratio, double temp, double press,
                                              double delt, float f[kIn]) {
  f[0] = static_cast<float>(log(r));
  f[1] = static_cast<float>(log(fmax(msol, 1e-300)));
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


// LDS of an inference-only forward pass of one wave
struct InferLds {
  static constexpr bool kHasXs = false;
  uint16_t xb[64][kInPad];
  uint16_t h1[64][kH1Stride];
  float ys[64];
};

}  // namespace cleo
