// mlp.hip -- learned warm start for the condensation Newton solve (gfx950).
//
// Network: x[6] -> 64 -> 64 -> 1, SiLU, bf16 weights, fp32 accumulation on MFMA
// (v_mfma_f32_32x32x8_bf16).  Inputs per drop: (log r, log msol, s_ratio, T, p, delt), each
// normalised by a (mean, std) pair; output: log(r_new / r), normalised likewise.  The guess
// handed to Newton is z = r^2 exp(2 y).
//
// One wave (64 lanes) owns 64 samples, the M dimension of every MFMA tile.  MFMA lane layout
// (32x32x8, bf16 operands): A[i][k] from lane (i + 32*(k/4)) element k%4; B[k][j] from lane
// (j + 32*(k/4)) element k%4; D[i][j] in lane (j + 32*((i/4)%2)) accumulator 4*(i/8) + i%4.
//
// Training: each wave computes the loss gradient of its 64 samples; the per-wave partial gradients
// are summed in a fixed order, in chunks of 32 waves and then the chunk sums in chunk order (the
// second kernel also applies Adam), so a run is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "mlp.hpp"
#include "mlp_device.hpp"
#include "sdm_kernels.hpp"

namespace cleo {

struct WaveLds {
  static constexpr bool kHasXs = true;
  float xs[64][kInPad];           // normalised inputs (fp32)
  uint16_t xb[64][kInPad];        // same, bf16 (A operand of layer 1)
  uint16_t h1[64][kH1Stride];     // SiLU(z1), bf16 (A operand of layer 2, B operand of dW2)
  uint16_t dz2[64][kH1Stride];    // dL/dz2, bf16
  float ys[64];
};

struct GuessArgs {
  SoA s;
  const uint32_t* off;
  const uint32_t* nact;
  GState st;
  Solute sol;
  double delt;
  uint32_t ngbx;
  uint64_t ntot;
  MLPWeights W;
  double* warm;
};

// learned guess z for every active in-domain drop (NaN elsewhere), 64 drops per network pass;
// blocks of four waves loop over the drops
__global__ void __launch_bounds__(256) k_mlp_guess(GuessArgs a) {
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
    float f[kIn] = {0, 0, 0, 0, 0, 0};
    double r = 0.0;
    if (k < ndom) {
      const uint32_t g = a.s.gbx[k];
      if (g < a.ngbx && k < a.off[g] + a.nact[g]) {
        valid = true;
        r = a.s.r[k];
        const double temp = a.st.temp[g], press = a.st.press[g];
        const double psat = saturation_pressure(temp);
        const double s_ratio = supersaturation_ratio(press, a.st.qv[g], psat);
        drop_features(r, a.s.msol[k], s_ratio, temp, press, a.delt, f);
      }
    }
    float xn[kIn];
    infer_normalise(K, f, valid, xn);
    const float yn = mlp_infer(F, K, xn);
    if (k < ndom) a.warm[k] = valid ? learned_z(K, r, yn) : __longlong_as_double(0x7FF8000000000000ll);
  }
}

struct SampleArgs {
  SoA s;
  const uint32_t* off;
  const uint32_t* nact;
  GState st;
  Solute sol;
  NewtonCfg cfg;
  uint32_t ngbx;
  float* x;        // [n][8]
  float* y;        // [n]
  uint32_t* flag;  // [n]
};

// (features, target) of every active drop from the plain solver, State and drops unchanged
__global__ void __launch_bounds__(256) k_mlp_samples(SampleArgs a) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint32_t ndom = a.off[a.ngbx];
  if (k >= ndom) return;
  const uint32_t g = a.s.gbx[k];
  uint32_t ok = 0;
  if (g < a.ngbx && k < a.off[g] + a.nact[g]) {
    const double r = a.s.r[k], msol = a.s.msol[k];
    const double temp = a.st.temp[g], press = a.st.press[g];
    const double psat = saturation_pressure(temp);
    const double s_ratio = supersaturation_ratio(press, a.st.qv[g], psat);
    const double ffactor = diffusion_factor(press, temp, psat);
    double akoh, bkoh;
    kohler_factors(msol, temp, a.sol, akoh, bkoh);
    SolveOut o{0, false, false, false, false};
    const double newr = solve_condensation(a.cfg, s_ratio, akoh, bkoh, rcp_nr(akoh), ffactor, r,
                                           __longlong_as_double(0x7FF8000000000000ll), 0u, o);
    if (!o.substep && !o.maxed && newr > 0.0) {
      float f[kIn];
      drop_features(r, msol, s_ratio, temp, press, a.cfg.delt, f);
      for (int j = 0; j < kIn; ++j) a.x[k * kInPad + j] = f[j];
      a.x[k * kInPad + 6] = 0.0f;
      a.x[k * kInPad + 7] = 0.0f;
      a.y[k] = static_cast<float>(log(newr) - log(r));
      ok = 1;
    }
  }
  a.flag[k] = ok;
}

__global__ void k_compact_samples(const float* x, const float* y, const uint32_t* flag, const uint32_t* pos,
                                  uint64_t n, uint64_t cap, float* xo, float* yo) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= n || !flag[k]) return;
  const uint64_t d = pos[k];
  if (d >= cap) return;
  for (int j = 0; j < kInPad; ++j) xo[d * kInPad + j] = x[k * kInPad + j];
  yo[d] = y[k];
}

struct TrainArgs {
  const float* x;        // [n][8]
  const float* y;        // [n]
  const uint32_t* perm;  // sample order
  uint64_t first;        // batch start in `perm`
  uint32_t count;        // samples in batch
  float gscale;          // 2 / batch (mean squared error)
  MLPWeights W;
  float* partial;        // [waves][NPARAM + 1]
};

// loss gradient of 64 samples (one wave) -> partial[wave]
__global__ void __launch_bounds__(64) k_mlp_grad(TrainArgs a) {
  __shared__ WaveLds L;
  const int lane = lane_id();
  const int col = lane % 32;
  const uint32_t s0 = blockIdx.x * 64;
  const bool valid = s0 + lane < a.count;
  float f[kIn] = {0, 0, 0, 0, 0, 0};
  float t = 0.0f;
  if (valid) {
    const uint64_t idx = a.perm ? a.perm[a.first + s0 + lane] : a.first + s0 + lane;
    for (int j = 0; j < kIn; ++j) f[j] = a.x[idx * kInPad + j];
    t = (a.y[idx] - a.W.p[P_YMEAN]) / a.W.p[P_YSTD];
  }
  load_normalised(L, a.W.p, f, valid);
  f32x16 z1[2][2], z2[2][2];
  const float yp = mlp_forward_wave(a.W, L, z1, z2);
  const float dy = valid ? a.gscale * (yp - t) : 0.0f;
  float* out = a.partial + static_cast<uint64_t>(blockIdx.x) * (NPARAM + 1);
  // loss partial: sum of squared errors
  float se = valid ? (yp - t) * (yp - t) : 0.0f;
  for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o);
  L.ys[lane] = dy;
  wave_sync();
  // dW3, db3, dz2
  float gw3[2] = {0.0f, 0.0f};
  float gb2[2] = {0.0f, 0.0f};
  const float w3v[2] = {a.W.p[P_W3 + col], a.W.p[P_W3 + 32 + col]};
  for (int mi = 0; mi < 2; ++mi) {
    for (int v = 0; v < 16; ++v) {
      const int s = mi * 32 + drow(v, lane);
      const float dys = L.ys[s];
      for (int ni = 0; ni < 2; ++ni) {
        const float z = z2[mi][ni][v];
        gw3[ni] += dys * silu(z);
        const float d = dys * w3v[ni] * dsilu(z);
        gb2[ni] += d;
        L.dz2[s][ni * 32 + col] = f2bf(d);
      }
    }
  }
  for (int ni = 0; ni < 2; ++ni) {
    gw3[ni] += __shfl_xor(gw3[ni], 32);
    gb2[ni] += __shfl_xor(gb2[ni], 32);
  }
  float db3 = dy;
  for (int o = 32; o > 0; o >>= 1) db3 += __shfl_xor(db3, o);
  if (lane < 32) {
    for (int ni = 0; ni < 2; ++ni) {
      out[P_W3 + ni * 32 + col] = gw3[ni];
      out[P_B2 + ni * 32 + col] = gb2[ni];
    }
  }
  if (lane == 0) {
    out[P_B3] = db3;
    out[NPARAM] = se;
  }
  wave_sync();
  const int kq = 4 * (lane / 32);
  // dW2[o][i] = sum_s dz2[s][o] h1[s][i]  (A = dz2^T, B = h1, K = samples)
  for (int mi = 0; mi < 2; ++mi) {
    for (int ni = 0; ni < 2; ++ni) {
      f32x16 acc = {};
      for (int kk = 0; kk < 8; ++kk) {
        const int sb = kk * 8 + kq;
        const bf16x4 av = {static_cast<short>(L.dz2[sb][mi * 32 + col]), static_cast<short>(L.dz2[sb + 1][mi * 32 + col]),
                           static_cast<short>(L.dz2[sb + 2][mi * 32 + col]), static_cast<short>(L.dz2[sb + 3][mi * 32 + col])};
        const bf16x4 bv = {static_cast<short>(L.h1[sb][ni * 32 + col]), static_cast<short>(L.h1[sb + 1][ni * 32 + col]),
                           static_cast<short>(L.h1[sb + 2][ni * 32 + col]), static_cast<short>(L.h1[sb + 3][ni * 32 + col])};
        acc = mfma(av, bv, acc);
      }
      for (int v = 0; v < 16; ++v) out[P_W2 + (mi * 32 + drow(v, lane)) * kHid + ni * 32 + col] = acc[v];
    }
  }
  // dh1[s][i] = sum_o dz2[s][o] W2[o][i]  (A = dz2, B = W2 via its transpose copy), dz1 = dh1 * SiLU'(z1)
  float gb1[2] = {0.0f, 0.0f};
  float gw1[2][kIn] = {};
  for (int mi = 0; mi < 2; ++mi) {
    for (int ni = 0; ni < 2; ++ni) {
      f32x16 acc = {};
      for (int kk = 0; kk < 8; ++kk) {
        const uint16_t* ar = &L.dz2[mi * 32 + col][kk * 8 + kq];
        const bf16x4 av = {static_cast<short>(ar[0]), static_cast<short>(ar[1]), static_cast<short>(ar[2]),
                           static_cast<short>(ar[3])};
        const uint16_t* br = a.W.w2t + (ni * 32 + col) * kHid + kk * 8 + kq;
        const bf16x4 bv = {static_cast<short>(br[0]), static_cast<short>(br[1]), static_cast<short>(br[2]),
                           static_cast<short>(br[3])};
        acc = mfma(av, bv, acc);
      }
      for (int v = 0; v < 16; ++v) {
        const int s = mi * 32 + drow(v, lane);
        const float d = acc[v] * dsilu(z1[mi][ni][v]);
        gb1[ni] += d;
        for (int j = 0; j < kIn; ++j) gw1[ni][j] += d * L.xs[s][j];
      }
    }
  }
  for (int ni = 0; ni < 2; ++ni) {
    gb1[ni] += __shfl_xor(gb1[ni], 32);
    for (int j = 0; j < kIn; ++j) gw1[ni][j] += __shfl_xor(gw1[ni][j], 32);
  }
  if (lane < 32) {
    for (int ni = 0; ni < 2; ++ni) {
      out[P_B1 + ni * 32 + col] = gb1[ni];
      for (int j = 0; j < kIn; ++j) out[P_W1 + (ni * 32 + col) * kIn + j] = gw1[ni][j];
    }
  }
}

// first level of the gradient sum: chunk c of kReduceChunk waves, summed in wave order, for every
// parameter (and the loss, entry NPARAM); one thread per (parameter, chunk)
constexpr uint32_t kReduceChunk = 32;
__global__ void k_mlp_reduce_chunks(const float* partial, uint32_t nwaves, float* chunks) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > NPARAM) return;
  const uint32_t w0 = blockIdx.y * kReduceChunk, w1 = min(w0 + kReduceChunk, nwaves);
  float s = 0.0f;
  for (uint32_t w = w0; w < w1; ++w) s += partial[static_cast<uint64_t>(w) * (NPARAM + 1) + p];
  chunks[static_cast<uint64_t>(blockIdx.y) * (NPARAM + 1) + p] = s;
}

// grad[p] = the chunk sums added in chunk order (a fixed order: training is deterministic); then
// Adam on the fp32 master weights
__global__ void k_mlp_reduce_adam(const float* chunks, uint32_t nchunks, float* grad, float* loss, float* master,
                                  float* m, float* v, int32_t apply, float lr, float b1, float b2, float eps,
                                  float bc1, float bc2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > NPARAM) return;
  float s = 0.0f;
  for (uint32_t w = 0; w < nchunks; ++w) s += chunks[static_cast<uint64_t>(w) * (NPARAM + 1) + p];
  if (p == NPARAM) {
    *loss = s;
    return;
  }
  grad[p] = s;
  if (!apply) return;
  const float mm = b1 * m[p] + (1.0f - b1) * s;
  const float vv = b2 * v[p] + (1.0f - b2) * s * s;
  m[p] = mm;
  v[p] = vv;
  master[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
}

// bf16 copies of W1 (padded), W2 and W2^T from the fp32 master blob
__global__ void k_mlp_pack(const float* master, uint16_t* w1p, uint16_t* w2, uint16_t* w2t) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < kHid * kInPad) {
    const int o = t / kInPad, k = t % kInPad;
    w1p[t] = k < kIn ? f2bf(master[P_W1 + o * kIn + k]) : 0;
  }
  if (t < kHid * kHid) {
    const int o = t / kHid, i = t % kHid;
    const uint16_t b = f2bf(master[P_W2 + o * kHid + i]);
    w2[o * kHid + i] = b;
    w2t[i * kHid + o] = b;
  }
}

// per-column sums of x (6 columns) and y in double, one partial per block
__global__ void k_mlp_moments(const float* x, const float* y, uint64_t n, double* part) {
  __shared__ double sh[256][14];
  double acc[14] = {};
  for (uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    for (int j = 0; j < kIn; ++j) {
      const double v = x[k * kInPad + j];
      acc[j] += v;
      acc[7 + j] += v * v;
    }
    acc[6] += y[k];
    acc[13] += static_cast<double>(y[k]) * y[k];
  }
  for (int j = 0; j < 14; ++j) sh[threadIdx.x][j] = acc[j];
  __syncthreads();
  if (threadIdx.x < 14) {
    double s = 0.0;
    for (int t = 0; t < 256; ++t) s += sh[t][threadIdx.x];
    part[blockIdx.x * 14 + threadIdx.x] = s;
  }
}

struct ForwardArgs {
  const float* x;
  float* y;
  uint64_t n;
  MLPWeights W;
};
// forward pass (the inference pass of the guess kernels): y[k] = network(x[k]) in target units
__global__ void __launch_bounds__(256) k_mlp_forward(ForwardArgs a) {
  __shared__ InferConst K;
  infer_const_load(K, a.W.p);
  __syncthreads();
  const InferFrags F = infer_frags(a.W);
  for (uint64_t k0 = static_cast<uint64_t>(blockIdx.x) * blockDim.x + (threadIdx.x & ~63u); k0 < a.n;
       k0 += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t k = k0 + lane_id();
    float f[kIn] = {0, 0, 0, 0, 0, 0};
    const bool valid = k < a.n;
    if (valid)
      for (int j = 0; j < kIn; ++j) f[j] = a.x[k * kInPad + j];
    float xn[kIn];
    infer_normalise(K, f, valid, xn);
    const float yn = mlp_infer(F, K, xn);
    if (valid) a.y[k] = yn * K.ystd + K.ymean;
  }
}

// ------------------------------------------------------------------------------------------------
struct MLPDev {
  bool ready = false;
  float* master = nullptr;  // NBLOB fp32
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  float* grad = nullptr;
  float* loss = nullptr;
  uint16_t* w1p = nullptr;
  uint16_t* w2 = nullptr;
  uint16_t* w2t = nullptr;
  double* warm = nullptr;   // maxn
  float* partial = nullptr;
  uint32_t partial_waves = 0;
  MLPWeights weights() const { return MLPWeights{w1p, w2, w2t, master}; }
};

// per-wave partial gradients, followed by room for the chunk sums of k_mlp_reduce_chunks
static uint32_t reduce_nchunks(uint32_t waves) { return (waves + kReduceChunk - 1) / kReduceChunk; }
static hipError_t ensure_partial(MLPDev* m, uint32_t waves) {
  if (m->partial_waves >= waves) return hipSuccess;
  hipFree(m->partial);
  m->partial = nullptr;
  m->partial_waves = 0;
  const size_t rows = static_cast<size_t>(waves) + reduce_nchunks(waves);
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&m->partial), rows * (NPARAM + 1) * 4);
  if (e == hipSuccess) m->partial_waves = waves;
  return e;
}
// the two-level gradient sum of `waves` per-wave partials, then Adam when `apply`
static void launch_reduce_adam(hipStream_t s, MLPDev* m, uint32_t waves, int32_t apply, float lr, float b1, float b2,
                               float eps, float bc1, float bc2) {
  const uint32_t nch = reduce_nchunks(waves);
  float* ch = m->partial + static_cast<uint64_t>(m->partial_waves) * (NPARAM + 1);
  hipLaunchKernelGGL(k_mlp_reduce_chunks, dim3((NPARAM + 256) / 256, nch), dim3(256), 0, s,
                     static_cast<const float*>(m->partial), waves, ch);
  hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((NPARAM + 256) / 256), dim3(256), 0, s, static_cast<const float*>(ch),
                     nch, m->grad, m->loss, m->master, m->adam_m, m->adam_v, apply, lr, b1, b2, eps, bc1, bc2);
}

hipError_t mlp_create(cleo_ctx* c) {
  MLPDev* m = new MLPDev();
  c->mlp = m;
  hipError_t e;
  if ((e = hipMalloc(reinterpret_cast<void**>(&m->master), NBLOB * 4)) != hipSuccess) return e;
  if ((e = hipMalloc(reinterpret_cast<void**>(&m->adam_m), NBLOB * 4)) != hipSuccess) return e;
  if ((e = hipMalloc(reinterpret_cast<void**>(&m->adam_v), NBLOB * 4)) != hipSuccess) return e;
  if ((e = hipMalloc(reinterpret_cast<void**>(&m->grad), NBLOB * 4)) != hipSuccess) return e;
  if ((e = hipMalloc(reinterpret_cast<void**>(&m->loss), 16)) != hipSuccess) return e;
  if ((e = hipMalloc(reinterpret_cast<void**>(&m->w1p), kHid * kInPad * 2)) != hipSuccess) return e;
  if ((e = hipMalloc(reinterpret_cast<void**>(&m->w2), kHid * kHid * 2)) != hipSuccess) return e;
  if ((e = hipMalloc(reinterpret_cast<void**>(&m->w2t), kHid * kHid * 2)) != hipSuccess) return e;
  return hipSuccess;
}

void mlp_destroy(cleo_ctx* c) {
  MLPDev* m = c->mlp;
  if (!m) return;
  for (void* p : {static_cast<void*>(m->master), static_cast<void*>(m->adam_m), static_cast<void*>(m->adam_v),
                  static_cast<void*>(m->grad), static_cast<void*>(m->loss), static_cast<void*>(m->w1p),
                  static_cast<void*>(m->w2), static_cast<void*>(m->w2t), static_cast<void*>(m->warm),
                  static_cast<void*>(m->partial)})
    if (p) hipFree(p);
  delete m;
  c->mlp = nullptr;
}

int mlp_check_ready(cleo_ctx* c) {
  if (!c->mlp || !c->mlp->ready) {
    c->msg = "MLP weights not set (cleo_mlp_set_weights or cleo_mlp_train)";
    return CLEO_E_STATE;
  }
  MLPDev* m = c->mlp;
  if (!m->warm) {  // allocated here, outside any graph capture
    if (hipMalloc(reinterpret_cast<void**>(&m->warm), std::max<uint64_t>(c->maxn, 1) * 8) != hipSuccess) {
      c->msg = "hipMalloc of the guess buffer failed";
      return CLEO_E_NOMEM;
    }
  }
  return CLEO_OK;
}

int mlp_enqueue_guess(cleo_ctx* c, const cleo_cond_params& p, const double** warm) {
  if (!c->mlp || !c->mlp->ready || !c->mlp->warm) return mlp_check_ready(c);
  MLPDev* m = c->mlp;
  GuessArgs a{};
  a.s = c->buf[c->cur];
  a.off = c->off;
  a.nact = c->nact;
  a.st = c->st;
  a.sol = c->solute;
  a.delt = static_cast<double>(p.interval) / 100000.0;
  a.ngbx = c->grid.ngbx;
  a.ntot = c->ntot;
  a.W = m->weights();
  a.warm = m->warm;
  const uint64_t nb = std::min<uint64_t>(8192, (c->ntot + 255) / 256);
  if (nb > 0) hipLaunchKernelGGL(k_mlp_guess, dim3(static_cast<unsigned>(nb)), dim3(256), 0, c->stream, a);
  *warm = m->warm;
  return hipGetLastError() == hipSuccess ? CLEO_OK : CLEO_E_HIP;
}

MLPWeights mlp_weights(cleo_ctx* c) { return c->mlp->weights(); }
double* mlp_warm_buffer(cleo_ctx* c) { return c->mlp->warm; }

}  // namespace cleo

using namespace cleo;

#define MLP_TRY(c, expr)                                                   \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess) {                                                \
      (c)->msg = std::string(#expr) + ": " + hipGetErrorString(e_);        \
      return CLEO_E_HIP;                                                   \
    }                                                                      \
  } while (0)

static void launch_pack(cleo_ctx* c) {
  MLPDev* m = c->mlp;
  hipLaunchKernelGGL(k_mlp_pack, dim3(16), dim3(256), 0, c->stream, m->master, m->w1p, m->w2, m->w2t);
}

extern "C" {

int cleo_mlp_set_weights(cleo_ctx* c, const float* blob) {
  if (!c || !blob) return CLEO_E_ARG;
  for (int j = 0; j < kIn; ++j)
    if (!(blob[P_XSTD + j] > 0.0f)) return CLEO_E_ARG;
  if (!(blob[P_YSTD] > 0.0f)) return CLEO_E_ARG;
  MLP_TRY(c, hipStreamSynchronize(c->stream));
  MLP_TRY(c, ctx_copy(c, c->mlp->master, blob, NBLOB * 4, hipMemcpyHostToDevice));
  launch_pack(c);
  MLP_TRY(c, hipGetLastError());
  MLP_TRY(c, hipStreamSynchronize(c->stream));
  c->mlp->ready = true;
  return mlp_check_ready(c);
}

int cleo_mlp_get_weights(cleo_ctx* c, float* blob) {
  if (!c || !blob) return CLEO_E_ARG;
  MLP_TRY(c, hipStreamSynchronize(c->stream));
  MLP_TRY(c, ctx_copy(c, blob, c->mlp->master, NBLOB * 4, hipMemcpyDeviceToHost));
  return CLEO_OK;
}

int cleo_mlp_make_samples(cleo_ctx* c, const cleo_cond_params* p, float* x_out, float* y_out, uint64_t cap,
                          uint64_t* nwritten) {
  if (!c || !p || !nwritten) return CLEO_E_ARG;
  const uint64_t n = c->ntot;
  *nwritten = 0;
  if (n == 0) return CLEO_OK;
  materialize_order(c);  // samples in drop order
  float *x = nullptr, *y = nullptr;
  uint32_t *flag = nullptr, *pos = nullptr;
  MLP_TRY(c, hipMalloc(reinterpret_cast<void**>(&x), n * kInPad * 4));
  MLP_TRY(c, hipMalloc(reinterpret_cast<void**>(&y), n * 4));
  MLP_TRY(c, hipMalloc(reinterpret_cast<void**>(&flag), n * 4));
  MLP_TRY(c, hipMalloc(reinterpret_cast<void**>(&pos), (n + 1) * 4));
  MLP_TRY(c, hipMemsetAsync(flag, 0, n * 4, c->stream));
  SampleArgs a{};
  a.s = c->buf[c->cur];
  a.off = c->off;
  a.nact = c->nact;
  a.st = c->st;
  a.sol = c->solute;
  a.cfg = NewtonCfg{static_cast<double>(p->interval) / 100000.0, p->minsubdelt_s / dimless::TIME0, p->rtol,
                    p->atol, p->maxniters, 0u};
  a.ngbx = c->grid.ngbx;
  a.x = x;
  a.y = y;
  a.flag = flag;
  hipLaunchKernelGGL(k_mlp_samples, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, c->stream, a);
  // exclusive positions of the kept samples
  const size_t need = scan_tmp_words(n);
  uint32_t* saved_tmp = c->scan_tmp;
  size_t saved_n = c->scan_tmp_n;
  uint32_t* tmp = nullptr;
  if (need > c->scan_tmp_n) {
    MLP_TRY(c, hipMalloc(reinterpret_cast<void**>(&tmp), need * 4));
    c->scan_tmp = tmp;
    c->scan_tmp_n = need;
  }
  scan_u32(c, flag, n, 1u, 0u, pos, nullptr);
  uint32_t total = 0;
  MLP_TRY(c, hipMemcpyAsync(&total, pos + n, 4, hipMemcpyDeviceToHost, c->stream));
  MLP_TRY(c, hipStreamSynchronize(c->stream));
  const uint64_t keep = std::min<uint64_t>(total, cap);
  if (x_out && y_out && keep > 0)
    hipLaunchKernelGGL(k_compact_samples, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, c->stream, x, y,
                       flag, pos, n, cap, x_out, y_out);
  MLP_TRY(c, hipGetLastError());
  MLP_TRY(c, hipStreamSynchronize(c->stream));
  if (tmp) {
    hipFree(tmp);
    c->scan_tmp = saved_tmp;
    c->scan_tmp_n = saved_n;
  }
  hipFree(x);
  hipFree(y);
  hipFree(flag);
  hipFree(pos);
  *nwritten = keep;
  return CLEO_OK;
}

// loss and gradient of the current weights on n samples (device x, y); grad_host: NPARAM floats
int cleo_mlp_loss_grad(cleo_ctx* c, const float* x, const float* y, uint64_t n, float* loss, float* grad_host) {
  if (!c || !x || !y || n == 0 || n > (1ull << 31)) return CLEO_E_ARG;
  const int rc = mlp_check_ready(c);
  if (rc != CLEO_OK) return rc;
  MLPDev* m = c->mlp;
  const uint32_t waves = static_cast<uint32_t>((n + 63) / 64);
  MLP_TRY(c, ensure_partial(m, waves));
  TrainArgs a{x, y, nullptr, 0, static_cast<uint32_t>(n), 2.0f / static_cast<float>(n), m->weights(), m->partial};
  hipLaunchKernelGGL(k_mlp_grad, dim3(waves), dim3(64), 0, c->stream, a);
  launch_reduce_adam(c->stream, m, waves, 0, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f);
  MLP_TRY(c, hipGetLastError());
  MLP_TRY(c, hipStreamSynchronize(c->stream));
  float l = 0.f;
  MLP_TRY(c, ctx_copy(c, &l, m->loss, 4, hipMemcpyDeviceToHost));
  if (loss) *loss = l / static_cast<float>(n);
  if (grad_host) MLP_TRY(c, ctx_copy(c, grad_host, m->grad, NPARAM * 4, hipMemcpyDeviceToHost));
  return CLEO_OK;
}

int cleo_mlp_forward(cleo_ctx* c, const float* x, float* y, uint64_t n) {
  if (!c || !x || !y) return CLEO_E_ARG;
  const int rc = mlp_check_ready(c);
  if (rc != CLEO_OK) return rc;
  if (n == 0) return CLEO_OK;
  ForwardArgs a{x, y, n, c->mlp->weights()};
  const uint64_t nb = std::min<uint64_t>(8192, (n + 255) / 256);
  hipLaunchKernelGGL(k_mlp_forward, dim3(static_cast<unsigned>(nb)), dim3(256), 0, c->stream, a);
  MLP_TRY(c, hipGetLastError());
  MLP_TRY(c, hipStreamSynchronize(c->stream));
  return CLEO_OK;
}

// Adam training on device-resident samples; initialises weights (He-uniform, seeded) and the
// normalisation from the sample moments unless weights were set before and params->seed == 0.
int cleo_mlp_train(cleo_ctx* c, const float* x, const float* y, uint64_t n, const cleo_mlp_train_params* tp,
                   float* final_loss) {
  if (!c || !x || !y || !tp || n < 64 || tp->batch == 0 || tp->epochs == 0) return CLEO_E_ARG;
  MLPDev* m = c->mlp;
  // normalisation from the sample moments
  const unsigned mb = 256;
  double* part = nullptr;
  MLP_TRY(c, hipMalloc(reinterpret_cast<void**>(&part), mb * 14 * 8));
  hipLaunchKernelGGL(k_mlp_moments, dim3(mb), dim3(256), 0, c->stream, x, y, n, part);
  std::vector<double> hp(mb * 14);
  MLP_TRY(c, hipMemcpyAsync(hp.data(), part, mb * 14 * 8, hipMemcpyDeviceToHost, c->stream));
  MLP_TRY(c, hipStreamSynchronize(c->stream));
  hipFree(part);
  double mom[14] = {};
  for (unsigned b = 0; b < mb; ++b)
    for (int j = 0; j < 14; ++j) mom[j] += hp[b * 14 + j];
  std::vector<float> blob(NBLOB, 0.0f);
  const bool fresh = !(m->ready && tp->seed == 0);
  if (!fresh) MLP_TRY(c, ctx_copy(c, blob.data(), m->master, NBLOB * 4, hipMemcpyDeviceToHost));
  if (fresh) {
    for (int j = 0; j < kIn; ++j) {
      const double mu = mom[j] / n;
      const double var = std::max(mom[7 + j] / n - mu * mu, 0.0);
      blob[P_XMEAN + j] = static_cast<float>(mu);
      blob[P_XSTD + j] = static_cast<float>(std::max(std::sqrt(var), 1e-6));
    }
    const double ym = mom[6] / n;
    blob[P_YMEAN] = static_cast<float>(ym);
    blob[P_YSTD] = static_cast<float>(std::max(std::sqrt(std::max(mom[13] / n - ym * ym, 0.0)), 1e-9));
    // seeded He-uniform initialisation (splitmix64 stream)
    uint64_t st = tp->seed ? tp->seed : 0x9E3779B97F4A7C15ull;
    auto urand = [&st]() {
      uint64_t z = (st += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      return static_cast<double>(z >> 11) * 0x1.0p-53;
    };
    auto fill = [&](int off, int cnt, int fan_in) {
      const double lim = std::sqrt(6.0 / fan_in);
      for (int i = 0; i < cnt; ++i) blob[off + i] = static_cast<float>((2.0 * urand() - 1.0) * lim);
    };
    fill(P_W1, kHid * kIn, kIn);
    fill(P_W2, kHid * kHid, kHid);
    fill(P_W3, kHid, kHid);
  }
  MLP_TRY(c, ctx_copy(c, m->master, blob.data(), NBLOB * 4, hipMemcpyHostToDevice));
  MLP_TRY(c, ctx_fill(c, m->adam_m, 0, NBLOB * 4));
  MLP_TRY(c, ctx_fill(c, m->adam_v, 0, NBLOB * 4));
  launch_pack(c);
  m->ready = true;
  // sample order: one seeded permutation, batches visited in a seeded order each epoch
  std::vector<uint32_t> perm(n);
  for (uint64_t i = 0; i < n; ++i) perm[i] = static_cast<uint32_t>(i);
  uint64_t st = tp->seed ^ 0xD1B54A32D192ED03ull;
  auto next = [&st]() {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
  };
  for (uint64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[next() % (i + 1)]);
  uint32_t* dperm = nullptr;
  MLP_TRY(c, hipMalloc(reinterpret_cast<void**>(&dperm), n * 4));
  MLP_TRY(c, ctx_copy(c, dperm, perm.data(), n * 4, hipMemcpyHostToDevice));
  const uint64_t nbatch = n / tp->batch ? n / tp->batch : 1;
  const uint32_t bsz = static_cast<uint32_t>(std::min<uint64_t>(tp->batch, n));
  const uint32_t waves = (bsz + 63) / 64;
  MLP_TRY(c, ensure_partial(m, waves));
  std::vector<uint64_t> order(nbatch);
  uint64_t step = 0;
  for (uint32_t ep = 0; ep < tp->epochs; ++ep) {
    for (uint64_t b = 0; b < nbatch; ++b) order[b] = b;
    for (uint64_t b = nbatch - 1; b > 0; --b) std::swap(order[b], order[next() % (b + 1)]);
    for (uint64_t bi = 0; bi < nbatch; ++bi) {
      ++step;
      TrainArgs a{x, y, dperm, order[bi] * bsz, bsz, 2.0f / static_cast<float>(bsz), m->weights(), m->partial};
      hipLaunchKernelGGL(k_mlp_grad, dim3(waves), dim3(64), 0, c->stream, a);
      const float bc1 = 1.0f - std::pow(tp->beta1, static_cast<float>(step));
      const float bc2 = 1.0f - std::pow(tp->beta2, static_cast<float>(step));
      launch_reduce_adam(c->stream, m, waves, 1, tp->lr, tp->beta1, tp->beta2, tp->eps, bc1, bc2);
      launch_pack(c);
    }
  }
  MLP_TRY(c, hipGetLastError());
  MLP_TRY(c, hipStreamSynchronize(c->stream));
  hipFree(dperm);
  const int rc = mlp_check_ready(c);
  if (rc != CLEO_OK) return rc;
  if (final_loss) {
    float l = 0.0f;
    MLP_TRY(c, ctx_copy(c, &l, m->loss, 4, hipMemcpyDeviceToHost));
    *final_loss = l / static_cast<float>(bsz);
  }
  return CLEO_OK;
}

}  // extern "C"
