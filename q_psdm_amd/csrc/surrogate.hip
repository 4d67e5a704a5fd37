// Learned source-leg surrogate: an MLP 5 -> 128 -> 128 -> 128 -> 128 -> 3
// (SiLU) mapping (sx, sy, x, y, z) to (t, t*, log A) of the straight ray from
// the surface source to the point, as residuals over the laterally
// homogeneous (anomaly-free) part of the model, which is evaluated in closed
// form.  BF16 operands on MFMA (v_mfma_f32_32x32x16_bf16), FP32 accumulation,
// FP32 master weights with Adam.
//
// Register layout.  Activations are kept transposed, units x points: a layer
// is D[out][pt] = W[out][:] . H[:][pt], with the weights as the MFMA A operand
// and a 32-point group as the B operand.  The accumulator of a 32x32 block
// holds, in lane l (half h = l / 32) and register i, unit
// 32 nb + 8 (i / 4) + 4 h + i % 4 of point l % 32.  Those 16 values become
// two k-steps of the next layer's B operand without any data movement, by
// numbering the next layer's contraction index in that order ("virtual k"):
// the packed weight arrays below hold each A operand chunk (64 lanes x 8
// bf16) with its columns permuted accordingly, forward (W) and backward
// (W^T) copies both, regenerated from the FP32 master after every update.
//
// Kernels: sample generator (FP64 ray integrals), forward (evaluation,
// training with stored activations, per-shot source tables, validity mask),
// backward through the layers, weight-gradient GEMM split over sample
// chunks (fixed-order reduction, no atomics), Adam + repacking.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "image_dev.h"
#include "ray_dev.h"

namespace qk = qpsdm_dev;

namespace {

typedef __bf16 bf16;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

constexpr int kH = 128, kIn = 5, kOut = 3;
constexpr int kKS = kH / 16;  // k-steps over a hidden layer
constexpr int kMB = kH / 32;  // 32-row blocks of a hidden layer
constexpr int kChunk = 512;   // bf16 per (block, k-step) operand chunk: 64 lanes x 8

// FP32 master parameters: W1 [128][5], b1, W2..W4 [128][128], b2..b4, W5 [3][128], b5.
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH * kH, oW4 = oB3 + kH, oB4 = oW4 + kH * kH, oW5 = oB4 + kH, oB5 = oW5 + kOut * kH,
              kNP = oB5 + kOut;
// Packed forward operands (bf16): W1 (4 blocks x 1 k-step), W2..W4, W5 (1 block x 8 k-steps).
constexpr int fW1 = 0, fW2 = fW1 + kMB * kChunk, fW3 = fW2 + kMB * kKS * kChunk, fW4 = fW3 + kMB * kKS * kChunk,
              fW5 = fW4 + kMB * kKS * kChunk, kNF = fW5 + kKS * kChunk;
// Packed backward operands: W5^T (4 blocks x 1 k-step), W4^T, W3^T, W2^T.
constexpr int bW5 = 0, bW4 = bW5 + kMB * kChunk, bW3 = bW4 + kMB * kKS * kChunk, bW2 = bW3 + kMB * kKS * kChunk,
              kNB = bW2 + kMB * kKS * kChunk;
constexpr int kNBias = 4 * kH + kOut;  // b1..b4 then b5, FP32
// Weight-gradient blocks (32 x 32): layer 1 (4), layers 2-4 (16 each), layer 5 (4).
constexpr int kNBlk = 4 + 3 * 16 + 4;
constexpr int kGradChunk = 2048;  // samples per weight-gradient work item

__host__ __device__ inline int vk_unit(int s, int h, int j) {
  return 32 * (s >> 1) + 16 * (s & 1) + 8 * (j >> 2) + 4 * h + (j & 3);
}
__host__ __device__ inline void vk_slot(int u, int *s, int *h, int *j) {
  const int r = u & 31;
  *s = 2 * (u >> 5) + (r >> 4);
  *h = (r >> 2) & 1;
  *j = 4 * ((r >> 3) & 1) + (r & 3);
}
__host__ __device__ inline int d_row(int i, int h) { return (i >> 2) * 8 + 4 * h + (i & 3); }

__host__ __device__ inline int fpack(int l, int o, int i) {
  if (l == 1) return fW1 + (o >> 5) * kChunk + (o & 31) * 8 + i;
  int s, h, j;
  vk_slot(i, &s, &h, &j);
  if (l == 5) return fW5 + s * kChunk + (32 * h + o) * 8 + j;
  const int base = l == 2 ? fW2 : (l == 3 ? fW3 : fW4);
  return base + ((o >> 5) * kKS + s) * kChunk + (32 * h + (o & 31)) * 8 + j;
}
__host__ __device__ inline int bpack(int l, int o, int i) {  // l in 2..5
  if (l == 5) return bW5 + (i >> 5) * kChunk + (i & 31) * 8 + o;
  int s, h, j;
  vk_slot(o, &s, &h, &j);
  const int base = l == 4 ? bW4 : (l == 3 ? bW3 : bW2);
  return base + ((i >> 5) * kKS + s) * kChunk + (32 * h + (i & 31)) * 8 + j;
}

// parameter index -> (layer 1..5, row o, column i); i = -1 for a bias
__host__ __device__ inline void param_decode(int p, int *l, int *o, int *i) {
  if (p < oB1) { *l = 1; *o = p / kIn; *i = p % kIn; return; }
  if (p < oW2) { *l = 1; *o = p - oB1; *i = -1; return; }
  if (p < oB2) { *l = 2; *o = (p - oW2) / kH; *i = (p - oW2) % kH; return; }
  if (p < oW3) { *l = 2; *o = p - oB2; *i = -1; return; }
  if (p < oB3) { *l = 3; *o = (p - oW3) / kH; *i = (p - oW3) % kH; return; }
  if (p < oW4) { *l = 3; *o = p - oB3; *i = -1; return; }
  if (p < oB4) { *l = 4; *o = (p - oW4) / kH; *i = (p - oW4) % kH; return; }
  if (p < oW5) { *l = 4; *o = p - oB4; *i = -1; return; }
  if (p < oB5) { *l = 5; *o = (p - oW5) / kH; *i = (p - oW5) % kH; return; }
  *l = 5; *o = p - oB5; *i = -1;
}
__host__ __device__ inline int bias_slot(int l, int o) { return (l - 1) * kH + o; }
// weight-gradient block of W_l[o][i] and its (row, column) inside the block
__host__ __device__ inline int grad_block(int l, int o, int i) {
  if (l == 1) return o >> 5;
  if (l == 5) return 52 + (i >> 5);
  return 4 + (l - 2) * 16 + (o >> 5) * 4 + (i >> 5);
}

// Normalisation and the closed-form baseline (the model without its anomaly).
struct Norm {
  float lo[kIn], sc[kIn];  // input v -> (v - lo) * sc - 1
  float osc[kOut];         // residual scale of (t, t*, log A)
  float lw[kOut];          // loss weights
};
struct Base {
  int kind;
  float v0, kgrad, q0;
  int nlayer;
  float layer_dz;
  float vl[16], ql[16];
};

// Straight-ray (t, t*, -log sigma) through the anomaly-free model, FP32:
// piecewise exact per layer (the FP64 integrator's layer pieces with the
// anomaly term removed), closed form for linear v(z).
__device__ inline void base_leg(const Base &b, float sx, float sy, float x, float y, float z, float *t,
                                float *ts, float *la) {
  const float hx = x - sx, hy = y - sy;
  const float L = sqrtf(hx * hx + hy * hy + z * z);
  if (b.kind == 0) {
    const float kz = b.kgrad * z;
    float tt, sg;
    if (kz > 1e-6f * b.v0) {
      tt = L / kz * log1pf(kz / b.v0);
      sg = L * (b.v0 + 0.5f * kz);
    } else {
      tt = L / b.v0;
      sg = L * b.v0;
    }
    *t = tt;
    *ts = tt / b.q0;
    *la = -logf(fmaxf(sg, 1e-30f));
    return;
  }
  float at = 0.f, aq = 0.f, av = 0.f, prev = 0.f;
  for (int l = 0; l < b.nlayer; ++l) {
    float nxt = 1.f;
    if (l + 1 < b.nlayer) {
      const float zb = (float)(l + 1) * b.layer_dz;
      if (zb < z) nxt = zb / z;
    }
    if (nxt > prev) {
      const float zm = 0.5f * (prev + nxt) * z;
      int li = (int)(zm / b.layer_dz);
      li = li >= b.nlayer ? b.nlayer - 1 : (li < 0 ? 0 : li);
      const float d = nxt - prev, v = b.vl[li];
      at += d / v;
      aq += d / (v * b.ql[li]);
      av += d * v;
    }
    prev = nxt;
    if (nxt >= 1.f) break;
  }
  *t = at * L;
  *ts = aq * L;
  *la = -logf(fmaxf(av * L, 1e-30f));
}

// SiLU and its derivative with the hardware reciprocal (the epilogue, not
// the matrix product, bounds these kernels: an IEEE division is a
// multi-instruction sequence per value).
__device__ inline float sigm(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigm(z); }
__device__ inline float dsilu(float z) {
  const float s = sigm(z);
  return s * (1.0f + z * (1.0f - s));
}

__device__ inline f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ inline bf16x8 ld8(const bf16 *p) { return *(const bf16x8 *)p; }

// B operand of a 128-wide activation for one 32-point group.
struct Acts {
  bf16x8 k[kKS];
};

// Stores of a training forward (unit-major [unit][n]) through buffer
// resources: the row of a store is a wave-uniform scalar offset, the point a
// lane offset, so no 64-bit addresses are kept per store.
struct TrainStore {
  __amdgpu_buffer_rsrc_t rZ;  // float [4][128][n]
  __amdgpu_buffer_rsrc_t rH;  // bf16 [4][128][n]
  int n, p0;                  // batch size, first point of the wave's 64
};

__device__ inline unsigned short bf16_bits(bf16 v) { return __builtin_bit_cast(unsigned short, v); }

template <bool STORE>
__device__ inline void epilogue(const f32x16 &c, int mb, const float *bias, Acts &out, int layer,
                                const TrainStore &st, int g) {
  const int lane = threadIdx.x & 63, h = lane >> 5;
  const int vp = 4 * h * st.n + st.p0 + 32 * g + (lane & 31);  // lane part of the element index
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int u = 32 * mb + d_row(i, h);
    const float z = c[i] + bias[u];
    const float hv = silu(z);
    const bf16 hb = (bf16)hv;
    out.k[2 * mb + (i >> 3)][i & 7] = hb;
    if (STORE) {
      const int row = layer * kH + 32 * mb + (i >> 2) * 8 + (i & 3);  // wave-uniform part
      __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(z), st.rZ, vp * 4, row * st.n * 4, 0);
      __builtin_amdgcn_raw_buffer_store_b16(bf16_bits(hb), st.rH, vp * 2, row * st.n * 2, 0);
    }
  }
}

template <bool STORE>
__device__ inline void hidden_layer(const bf16 *W, const float *bias, const Acts (&in)[2], Acts (&out)[2],
                                    int layer, const TrainStore &st) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int mb = 0; mb < kMB; ++mb) {
    f32x16 c0 = {}, c1 = {};
#pragma unroll
    for (int s = 0; s < kKS; ++s) {
      const bf16x8 w = ld8(W + (mb * kKS + s) * kChunk + lane * 8);
      c0 = mfma(w, in[0].k[s], c0);
      c1 = mfma(w, in[1].k[s], c1);
    }
    epilogue<STORE>(c0, mb, bias, out[0], layer, st, 0);
    epilogue<STORE>(c1, mb, bias, out[1], layer, st, 1);
    __builtin_amdgcn_sched_barrier(0);  // keep blocks apart: bounds the live registers
  }
}

// Forward pass of two 32-point groups; x[g] is the B operand of the
// normalised inputs (lanes of half 0 carry features 0..4).  r[g][o] is the
// normalised residual output, valid in lanes of half 0.
template <bool STORE>
__device__ inline void forward64(const bf16 *W, const float *bias, const bf16x8 (&x)[2], float (&r)[2][kOut],
                                 const TrainStore &st) {
  const int lane = threadIdx.x & 63;
  Acts a[2], b[2];
#pragma unroll
  for (int mb = 0; mb < kMB; ++mb) {
    const bf16x8 w = ld8(W + fW1 + mb * kChunk + lane * 8);
    f32x16 c0 = {}, c1 = {};
    c0 = mfma(w, x[0], c0);
    c1 = mfma(w, x[1], c1);
    epilogue<STORE>(c0, mb, bias, a[0], 0, st, 0);
    epilogue<STORE>(c1, mb, bias, a[1], 0, st, 1);
    __builtin_amdgcn_sched_barrier(0);
  }
  hidden_layer<STORE>(W + fW2, bias + kH, a, b, 1, st);
  hidden_layer<STORE>(W + fW3, bias + 2 * kH, b, a, 2, st);
  hidden_layer<STORE>(W + fW4, bias + 3 * kH, a, b, 3, st);
  f32x16 c0 = {}, c1 = {};
#pragma unroll
  for (int s = 0; s < kKS; ++s) {
    const bf16x8 w = ld8(W + fW5 + s * kChunk + lane * 8);
    c0 = mfma(w, b[0].k[s], c0);
    c1 = mfma(w, b[1].k[s], c1);
  }
#pragma unroll
  for (int o = 0; o < kOut; ++o) {
    r[0][o] = c0[o] + bias[4 * kH + o];
    r[1][o] = c1[o] + bias[4 * kH + o];
  }
}

__device__ inline bf16x8 feature_operand(const Norm &N, const float (&v)[kIn]) {
  const int h = (threadIdx.x & 63) >> 5;
  bf16x8 x;
#pragma unroll
  for (int j = 0; j < 8; ++j) x[j] = (bf16)0.0f;
  if (h == 0) {
#pragma unroll
    for (int f = 0; f < kIn; ++f) x[f] = (bf16)((v[f] - N.lo[f]) * N.sc[f] - 1.0f);
  }
  return x;
}

constexpr int kThreads = 256;  // 4 waves; one work-group per CU (weights in LDS)

__device__ inline void load_weights(bf16 *lw, float *lb, const bf16 *w, int nw, const float *bias) {
  for (int i = threadIdx.x; i < nw / 8; i += blockDim.x) ((bf16x8 *)lw)[i] = ((const bf16x8 *)w)[i];
  if (bias)
    for (int i = threadIdx.x; i < kNBias; i += blockDim.x) lb[i] = bias[i];
  __syncthreads();
}

// ---- forward kernels ---------------------------------------------------

struct FwdArgs {
  const bf16 *wf;
  const float *bias;
  Norm N;
  Base B;
};

// Evaluation on points in physical units: out [n][3] physical.
__global__ __launch_bounds__(kThreads) void eval_kernel(FwdArgs A, const float *__restrict__ X, int n,
                                                        float *__restrict__ out) {
  __shared__ __attribute__((aligned(16))) bf16 lw[kNF];
  __shared__ float lb[kNBias];
  load_weights(lw, lb, A.wf, kNF, A.bias);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ngroups = (n + 63) / 64;
  const TrainStore st{};
  for (int grp = blockIdx.x * 4 + wave; grp < ngroups; grp += gridDim.x * 4) {
    bf16x8 x[2];
    float v[2][kIn];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int p = grp * 64 + 32 * g + (lane & 31);
#pragma unroll
      for (int f = 0; f < kIn; ++f) v[g][f] = p < n ? X[(size_t)p * kIn + f] : 0.f;
      x[g] = feature_operand(A.N, v[g]);
    }
    float r[2][kOut];
    forward64<false>(lw, lb, x, r, st);
    if (lane < 32) {
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        const int p = grp * 64 + 32 * g + lane;
        if (p >= n) continue;
        float bt, bs, ba;
        base_leg(A.B, v[g][0], v[g][1], v[g][2], v[g][3], v[g][4], &bt, &bs, &ba);
        out[(size_t)p * kOut + 0] = bt + r[g][0] * A.N.osc[0];
        out[(size_t)p * kOut + 1] = bs + r[g][1] * A.N.osc[1];
        out[(size_t)p * kOut + 2] = ba + r[g][2] * A.N.osc[2];
      }
    }
  }
}

// Training forward on n points (n % 64 == 0): stores the bf16 inputs
// Xb [5][n], pre-activations Z, activations H, the loss gradient G [3][n]
// (bf16) and adds the loss to *loss.
__global__ __launch_bounds__(kThreads) void train_fwd_kernel(FwdArgs A, const float *__restrict__ X,
                                                             const float *__restrict__ T, int n, bf16 *Xb,
                                                             float *Z, bf16 *H, bf16 *G, float *loss) {
  __shared__ __attribute__((aligned(16))) bf16 lw[kNF];
  __shared__ float lb[kNBias];
  load_weights(lw, lb, A.wf, kNF, A.bias);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ngroups = n / 64;
  float lsum = 0.f;
  for (int grp = blockIdx.x * 4 + wave; grp < ngroups; grp += gridDim.x * 4) {
    bf16x8 x[2];
    float v[2][kIn];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
      const int p = grp * 64 + 32 * g + (lane & 31);
#pragma unroll
      for (int f = 0; f < kIn; ++f) v[g][f] = X[(size_t)p * kIn + f];
      x[g] = feature_operand(A.N, v[g]);
      if (lane < 32) {
#pragma unroll
        for (int f = 0; f < kIn; ++f) Xb[(size_t)f * n + p] = x[g][f];
      }
    }
    const TrainStore st{qk::make_rsrc(Z, (unsigned)(sizeof(float) * 4 * kH * (size_t)n)),
                        qk::make_rsrc(H, (unsigned)(sizeof(bf16) * 4 * kH * (size_t)n)), n, grp * 64};
    float r[2][kOut];
    forward64<true>(lw, lb, x, r, st);
    if (lane < 32) {
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        const int p = grp * 64 + 32 * g + lane;
        float bt[kOut];
        base_leg(A.B, v[g][0], v[g][1], v[g][2], v[g][3], v[g][4], &bt[0], &bt[1], &bt[2]);
#pragma unroll
        for (int o = 0; o < kOut; ++o) {
          const float R = (T[(size_t)p * kOut + o] - bt[o]) / A.N.osc[o];
          const float d = r[g][o] - R;
          lsum += A.N.lw[o] * d * d;
          G[(size_t)o * n + p] = (bf16)(2.0f * A.N.lw[o] * d / (float)n);
        }
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off, 64);
  if (lane == 0) atomicAdd(loss, lsum / (float)n);
}

// ---- backward through the layers -----------------------------------------

// Buffers of one layer's Z (float) and dZ (bf16), unit-major [128][n].
struct LayerBufs {
  __amdgpu_buffer_rsrc_t rZ, rD;
};

// The pre-activations of one 32-row block for the lane's point p, loaded
// before the block's MFMAs so that their latency hides behind them.
__device__ inline void bwd_load_z(float (&z)[16], int mb, const LayerBufs &lb, int n, int p) {
  const int h = (threadIdx.x & 63) >> 5;
  const int vp = 4 * h * n + p;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int row = 32 * mb + (i >> 2) * 8 + (i & 3);
    z[i] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(lb.rZ, vp * 4, row * n * 4, 0));
  }
}

template <bool PACK>
__device__ inline void bwd_epilogue(const f32x16 &c, const float (&z)[16], int mb, const LayerBufs &lb, Acts &out,
                                    int n, int p) {
  const int h = (threadIdx.x & 63) >> 5;
  const int vp = 4 * h * n + p;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int row = 32 * mb + (i >> 2) * 8 + (i & 3);
    const float dz = c[i] * dsilu(z[i]);
    const bf16 db = (bf16)dz;
    __builtin_amdgcn_raw_buffer_store_b16(bf16_bits(db), lb.rD, vp * 2, row * n * 2, 0);
    if (PACK) out.k[2 * mb + (i >> 3)][i & 7] = db;
  }
}

template <bool PACK>
__device__ inline void bwd_hidden(const bf16 *Wt, const Acts (&in)[2], Acts (&out)[2], const LayerBufs &lb,
                                  int n, int p0) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int mb = 0; mb < kMB; ++mb) {
    float z0[16], z1[16];
    bwd_load_z(z0, mb, lb, n, p0 + (lane & 31));
    bwd_load_z(z1, mb, lb, n, p0 + 32 + (lane & 31));
    f32x16 c0 = {}, c1 = {};
#pragma unroll
    for (int s = 0; s < kKS; ++s) {
      const bf16x8 w = ld8(Wt + (mb * kKS + s) * kChunk + lane * 8);
      c0 = mfma(w, in[0].k[s], c0);
      c1 = mfma(w, in[1].k[s], c1);
    }
    bwd_epilogue<PACK>(c0, z0, mb, lb, out[0], n, p0 + (lane & 31));
    bwd_epilogue<PACK>(c1, z1, mb, lb, out[1], n, p0 + 32 + (lane & 31));
    __builtin_amdgcn_sched_barrier(0);
  }
}

// dZ_l = (W_{l+1}^T dZ_{l+1}) * SiLU'(Z_l), l = 4..1, from G = dL/dY.
__global__ __launch_bounds__(kThreads) void train_bwd_kernel(const bf16 *wb, int n, const bf16 *__restrict__ G,
                                                             float *Z, bf16 *dZ) {
  __shared__ __attribute__((aligned(16))) bf16 lw[kNB];
  load_weights(lw, nullptr, wb, kNB, nullptr);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5;
  const int ngroups = n / 64;
  const size_t L = (size_t)kH * n;
  LayerBufs lb[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    lb[k].rZ = qk::make_rsrc(Z + k * L, (unsigned)(sizeof(float) * L));
    lb[k].rD = qk::make_rsrc(dZ + k * L, (unsigned)(sizeof(bf16) * L));
  }
  for (int grp = blockIdx.x * 4 + wave; grp < ngroups; grp += gridDim.x * 4) {
    const int p0 = grp * 64;
    bf16x8 gb[2];
#pragma unroll
    for (int g = 0; g < 2; ++g) {
#pragma unroll
      for (int j = 0; j < 8; ++j) gb[g][j] = (bf16)0.0f;
      if (h == 0) {
#pragma unroll
        for (int o = 0; o < kOut; ++o) gb[g][o] = G[(size_t)o * n + p0 + 32 * g + (lane & 31)];
      }
    }
    Acts a[2], b[2];
#pragma unroll
    for (int mb = 0; mb < kMB; ++mb) {
      float z0[16], z1[16];
      bwd_load_z(z0, mb, lb[3], n, p0 + (lane & 31));
      bwd_load_z(z1, mb, lb[3], n, p0 + 32 + (lane & 31));
      const bf16x8 w = ld8(lw + bW5 + mb * kChunk + lane * 8);
      f32x16 c0 = {}, c1 = {};
      c0 = mfma(w, gb[0], c0);
      c1 = mfma(w, gb[1], c1);
      bwd_epilogue<true>(c0, z0, mb, lb[3], a[0], n, p0 + (lane & 31));
      bwd_epilogue<true>(c1, z1, mb, lb[3], a[1], n, p0 + 32 + (lane & 31));
      __builtin_amdgcn_sched_barrier(0);
    }
    bwd_hidden<true>(lw + bW4, a, b, lb[2], n, p0);
    bwd_hidden<true>(lw + bW3, b, a, lb[1], n, p0);
    bwd_hidden<false>(lw + bW2, a, b, lb[0], n, p0);
  }
}

// ---- weight gradients ------------------------------------------------------

// One wave per (32x32 block of one layer's dW, chunk of kGradChunk samples):
// dW[o][i] = sum_s dZ[o][s] H[i][s]; partial [chunk][block][32][32] and the
// bias partial sum_s dZ[o][s] [chunk][layer][128].
__global__ __launch_bounds__(64) void wgrad_kernel(int n, const bf16 *__restrict__ Xb, const bf16 *__restrict__ H,
                                                   const bf16 *__restrict__ dZ, const bf16 *__restrict__ G,
                                                   float *part, float *bpart) {
  const int blk = blockIdx.x, chunk = blockIdx.y;
  const int lane = threadIdx.x, h = lane >> 5, m = lane & 31;
  int l, mb, nb;
  if (blk < 4) {
    l = 1, mb = blk, nb = 0;
  } else if (blk < 52) {
    l = 2 + (blk - 4) / 16, mb = ((blk - 4) % 16) / 4, nb = (blk - 4) % 4;
  } else {
    l = 5, mb = 0, nb = blk - 52;
  }
  const size_t L = (size_t)kH * n;
  const bf16 *Ar;
  bool avalid;
  if (l == 5) {
    Ar = G + (size_t)min(m, kOut - 1) * n;  // rows past the outputs read row 2 and are zeroed
    avalid = m < kOut;
  } else {
    Ar = dZ + (size_t)(l - 1) * L + (size_t)(32 * mb + m) * n;
    avalid = true;
  }
  const bf16 *Br;
  bool bvalid;
  if (l == 1) {
    Br = Xb + (size_t)min(m, kIn - 1) * n;
    bvalid = m < kIn;
  } else {
    Br = H + (size_t)(l - 2) * L + (size_t)(32 * nb + m) * n;
    bvalid = true;
  }
  const int s0 = chunk * kGradChunk;
  const int s1 = min(n, s0 + kGradChunk);
  f32x16 acc = {};
  float bsum = 0.f;
  bf16x8 zero;
#pragma unroll
  for (int j = 0; j < 8; ++j) zero[j] = (bf16)0.0f;
  for (int s = s0; s < s1; s += 16) {
    bf16x8 a = ld8(Ar + s + 8 * h), b = ld8(Br + s + 8 * h);
    if (!avalid) a = zero;
    if (!bvalid) b = zero;
    acc = mfma(a, b, acc);
    if (nb == 0) {
#pragma unroll
      for (int j = 0; j < 8; ++j) bsum += (float)a[j];
    }
  }
  float *o = part + ((size_t)chunk * kNBlk + blk) * 1024;
#pragma unroll
  for (int i = 0; i < 16; ++i) o[d_row(i, h) * 32 + m] = acc[i];
  if (nb == 0) {
    bsum += __shfl_xor(bsum, 32, 64);
    if (h == 0 && avalid) bpart[((size_t)chunk * 5 + (l - 1)) * kH + 32 * mb + m] = bsum;
  }
}

// ---- Adam, repacking ---------------------------------------------------------

struct AdamArgs {
  float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
  int update;                     // 0: gradient only (and repack)
  int nchunk;
};

__global__ void adam_kernel(AdamArgs a, float *w, float *m1, float *m2, const float *part, const float *bpart,
                            float *grad_out, bf16 *wf, bf16 *wb, float *bias) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNP) return;
  int l, o, i;
  param_decode(p, &l, &o, &i);
  if (part) {
    float g = 0.f;
    if (i < 0) {
      for (int c = 0; c < a.nchunk; ++c) g += bpart[((size_t)c * 5 + (l - 1)) * kH + o];
    } else {
      const int blk = grad_block(l, o, i);
      const int r = l == 5 ? o : (o & 31), col = l == 1 ? i : (i & 31);
      for (int c = 0; c < a.nchunk; ++c) g += part[((size_t)c * kNBlk + blk) * 1024 + r * 32 + col];
    }
    if (grad_out) grad_out[p] = g;
    if (a.update) {
      const float mm = a.b1 * m1[p] + (1.f - a.b1) * g;
      const float vv = a.b2 * m2[p] + (1.f - a.b2) * g * g;
      m1[p] = mm;
      m2[p] = vv;
      w[p] -= a.lr * (mm / a.c1) / (sqrtf(vv / a.c2) + a.eps);
    }
  }
  const float v = w[p];
  if (i < 0) {
    bias[bias_slot(l, o)] = v;
  } else {
    wf[fpack(l, o, i)] = (bf16)v;
    if (l >= 2) wb[bpack(l, o, i)] = (bf16)v;
  }
}

// ---- training samples --------------------------------------------------------

__device__ inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ inline double unif(uint64_t seed, uint64_t idx, int k) {
  return (double)(splitmix64(seed ^ splitmix64(idx * 8 + k)) >> 11) * (1.0 / 9007199254740992.0);
}

// Uniform points in the input box; exact (t, t*, -log sigma) by the FP64
// straight-ray integrals of the full model.
__global__ void samples_kernel(qk::ModelDev m, Norm N, uint64_t seed, long long offset, int n, float *X,
                               float *T) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double v[kIn];
    for (int f = 0; f < kIn; ++f) {
      const double lo = N.lo[f], hi = N.lo[f] + 2.0 / N.sc[f];
      v[f] = lo + (hi - lo) * unif(seed, (uint64_t)(offset + i), f);
    }
    double t, ts, sg;
    qk::ray_integrals(m, v[0], v[1], v[2], v[3], v[4], &t, &ts, &sg);
    for (int f = 0; f < kIn; ++f) X[(size_t)i * kIn + f] = (float)v[f];
    T[(size_t)i * kOut + 0] = (float)t;
    T[(size_t)i * kOut + 1] = (float)ts;
    T[(size_t)i * kOut + 2] = (float)(-log(sg));
  }
}

// ---- validity mask against the node tables ----------------------------------

// Surrogate at each node's own source position over that node's FP32 table;
// max |dt| per 8^3 brick into err (float bits, zero-initialised).
__global__ __launch_bounds__(kThreads) void mask_err_kernel(FwdArgs A, qpsdm_tgeom g, const float2 *__restrict__ tab,
                                                            unsigned *err) {
  __shared__ __attribute__((aligned(16))) bf16 lw[kNF];
  __shared__ float lb[kNBias];
  load_weights(lw, lb, A.wf, kNF, A.bias);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long per = (long long)g.nv * g.nu * g.nw;
  const long long total = per * g.nnx * g.nny;
  const long long ngroups = (total + 63) / 64;
  const int nbv = (g.nv + 7) / 8, nbu = (g.nu + 7) / 8, nbw = (g.nw + 7) / 8;
  const TrainStore st{};
  for (long long grp = (long long)blockIdx.x * 4 + wave; grp < ngroups; grp += (long long)gridDim.x * 4) {
    bf16x8 x[2];
    float v[2][kIn];
    long long idx[2];
#pragma unroll
    for (int gg = 0; gg < 2; ++gg) {
      long long q = grp * 64 + 32 * gg + (lane & 31);
      idx[gg] = q;
      if (q >= total) q = total - 1;
      const long long node = q / per;
      long long r = q - node * per;
      const int iw = (int)(r % g.nw);
      r /= g.nw;
      const int iu = (int)(r % g.nu), iv = (int)(r / g.nu);
      const float px = g.n0x + (float)(node % g.nnx) * g.ndx, py = g.n0y + (float)(node / g.nnx) * g.ndy;
      v[gg][0] = px;
      v[gg][1] = py;
      v[gg][2] = px + g.u0 + (float)iu * g.du;
      v[gg][3] = py + g.v0 + (float)iv * g.dv;
      v[gg][4] = (float)iw * g.dw;
      x[gg] = feature_operand(A.N, v[gg]);
    }
    float r[2][kOut];
    forward64<false>(lw, lb, x, r, st);
    if (lane < 32) {
#pragma unroll
      for (int gg = 0; gg < 2; ++gg) {
        const long long q = idx[gg];
        if (q >= total) continue;
        float bt, bs, ba;
        base_leg(A.B, v[gg][0], v[gg][1], v[gg][2], v[gg][3], v[gg][4], &bt, &bs, &ba);
        const float t = bt + r[gg][0] * A.N.osc[0];
        // judged only inside the training box: the source tables query
        // nothing else (their sub-grid stays within the box's margin)
        bool inside = true;
#pragma unroll
        for (int f = 0; f < kIn; ++f) {
          const float u = (v[gg][f] - A.N.lo[f]) * A.N.sc[f] - 1.0f;
          inside = inside && u >= -1.0001f && u <= 1.0001f;
        }
        if (!inside) continue;
        const float e = fabsf(t - tab[q].x);
        const long long node = q / per;
        long long rr = q - node * per;
        const int iw = (int)(rr % g.nw);
        rr /= g.nw;
        const int iu = (int)(rr % g.nu), iv = (int)(rr / g.nu);
        const long long brick = ((node * nbv + iv / 8) * nbu + iu / 8) * nbw + iw / 8;
        atomicMax(err + brick, __float_as_uint(e));
      }
    }
  }
}

// The same judge at off-node sources: for every brick, the network against
// the FP64 ray integrals at the brick's eight corner points from sources
// displaced from the node by (+-0.45 ndx, +-0.45 ndy) and (+-0.45 ndx, 0),
// (0, +-0.45 ndy) (inside the node's cell, where its shots lie); max |dt|
// per brick into err (atomic max on the float bits, as mask_err_kernel).
constexpr int kOffSrc = 8, kCorners = 8;

__global__ __launch_bounds__(kThreads) void mask_offnode_kernel(FwdArgs A, qk::ModelDev M, qpsdm_tgeom g,
                                                                unsigned *err) {
  __shared__ __attribute__((aligned(16))) bf16 lw[kNF];
  __shared__ float lb[kNBias];
  load_weights(lw, lb, A.wf, kNF, A.bias);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nbv = (g.nv + 7) / 8, nbu = (g.nu + 7) / 8, nbw = (g.nw + 7) / 8;
  const long long nbrick = (long long)g.nnx * g.nny * nbv * nbu * nbw;
  const long long total = nbrick * kOffSrc * kCorners;
  const long long ngroups = (total + 63) / 64;
  const float ox[kOffSrc] = {0.45f, -0.45f, 0.45f, -0.45f, 0.45f, -0.45f, 0.0f, 0.0f};
  const float oy[kOffSrc] = {0.45f, 0.45f, -0.45f, -0.45f, 0.0f, 0.0f, 0.45f, -0.45f};
  const TrainStore st{};
  for (long long grp = (long long)blockIdx.x * 4 + wave; grp < ngroups; grp += (long long)gridDim.x * 4) {
    bf16x8 x[2];
    float v[2][kIn];
    long long idx[2], brick[2];
#pragma unroll
    for (int gg = 0; gg < 2; ++gg) {
      long long q = grp * 64 + 32 * gg + (lane & 31);
      idx[gg] = q;
      if (q >= total) q = total - 1;
      const int corner = (int)(q % kCorners);
      const int pos = (int)((q / kCorners) % kOffSrc);
      const long long b = q / (kCorners * kOffSrc);
      brick[gg] = b;
      long long r = b;
      const int bw = (int)(r % nbw);
      r /= nbw;
      const int bu = (int)(r % nbu);
      r /= nbu;
      const int bv = (int)(r % nbv);
      const int node = (int)(r / nbv);
      const int iu = min(bu * 8 + ((corner & 1) ? 7 : 0), g.nu - 1);
      const int iv = min(bv * 8 + ((corner & 2) ? 7 : 0), g.nv - 1);
      const int iw = min(bw * 8 + ((corner & 4) ? 7 : 0), g.nw - 1);
      const float sx = g.n0x + (float)(node % g.nnx) * g.ndx + ox[pos] * g.ndx;
      const float sy = g.n0y + (float)(node / g.nnx) * g.ndy + oy[pos] * g.ndy;
      v[gg][0] = sx;
      v[gg][1] = sy;
      v[gg][2] = sx + g.u0 + (float)iu * g.du;
      v[gg][3] = sy + g.v0 + (float)iv * g.dv;
      v[gg][4] = (float)iw * g.dw;
      x[gg] = feature_operand(A.N, v[gg]);
    }
    float r[2][kOut];
    forward64<false>(lw, lb, x, r, st);
    if (lane < 32) {
#pragma unroll
      for (int gg = 0; gg < 2; ++gg) {
        if (idx[gg] >= total) continue;
        bool inside = true;
#pragma unroll
        for (int f = 0; f < kIn; ++f) {
          const float u = (v[gg][f] - A.N.lo[f]) * A.N.sc[f] - 1.0f;
          inside = inside && u >= -1.0001f && u <= 1.0001f;
        }
        if (!inside) continue;
        float bt, bs, ba;
        base_leg(A.B, v[gg][0], v[gg][1], v[gg][2], v[gg][3], v[gg][4], &bt, &bs, &ba);
        double t, ts;
        qk::ray_integrals(M, v[gg][0], v[gg][1], v[gg][2], v[gg][3], v[gg][4], &t, &ts, nullptr);
        const float e = fabsf(bt + r[gg][0] * A.N.osc[0] - (float)t);
        atomicMax(err + brick[gg], __float_as_uint(e));
      }
    }
  }
}

__global__ void mask_threshold_kernel(const unsigned *err, long long nb, float tol, unsigned char *mask) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (long long)gridDim.x * blockDim.x)
    mask[i] = __uint_as_float(err[i]) <= tol ? 1 : 0;
}

// ---- per-shot source tables ---------------------------------------------------

struct SrcGeo {
  qpsdm_tgeom g;   // node tables
  int nus, nvs;    // per-shot sub-grid size
  long long stride;  // float2 per shot table
};

// Header of each shot: first valid trace's source, sub-grid relative origin.
__global__ void shot_src_kernel(SrcGeo S, const float4 *__restrict__ geom, const qpsdm_shot *__restrict__ shots,
                                int nshot, float4 *hdr) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nshot) return;
  const qpsdm_shot sh = shots[r];
  float sx = NAN, sy = NAN;
  for (int t = sh.tr0; t < sh.tr0 + sh.ntr; ++t) {
    const float4 gg = geom[t];
    if (gg.x == gg.x) {
      sx = gg.x;
      sy = gg.y;
      break;
    }
  }
  float su0 = 0.f, sv0 = 0.f;
  if (sx == sx) {
    const int iu0 = min(max((int)floorf((-sx - S.g.u0) / S.g.du), 0), S.g.nu - S.nus);
    const int iv0 = min(max((int)floorf((-sy - S.g.v0) / S.g.dv), 0), S.g.nv - S.nvs);
    su0 = S.g.u0 + (float)iu0 * S.g.du;
    sv0 = S.g.v0 + (float)iv0 * S.g.dv;
  }
  hdr[r] = make_float4(sx, sy, su0, sv0);
}

// (t, t*) of a source at (sx, sy) at relative node (jv, ju, jw): the node
// tables of the (up to) four surface nodes around the source, each read at
// the same relative node, blended bilinearly in the source position (one
// node per axis where the node grid has one).  Exact for a laterally
// invariant model; elsewhere the lateral variation between nodes is
// interpolated instead of frozen at the nearest node.
__device__ inline float2 node_blend(const qpsdm_tgeom &g, const float2 *__restrict__ tab, float sx, float sy, int jv,
                                    int ju, int jw) {
  const float fx = fminf(fmaxf((sx - g.n0x) / g.ndx, 0.0f), (float)(g.nnx - 1));
  const float fy = fminf(fmaxf((sy - g.n0y) / g.ndy, 0.0f), (float)(g.nny - 1));
  const int ix = min((int)fx, max(g.nnx - 2, 0)), iy = min((int)fy, max(g.nny - 2, 0));
  const float a = g.nnx > 1 ? fx - (float)ix : 0.0f, b = g.nny > 1 ? fy - (float)iy : 0.0f;
  const int ix1 = min(ix + 1, g.nnx - 1), iy1 = min(iy + 1, g.nny - 1);
  const long long per = (long long)g.nv * g.nu * g.nw;
  const long long cell = ((long long)jv * g.nu + ju) * g.nw + jw;
  const float2 v00 = tab[(long long)(iy * g.nnx + ix) * per + cell];
  const float2 v01 = tab[(long long)(iy * g.nnx + ix1) * per + cell];
  const float2 v10 = tab[(long long)(iy1 * g.nnx + ix) * per + cell];
  const float2 v11 = tab[(long long)(iy1 * g.nnx + ix1) * per + cell];
  const float lx = fmaf(a, v01.x - v00.x, v00.x), ly = fmaf(a, v01.y - v00.y, v00.y);
  const float hx = fmaf(a, v11.x - v10.x, v10.x), hy = fmaf(a, v11.y - v10.y, v10.y);
  return make_float2(fmaf(b, hx - lx, lx), fmaf(b, hy - ly, ly));
}

// Table path of the per-shot source tables: every value from node_blend.
__global__ void src_nodes_kernel(SrcGeo S, const float4 *__restrict__ hdr, int nshot, const float2 *__restrict__ tab,
                                 float2 *out) {
  const qpsdm_tgeom &g = S.g;
  const long long total = S.stride * nshot;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const int shot = (int)(q / S.stride);
    long long r = q - (long long)shot * S.stride;
    const int jw = (int)(r % g.nw);
    r /= g.nw;
    const int iu = (int)(r % S.nus), iv = (int)(r / S.nus);
    const float4 hd = hdr[shot];
    if (!(hd.x == hd.x)) {  // a shot with no valid trace: never read
      out[q] = make_float2(0.f, 0.f);
      continue;
    }
    const int ju = min(max((int)rintf((hd.z - g.u0) / g.du) + iu, 0), g.nu - 1);
    const int jv = min(max((int)rintf((hd.w - g.v0) / g.dv) + iv, 0), g.nv - 1);
    out[q] = node_blend(g, tab, hd.x, hd.y, jv, ju, jw);
  }
}

// Surrogate (t, t*) on each shot's sub-grid [nvs][nus][nw]; where the
// nearest node's brick failed the validity check, the table builder's value
// at the shot's own source (the FP64 straight-ray integrals the node tables
// and the training samples come from): a node table read off-node is off by
// the lateral variation between nodes, far more than the network.
__global__ __launch_bounds__(kThreads) void src_table_kernel(FwdArgs A, qk::ModelDev M, SrcGeo S,
                                                             const float4 *__restrict__ hdr, int nshot,
                                                             const unsigned char *__restrict__ mask,
                                                             float2 *out) {
  __shared__ __attribute__((aligned(16))) bf16 lw[kNF];
  __shared__ float lb[kNBias];
  load_weights(lw, lb, A.wf, kNF, A.bias);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const qpsdm_tgeom &g = S.g;
  const long long total = S.stride * nshot;
  const long long ngroups = (total + 63) / 64;
  const int nbv = (g.nv + 7) / 8, nbu = (g.nu + 7) / 8, nbw = (g.nw + 7) / 8;
  const TrainStore st{};
  for (long long grp = (long long)blockIdx.x * 4 + wave; grp < ngroups; grp += (long long)gridDim.x * 4) {
    bf16x8 x[2];
    float v[2][kIn];
    long long idx[2];
    int node[2], ju[2], jv[2], jw[2];
#pragma unroll
    for (int gg = 0; gg < 2; ++gg) {
      long long q = grp * 64 + 32 * gg + (lane & 31);
      idx[gg] = q;
      if (q >= total) q = total - 1;
      const int shot = (int)(q / S.stride);
      long long r = q - (long long)shot * S.stride;
      jw[gg] = (int)(r % g.nw);
      r /= g.nw;
      const int iu = (int)(r % S.nus), iv = (int)(r / S.nus);
      const float4 hd = hdr[shot];
      const bool live = hd.x == hd.x;
      const float sx = live ? hd.x : 0.f, sy = live ? hd.y : 0.f;
      const int iu0 = live ? (int)rintf((hd.z - g.u0) / g.du) : 0;
      const int iv0 = live ? (int)rintf((hd.w - g.v0) / g.dv) : 0;
      ju[gg] = min(max(iu0 + iu, 0), g.nu - 1);
      jv[gg] = min(max(iv0 + iv, 0), g.nv - 1);
      // nearest node as the imaging kernel rounds it
      node[gg] = min(max((int)floorf((sx - g.n0x) * (1.0f / g.ndx) + 0.5f), 0), g.nnx - 1) +
                 g.nnx * min(max((int)floorf((sy - g.n0y) * (1.0f / g.ndy) + 0.5f), 0), g.nny - 1);
      v[gg][0] = sx;
      v[gg][1] = sy;
      v[gg][2] = sx + hd.z + (float)iu * g.du;
      v[gg][3] = sy + hd.w + (float)iv * g.dv;
      v[gg][4] = (float)jw[gg] * g.dw;
      x[gg] = feature_operand(A.N, v[gg]);
    }
    float r[2][kOut];
    forward64<false>(lw, lb, x, r, st);
    if (lane < 32) {
#pragma unroll
      for (int gg = 0; gg < 2; ++gg) {
        const long long q = idx[gg];
        if (q >= total) continue;
        float2 val;
        const long long brick = (((long long)node[gg] * nbv + jv[gg] / 8) * nbu + ju[gg] / 8) * nbw + jw[gg] / 8;
        if (mask && !mask[brick]) {
          double t, ts;
          qk::ray_integrals(M, v[gg][0], v[gg][1], v[gg][2], v[gg][3], v[gg][4], &t, &ts, nullptr);
          val = make_float2((float)t, (float)ts);
        } else {
          float bt, bs, ba;
          base_leg(A.B, v[gg][0], v[gg][1], v[gg][2], v[gg][3], v[gg][4], &bt, &bs, &ba);
          val = make_float2(bt + r[gg][0] * A.N.osc[0], bs + r[gg][1] * A.N.osc[1]);
        }
        out[q] = val;
      }
    }
  }
}

int grid_for(long long ngroups) {
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const long long need = (ngroups + 3) / 4;
  return (int)std::max(1LL, std::min(need, (long long)ncu * 2));
}

}  // namespace

// ---- host object ----------------------------------------------------------------

struct qpsdm_surrogate {
  qk::ModelDev model;
  Norm N;
  Base B;
  float *w = nullptr, *m1 = nullptr, *m2 = nullptr;  // FP32 master + Adam moments [kNP]
  bf16 *wf = nullptr, *wb = nullptr;                 // packed operands
  float *bias = nullptr;                             // [kNBias]
  int step = 0;
  // training scratch, sized for cap samples
  int cap = 0;
  bf16 *Xb = nullptr, *H = nullptr, *dZ = nullptr, *G = nullptr;
  float *Z = nullptr, *part = nullptr, *bpart = nullptr, *loss = nullptr;
};

namespace {

FwdArgs fwd_args(const qpsdm_surrogate *s) { return FwdArgs{s->wf, s->bias, s->N, s->B}; }

void free_scratch(qpsdm_surrogate *s) {
  (void)hipFree(s->Xb);
  (void)hipFree(s->H);
  (void)hipFree(s->dZ);
  (void)hipFree(s->G);
  (void)hipFree(s->Z);
  (void)hipFree(s->part);
  (void)hipFree(s->bpart);
  (void)hipFree(s->loss);
  s->Xb = s->H = s->dZ = s->G = nullptr;
  s->Z = s->part = s->bpart = s->loss = nullptr;
  s->cap = 0;
}

int ensure_scratch(qpsdm_surrogate *s, int n) {
  if (n <= s->cap) return QPSDM_OK;
  free_scratch(s);
  const size_t L = (size_t)kH * n;
  const int nchunk = (n + kGradChunk - 1) / kGradChunk;
  QPSDM_HIP(hipMalloc(&s->Xb, sizeof(bf16) * kIn * n));
  QPSDM_HIP(hipMalloc(&s->H, sizeof(bf16) * 4 * L));
  QPSDM_HIP(hipMalloc(&s->dZ, sizeof(bf16) * 4 * L));
  QPSDM_HIP(hipMalloc(&s->G, sizeof(bf16) * kOut * n));
  QPSDM_HIP(hipMalloc(&s->Z, sizeof(float) * 4 * L));
  QPSDM_HIP(hipMalloc(&s->part, sizeof(float) * (size_t)nchunk * kNBlk * 1024));
  QPSDM_HIP(hipMalloc(&s->bpart, sizeof(float) * (size_t)nchunk * 5 * kH));
  QPSDM_HIP(hipMalloc(&s->loss, sizeof(float)));
  s->cap = n;
  return QPSDM_OK;
}

int repack(qpsdm_surrogate *s, hipStream_t st) {
  AdamArgs a{};
  hipLaunchKernelGGL(adam_kernel, dim3((kNP + 255) / 256), dim3(256), 0, st, a, s->w, s->m1, s->m2,
                     (const float *)nullptr, (const float *)nullptr, (float *)nullptr, s->wf, s->wb, s->bias);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

// forward + backward + gradient reduction on n samples; update != 0 also
// applies one Adam step with learning rate lr.
int step_once(qpsdm_surrogate *s, const float *X, const float *T, int n, float *grad_out, int update, float lr,
              hipStream_t st) {
  int rc = ensure_scratch(s, n);
  if (rc) return rc;
  const int ngroups = n / 64;
  const int nchunk = (n + kGradChunk - 1) / kGradChunk;
  hipLaunchKernelGGL(train_fwd_kernel, dim3(grid_for(ngroups)), dim3(kThreads), 0, st, fwd_args(s), X, T, n, s->Xb,
                     s->Z, s->H, s->G, s->loss);
  hipLaunchKernelGGL(train_bwd_kernel, dim3(grid_for(ngroups)), dim3(kThreads), 0, st, (const bf16 *)s->wb, n,
                     (const bf16 *)s->G, s->Z, s->dZ);
  hipLaunchKernelGGL(wgrad_kernel, dim3(kNBlk, nchunk), dim3(64), 0, st, n, (const bf16 *)s->Xb,
                     (const bf16 *)s->H, (const bf16 *)s->dZ, (const bf16 *)s->G, s->part, s->bpart);
  AdamArgs a;
  a.lr = lr;
  a.b1 = 0.9f;
  a.b2 = 0.999f;
  a.eps = 1e-8f;
  a.update = update;
  a.nchunk = nchunk;
  if (update) s->step += 1;
  const int t = std::max(s->step, 1);
  a.c1 = (float)(1.0 - pow(0.9, t));
  a.c2 = (float)(1.0 - pow(0.999, t));
  hipLaunchKernelGGL(adam_kernel, dim3((kNP + 255) / 256), dim3(256), 0, st, a, s->w, s->m1, s->m2,
                     (const float *)s->part, (const float *)s->bpart, grad_out, s->wf, s->wb, s->bias);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

uint64_t host_mix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

int src_geo(const qpsdm_image_params *ip, SrcGeo *S) {
  QPSDM_REQUIRE(ip, "surrogate: null image params");
  const qpsdm_tgeom &g = ip->tg;
  QPSDM_REQUIRE(g.nu >= 2 && g.nv >= 2 && g.nw >= 2 && g.du > 0 && g.dv > 0 && g.dw > 0, "surrogate: bad tables");
  S->g = g;
  qk::src_subgrid(ip, &S->nus, &S->nvs);
  S->stride = (long long)S->nvs * S->nus * g.nw;
  return QPSDM_OK;
}

}  // namespace

extern "C" int qpsdm_surrogate_nparams(void) { return kNP; }

extern "C" int qpsdm_surrogate_create(const qpsdm_model *m, const float box_lo[5], const float box_hi[5],
                                      const float out_scale[3], unsigned long long seed, qpsdm_surrogate **out) {
  QPSDM_REQUIRE(out && box_lo && box_hi && out_scale, "surrogate_create: null args");
  QPSDM_REQUIRE(qk::model_ok(m), "surrogate_create: bad model");
  qpsdm_surrogate *s = new qpsdm_surrogate();
  s->model = qk::make_model(*m);
  for (int f = 0; f < kIn; ++f) {
    if (!(box_hi[f] > box_lo[f])) {
      delete s;
      return qpsdm::set_error(QPSDM_EINVAL, "surrogate_create: empty input box on feature %d", f);
    }
    s->N.lo[f] = box_lo[f];
    s->N.sc[f] = 2.0f / (box_hi[f] - box_lo[f]);
  }
  for (int o = 0; o < kOut; ++o) {
    if (!(out_scale[o] > 0)) {
      delete s;
      return qpsdm::set_error(QPSDM_EINVAL, "surrogate_create: output scale %d must be > 0", o);
    }
    s->N.osc[o] = out_scale[o];
    s->N.lw[o] = 1.0f;
  }
  s->B.kind = m->kind;
  s->B.v0 = (float)m->v0;
  s->B.kgrad = (float)m->kgrad;
  s->B.q0 = (float)m->q0;
  s->B.nlayer = m->nlayer;
  s->B.layer_dz = (float)m->layer_dz;
  for (int i = 0; i < 16; ++i) s->B.vl[i] = (float)m->vl[i], s->B.ql[i] = (float)m->ql[i];
  // Xavier-uniform weights (seeded), zero biases, the output layer scaled down
  std::vector<float> w(kNP, 0.0f);
  uint64_t st = seed;
  auto uni = [&]() { return (double)(host_mix(st++) >> 11) * (1.0 / 9007199254740992.0); };
  auto fill = [&](int off, int fo, int fi, double gain) {
    const double a = gain * sqrt(6.0 / (fi + fo));
    for (int k = 0; k < fo * fi; ++k) w[off + k] = (float)((2.0 * uni() - 1.0) * a);
  };
  fill(oW1, kH, kIn, 1.0);
  fill(oW2, kH, kH, 1.0);
  fill(oW3, kH, kH, 1.0);
  fill(oW4, kH, kH, 1.0);
  fill(oW5, kOut, kH, 0.1);
  auto fail = [&](int code) {
    (void)hipFree(s->w);
    (void)hipFree(s->m1);
    (void)hipFree(s->m2);
    (void)hipFree(s->wf);
    (void)hipFree(s->wb);
    (void)hipFree(s->bias);
    delete s;
    return code;
  };
  if (hipMalloc(&s->w, sizeof(float) * kNP) || hipMalloc(&s->m1, sizeof(float) * kNP) ||
      hipMalloc(&s->m2, sizeof(float) * kNP) || hipMalloc(&s->wf, sizeof(bf16) * kNF) ||
      hipMalloc(&s->wb, sizeof(bf16) * kNB) || hipMalloc(&s->bias, sizeof(float) * kNBias))
    return fail(qpsdm::set_error(QPSDM_EHIP, "surrogate_create: device allocation failed"));
  if (hipMemcpy(s->w, w.data(), sizeof(float) * kNP, hipMemcpyHostToDevice) ||
      hipMemset(s->m1, 0, sizeof(float) * kNP) || hipMemset(s->m2, 0, sizeof(float) * kNP) ||
      hipMemset(s->wf, 0, sizeof(bf16) * kNF) || hipMemset(s->wb, 0, sizeof(bf16) * kNB) ||
      hipMemset(s->bias, 0, sizeof(float) * kNBias))
    return fail(qpsdm::set_error(QPSDM_EHIP, "surrogate_create: device initialisation failed"));
  if (repack(s, nullptr) || hipDeviceSynchronize() != hipSuccess)
    return fail(qpsdm::set_error(QPSDM_EHIP, "surrogate_create: packing failed"));
  *out = s;
  return QPSDM_OK;
}

extern "C" int qpsdm_surrogate_destroy(qpsdm_surrogate *s) {
  if (!s) return QPSDM_OK;
  free_scratch(s);
  (void)hipFree(s->w);
  (void)hipFree(s->m1);
  (void)hipFree(s->m2);
  (void)hipFree(s->wf);
  (void)hipFree(s->wb);
  (void)hipFree(s->bias);
  delete s;
  return QPSDM_OK;
}

extern "C" int qpsdm_surrogate_get_params(const qpsdm_surrogate *s, float *host) {
  QPSDM_REQUIRE(s && host, "surrogate_get_params: null args");
  QPSDM_HIP(hipMemcpy(host, s->w, sizeof(float) * kNP, hipMemcpyDeviceToHost));
  return QPSDM_OK;
}

extern "C" int qpsdm_surrogate_set_params(qpsdm_surrogate *s, const float *host) {
  QPSDM_REQUIRE(s && host, "surrogate_set_params: null args");
  QPSDM_HIP(hipMemcpy(s->w, host, sizeof(float) * kNP, hipMemcpyHostToDevice));
  int rc = repack(s, nullptr);
  if (rc) return rc;
  QPSDM_HIP(hipDeviceSynchronize());
  return QPSDM_OK;
}

extern "C" int qpsdm_surrogate_samples(const qpsdm_surrogate *s, unsigned long long seed, long long offset, int n,
                                       float *X, float *T, void *stream) {
  QPSDM_REQUIRE(s && X && T && n >= 0, "surrogate_samples: bad args");
  if (n == 0) return QPSDM_OK;
  hipLaunchKernelGGL(samples_kernel, dim3(std::min((n + 255) / 256, 8192)), dim3(256), 0, (hipStream_t)stream,
                     s->model, s->N, (uint64_t)seed, offset, n, X, T);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

extern "C" int qpsdm_surrogate_eval(const qpsdm_surrogate *s, const float *X, int n, float *out, void *stream) {
  QPSDM_REQUIRE(s && X && out && n >= 0, "surrogate_eval: bad args");
  if (n == 0) return QPSDM_OK;
  const int ngroups = (n + 63) / 64;
  hipLaunchKernelGGL(eval_kernel, dim3(grid_for(ngroups)), dim3(kThreads), 0, (hipStream_t)stream, fwd_args(s), X,
                     n, out);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

extern "C" int qpsdm_surrogate_grad(qpsdm_surrogate *s, const float *X, const float *T, int n, float *grad,
                                    float *loss_host, void *stream) {
  QPSDM_REQUIRE(s && X && T && grad, "surrogate_grad: null args");
  QPSDM_REQUIRE(n > 0 && n % 64 == 0 && n <= (1 << 20), "surrogate_grad: n must be a multiple of 64 in [64, 2^20]");
  const hipStream_t st = (hipStream_t)stream;
  int rc = ensure_scratch(s, n);
  if (rc) return rc;
  QPSDM_HIP(hipMemsetAsync(s->loss, 0, sizeof(float), st));
  rc = step_once(s, X, T, n, grad, 0, 0.f, st);
  if (rc) return rc;
  if (loss_host) {
    QPSDM_HIP(hipMemcpyAsync(loss_host, s->loss, sizeof(float), hipMemcpyDeviceToHost, st));
    QPSDM_HIP(hipStreamSynchronize(st));
  }
  return QPSDM_OK;
}

extern "C" int qpsdm_surrogate_train(qpsdm_surrogate *s, const float *X, const float *T, long long nsamples,
                                     int steps, int batch, float lr0, float lr1, float *loss_host, void *stream) {
  QPSDM_REQUIRE(s && X && T, "surrogate_train: null args");
  QPSDM_REQUIRE(batch > 0 && batch % 64 == 0 && batch <= (1 << 20) && nsamples >= batch,
                "surrogate_train: batch must be a multiple of 64 in [64, 2^20] and at most the sample count");
  QPSDM_REQUIRE(steps >= 0 && lr0 > 0 && lr1 > 0, "surrogate_train: bad schedule");
  const hipStream_t st = (hipStream_t)stream;
  int rc = ensure_scratch(s, batch);
  if (rc) return rc;
  const long long nb = nsamples / batch;
  float *hist = nullptr;
  if (loss_host && steps > 0) QPSDM_HIP(hipMalloc(&hist, sizeof(float) * steps));
  for (int k = 0; k < steps; ++k) {
    // cosine decay from lr0 to lr1
    const float lr = lr1 + 0.5f * (lr0 - lr1) * (1.0f + cosf((float)M_PI * k / std::max(steps - 1, 1)));
    const long long b = k % nb;
    float *lacc = hist ? hist + k : s->loss;
    QPSDM_HIP(hipMemsetAsync(lacc, 0, sizeof(float), st));
    float *keep = s->loss;
    s->loss = lacc;
    rc = step_once(s, X + b * batch * kIn, T + b * batch * kOut, batch, nullptr, 1, lr, st);
    s->loss = keep;
    if (rc) {
      (void)hipFree(hist);
      return rc;
    }
  }
  if (hist) {
    QPSDM_HIP(hipMemcpyAsync(loss_host, hist, sizeof(float) * steps, hipMemcpyDeviceToHost, st));
    QPSDM_HIP(hipStreamSynchronize(st));
    (void)hipFree(hist);
  }
  return QPSDM_OK;
}

extern "C" long long qpsdm_surrogate_mask_bricks(const qpsdm_tgeom *g) {
  if (!g || g->nnx < 1 || g->nny < 1 || g->nu < 2 || g->nv < 2 || g->nw < 2) return 0;
  return (long long)g->nnx * g->nny * ((g->nv + 7) / 8) * ((g->nu + 7) / 8) * ((g->nw + 7) / 8);
}

extern "C" int qpsdm_surrogate_mask(const qpsdm_surrogate *s, const qpsdm_tgeom *g, const float *tab, float tol,
                                    unsigned char *mask, float *err_dev, void *stream) {
  QPSDM_REQUIRE(s && g && tab && mask && err_dev, "surrogate_mask: null args");
  const long long nb = qpsdm_surrogate_mask_bricks(g);
  QPSDM_REQUIRE(nb > 0, "surrogate_mask: bad table geometry");
  const hipStream_t st = (hipStream_t)stream;
  QPSDM_HIP(hipMemsetAsync(err_dev, 0, sizeof(float) * nb, st));
  const long long total = (long long)g->nv * g->nu * g->nw * g->nnx * g->nny;
  hipLaunchKernelGGL(mask_err_kernel, dim3(grid_for((total + 63) / 64)), dim3(kThreads), 0, st, fwd_args(s), *g,
                     (const float2 *)tab, (unsigned *)err_dev);
  const long long off = nb * kOffSrc * kCorners;
  hipLaunchKernelGGL(mask_offnode_kernel, dim3(grid_for((off + 63) / 64)), dim3(kThreads), 0, st, fwd_args(s),
                     s->model, *g, (unsigned *)err_dev);
  hipLaunchKernelGGL(mask_threshold_kernel, dim3(1024), dim3(256), 0, st, (const unsigned *)err_dev, nb, tol, mask);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

extern "C" size_t qpsdm_surrogate_src_bytes(const qpsdm_image_params *ip, int nshot) {
  SrcGeo S;
  if (nshot < 0 || src_geo(ip, &S)) return 0;
  return (size_t)nshot * (16 + (size_t)S.stride * sizeof(float2));
}

extern "C" int qpsdm_surrogate_src_tables(const qpsdm_surrogate *s, const qpsdm_image_params *ip, const float *geom,
                                          const qpsdm_shot *shots, int nshot, const float *tab,
                                          const unsigned char *mask, void *src, void *stream) {
  QPSDM_REQUIRE(s && geom && shots && src && nshot >= 0, "surrogate_src_tables: null args");
  (void)tab;  // the fallback is computed at the source; the node tables serve qpsdm_surrogate_mask
  SrcGeo S;
  int rc = src_geo(ip, &S);
  if (rc) return rc;
  if (nshot == 0) return QPSDM_OK;
  const hipStream_t st = (hipStream_t)stream;
  float4 *hdr = (float4 *)src;
  float2 *tabs = (float2 *)((char *)src + (size_t)nshot * 16);
  hipLaunchKernelGGL(shot_src_kernel, dim3((nshot + 63) / 64), dim3(64), 0, st, S, (const float4 *)geom, shots,
                     nshot, hdr);
  const long long total = S.stride * nshot;
  hipLaunchKernelGGL(src_table_kernel, dim3(grid_for((total + 63) / 64)), dim3(kThreads), 0, st, fwd_args(s),
                     s->model, S, (const float4 *)hdr, nshot, mask, tabs);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

extern "C" int qpsdm_src_tables_nodes(const qpsdm_image_params *ip, const float *geom, const qpsdm_shot *shots,
                                      int nshot, const float *tab, void *src, void *stream) {
  QPSDM_REQUIRE(geom && shots && tab && src && nshot >= 0, "src_tables_nodes: null args");
  SrcGeo S;
  int rc = src_geo(ip, &S);
  if (rc) return rc;
  if (nshot == 0) return QPSDM_OK;
  const hipStream_t st = (hipStream_t)stream;
  float4 *hdr = (float4 *)src;
  float2 *tabs = (float2 *)((char *)src + (size_t)nshot * 16);
  hipLaunchKernelGGL(shot_src_kernel, dim3((nshot + 63) / 64), dim3(64), 0, st, S, (const float4 *)geom, shots,
                     nshot, hdr);
  const long long total = S.stride * nshot;
  const int grid = (int)std::min<long long>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(src_nodes_kernel, dim3(grid), dim3(256), 0, st, S, (const float4 *)hdr, nshot,
                     (const float2 *)tab, tabs);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}
