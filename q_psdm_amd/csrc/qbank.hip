// Kernel (a): batched per-trace Q-compensation filter bank for gfx950.
//
// One work-group per trace.  In LDS:
//   1. t^2 gain, full-trace raised-cosine taper and the two tapered windows
//      of the reference's pre-processing (src/q-psdm.cu:2173-2216,
//      src/psdmpkg.cc:208-231), packed as one complex sequence
//      z = seg1 + i seg2;
//   2. forward mixed-radix Stockham FFT of length nsb (radices 2,3,4,5,7,
//      11,13 cover every reference length 2*divisor(720720)), split into
//      the two segment spectra and multiplied by -i*omega on bins nf1..nf4
//      (the reference's re = w Im, im = -w Re);
//   3. for each t* level j: filter H_j (FP64-generated on the host: high-cut
//      taper, gain-limited exp(w t*/2), Kolsky dispersion phase), pack both
//      segments as one Hermitian pair, inverse FFT, and write the level's
//      time trace, segment 1 up to tdmid and segment 2 after it.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <math.h>

#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "common.h"
#include "pair_bank.h"

namespace {

constexpr int kThreads = 256;

// Passes of each radix, 4 bits per radix in the order 4, 2, 3, 5, 7, 11,
// 13 (the order they run in).
struct Plan {
  unsigned long long code;
};

struct BankDev {
  int nt;
  float dt;
  int nsb, itb1, itb2, nf1, nww;
  float tdmid;
  int K;
  long long row;        // output elements between consecutive traces
  int lead;             // output elements of the row header before level 0
  Plan plan;
  const float2 *tw;     // [nsb] exp(-2 pi i a / nsb)
  const float2 *H;      // [K][nww]
  const float *wtab;    // [nww] omega of each kept bin
  const float *taper;   // [200] raised-cosine coefficients
};

__device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// cos and sin of 2 pi m / R, m = 0 .. R-1 (FP64 values rounded to FP32), for
// the odd-radix butterflies: compile-time constants, so the butterflies
// take them as instruction literals instead of registers or LDS reads.
template <int R>
struct DftConst;
template <>
struct DftConst<3> {
  static constexpr float c[3] = {1.0f, -0.5f, -0.5f};
  static constexpr float s[3] = {0.0f, 0.866025388f, -0.866025388f};
};
template <>
struct DftConst<5> {
  static constexpr float c[5] = {1.0f, 0.309017003f, -0.809017003f, -0.809017003f, 0.309017003f};
  static constexpr float s[5] = {0.0f, 0.95105654f, 0.587785244f, -0.587785244f, -0.95105654f};
};
template <>
struct DftConst<7> {
  static constexpr float c[7] = {1.0f, 0.623489797f, -0.222520933f, -0.90096885f, -0.90096885f, -0.222520933f, 0.623489797f};
  static constexpr float s[7] = {0.0f, 0.781831503f, 0.974927902f, 0.433883727f, -0.433883727f, -0.974927902f, -0.781831503f};
};
template <>
struct DftConst<11> {
  static constexpr float c[11] = {1.0f, 0.841253519f, 0.415415019f, -0.142314836f, -0.654860735f, -0.959492981f, -0.959492981f, -0.654860735f, -0.142314836f, 0.415415019f, 0.841253519f};
  static constexpr float s[11] = {0.0f, 0.540640831f, 0.909631968f, 0.989821434f, 0.755749583f, 0.281732559f, -0.281732559f, -0.755749583f, -0.989821434f, -0.909631968f, -0.540640831f};
};
template <>
struct DftConst<13> {
  static constexpr float c[13] = {1.0f, 0.885456026f, 0.568064749f, 0.120536678f, -0.3546049f, -0.748510778f, -0.970941842f, -0.970941842f, -0.748510778f, -0.3546049f, 0.120536678f, 0.568064749f, 0.885456026f};
  static constexpr float s[13] = {0.0f, 0.46472317f, 0.822983861f, 0.992708862f, 0.935016215f, 0.663122654f, 0.239315659f, -0.239315659f, -0.663122654f, -0.935016215f, -0.992708862f, -0.822983861f, -0.46472317f};
};

// One Stockham pass of radix R (SIGN -1 forward, +1 inverse).  The
// butterfly index b splits into (b / Ns, b % Ns) by a float reciprocal:
// floor((b + 1/2) / Ns) is exact for b < 2^21 (the half keeps the product
// away from integers by more than its rounding error).  Radices 2 and 4 are
// written out; odd radices use the symmetric form of the DFT: with
// a_q = v_q + v_{R-q}, d_q = v_q - v_{R-q} and W^{pq} = c + i s, outputs p
// and R - p are v_0 + sum_q a_q c  +/-  i sum_q d_q s (q = 1 .. (R-1)/2),
// about a third of the naive product count and (R-1)/2 complex pairs live
// instead of R inputs and R twiddles.
template <int R, int SIGN>
__device__ void stockham_pass(const float2 *__restrict__ in, float2 *__restrict__ out,
                              const float2 *__restrict__ tw, int N, int Ns) {
  const int nb = N / R;
  const int step = N / (Ns * R);
  const float inv_ns = 1.0f / (float)Ns;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    const int hi = (int)(((float)b + 0.5f) * inv_ns);
    const int k = b - hi * Ns;
    float2 v[R];
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = in[b + q * nb];
#pragma unroll
    for (int q = 1; q < R; ++q) {
      float2 w = tw[q * k * step];
      v[q] = cmul(v[q], SIGN > 0 ? cconj(w) : w);
    }
    float2 *o = out + hi * Ns * R + k;
    if constexpr (R == 2) {
      o[0] = make_float2(v[0].x + v[1].x, v[0].y + v[1].y);
      o[Ns] = make_float2(v[0].x - v[1].x, v[0].y - v[1].y);
    } else if constexpr (R == 4) {
      const float2 s02 = make_float2(v[0].x + v[2].x, v[0].y + v[2].y);
      const float2 d02 = make_float2(v[0].x - v[2].x, v[0].y - v[2].y);
      const float2 s13 = make_float2(v[1].x + v[3].x, v[1].y + v[3].y);
      const float2 d13 = make_float2(v[1].x - v[3].x, v[1].y - v[3].y);
      // W d13 with W = -i (forward) or +i (inverse)
      const float2 wd = SIGN > 0 ? make_float2(-d13.y, d13.x) : make_float2(d13.y, -d13.x);
      o[0] = make_float2(s02.x + s13.x, s02.y + s13.y);
      o[Ns] = make_float2(d02.x + wd.x, d02.y + wd.y);
      o[2 * Ns] = make_float2(s02.x - s13.x, s02.y - s13.y);
      o[3 * Ns] = make_float2(d02.x - wd.x, d02.y - wd.y);
    } else {
      constexpr int H = (R - 1) / 2;
      float2 a[H], d[H];
#pragma unroll
      for (int q = 1; q <= H; ++q) {
        a[q - 1] = make_float2(v[q].x + v[R - q].x, v[q].y + v[R - q].y);
        d[q - 1] = make_float2(v[q].x - v[R - q].x, v[q].y - v[R - q].y);
      }
      float2 y0 = v[0];
#pragma unroll
      for (int q = 0; q < H; ++q) y0 = make_float2(y0.x + a[q].x, y0.y + a[q].y);
      o[0] = y0;
#pragma unroll
      for (int p = 1; p <= H; ++p) {
        float2 A = v[0], B = make_float2(0.f, 0.f);
#pragma unroll
        for (int q = 1; q <= H; ++q) {
          constexpr float sg = SIGN > 0 ? 1.0f : -1.0f;
          const float c = DftConst<R>::c[(p * q) % R], s = sg * DftConst<R>::s[(p * q) % R];
          A.x = fmaf(a[q - 1].x, c, A.x);
          A.y = fmaf(a[q - 1].y, c, A.y);
          B.x = fmaf(d[q - 1].x, s, B.x);
          B.y = fmaf(d[q - 1].y, s, B.y);
        }
        // outputs A + i B and A - i B, i B = (-B.y, B.x)
        o[p * Ns] = make_float2(A.x - B.y, A.y + B.x);
        o[(R - p) * Ns] = make_float2(A.x + B.y, A.y - B.x);
      }
    }
  }
}

template <int R, int SIGN>
__device__ inline void passes(float2 *&a, float2 *&b, const float2 *tw, int N, int &Ns, int count) {
  for (int c = 0; c < count; ++c) {
    stockham_pass<R, SIGN>(a, b, tw, N, Ns);
    Ns *= R;
    __syncthreads();
    float2 *t = a;
    a = b;
    b = t;
  }
}

// Runs the plan: the passes of each radix in a loop of their own, in the
// order 4, 2, 3, 5, 7, 11, 13 (a switch over the radix inside one loop made
// the compiler hold every butterfly's values at once: 248 registers and
// stack spills).  Data starts in a; the result pointer is returned.
template <int SIGN>
__device__ float2 *fft_run(float2 *a, float2 *b, const float2 *tw, int N, unsigned long long code) {
  int Ns = 1;
  passes<4, SIGN>(a, b, tw, N, Ns, (int)(code & 15));
  passes<2, SIGN>(a, b, tw, N, Ns, (int)((code >> 4) & 15));
  passes<3, SIGN>(a, b, tw, N, Ns, (int)((code >> 8) & 15));
  passes<5, SIGN>(a, b, tw, N, Ns, (int)((code >> 12) & 15));
  passes<7, SIGN>(a, b, tw, N, Ns, (int)((code >> 16) & 15));
  passes<11, SIGN>(a, b, tw, N, Ns, (int)((code >> 20) & 15));
  passes<13, SIGN>(a, b, tw, N, Ns, (int)((code >> 24) & 15));
  return a;
}

// fetch_and_taper's factor for position i of an n-sample window: the
// reference multiplies by c[i] at step i and by c[n-1-i] at step n-1-i
// (src/psdmpkg.cc:224-229); the earlier step is applied first.
__device__ inline float taper_apply(float x, int i, int n, const float *c) {
  const int j = n - 1 - i;
  const bool a = i < 200, b = j < 200;
  if (i <= j) {
    if (a) x *= c[i];
    if (b) x *= c[j];
  } else {
    if (b) x *= c[j];
    if (a) x *= c[i];
  }
  return x;
}

__device__ inline float gain_t2(float x, int it, float dt) {
  const float t = it * dt;
  if (t < 0.1) return (float)((double)x * (0.1 * 0.1));
  return x * (t * t);
}

__device__ inline void store_out(float *p, float v) { *p = v; }
__device__ inline void store_out(__hip_bfloat16 *p, float v) { *p = __float2bfloat16(v); }
__device__ inline void store_out(__half *p, float v) { *p = __float2half(v); }

using qpsdm_dev::PairF16;
using qpsdm_dev::PairF32;

// Pair banks: each entry holds the sample and its forward difference.
template <typename OutT>
constexpr bool pair_out() {
  return std::is_same<OutT, PairF32>::value || std::is_same<OutT, PairF16>::value;
}
// Banks stored under a per-trace power-of-two scale with a row header.
template <typename OutT>
constexpr bool scaled_out() {
  return std::is_same<OutT, __half>::value || std::is_same<OutT, PairF16>::value;
}
__device__ inline void store_pair(PairF32 *p, float b0, float b1) { *p = PairF32{b0, b1 - b0}; }
__device__ inline void store_pair(PairF16 *p, float b0, float b1) {
  const __half2 h = __halves2half2(__float2half(b0), __float2half(b1 - b0));
  *p = PairF16{__builtin_bit_cast(unsigned, h)};
}
__device__ inline void store_zero(__half *p) { *p = __float2half(0.0f); }
__device__ inline void store_zero(PairF16 *p) { *p = PairF16{0u}; }

// Largest value over the work-group (every thread gets it).
__device__ inline float block_max(float m) {
  __shared__ float red[kThreads / 64];
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  m = red[0];
  for (int w = 1; w < kThreads / 64; ++w) m = fmaxf(m, red[w]);
  return m;
}

// MODE 0: spectra only (re/im of the reference pre-processing).
// MODE 1: full bank into OutT.  FP16 output is scaled per trace by
// 2^(15 - ex), ex the binary exponent (frexp) of a bound on the trace's
// largest |sample| over all levels: a level's samples come from an
// unnormalised inverse FFT of a spectrum whose nonzero bins are
// s1 + i s2 = (C1 + i C2) H at the kept bins and their mirror images, so
// |sample| <= sum over bins of |bin| <= 2 sum_k |H_jk| (|C1_k| + |C2_k|).
// Every scaled sample is then below 2^15, the power of two is exact both
// ways, and FP16's relative precision does not depend on how far below the
// bound the samples lie (they stay far above its subnormal range).  The
// row's 16-byte header holds {1/scale, scale, 0, 0}.
// Held to three waves per SIMD: the compiler's own schedule of the odd
// butterflies wants ~184 registers (two waves); at three it spills a few
// values outside the butterflies and runs 15 % faster, at four it spills
// inside them and runs slower.
template <int MODE, typename OutT>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(3))) void qbank_kernel(BankDev P, const float *__restrict__ traces,
                                                         OutT *__restrict__ bank,
                                                         float *__restrict__ re_out,
                                                         float *__restrict__ im_out) {
  extern __shared__ float2 lds[];
  const int N = P.nsb;
  float2 *A = lds;
  float2 *B = lds + N;
  float2 *tw = lds + 2 * N;
  float2 *C1 = tw + N;
  float2 *C2 = C1 + P.nww;
  __shared__ float tap[200];
  const int tr = blockIdx.x;
  const float *x = traces + (size_t)tr * P.nt;

  for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = P.tw[i];
  for (int i = threadIdx.x; i < 200; i += blockDim.x) tap[i] = P.taper[i];
  float *d = reinterpret_cast<float *>(B);  // gained + tapered trace, aliases B
  __syncthreads();
  for (int it = threadIdx.x; it < P.nt; it += blockDim.x)
    d[it] = taper_apply(gain_t2(x[it], it, P.dt), it, P.nt, tap);
  __syncthreads();
  for (int n = threadIdx.x; n < N; n += blockDim.x) {
    const int i1 = P.itb1 + n, i2 = P.itb2 + n;
    const float s1 = (i1 >= 0 && i1 < P.nt) ? d[i1] : 0.0f;
    const float s2 = (i2 >= 0 && i2 < P.nt) ? d[i2] : 0.0f;
    A[n] = make_float2(taper_apply(s1, n, N, tap), taper_apply(s2, n, N, tap));
  }
  __syncthreads();
  const float2 *Z = fft_run<-1>(A, B, tw, N, P.plan.code);
  for (int k = threadIdx.x; k < P.nww; k += blockDim.x) {
    const int m = P.nf1 + k;
    const float2 zp = Z[m], zm = cconj(Z[N - m]);
    const float2 x1 = make_float2(0.5f * (zp.x + zm.x), 0.5f * (zp.y + zm.y));
    const float2 x2 = make_float2(0.5f * (zp.y - zm.y), -0.5f * (zp.x - zm.x));
    const float w = P.wtab[k];
    C1[k] = make_float2(x1.y * w, -x1.x * w);
    C2[k] = make_float2(x2.y * w, -x2.x * w);
    if (MODE == 0) {
      re_out[(size_t)tr * 2 * P.nww + k] = C1[k].x;
      im_out[(size_t)tr * 2 * P.nww + k] = C1[k].y;
      re_out[(size_t)tr * 2 * P.nww + P.nww + k] = C2[k].x;
      im_out[(size_t)tr * 2 * P.nww + P.nww + k] = C2[k].y;
    }
  }
  if (MODE == 0) return;
  __syncthreads();
  // level j: filter both windows' spectra, inverse FFT as one Hermitian
  // pair; sample it of the level is then level_sample(Y, it)
  auto level = [&](int j) {
    const float2 *Hj = P.H + (size_t)j * P.nww;
    for (int n = threadIdx.x; n < N; n += blockDim.x) {
      float2 v = make_float2(0.f, 0.f);
      const int kp = n - P.nf1, km = N - n - P.nf1;
      if (kp >= 0 && kp < P.nww) {
        const float2 h = Hj[kp];
        const float2 s1 = cmul(C1[kp], h), s2 = cmul(C2[kp], h);
        v = make_float2(s1.x - s2.y, s1.y + s2.x);  // s1 + i s2
      } else if (km >= 0 && km < P.nww) {
        const float2 h = Hj[km];
        const float2 s1 = cconj(cmul(C1[km], h)), s2 = cconj(cmul(C2[km], h));
        v = make_float2(s1.x - s2.y, s1.y + s2.x);
      }
      A[n] = v;
    }
    __syncthreads();
    return fft_run<+1>(A, B, tw, N, P.plan.code);
  };
  auto level_sample = [&](const float2 *Y, int it) {
    const bool s2 = it * P.dt > P.tdmid;
    const int n = it - (s2 ? P.itb2 : P.itb1);
    return (n >= 0 && n < N) ? (s2 ? Y[n].y : Y[n].x) : 0.0f;
  };
  float scale = 1.0f;
  if constexpr (scaled_out<OutT>()) {
    // eight lanes per level, each over every eighth bin, then summed across
    // the eight (kThreads is a multiple of 64, so every pass is wave-uniform)
    float m = 0.0f;
    for (int base = 0; base < P.K * 8; base += kThreads) {
      const int p = base + threadIdx.x;
      float b = 0.0f;
      if (p < P.K * 8) {
        const float2 *Hj = P.H + (size_t)(p >> 3) * P.nww;
        for (int k = p & 7; k < P.nww; k += 8)
          b = fmaf(hypotf(Hj[k].x, Hj[k].y), hypotf(C1[k].x, C1[k].y) + hypotf(C2[k].x, C2[k].y), b);
      }
      b += __shfl_xor(b, 1, 64);
      b += __shfl_xor(b, 2, 64);
      b += __shfl_xor(b, 4, 64);
      m = fmaxf(m, 2.0f * b);
    }
    m = block_max(m);
    m *= 1.001f;  // margin for the rounding of the bound's own sum
    int ex = 0;
    if (m > 0.0f && m <= 3.4e38f) {
      (void)frexpf(m, &ex);  // m in [2^(ex-1), 2^ex)
      ex = max(-100, min(ex, 100));
    }
    // pair banks store differences of two samples too: one bit more headroom
    const int top = pair_out<OutT>() ? 14 : 15;
    scale = ldexpf(1.0f, top - ex);
    if (threadIdx.x == 0) {
      float *h = reinterpret_cast<float *>(bank + (size_t)tr * P.row);
      h[0] = ldexpf(1.0f, ex - top);
      h[1] = scale;
      h[2] = 0.0f;
      h[3] = 0.0f;
    }
    // the row's tail padding (to a 16-byte multiple) is written too, so the
    // output is fully determined by the input
    const long long used = P.lead + (long long)P.K * P.nt;
    for (long long i = used + threadIdx.x; i < P.row; i += blockDim.x) store_zero(bank + (size_t)tr * P.row + i);
  }
  for (int j = 0; j < P.K; ++j) {
    const float2 *Y = level(j);
    OutT *o = bank + (size_t)tr * P.row + P.lead + (size_t)j * P.nt;
    if constexpr (pair_out<OutT>()) {
      // (b[i], b[i+1] - b[i]); entries 0 and nt-1 zero
      for (int it = threadIdx.x; it < P.nt; it += blockDim.x) {
        const bool edge = it == 0 || it == P.nt - 1;
        const float b0 = edge ? 0.0f : level_sample(Y, it) * scale;
        const float b1 = edge ? 0.0f : level_sample(Y, it + 1) * scale;
        store_pair(o + it, b0, b1);
      }
    } else {
      for (int it = threadIdx.x; it < P.nt; it += blockDim.x) store_out(o + it, level_sample(Y, it) * scale);
    }
    __syncthreads();
  }
}

// ---- host side --------------------------------------------------------
struct Coefs {
  float2 *tw = nullptr;
  float2 *H = nullptr;
  float *wtab = nullptr;
  float *taper = nullptr;
  Plan plan{};
};

std::mutex g_mu;
std::map<std::vector<double>, Coefs> g_cache;

bool make_plan(int n, Plan *pl) {
  pl->code = 0;
  const int order[] = {4, 2, 3, 5, 7, 11, 13};
  for (int i = 0; i < 7; ++i) {
    const int r = order[i];
    unsigned long long c = 0;
    while (n % r == 0 && c < 15) {
      if (r == 2 && n % 4 == 0) break;
      ++c;
      n /= r;
    }
    pl->code |= c << (4 * i);
  }
  return n == 1;
}

__host__ __device__ inline double gain_limit(double x, double G) {
  const double L = log(G);
  if (x <= L) return exp(x);
  if (x <= L + 0.2) {
    const double c1 = G * (1.0 - L - 2.5 * L * L), c2 = G * (1.0 + 5.0 * L), c3 = -2.5 * G;
    return c1 + c2 * x + c3 * x * x;
  }
  return 1.1 * G;
}

// Arguments of the filter coefficients, shared by the host generator (the
// cached coefficients) and the device one (the no-allocation entry point).
struct CoefArgs {
  int N, K, nww, nf1, nf3, nf4;
  double dwx, wr, dtstar, threshold;
};

inline CoefArgs coef_args(const qpsdm_bank_params *p) {
  const qpsdm_window &w = p->win;
  CoefArgs a;
  a.N = w.nsb;
  a.K = p->K;
  a.nww = w.nww;
  a.nf1 = w.nf1;
  a.nf3 = w.nf3;
  a.nf4 = w.nf4;
  a.dwx = 2.0 * M_PI / ((double)w.nsb * p->dt);
  a.wr = 2.0 * M_PI * p->fc;
  a.dtstar = (double)p->dtstar;
  a.threshold = (double)p->threshold;
  return a;
}

// Filter of level j at kept bin k, in FP64: high cut (the reference's
// derivative filter shape, src/q-psdm.cu:1236-1247), gain limiter,
// dispersion.
__host__ __device__ inline float2 filter_coef(const CoefArgs &a, int j, int k) {
  const int m = a.nf1 + k;
  double hc = 1.0;
  if (m > a.nf4) {
    hc = 0.0;
  } else if (m >= a.nf3 && a.nf4 > a.nf3) {
    const double c = cos((double)(m - a.nf3) / (a.nf4 - a.nf3) * M_PI * 0.5);
    hc = c * c;
  }
  const double om = m * a.dwx, ts = j * a.dtstar;
  const double amp = hc * gain_limit(om * ts * 0.5, a.threshold);
  const double ph = -om * ts * log(om / a.wr) / M_PI;
  return make_float2((float)(amp * cos(ph)), (float)(amp * sin(ph)));
}

__host__ __device__ inline float2 twiddle(int a, int N) {
  return make_float2((float)cos(2.0 * M_PI * a / N), (float)(-sin(2.0 * M_PI * a / N)));
}

__host__ __device__ inline float taper_coef(int i) { return 0.5 - 0.5 * cosf(M_PI * i / 200.0); }

int get_coefs(const qpsdm_bank_params *p, Coefs **out) {
  const qpsdm_window &w = p->win;
  std::vector<double> key = {(double)p->nt, p->dt, (double)w.nsb, (double)w.nf1, (double)w.nww,
                             (double)w.nf3, (double)w.nf4, w.dw, (double)p->K, p->dtstar, p->fc,
                             p->threshold};
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_cache.find(key);
  if (it != g_cache.end()) {
    *out = &it->second;
    return QPSDM_OK;
  }
  Coefs c;
  if (!make_plan(w.nsb, &c.plan))
    return qpsdm::set_error(QPSDM_EINVAL, "nsb=%d has a prime factor above 13", w.nsb);
  const int N = w.nsb, K = p->K, nww = w.nww;
  std::vector<float2> tw(N), H((size_t)K * nww);
  std::vector<float> wt(nww), tap(200);
  for (int a = 0; a < N; ++a) tw[a] = twiddle(a, N);
  const CoefArgs ca = coef_args(p);
  for (int j = 0; j < K; ++j)
    for (int k = 0; k < nww; ++k) H[(size_t)j * nww + k] = filter_coef(ca, j, k);
  // the reference's float accumulation of omega (src/q-psdm.cu:2186-2191)
  float wv = w.nf1 * w.dw;
  for (int k = 0; k < nww; ++k, wv += w.dw) wt[k] = wv;
  for (int i = 0; i < 200; ++i) tap[i] = taper_coef(i);
  QPSDM_HIP(hipMalloc(&c.tw, sizeof(float2) * N));
  QPSDM_HIP(hipMalloc(&c.H, sizeof(float2) * H.size()));
  QPSDM_HIP(hipMalloc(&c.wtab, sizeof(float) * nww));
  QPSDM_HIP(hipMalloc(&c.taper, sizeof(float) * 200));
  QPSDM_HIP(hipMemcpy(c.tw, tw.data(), sizeof(float2) * N, hipMemcpyHostToDevice));
  QPSDM_HIP(hipMemcpy(c.H, H.data(), sizeof(float2) * H.size(), hipMemcpyHostToDevice));
  QPSDM_HIP(hipMemcpy(c.wtab, wt.data(), sizeof(float) * nww, hipMemcpyHostToDevice));
  QPSDM_HIP(hipMemcpy(c.taper, tap.data(), sizeof(float) * 200, hipMemcpyHostToDevice));
  auto res = g_cache.emplace(key, c);
  *out = &res.first->second;
  return QPSDM_OK;
}

int check_params(const qpsdm_bank_params *p) {
  QPSDM_REQUIRE(p, "bank: null params");
  const qpsdm_window &w = p->win;
  QPSDM_REQUIRE(p->nt > 1 && p->dt > 0 && p->K >= 1, "bank: bad nt/dt/K");
  QPSDM_REQUIRE(w.nsb >= 4 && w.nsb <= 8192, "bank: nsb=%d outside [4, 8192]", w.nsb);
  QPSDM_REQUIRE(2 * w.nsb >= p->nt, "bank: trace longer than 2*nsb");
  QPSDM_REQUIRE(w.nf1 >= 1 && w.nww >= 1 && w.nf1 + w.nww - 1 < w.nsb / 2,
                "bank: kept bins must lie strictly between DC and Nyquist");
  return QPSDM_OK;
}

// One work-group holds a trace's windows, their spectra and the twiddles in
// LDS: 8 (3 nsb + 2 nww) bytes plus the static taper table.  Past the
// default per-work-group limit the kernel is opted in to the device's larger
// limit where there is one (hipDeviceAttributeSharedMemPerBlockOptin); past
// that, an error names the limit instead of a launch that fails.
template <typename Kern>
int fit_kernel_lds(Kern kern, size_t lds, int nsb) {
  int dev = 0, base = 0, optin = 0;
  QPSDM_HIP(hipGetDevice(&dev));
  QPSDM_HIP(hipDeviceGetAttribute(&base, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
  if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess) {
    (void)hipGetLastError();
    optin = 0;
  }
  const size_t total = lds + 1024;  // static LDS of the kernel (taper table, block reduction)
  const size_t limit = (size_t)(optin > base ? optin : base);
  QPSDM_REQUIRE(total <= limit,
                "bank: window of nsb=%d needs %zu bytes of LDS per work-group; this device allows %zu", nsb, total,
                limit);
  if (total > (size_t)base)
    QPSDM_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds));
  return QPSDM_OK;
}

// spectra selects the spectra-only kernel; else dtype the bank's.
int fit_lds(size_t lds, int nsb, bool spectra, int dtype) {
  if (spectra) return fit_kernel_lds(qbank_kernel<0, float>, lds, nsb);
  switch (dtype) {
    case QPSDM_F32: return fit_kernel_lds(qbank_kernel<1, float>, lds, nsb);
    case QPSDM_BF16: return fit_kernel_lds(qbank_kernel<1, __hip_bfloat16>, lds, nsb);
    case QPSDM_F16: return fit_kernel_lds(qbank_kernel<1, __half>, lds, nsb);
    case QPSDM_F32P: return fit_kernel_lds(qbank_kernel<1, PairF32>, lds, nsb);
    case QPSDM_F16P: return fit_kernel_lds(qbank_kernel<1, PairF16>, lds, nsb);
    default: return qpsdm::set_error(QPSDM_EINVAL, "bank: dtype %d not F32/BF16/F16/F32P/F16P", dtype);
  }
}

int launch_coefs(const qpsdm_bank_params *p, const Coefs *c, const float *traces, int ntr, void *bank, int dtype,
                 float *re, float *im, void *stream);

int launch(const qpsdm_bank_params *p, const float *traces, int ntr, void *bank, int dtype,
           float *re, float *im, void *stream) {
  int rc = check_params(p);
  if (rc) return rc;
  QPSDM_REQUIRE((traces || ntr == 0) && ntr >= 0, "bank: bad traces");
  if (ntr == 0) return QPSDM_OK;
  Coefs *c = nullptr;
  rc = get_coefs(p, &c);
  if (rc) return rc;
  return launch_coefs(p, c, traces, ntr, bank, dtype, re, im, stream);
}

int launch_coefs(const qpsdm_bank_params *p, const Coefs *c, const float *traces, int ntr, void *bank, int dtype,
                 float *re, float *im, void *stream) {
  int rc = QPSDM_OK;
  BankDev P;
  P.nt = p->nt;
  P.dt = p->dt;
  P.nsb = p->win.nsb;
  P.itb1 = p->win.itb1;
  P.itb2 = p->win.itb2;
  P.nf1 = p->win.nf1;
  P.nww = p->win.nww;
  P.tdmid = p->win.tdmid;
  P.K = p->K;
  P.row = (long long)p->K * p->nt;
  P.lead = 0;
  if (dtype == QPSDM_F16) {
    P.row = (long long)(qpsdm_bank_row_bytes(p->K, p->nt, QPSDM_F16) / 2);
    P.lead = 8;  // the 16-byte header
  }
  if (dtype == QPSDM_F16P) {
    P.row = (long long)(qpsdm_bank_row_bytes(p->K, p->nt, QPSDM_F16P) / 4);
    P.lead = 4;
  }
  P.plan = c->plan;
  P.tw = c->tw;
  P.H = c->H;
  P.wtab = c->wtab;
  P.taper = c->taper;
  const size_t lds = sizeof(float2) * (3 * (size_t)P.nsb + 2 * P.nww);
  rc = fit_lds(lds, P.nsb, re != nullptr, dtype);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (re) {
    hipLaunchKernelGGL((qbank_kernel<0, float>), dim3(ntr), dim3(kThreads), lds, s, P, traces,
                       (float *)nullptr, re, im);
  } else if (dtype == QPSDM_F32) {
    hipLaunchKernelGGL((qbank_kernel<1, float>), dim3(ntr), dim3(kThreads), lds, s, P, traces,
                       (float *)bank, (float *)nullptr, (float *)nullptr);
  } else if (dtype == QPSDM_BF16) {
    hipLaunchKernelGGL((qbank_kernel<1, __hip_bfloat16>), dim3(ntr), dim3(kThreads), lds, s, P,
                       traces, (__hip_bfloat16 *)bank, (float *)nullptr, (float *)nullptr);
  } else if (dtype == QPSDM_F16) {
    hipLaunchKernelGGL((qbank_kernel<1, __half>), dim3(ntr), dim3(kThreads), lds, s, P, traces,
                       (__half *)bank, (float *)nullptr, (float *)nullptr);
  } else if (dtype == QPSDM_F32P) {
    hipLaunchKernelGGL((qbank_kernel<1, PairF32>), dim3(ntr), dim3(kThreads), lds, s, P, traces,
                       (PairF32 *)bank, (float *)nullptr, (float *)nullptr);
  } else if (dtype == QPSDM_F16P) {
    hipLaunchKernelGGL((qbank_kernel<1, PairF16>), dim3(ntr), dim3(kThreads), lds, s, P, traces,
                       (PairF16 *)bank, (float *)nullptr, (float *)nullptr);
  } else {
    return qpsdm::set_error(QPSDM_EINVAL, "bank: dtype %d not F32/BF16/F16/F32P/F16P", dtype);
  }
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

}  // namespace

// Coefficients of a parameter set uploaded ahead of use (so that a later
// launch can be captured into a graph without allocations or copies).
int qpsdm_qbank_prepare(const qpsdm_bank_params *p) {
  int rc = check_params(p);
  if (rc) return rc;
  Coefs *c = nullptr;
  return get_coefs(p, &c);
}

extern "C" int qpsdm_trace_spectra(const qpsdm_bank_params *p, const float *traces, int ntr,
                                   float *re, float *im, void *stream) {
  QPSDM_REQUIRE((re && im) || ntr == 0, "spectra: null outputs");
  return launch(p, traces, ntr, nullptr, QPSDM_F32, re, im, stream);
}

extern "C" size_t qpsdm_bank_row_bytes(int K, int nt, int dtype) {
  if (K <= 0 || nt <= 0) return 0;
  const size_t n = (size_t)K * nt;
  switch (dtype) {
    case QPSDM_F32: return n * 4;
    case QPSDM_BF16: return n * 2;
    case QPSDM_F16: return 16 + (n * 2 + 15) / 16 * 16;
    case QPSDM_F32P: return n * 8;
    case QPSDM_F16P: return 16 + (n * 4 + 15) / 16 * 16;
    default: return 0;
  }
}

extern "C" int qpsdm_qbank(const qpsdm_bank_params *p, const float *traces, int ntr, void *bank,
                           int dtype, void *stream) {
  QPSDM_REQUIRE(bank || ntr == 0, "bank: null output");
  return launch(p, traces, ntr, bank, dtype, nullptr, nullptr, stream);
}

// ---- the entry point that allocates nothing ------------------------------
// qpsdm_qbank keeps each parameter set's coefficients in device memory it
// allocates on the set's first call.  qpsdm_qbank_noalloc instead generates
// them on the device, in FP64, into `scratch` (a file-scope device array of
// fixed size, one copy per GPU), on the call's stream before the bank
// kernel reads them: no allocation, no host-to-device copy, so the call can
// be captured into a graph.  The array keeps the last call's coefficients;
// calls that overlap on one GPU race for it.
namespace {

constexpr int kScrN = 8192;      // twiddles: nsb <= 8192 (the bank's own limit)
constexpr int kScrKW = 131072;   // filters: K * nww <= 131072
constexpr int kScrW = 4096;      // omega per kept bin: nww <= 4096
constexpr int kScrTaper = 200;   // raised-cosine taper
constexpr size_t kScrFloat2 = (size_t)kScrN + kScrKW + (kScrW + kScrTaper + 1) / 2;

__device__ __align__(256) float2 scratch[kScrFloat2];

std::mutex g_scr_mu;
void *g_scr_ptr[64] = {};

// The calling device's copy of scratch, looked up once per device with that
// device current (a host-side name of a __device__ array is not a device
// address).
int scratch_base(float2 **out) {
  int dev = 0;
  QPSDM_HIP(hipGetDevice(&dev));
  QPSDM_REQUIRE(dev >= 0 && dev < 64, "qbank_noalloc: device %d beyond the 64 this library tracks", dev);
  std::lock_guard<std::mutex> lk(g_scr_mu);
  if (!g_scr_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return qpsdm::set_error(QPSDM_EHIP, "qbank_noalloc: hipGetSymbolAddress(scratch) failed: %s",
                              hipGetErrorString(e));
    g_scr_ptr[dev] = ptr;
  }
  *out = (float2 *)g_scr_ptr[dev];
  return QPSDM_OK;
}

__global__ void coef_kernel(CoefArgs a, float nf1dw, float dw, float2 *tw, float2 *H, float *wtab,
                            float *taper) {
  const int nh = a.K * a.nww;
  const int total = a.N + nh + kScrTaper;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    if (i < a.N) {
      tw[i] = twiddle(i, a.N);
    } else if (i < a.N + nh) {
      const int e = i - a.N;
      H[e] = filter_coef(a, e / a.nww, e % a.nww);
    } else {
      taper[i - a.N - nh] = taper_coef(i - a.N - nh);
    }
  }
  // the reference's float accumulation of omega, in order (one lane)
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    float wv = nf1dw;
    for (int k = 0; k < a.nww; ++k, wv += dw) wtab[k] = wv;
  }
}

}  // namespace

extern "C" int qpsdm_qbank_noalloc(const qpsdm_bank_params *p, const float *traces, int ntr, void *bank, int dtype,
                                   void *stream) {
  int rc = check_params(p);
  if (rc) return rc;
  QPSDM_REQUIRE(bank || ntr == 0, "bank: null output");
  QPSDM_REQUIRE((traces || ntr == 0) && ntr >= 0, "bank: bad traces");
  const qpsdm_window &w = p->win;
  if (w.nsb > kScrN || (size_t)p->K * w.nww > (size_t)kScrKW || w.nww > kScrW)
    return qpsdm::set_error(QPSDM_ERANGE,
                            "qbank_noalloc: nsb=%d, K*nww=%lld, nww=%d beyond the scratch limits %d, %d, %d", w.nsb,
                            (long long)p->K * w.nww, w.nww, kScrN, kScrKW, kScrW);
  if (ntr == 0) return QPSDM_OK;
  Coefs c;
  if (!make_plan(w.nsb, &c.plan))
    return qpsdm::set_error(QPSDM_EINVAL, "nsb=%d has a prime factor above 13", w.nsb);
  float2 *base = nullptr;
  rc = scratch_base(&base);
  if (rc) return rc;
  c.tw = base;
  c.H = base + kScrN;
  c.wtab = (float *)(base + kScrN + kScrKW);
  c.taper = c.wtab + kScrW;
  const CoefArgs a = coef_args(p);
  const int total = a.N + a.K * a.nww + kScrTaper;
  hipLaunchKernelGGL(coef_kernel, dim3((total + 255) / 256), dim3(256), 0, (hipStream_t)stream, a,
                     (float)(w.nf1 * w.dw), w.dw, c.tw, c.H, c.wtab, c.taper);
  QPSDM_HIP(hipGetLastError());
  return launch_coefs(p, &c, traces, ntr, bank, dtype, nullptr, nullptr, stream);
}
