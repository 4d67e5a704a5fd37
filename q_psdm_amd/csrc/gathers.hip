// Dip-angle gathers at selected image columns, for gfx950.
//
// The reference declares them (options opcrp, cdps, aspect, dipx, dipy,
// ddipx, ddipy at src/q-psdm.cu:651-702, bin counts ndipx = (int)(dipx /
// ddipx + 0.5) at src/q-psdm.cu:828-829, device buffers gthrxd/gthryd at
// src/q-psdm.cu:211-212) but never fills them.  Here every update of the
// imaging sum at a selected column -- the same term the image receives --
// is also added to the bin of its bisector's dip angle, theta = atan(X / Z)
// in x and atan(Y / Z) in y, bins of width ddip centred on multiples of
// ddip from -ndip ddip to +ndip ddip.  Bin boundaries are the tangents of
// the half-bin angles, computed in FP64 on the host, so a point's bin is
// decided by comparing its FP32 tangent with them (no arc tangent on the
// device).  Summed over bins, a column's gather is its image column when
// the bins span the dip aperture.
//
// One wave per (column, group of 64 depths), one depth per lane, traces in
// order: the column's table columns, distances' horizontal parts and trace
// row are wave-uniform; each lane owns its depth in an LDS gather [bin][64],
// so bins need no atomics and every bin sums its updates in trace order.
// The per-point arithmetic is the FP32 imaging path's, in its order, with
// no contraction beyond the explicit fmaf calls.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <math.h>

#include <vector>

#include "common.h"
#include "image_dev.h"

namespace qk = qpsdm_dev;

namespace {

constexpr int kMaxBins = 255;     // per axis
constexpr int kMaxBinsBoth = 250; // both axes together (LDS: 64 lanes x 4 bytes each)

struct Edges {
  float x[kMaxBins + 1], y[kMaxBins + 1];  // bin b covers [e[b], e[b + 1])
};

struct GatherDev {
  int nbx, nby;  // bins per axis; 0 = that axis not gathered
};

// Table value at depth node w of a wave-uniform column, w per lane.
__device__ inline float2 at_lane(const qk::TabF32 &t, const qk::TabF32::Col &c, int w) {
  const int vo = c.o + w * 8, du = t.nw * 8, dv = t.nu * t.nw * 8;
  const qk::TabF32::Raw q{qk::ld_f2(c.r, vo, 0), qk::ld_f2(c.r, vo + du, 0), qk::ld_f2(c.r, vo + dv, 0),
                          qk::ld_f2(c.r, vo + dv + du, 0)};
  return t.reduce(c, q);
}
__device__ inline float2 at_lane(const qk::TabU16X &t, const qk::TabU16X::Col &c, int w) {
  const int sb = (w >> 3) * qk::kXBrickBytes, vo = c.o + sb + (w & 7) * 4;
  qk::TabU16X::Raw q;
  q.h = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(c.r, c.base + sb, 0, 0));
  q.q00 = __builtin_amdgcn_raw_buffer_load_b32(c.r, vo, 0, 0);
  q.q01 = __builtin_amdgcn_raw_buffer_load_b32(c.r, vo + 32, 0, 0);
  q.q10 = __builtin_amdgcn_raw_buffer_load_b32(c.r, vo + 9 * 32, 0, 0);
  q.q11 = __builtin_amdgcn_raw_buffer_load_b32(c.r, vo + 10 * 32, 0, 0);
  return t.reduce(c, q);
}

__device__ inline float2 at_lane(const qk::TabU16Q &t, const qk::TabU16Q::Col &c, int w) {
  const int sb = (w >> 3) * qk::kQBrickBytes;
  qk::TabU16Q::Raw q;
  q.h = qk::TabU16Q::header(c, sb);
  const uint4 v = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(c.r, c.o, sb + (w & 7) * 16, 0));
  q.q00 = v.x, q.q01 = v.y, q.q10 = v.z, q.q11 = v.w;
  return t.reduce(c, q);
}

// Bin of tangent v among n bins with boundaries e[0..n] (LDS), -1 outside.
__device__ inline int bin_of(float v, const float *e, int n) {
  if (!(v >= e[0] && v < e[n])) return -1;
  int lo = 0, hi = n;  // e[lo] <= v < e[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (v >= e[mid])
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <typename BankT, typename Tab>
__global__ __launch_bounds__(64) void gather_kernel(qk::ImgDev P, GatherDev G, Edges E, Tab tab,
                                                    const int2 *__restrict__ cols, float *__restrict__ gx,
                                                    float *__restrict__ gy, const float4 *__restrict__ geom,
                                                    const BankT *__restrict__ bank) {
#pragma clang fp contract(off)
  extern __shared__ float lds[];  // edges x [nbx + 1], y [nby + 1], then gathers [nbx + nby][64]
  float *ex = lds, *ey = lds + G.nbx + 1, *acc = ey + G.nby + 1;
  const int lane = threadIdx.x;
  const int iz = blockIdx.y * 64 + lane;
  for (int i = lane; i <= G.nbx; i += 64) ex[i] = E.x[i];
  for (int i = lane; i <= G.nby; i += 64) ey[i] = E.y[i];
  for (int b = 0; b < G.nbx + G.nby; ++b) acc[b * 64 + lane] = 0.0f;
  __syncthreads();
  const int2 c = cols[blockIdx.x];
  const float x = (float)c.x * P.dx, y = (float)c.y * P.dy, z = (float)iz * P.dz;
  const bool zok = iz >= 1 && iz < P.nz;
  const qk::Dip &D = P.dip;
  const float fw = fminf(z * P.inv_dw, (float)(P.nw - 1));
  const int w = min((int)fw, P.nw - 2);
  const float cw = fw - (float)w;
  const float inv_dt = P.inv_dt, inv_dts = P.inv_dtstar;
  auto sat01 = [](float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); };
  for (int t = 0; t < P.ntr; ++t) {
    const float4 g = geom[P.tr0 + t];
    if (g.x != g.x) continue;  // rejected by the extent check
    const float dsx = g.x - x, dsy = g.y - y, dgx = g.z - x, dgy = g.w - y;
    const float hs2 = dsx * dsx + dsy * dsy, hg2 = dgx * dgx + dgy * dgy;
    // source leg
    auto cs = qk::column_setup(P, tab, g.x, g.y, x, y);
    const float2 slo = at_lane(tab, cs, w), shi = at_lane(tab, cs, w + 1);
    float ts = fmaf(cw, shi.x - slo.x, slo.x);
    const float tss = fmaf(cw, shi.y - slo.y, slo.y);
    ts = fmaxf(ts, P.tmin);
    const float d2s = hs2 + z * z;
    const float rds = 1.0f / sqrtf(d2s), ds = d2s * rds;
    const float tsn = ts * inv_dt, tsl = fmaf(tss, inv_dts, 0.5f);
    const float A = ds * ds * ds * (1.0f / ts);
    // receiver leg and bisector (scaled by ds); dip taper as clamped ramps
    const float rdg = 1.0f / sqrtf(hg2 + z * z);
    const float q = rdg * ds;
    const float X = fmaf(dgx, q, dsx), Y = fmaf(dgy, q, dsy), Z = fmaf(z, q, z);
    const float rz = 1.0f / Z, tanx = X * rz, tany = Y * rz;
    const float wx = fminf(sat01(fmaf(tanx, D.rx1, D.cx1)), sat01(fmaf(tanx, -D.rx2, D.cx2)));
    const float wy = fminf(sat01(fmaf(tany, D.ry1, D.cy1)), sat01(fmaf(tany, -D.ry2, D.cy2)));
    const float wt = wx * wy;
    auto cg = qk::column_setup(P, tab, g.z, g.w, x, y);
    const float2 glo = at_lane(tab, cg, w), ghi = at_lane(tab, cg, w + 1);
    float unit;
    const __amdgpu_buffer_rsrc_t br = qk::bank_row(bank, (size_t)t, P.K, P.nt, &unit);
    if (!(zok && wt > 0.0f)) continue;
    float tg = fmaf(cw, ghi.x - glo.x, glo.x);
    const float tsg = fmaf(cw, ghi.y - glo.y, glo.y);
    tg = fmaxf(tg, P.tmin);
    const float f = fmaf(tg, inv_dt, tsn);
    const int i0 = (int)f;
    if (!(i0 > 0 && i0 + 1 < P.nt)) continue;
    int j = (int)fmaf(tsg, inv_dts, tsl);
    j = j < 0 ? 0 : (j > P.K - 1 ? P.K - 1 : j);
    const float fr = f - (float)i0;
    const float2 s = qk::bank_pair(br, j * P.nt + i0, bank);
    const float tr = A * (rdg * rdg * rdg) * tg;
    const float W = (wt * unit) * (tr + 1.0f / tr), sv = fmaf(fr, s.y - s.x, s.x);
    if (G.nbx) {
      const int b = bin_of(tanx, ex, G.nbx);
      if (b >= 0) acc[b * 64 + lane] = fmaf(W, sv, acc[b * 64 + lane]);
    }
    if (G.nby) {
      const int b = bin_of(tany, ey, G.nby);
      if (b >= 0) acc[(G.nbx + b) * 64 + lane] = fmaf(W, sv, acc[(G.nbx + b) * 64 + lane]);
    }
  }
  if (!zok) return;
  // gathers [col][bin][nz] accumulate across launches, like the image
  for (int b = 0; b < G.nbx; ++b) gx[((size_t)blockIdx.x * G.nbx + b) * P.nz + iz] += acc[b * 64 + lane];
  for (int b = 0; b < G.nby; ++b)
    gy[((size_t)blockIdx.x * G.nby + b) * P.nz + iz] += acc[(G.nbx + b) * 64 + lane];
}

template <typename BankT, typename Tab>
int launch(const qk::ImgDev &P, const GatherDev &G, const Edges &E, const Tab &tab, const int2 *cols, int ncol,
           float *gx, float *gy, const float4 *geom, const BankT *bank, hipStream_t s) {
  const size_t lds = (size_t)(G.nbx + 1 + G.nby + 1) * 4 + (size_t)(G.nbx + G.nby) * 64 * 4;
  const dim3 grid(ncol, (P.nz + 63) / 64);
  hipLaunchKernelGGL((gather_kernel<BankT, Tab>), grid, dim3(64), lds, s, P, G, E, tab, cols, gx, gy, geom, bank);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

template <typename Tab>
int launch_bank(const qk::ImgDev &P, const GatherDev &G, const Edges &E, const Tab &tab, const int2 *cols,
                int ncol, float *gx, float *gy, const float4 *geom, const void *bank, int bank_dtype,
                hipStream_t s) {
  switch (bank_dtype) {
    case QPSDM_F32: return launch(P, G, E, tab, cols, ncol, gx, gy, geom, (const float *)bank, s);
    case QPSDM_BF16: return launch(P, G, E, tab, cols, ncol, gx, gy, geom, (const __hip_bfloat16 *)bank, s);
    case QPSDM_F16: return launch(P, G, E, tab, cols, ncol, gx, gy, geom, (const __half *)bank, s);
    default: return qpsdm::set_error(QPSDM_EINVAL, "gathers: bank dtype %d", bank_dtype);
  }
}

// Boundaries of n = 2 ndip + 1 bins of width ddip degrees centred on
// k ddip, k = -ndip .. ndip: e[b] = tan((b - ndip - 1/2) ddip).
void edges(int ndip, double ddip, float *e) {
  const double rad = 3.14159265358979323846 / 180.0;
  for (int b = 0; b <= 2 * ndip + 1; ++b) e[b] = (float)tan(((double)(b - ndip) - 0.5) * ddip * rad);
}

}  // namespace

extern "C" int qpsdm_gather_bins(const qpsdm_gather_params *g, int *nbx, int *nby) {
  QPSDM_REQUIRE(g && nbx && nby, "gather_bins: null args");
  QPSDM_REQUIRE(g->aspect >= 1 && g->aspect <= 4, "gather_bins: aspect %d (1 all, 2 x, 3 y, 4 none)", g->aspect);
  const bool wx = g->aspect == 1 || g->aspect == 2, wy = g->aspect == 1 || g->aspect == 3;
  int n[2] = {0, 0};
  const float dip[2] = {g->dipx, g->dipy}, dd[2] = {g->ddipx, g->ddipy};
  const bool want[2] = {wx, wy};
  for (int a = 0; a < 2; ++a) {
    if (!want[a]) continue;
    QPSDM_REQUIRE(dd[a] > 0 && dip[a] >= 0, "gather_bins: dip %g and bin width %g degrees", dip[a], dd[a]);
    const int ndip = (int)(dip[a] / dd[a] + 0.5);  // src/q-psdm.cu:828-829
    QPSDM_REQUIRE(2 * ndip + 1 <= kMaxBins, "gather_bins: %d bins on one axis (at most %d)", 2 * ndip + 1,
                  kMaxBins);
    QPSDM_REQUIRE((ndip + 0.5) * dd[a] < 90.0, "gather_bins: bins reach %g degrees (below 90 required)",
                  (ndip + 0.5) * dd[a]);
    n[a] = 2 * ndip + 1;
  }
  QPSDM_REQUIRE(n[0] + n[1] <= kMaxBinsBoth, "gather_bins: %d bins on both axes (at most %d)", n[0] + n[1],
                kMaxBinsBoth);
  *nbx = n[0];
  *nby = n[1];
  return QPSDM_OK;
}

extern "C" int qpsdm_gathers_shot(const qpsdm_image_params *p, const qpsdm_gather_params *g, const int *cols_dev,
                                  int ncol, float *gx_dev, float *gy_dev, const float *geom_dev, int tr0, int ntr,
                                  const void *bank_dev, int bank_dtype, const void *tab_dev, int table_dtype,
                                  void *stream) {
  qk::ImgDev P;
  QPSDM_REQUIRE(p, "gathers: null params");
  const qpsdm_image_params pe = qk::exact_form(p);
  int rc = qk::make_dev(&pe, nullptr, tr0, ntr, &P);
  if (rc) return rc;
  GatherDev G;
  rc = qpsdm_gather_bins(g, &G.nbx, &G.nby);
  if (rc) return rc;
  QPSDM_REQUIRE(ncol >= 0, "gathers: ncol %d", ncol);
  if (ncol == 0 || ntr == 0 || G.nbx + G.nby == 0) return QPSDM_OK;
  QPSDM_REQUIRE(cols_dev && geom_dev && bank_dev && tab_dev, "gathers: null buffer");
  QPSDM_REQUIRE(!G.nbx || gx_dev, "gathers: aspect asks for x gathers and gx is NULL");
  QPSDM_REQUIRE(!G.nby || gy_dev, "gathers: aspect asks for y gathers and gy is NULL");
  Edges E = {};
  if (G.nbx) edges(G.nbx / 2, g->ddipx, E.x);
  if (G.nby) edges(G.nby / 2, g->ddipy, E.y);
  const hipStream_t s = (hipStream_t)stream;
  const int2 *cols = (const int2 *)cols_dev;
  const float4 *g4 = (const float4 *)geom_dev;
  switch (table_dtype) {
    case QPSDM_F32:
      return launch_bank(P, G, E, qk::make_tab_f32(p->tg, tab_dev), cols, ncol, gx_dev, gy_dev, g4, bank_dev,
                         bank_dtype, s);
    case QPSDM_U16X:
      return launch_bank(P, G, E, qk::make_tab_u16x(p->tg, tab_dev), cols, ncol, gx_dev, gy_dev, g4, bank_dev,
                         bank_dtype, s);
    case QPSDM_U16Q:
      QPSDM_REQUIRE(qk::u16q_fits(p->tg), "gathers: U16Q node of %zu bytes (below 2 GiB)", qk::u16q_node_bytes(p->tg));
      return launch_bank(P, G, E, qk::make_tab_u16q(p->tg, tab_dev), cols, ncol, gx_dev, gy_dev, g4, bank_dev,
                         bank_dtype, s);
    default:
      return qpsdm::set_error(QPSDM_EINVAL, "gathers: table dtype %d (QPSDM_F32, QPSDM_U16X or QPSDM_U16Q)",
                              table_dtype);
  }
}
