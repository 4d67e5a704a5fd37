// Kernel (b): Kirchhoff imaging sum for gfx950, one launch per run of
// traces (the per-tile work is image_tile, image_tile.h).
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "common.h"
#include "image_nodes.h"

namespace qk = qpsdm_dev;

namespace qpsdm_dev {

int node_range_ni(const qpsdm_image_params *p, int ni, ImgDev *o, int align);

// Node range of the node form (ImgDev nf_*), with the device's float rule
// for a point's interval: intervals from the first node at or below z_exact
// (node 1 at least: the surface node has no bisector) to the last interval
// whose points lie within the tables' depth range; ni intervals per tile.
int node_range(const qpsdm_image_params *p, int ni, ImgDev *o) {
  const int nz = p->nz, nw = p->tg.nw;
  const float dz = p->dz, inv_dw = o->inv_dw;
  QPSDM_REQUIRE(p->z_exact >= 0.0f, "image: z_exact must be >= 0");
  QPSDM_REQUIRE(ni >= 2 && ni <= 4, "image: the node form takes tile_z = 2, 3 or 4 depth intervals per tile");
  return node_range_ni(p, ni, o, 1);
}

// The span form's node range (the node form's, its first interval rounded
// up to a multiple of 8) and its bands: band k starts
// at the first interval at or below span_z[k], rounded up to a multiple of 8
// intervals past the range's first interval ws (so that tiles of 2, 4 or 8
// intervals from a band's start never straddle two bands and every span
// length divides the band's offsets); intervals above the first band form a
// band of one-interval spans; a band that a later one starts at the same
// interval is dropped.
int span_range(const qpsdm_image_params *p, int ni, ImgDev *o) {
  QPSDM_REQUIRE(p->z_exact >= 0.0f, "image: z_exact must be >= 0");
  QPSDM_REQUIRE(ni == 4 || ni == 8, "image: the span form takes tile_z = 4 or 8 intervals per tile");
  QPSDM_REQUIRE(p->span_n >= 1 && p->span_n <= 4, "image: the span form takes 1 to 4 bands (span_n)");
  for (int k = 0; k < p->span_n; ++k) {
    const int S = p->span_s[k];
    QPSDM_REQUIRE(S == 1 || S == 2 || S == 4 || S == 8, "image: span_s[%d] = %d (1, 2, 4 or 8)", k, S);
    QPSDM_REQUIRE(p->span_z[k] >= 0.0f && (k == 0 || p->span_z[k] >= p->span_z[k - 1]),
                  "image: span_z must be >= 0 and ascending");
  }
  int rc = node_range_ni(p, ni, o, 8);
  if (rc) return rc;
  o->nf_ngroups = 0;
  if (o->nf_ps >= o->nf_pe) return QPSDM_OK;  // no node range: the exact body images every point
  const int ws = o->nf_ws, we = o->nf_we;
  int start[5], span[5], n = 0;
  for (int k = 0; k < p->span_n; ++k) {
    const double x = (double)p->span_z[k] / (double)p->tg.dw - (double)ws;
    const int b = std::min(ws + 8 * (int)std::ceil(std::max(0.0, x) / 8.0), we);
    if (n == 0 && b > ws) {
      start[n] = ws;
      span[n++] = 1;
    }
    if (n > 0 && start[n - 1] == b) --n;
    start[n] = b;
    span[n++] = p->span_s[k];
  }
  if (n > 0 && start[n - 1] == we) --n;
  o->sp_n = n;
  for (int k = 0; k < n; ++k) {
    o->sp_w[k] = start[k];
    o->sp_s[k] = span[k];
  }
  o->sp_w[n] = we;
  // the 3, 2 slot pattern: every interval of the range (band starts are 8
  // intervals apart, so the parity from ws is the parity within a tile)
  auto fits = [&](int w) {
    const int a = std::max(first_point(w, p->nz, p->dz, o->inv_dw, p->tg.nw), o->nf_ps);
    const int b = std::min(first_point(w + 1, p->nz, p->dz, o->inv_dw, p->tg.nw), o->nf_pe);
    return b - a <= (((w - ws) & 1) ? 2 : 3);
  };
  int pat = 1;
  for (int w = ws; w < we - 1 && pat; ++w) pat = fits(w);
  o->sp_pat = pat;
  // (the last interval can hold the points down to the tables' last node:
  // one more than the pattern, which a tile whose last interval it is takes
  // in a third slot)
  const int last_ni = n > 0 && std::max(ni, o->sp_s[n - 1]) == 8 ? 8 : 4;
  o->sp_pat_last = !pat || n == 0 ? 0 : fits(we - 1) ? 1 : (we - o->sp_w[n - 1]) % last_ni == 0 ? 2 : 0;
  return QPSDM_OK;
}

// (align: the first interval rounded up to a multiple of it; the span
// form's is 8, so that a tile of 4 or 8 intervals from a band's start keeps
// all but its last table node in one 8-node depth brick)
int node_range_ni(const qpsdm_image_params *p, int ni, ImgDev *o, int align) {
  const int nz = p->nz, nw = p->tg.nw;
  const float dz = p->dz, inv_dw = o->inv_dw;
  int ws = std::max(1, (int)std::ceil((double)p->z_exact / (double)p->tg.dw));
  ws = (ws + align - 1) / align * align;
  int pe = nz;
  for (int iz = 0; iz < nz; ++iz) {
    const float z = (float)iz * dz;
    if (z * inv_dw > (float)(nw - 1)) {
      pe = iz;
      break;
    }
  }
  if (ws > nw - 2) return QPSDM_OK;
  const int ps = first_point(ws, nz, dz, inv_dw, nw);
  if (ps >= pe) return QPSDM_OK;
  float f;
  const int we = point_interval(pe - 1, dz, inv_dw, nw, &f) + 1;
  int pm = 0;
  for (int w = ws; w < we; ++w) {
    const int a = std::max(first_point(w, nz, dz, inv_dw, nw), ps);
    const int b = std::min(first_point(w + 1, nz, dz, inv_dw, nw), pe);
    pm = std::max(pm, b - a);
  }
  QPSDM_REQUIRE(pm <= 3, "image: the node form takes at most 3 image points per table depth interval (dw <= 3 dz), "
                         "found %d", pm);
  o->nf_ws = ws;
  o->nf_we = we;
  o->nf_ps = ps;
  o->nf_pe = pe;
  o->nf_pm = std::max(pm, 2);
  o->nf_ngroups = (we - ws + ni - 1) / ni;
  return QPSDM_OK;
}

}  // namespace qpsdm_dev

namespace {

// The launch's single run: the parameters' dip bounds, bank rows from 0.
struct OneRun {
  qk::Run run;
  __device__ qk::Run operator()(int) const { return run; }
};

inline OneRun one_run(const qk::ImgDev &P) {
  OneRun r;
  r.run.tr0 = P.tr0;
  r.run.ntr = P.ntr;
  r.run.brow0 = 0;
  r.run.dip = P.dip;
  r.run.src = nullptr;
  r.run.ssx = r.run.ssy = r.run.su0 = r.run.sv0 = 0.f;
  r.run.snu = r.run.snv = 0;
  return r;
}

// One wave per tile: blockIdx.x = column tile, blockIdx.y = depth group.
// RANGE: the depth groups of [zb, ze) (the node form's exact parts).
template <int ZP, typename BankT, typename Tab, bool RANGE>
__global__ __launch_bounds__(64) void image_kernel(qk::ImgDev P, Tab tab, float *__restrict__ img,
                                                   const float4 *__restrict__ geom,
                                                   const BankT *__restrict__ bank, OneRun run) {
  qk::image_tile<ZP, false, RANGE, false>(P, tab, img, geom, bank, blockIdx.x, blockIdx.y, run, 1);
}

// Node form: blockIdx.x = column tile, blockIdx.y = node group.
template <int NI, int PM, typename BankT, typename Tab>
__global__ __launch_bounds__(64) void image_nodes_kernel(qk::ImgDev P, Tab tab, float *__restrict__ img,
                                                         const float4 *__restrict__ geom,
                                                         const BankT *__restrict__ bank, OneRun run) {
  qk::image_tile_nodes<NI, PM, false>(P, tab, img, geom, bank, blockIdx.x, blockIdx.y, run, 1);
}

template <int ZP>
__global__ __launch_bounds__(64) void count_kernel(qk::ImgDev P, const float4 *__restrict__ geom,
                                                   unsigned long long *count, OneRun run) {
  const unsigned n = qk::count_tile<ZP>(P, geom, blockIdx.x, blockIdx.y, run, 1);
  unsigned long long v = n;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (threadIdx.x == 0 && v) atomicAdd(count, v);
}

__global__ void scale_kernel(float *p, size_t n, float s) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] *= s;
}

template <int ZP, bool RANGE = false, typename BankT, typename Tab>
int launch_t(const qk::ImgDev &P, const Tab &tab, float *img, const float4 *geom, const BankT *bank,
             hipStream_t s) {
  const int nxt = (P.nx + P.tile_x - 1) / P.tile_x, nyt = (P.ny + P.tile_y - 1) / P.tile_y;
  if (P.ze <= P.zb) return QPSDM_OK;
  const dim3 grid(nxt * nyt, (P.ze - P.zb + ZP - 1) / ZP);
  hipLaunchKernelGGL((image_kernel<ZP, BankT, Tab, RANGE>), grid, dim3(64), 0, s, P, tab, img, geom, bank,
                     one_run(P));
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

template <typename BankT, typename Tab>
int launch_tz(const qk::ImgDev &P, const Tab &tab, float *img, const float4 *geom,
              const BankT *bank, hipStream_t s) {
  switch (P.tile_z) {
    case 2: return launch_t<2>(P, tab, img, geom, bank, s);
    case 3: return launch_t<3>(P, tab, img, geom, bank, s);
    case 4: return launch_t<4>(P, tab, img, geom, bank, s);
    case 6: return launch_t<6>(P, tab, img, geom, bank, s);
    case 8: return launch_t<8>(P, tab, img, geom, bank, s);
    case 12: return launch_t<12>(P, tab, img, geom, bank, s);
    case 16: return launch_t<16>(P, tab, img, geom, bank, s);
    default:
      return qpsdm::set_error(QPSDM_EINVAL, "image: unsupported tile_z=%d (2, 3, 4, 6, 8, 12, 16)",
                              P.tile_z);
  }
}

template <int NI, int PM, typename BankT, typename Tab>
int launch_nodes_t(const qk::ImgDev &P, const Tab &tab, float *img, const float4 *geom, const BankT *bank,
                   hipStream_t s) {
  const int nxt = (P.nx + P.tile_x - 1) / P.tile_x, nyt = (P.ny + P.tile_y - 1) / P.tile_y;
  hipLaunchKernelGGL((image_nodes_kernel<NI, PM, BankT, Tab>), dim3(nxt * nyt, P.nf_ngroups), dim3(64), 0, s, P,
                     tab, img, geom, bank, one_run(P));
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

// The node form's storage types: FP32 or FP16 banks, FP32 or U16Q tables.
template <typename BankT, typename Tab>
constexpr bool node_form_types() {
  return (std::is_same<BankT, float>::value || std::is_same<BankT, __half>::value) &&
         (std::is_same<Tab, qk::TabF32>::value || std::is_same<Tab, qk::TabU16Q>::value);
}

// One run in the form the parameters ask for: the exact body over the whole
// depth range, or the node form -- the exact body (8 points per tile) above
// the node range and beyond the tables' last node, the node body between.
template <typename BankT, typename Tab>
int launch_form(const qk::ImgDev &P, const Tab &tab, float *img, const float4 *geom, const BankT *bank,
                hipStream_t s) {
  if (P.nf_ngroups == 0) return launch_tz(P, tab, img, geom, bank, s);
  if constexpr (!node_form_types<BankT, Tab>()) {
    return qpsdm::set_error(QPSDM_EINVAL, "image: the node form takes FP32 or FP16 banks and FP32 or U16Q tables");
  } else {
    qk::ImgDev E = P;
    E.tile_z = 8;
    E.zb = 0;
    E.ze = P.nf_ps;
    int rc = launch_t<8, true>(E, tab, img, geom, bank, s);
    E.zb = P.nf_pe;
    E.ze = P.nz;
    if (!rc) rc = launch_t<8, true>(E, tab, img, geom, bank, s);
    if (rc) return rc;
    switch (P.tile_z * 4 + P.nf_pm) {
      case 2 * 4 + 2: return launch_nodes_t<2, 2>(P, tab, img, geom, bank, s);
      case 2 * 4 + 3: return launch_nodes_t<2, 3>(P, tab, img, geom, bank, s);
      case 3 * 4 + 2: return launch_nodes_t<3, 2>(P, tab, img, geom, bank, s);
      case 3 * 4 + 3: return launch_nodes_t<3, 3>(P, tab, img, geom, bank, s);
      case 4 * 4 + 2: return launch_nodes_t<4, 2>(P, tab, img, geom, bank, s);
      case 4 * 4 + 3: return launch_nodes_t<4, 3>(P, tab, img, geom, bank, s);
      default:
        return qpsdm::set_error(QPSDM_EINVAL, "image: node form with %d intervals x %d points per tile", P.tile_z,
                                P.nf_pm);
    }
  }
}

template <typename Tab>
int launch_bank(const qk::ImgDev &P, const Tab &T, float *img, const float4 *g4, const void *bank, int bank_dtype,
                hipStream_t s) {
  switch (bank_dtype) {
    case QPSDM_F32: return launch_form(P, T, img, g4, (const float *)bank, s);
    case QPSDM_BF16: return launch_form(P, T, img, g4, (const __hip_bfloat16 *)bank, s);
    case QPSDM_F16: return launch_form(P, T, img, g4, (const __half *)bank, s);
    default: return qpsdm::set_error(QPSDM_EINVAL, "image: bank dtype %d", bank_dtype);
  }
}

}  // namespace

extern "C" int qpsdm_image_shot(const qpsdm_image_params *p, const qpsdm_launch *shape, float *img,
                                const float *geom, int tr0, int ntr, const void *bank,
                                int bank_dtype, const void *tab, int table_dtype, void *stream) {
  qk::ImgDev P;
  QPSDM_REQUIRE(p && p->weight_nodes != 2, "image: the span form (weight_nodes = 2) runs through qpsdm_image_batch");
  int rc = qk::make_dev(p, shape, tr0, ntr, &P);
  if (rc) return rc;
  QPSDM_REQUIRE(img && geom && bank && tab, "image: null buffer");
  if (ntr == 0) return QPSDM_OK;
  hipStream_t s = (hipStream_t)stream;
  const float4 *g4 = (const float4 *)geom;
  switch (table_dtype) {
    case QPSDM_F32: return launch_bank(P, qk::make_tab_f32(p->tg, tab), img, g4, bank, bank_dtype, s);
    case QPSDM_F16B: return launch_bank(P, qk::make_tab_b16<false>(p->tg, tab), img, g4, bank, bank_dtype, s);
    case QPSDM_U16B: return launch_bank(P, qk::make_tab_b16<true>(p->tg, tab), img, g4, bank, bank_dtype, s);
    case QPSDM_U16X: return launch_bank(P, qk::make_tab_u16x(p->tg, tab), img, g4, bank, bank_dtype, s);
    case QPSDM_U16Q:
      QPSDM_REQUIRE(qk::u16q_fits(p->tg), "image: U16Q node of %zu bytes (below 2 GiB)", qk::u16q_node_bytes(p->tg));
      return launch_bank(P, qk::make_tab_u16q(p->tg, tab), img, g4, bank, bank_dtype, s);
    default: break;
  }
  return qpsdm::set_error(QPSDM_EINVAL, "image: bad dtypes bank=%d table=%d", bank_dtype,
                          table_dtype);
}

// ---- table lookup operator -------------------------------------------------
template <typename Tab>
__global__ void lookup_kernel(qk::ImgDev P, Tab tab, float px, float py, const float *__restrict__ xyz, int n,
                              float2 *__restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  auto c = qk::column_setup(P, tab, px, py, x, y);
  float fz;
  const int w = qk::depth_node(P, z, &fz);
  const float2 lo = tab.at(c, w), hi = tab.at(c, w + 1);
  out[i] = make_float2(fmaf(fz, hi.x - lo.x, lo.x), fmaf(fz, hi.y - lo.y, lo.y));
}

template <typename Tab>
int launch_lookup(const qk::ImgDev &P, const Tab &tab, float px, float py, const float *xyz, int n, float *out,
                  hipStream_t s) {
  hipLaunchKernelGGL(lookup_kernel<Tab>, dim3((n + 255) / 256), dim3(256), 0, s, P, tab, px, py, xyz, n,
                     (float2 *)out);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

extern "C" int qpsdm_tables_lookup(const qpsdm_image_params *p, const void *tab, int table_dtype, float px,
                                   float py, const float *xyz, int n, float *out, void *stream) {
  qk::ImgDev P;
  QPSDM_REQUIRE(p, "tables_lookup: null params");
  const qpsdm_image_params pe = qk::exact_form(p);
  int rc = qk::make_dev(&pe, nullptr, 0, 0, &P);
  if (rc) return rc;
  QPSDM_REQUIRE(tab && xyz && out && n >= 0, "tables_lookup: null args");
  if (n == 0) return QPSDM_OK;
  const hipStream_t s = (hipStream_t)stream;
  switch (table_dtype) {
    case QPSDM_F32: return launch_lookup(P, qk::make_tab_f32(p->tg, tab), px, py, xyz, n, out, s);
    case QPSDM_F16B: return launch_lookup(P, qk::make_tab_b16<false>(p->tg, tab), px, py, xyz, n, out, s);
    case QPSDM_U16B: return launch_lookup(P, qk::make_tab_b16<true>(p->tg, tab), px, py, xyz, n, out, s);
    case QPSDM_U16X: return launch_lookup(P, qk::make_tab_u16x(p->tg, tab), px, py, xyz, n, out, s);
    case QPSDM_U16Q:
      QPSDM_REQUIRE(qk::u16q_fits(p->tg), "tables_lookup: U16Q node of %zu bytes (below 2 GiB)",
                    qk::u16q_node_bytes(p->tg));
      return launch_lookup(P, qk::make_tab_u16q(p->tg, tab), px, py, xyz, n, out, s);
    default: return qpsdm::set_error(QPSDM_EINVAL, "tables_lookup: table dtype %d", table_dtype);
  }
}

extern "C" int qpsdm_count_shot(const qpsdm_image_params *p, const float *geom, int tr0, int ntr,
                                unsigned long long *count, void *stream) {
  qpsdm_launch sh = {8, 8, 8};  // count kernel shape
  qk::ImgDev P;
  QPSDM_REQUIRE(p, "count: null params");
  const qpsdm_image_params pe = qk::exact_form(p);
  int rc = qk::make_dev(&pe, &sh, tr0, ntr, &P);
  if (rc) return rc;
  QPSDM_REQUIRE(geom && count, "count: null buffer");
  if (ntr == 0) return QPSDM_OK;
  const int nxt = (P.nx + 7) / 8, nyt = (P.ny + 7) / 8;
  hipLaunchKernelGGL(count_kernel<8>, dim3(nxt * nyt, (P.nz + 7) / 8), dim3(64), 0,
                     (hipStream_t)stream, P, (const float4 *)geom, count, one_run(P));
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

extern "C" int qpsdm_image_scale(float *img, size_t n, float s, void *stream) {
  QPSDM_REQUIRE(img, "scale: null image");
  if (n == 0) return QPSDM_OK;
  hipLaunchKernelGGL(scale_kernel, dim3(1024), dim3(256), 0, (hipStream_t)stream, img, n, s);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}
