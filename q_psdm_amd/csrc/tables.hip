// Traveltime / t* table kernels: straight-ray builder (FP64), FP16
// offset+scale brick quantisation, and the line blend of get_ttt_all.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "image_dev.h"
#include "ray_dev.h"

namespace {

__global__ void build_kernel(qpsdm_dev::ModelDev m, qpsdm_tgeom g, float2 *tab) {
  const size_t per = (size_t)g.nv * g.nu * g.nw;
  const size_t total = per * g.nnx * g.nny;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t node = i / per;
    size_t r = i - node * per;
    const int iw = (int)(r % g.nw);
    r /= g.nw;
    const int iu = (int)(r % g.nu);
    const int iv = (int)(r / g.nu);
    const double px = g.n0x + (double)(node % g.nnx) * g.ndx;
    const double py = g.n0y + (double)(node / g.nnx) * g.ndy;
    double t, ts;
    qpsdm_dev::ray_integrals(m, px, py, px + g.u0 + (double)iu * g.du, py + g.v0 + (double)iv * g.dv,
              (double)iw * g.dw, &t, &ts, nullptr);
    tab[i] = make_float2((float)t, (float)ts);
  }
}

// One work-group per brick: min/max of t and t* over the brick's valid
// cells, then value = off + scale * h with h in [0, 1] stored as FP16 or as
// a 16-bit unsigned fixed-point code.
template <bool UNORM>
__global__ __launch_bounds__(512) void quant_kernel(qpsdm_tgeom g, const float2 *tab,
                                                    unsigned char *out) {
  const int nbv = (g.nv + 7) / 8, nbu = (g.nu + 7) / 8, nbw = (g.nw + 7) / 8;
  const size_t nb_node = (size_t)nbv * nbu * nbw;
  const size_t b = blockIdx.x;
  const size_t node = b / nb_node;
  size_t r = b - node * nb_node;
  const int bw = (int)(r % nbw);
  r /= nbw;
  const int bu = (int)(r % nbu), bv = (int)(r / nbu);
  const int tid = threadIdx.x;
  const int iv = bv * 8 + (tid >> 6), iu = bu * 8 + ((tid >> 3) & 7), iw = bw * 8 + (tid & 7);
  const bool in = iv < g.nv && iu < g.nu && iw < g.nw;
  float2 v = make_float2(0.f, 0.f);
  if (in) v = tab[node * (size_t)g.nv * g.nu * g.nw + ((size_t)iv * g.nu + iu) * g.nw + iw];
  __shared__ float red[4][512];
  red[0][tid] = in ? v.x : INFINITY;
  red[1][tid] = in ? v.x : -INFINITY;
  red[2][tid] = in ? v.y : INFINITY;
  red[3][tid] = in ? v.y : -INFINITY;
  __syncthreads();
  for (int s = 256; s > 0; s >>= 1) {
    if (tid < s) {
      red[0][tid] = fminf(red[0][tid], red[0][tid + s]);
      red[1][tid] = fmaxf(red[1][tid], red[1][tid + s]);
      red[2][tid] = fminf(red[2][tid], red[2][tid + s]);
      red[3][tid] = fmaxf(red[3][tid], red[3][tid + s]);
    }
    __syncthreads();
  }
  const float t0 = red[0][0], ts = fmaxf(red[1][0] - t0, 1e-30f);
  const float s0 = red[2][0], ss = fmaxf(red[3][0] - s0, 1e-30f);
  unsigned char *o = out + b * qpsdm_dev::kBrickBytes;
  if (tid == 0) *(float4 *)o = make_float4(t0, ts, s0, ss);
  const float ht = in ? fminf(fmaxf((v.x - t0) / ts, 0.f), 1.f) : 0.f;
  const float hs = in ? fminf(fmaxf((v.y - s0) / ss, 0.f), 1.f) : 0.f;
  unsigned code;
  if (UNORM) {
    code = (unsigned)__float2uint_rn(ht * 65535.0f) | ((unsigned)__float2uint_rn(hs * 65535.0f) << 16);
  } else {
    code = __builtin_bit_cast(unsigned, __floats2half2_rn(ht, hs));
  }
  *(unsigned *)(o + 16 + tid * 4) = code;
}

__global__ void blend_kernel(const float *a, const float *b, size_t n, float c1, float c2, float *o) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    o[i] = __fadd_rn(__fmul_rn(a[i], c1), __fmul_rn(b[i], c2));  // no contraction: as the host blend
}

// One work-group per edge-sharing brick: min/max of t and t* over its
// 9 x 9 x 8 nodes (indices clamped to the table), offset = min, step =
// range / 65535, code = round((v - off) / step).  QUAD: the same header and
// codes written as the U16Q corner words of the brick's 8 x 8 x 8 cells.
template <bool QUAD>
__global__ __launch_bounds__(256) void quant_x_kernel(qpsdm_tgeom g, int nbv, int nbu, int nbw, const float2 *tab,
                                                      unsigned char *out) {
  constexpr int N = qpsdm_dev::kXCodes;
  const size_t nb_node = (size_t)nbv * nbu * nbw;
  const size_t b = blockIdx.x;
  const size_t node = b / nb_node;
  size_t r = b - node * nb_node;
  const int bw = (int)(r % nbw);
  r /= nbw;
  const int bu = (int)(r % nbu), bv = (int)(r / nbu);
  const float2 *T = tab + node * (size_t)g.nv * g.nu * g.nw;
  auto at = [&](int k) {
    const int lw = k & 7, lu = (k >> 3) % 9, lv = (k >> 3) / 9;
    const int v = min(8 * bv + lv, g.nv - 1), u = min(8 * bu + lu, g.nu - 1), w = min(8 * bw + lw, g.nw - 1);
    return T[((size_t)v * g.nu + u) * g.nw + w];
  };
  float mn0 = INFINITY, mx0 = -INFINITY, mn1 = INFINITY, mx1 = -INFINITY;
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    const float2 v = at(k);
    mn0 = fminf(mn0, v.x), mx0 = fmaxf(mx0, v.x), mn1 = fminf(mn1, v.y), mx1 = fmaxf(mx1, v.y);
  }
  __shared__ float red[4][256];
  red[0][threadIdx.x] = mn0, red[1][threadIdx.x] = mx0, red[2][threadIdx.x] = mn1, red[3][threadIdx.x] = mx1;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      red[0][threadIdx.x] = fminf(red[0][threadIdx.x], red[0][threadIdx.x + st]);
      red[1][threadIdx.x] = fmaxf(red[1][threadIdx.x], red[1][threadIdx.x + st]);
      red[2][threadIdx.x] = fminf(red[2][threadIdx.x], red[2][threadIdx.x + st]);
      red[3][threadIdx.x] = fmaxf(red[3][threadIdx.x], red[3][threadIdx.x + st]);
    }
    __syncthreads();
  }
  const float t0 = red[0][0], s0 = red[2][0];
  const float tst = fmaxf(red[1][0] - t0, 1e-30f) / 65535.0f, sst = fmaxf(red[3][0] - s0, 1e-30f) / 65535.0f;
  auto code = [&](int k) {
    const float2 v = at(k);
    const unsigned ct = min(__float2uint_rn(fmaxf((v.x - t0) / tst, 0.0f)), 65535u);
    const unsigned cs = min(__float2uint_rn(fmaxf((v.y - s0) / sst, 0.0f)), 65535u);
    return ct | (cs << 16);
  };
  if (!QUAD) {
    unsigned char *o = out + b * qpsdm_dev::kXBrickBytes;
    if (threadIdx.x == 0) *(float4 *)o = make_float4(t0, tst, s0, sst);
    for (int k = threadIdx.x; k < N; k += blockDim.x) *(unsigned *)(o + 16 + k * 4) = code(k);
  } else {
    // the steps divided by 65535, the sum of a cell's fixed-point bilinear
    // weights (a reader decodes a weighted code sum with one fma)
    unsigned char *o = out + b * qpsdm_dev::kQBrickBytes;
    if (threadIdx.x == 0) *(float4 *)o = make_float4(t0, tst * (1.0f / 65535.0f), s0, sst * (1.0f / 65535.0f));
    // cell (lv, lu) at depth lw: corners k = ((lv [+1]) * 9 + lu [+1]) * 8 + lw of the 9 x 9 x 8 nodes;
    // the word pairs each row's two t codes and two t* codes:
    // {t00 | t01 << 16, t10 | t11 << 16, s00 | s01 << 16, s10 | s11 << 16}
    for (int c = threadIdx.x; c < 8 * 8 * 8; c += blockDim.x) {
      const int lw = c & 7, lu = (c >> 3) & 7, lv = c >> 6;
      const int k00 = (lv * 9 + lu) * 8 + lw;
      const unsigned c00 = code(k00), c01 = code(k00 + 8), c10 = code(k00 + 72), c11 = code(k00 + 80);
      *(uint4 *)(o + 16 + c * 16) =
          make_uint4((c00 & 0xffffu) | (c01 << 16), (c10 & 0xffffu) | (c11 << 16), (c00 >> 16) | (c01 & 0xffff0000u),
                     (c10 >> 16) | (c11 & 0xffff0000u));
    }
  }
}

__global__ void copy_kernel(const float4 *a, float4 *b, size_t n4) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
    b[i] = a[i];
}

int check_geom(const qpsdm_tgeom *g) {
  QPSDM_REQUIRE(g && g->nnx >= 1 && g->nny >= 1 && g->nu >= 2 && g->nv >= 2 && g->nw >= 2,
                "tables: bad geometry");
  return QPSDM_OK;
}

}  // namespace

extern "C" int qpsdm_tables_build(const qpsdm_model *m, const qpsdm_tgeom *g, float *tab,
                                  void *stream) {
  int rc = check_geom(g);
  if (rc) return rc;
  QPSDM_REQUIRE(m && tab, "tables_build: null args");
  QPSDM_REQUIRE(qpsdm_dev::model_ok(m), "tables_build: bad model");
  const qpsdm_dev::ModelDev d = qpsdm_dev::make_model(*m);
  hipLaunchKernelGGL(build_kernel, dim3(4096), dim3(256), 0, (hipStream_t)stream, d, *g,
                     (float2 *)tab);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

extern "C" size_t qpsdm_tables_f16b_bytes(const qpsdm_tgeom *g) {
  if (check_geom(g)) return 0;
  return qpsdm_dev::b16_node_bytes(*g) * g->nnx * g->nny;
}

extern "C" size_t qpsdm_tables_bytes(const qpsdm_tgeom *g, int dtype) {
  if (check_geom(g)) return 0;
  switch (dtype) {
    case QPSDM_F32: return (size_t)g->nv * g->nu * g->nw * 8 * g->nnx * g->nny;
    case QPSDM_F16B:
    case QPSDM_U16B: return qpsdm_dev::b16_node_bytes(*g) * g->nnx * g->nny;
    case QPSDM_U16X: return qpsdm_dev::u16x_node_bytes(*g) * g->nnx * g->nny;
    case QPSDM_U16Q: return qpsdm_dev::u16q_node_bytes(*g) * g->nnx * g->nny;
    default: return 0;
  }
}

extern "C" int qpsdm_tables_quantize(const qpsdm_tgeom *g, const float *tab, void *out, int code,
                                     void *stream) {
  int rc = check_geom(g);
  if (rc) return rc;
  QPSDM_REQUIRE(tab && out, "quantize: null args");
  const size_t nb = (size_t)((g->nv + 7) / 8) * ((g->nu + 7) / 8) * ((g->nw + 7) / 8) * g->nnx *
                    g->nny;
  QPSDM_REQUIRE(nb < (1ull << 31), "quantize: too many bricks for one launch");
  QPSDM_REQUIRE(code == QPSDM_F16B || code == QPSDM_U16B || code == QPSDM_U16X || code == QPSDM_U16Q,
                "quantize: code must be F16B, U16B, U16X or U16Q");
  if (code == QPSDM_U16X || code == QPSDM_U16Q) {
    int nbv, nbu, nbw;
    qpsdm_dev::u16x_bricks(*g, &nbv, &nbu, &nbw);
    const size_t nx = (size_t)nbv * nbu * nbw * g->nnx * g->nny;
    QPSDM_REQUIRE(nx < (1ull << 31), "quantize: too many bricks for one launch");
    if (code == QPSDM_U16X)
      hipLaunchKernelGGL(quant_x_kernel<false>, dim3((unsigned)nx), dim3(256), 0, (hipStream_t)stream, *g, nbv,
                         nbu, nbw, (const float2 *)tab, (unsigned char *)out);
    else
      hipLaunchKernelGGL(quant_x_kernel<true>, dim3((unsigned)nx), dim3(256), 0, (hipStream_t)stream, *g, nbv,
                         nbu, nbw, (const float2 *)tab, (unsigned char *)out);
  } else if (code == QPSDM_U16B)
    hipLaunchKernelGGL(quant_kernel<true>, dim3((unsigned)nb), dim3(512), 0, (hipStream_t)stream, *g,
                       (const float2 *)tab, (unsigned char *)out);
  else
    hipLaunchKernelGGL(quant_kernel<false>, dim3((unsigned)nb), dim3(512), 0, (hipStream_t)stream, *g,
                       (const float2 *)tab, (unsigned char *)out);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

extern "C" int qpsdm_tables_blend(const float *a, const float *b, size_t n, float c, float *o,
                                  void *stream) {
  QPSDM_REQUIRE(a && b && o, "blend: null args");
  if (n == 0) return QPSDM_OK;
  // same float weights as get_ttt_all: c2 = c, c1 = 1 - c2
  const float c2 = c, c1 = 1.0f - c2;
  hipLaunchKernelGGL(blend_kernel, dim3(2048), dim3(256), 0, (hipStream_t)stream, a, b, n, c1, c2, o);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

extern "C" int qpsdm_copy_f32(const float *src, float *dst, size_t n, void *stream) {
  QPSDM_REQUIRE(src && dst, "copy: null args");
  QPSDM_REQUIRE(n % 4 == 0 && ((uintptr_t)src & 15) == 0 && ((uintptr_t)dst & 15) == 0,
                "copy: n must be a multiple of 4 and both buffers 16-byte aligned");
  if (n == 0) return QPSDM_OK;
  hipLaunchKernelGGL(copy_kernel, dim3(2048), dim3(256), 0, (hipStream_t)stream, (const float4 *)src,
                     (float4 *)dst, n / 4);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

// ---- reference table layout (Lxxx.tidx|qidx + L<n>.ttt|qqq) -----------------
namespace {

struct RefGeo {
  qpsdm_tgeom g;
  qpsdm_ref_layout L;
  int nxg, nzg;
  long long per_cell;  // nny nnx nxi0 nzi0 floats
  long long per_line;  // nxg nzg per_cell
};

int ref_geo(const qpsdm_tgeom *g, const qpsdm_ref_layout *L, RefGeo *o) {
  QPSDM_REQUIRE(g && L, "tables_ref: null args");
  QPSDM_REQUIRE(L->nlines >= 1 && L->ncol >= 1 && L->nz >= 1 && L->nxi0 >= 1 && L->nzi0 >= 1,
                "tables_ref: empty layout");
  QPSDM_REQUIRE(L->dx > 0 && L->dy > 0 && L->dz > 0, "tables_ref: bad spacing");
  QPSDM_REQUIRE(g->ndx >= 1 && g->ndx <= 65535 && g->ndy >= 1 && g->ndy <= 65535 && g->ndx == rintf(g->ndx) &&
                    g->ndy == rintf(g->ndy),
                "tables_ref: node spacing must be whole metres below 65536 (16-bit source-grid steps)");
  QPSDM_REQUIRE(g->nnx <= 65535 && g->nny <= 65535, "tables_ref: too many nodes for 16-bit source grids");
  // the layout's lattice is the node tables' own: every payload value is a
  // table entry, not an interpolation
  QPSDM_REQUIRE(L->dx == g->du && L->dy == g->dv && L->dz == g->dw,
                "tables_ref: layout spacing %g x %g x %g is not the tables' %g x %g x %g", L->dx, L->dy, L->dz, g->du,
                g->dv, g->dw);
  const double fu = ((double)L->x0 - g->n0x - g->u0) / g->du, fv = ((double)L->y0 - g->n0y - g->v0) / g->dv;
  QPSDM_REQUIRE(fabs(fu - rint(fu)) < 1e-3 && fabs(fv - rint(fv)) < 1e-3,
                "tables_ref: layout origin (%g, %g) is not on the tables' relative lattice", L->x0, L->y0);
  o->g = *g;
  o->L = *L;
  o->nxg = (L->ncol + L->nxi0 - 1) / L->nxi0;
  o->nzg = (L->nz + L->nzi0 - 1) / L->nzi0;
  o->per_cell = (long long)g->nny * g->nnx * L->nxi0 * L->nzi0;
  o->per_line = (long long)o->nxg * o->nzg * o->per_cell;
  return QPSDM_OK;
}

// Payload of line l: element o -> node table entry (component comp); a
// column, line or depth outside the node table's relative range is clamped.
__global__ void ref_pack_kernel(RefGeo R, const float2 *__restrict__ tab, int l, int comp, float *out) {
  const qpsdm_tgeom &g = R.g;
  const qpsdm_ref_layout &L = R.L;
  for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < R.per_line;
       o += (long long)gridDim.x * blockDim.x) {
    const long long cell = o / R.per_cell;
    long long r = o - cell * R.per_cell;
    const int izc = (int)(r % L.nzi0);
    r /= L.nzi0;
    const int ixc = (int)(r % L.nxi0);
    r /= L.nxi0;
    const int ixs = (int)(r % g.nnx), iys = (int)(r / g.nnx);
    const int ixg = (int)(cell / R.nzg), izg = (int)(cell % R.nzg);
    const int c = min(ixg * L.nxi0 + ixc, L.ncol - 1);
    const int k = min(izg * L.nzi0 + izc, L.nz - 1);
    const float x = L.x0 + (float)c * L.dx, y = L.y0 + (float)l * L.dy, z = (float)k * L.dz;
    const float px = g.n0x + (float)ixs * g.ndx, py = g.n0y + (float)iys * g.ndy;
    const int iu = min(max((int)rintf((x - px - g.u0) / g.du), 0), g.nu - 1);
    const int iv = min(max((int)rintf((y - py - g.v0) / g.dv), 0), g.nv - 1);
    const int iw = min(max((int)rintf(z / g.dw), 0), g.nw - 1);
    const float2 v = tab[(((long long)(iys * g.nnx + ixs) * g.nv + iv) * g.nu + iu) * g.nw + iw];
    out[o] = comp ? v.y : v.x;
  }
}

// Node table entries (component comp) from the lines' payloads (all lines
// [nlines][per_line] on the device), each at its point's nearest covered
// column, line and depth node.
__global__ void ref_unpack_kernel(RefGeo R, const float *__restrict__ lines, int comp, float2 *tab) {
  const qpsdm_tgeom &g = R.g;
  const qpsdm_ref_layout &L = R.L;
  const long long per = (long long)g.nv * g.nu * g.nw;
  const long long total = per * g.nnx * g.nny;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const int node = (int)(i / per);
    long long r = i - node * per;
    const int iw = (int)(r % g.nw);
    r /= g.nw;
    const int iu = (int)(r % g.nu), iv = (int)(r / g.nu);
    const int ixs = node % g.nnx, iys = node / g.nnx;
    const float px = g.n0x + (float)ixs * g.ndx, py = g.n0y + (float)iys * g.ndy;
    const float x = px + g.u0 + (float)iu * g.du, y = py + g.v0 + (float)iv * g.dv, z = (float)iw * g.dw;
    const int c = min(max((int)rintf((x - L.x0) / L.dx), 0), L.ncol - 1);
    const int l = min(max((int)rintf((y - L.y0) / L.dy), 0), L.nlines - 1);
    const int k = min(max((int)rintf(z / L.dz), 0), L.nz - 1);
    const int ixg = c / L.nxi0, ixc = c % L.nxi0, izg = k / L.nzi0, izc = k % L.nzi0;
    const long long o = (long long)l * R.per_line + (long long)(ixg * R.nzg + izg) * R.per_cell +
                        (((long long)iys * g.nnx + ixs) * L.nxi0 + ixc) * L.nzi0 + izc;
    float2 v = tab[i];
    if (comp)
      v.y = lines[o];
    else
      v.x = lines[o];
    tab[i] = v;
  }
}

const char *kIdxSuffix[2] = {"tidx", "qidx"};
const char *kDatSuffix[2] = {"ttt", "qqq"};

}  // namespace

extern "C" int qpsdm_tables_write_ref(const qpsdm_tgeom *g, const float *tab, const qpsdm_ref_layout *L,
                                      const char *dir) {
  RefGeo R;
  int rc = ref_geo(g, L, &R);
  if (rc) return rc;
  QPSDM_REQUIRE(tab && dir, "tables_write_ref: null args");
  qpsdm_tidx_header h;
  h.cdp1 = L->cdp1;
  h.cdp2 = L->cdp1 + L->ncol - 1;
  h.dx = L->dx;
  h.dz = L->dz;
  h.nzg = R.nzg;
  h.nxg = R.nxg;
  h.nzi0 = L->nzi0;
  h.nxi0 = L->nxi0;
  std::vector<qpsdm_ttt_desc> d((size_t)R.nxg * R.nzg);
  for (int ixg = 0; ixg < R.nxg; ++ixg)
    for (int izg = 0; izg < R.nzg; ++izg) {
      qpsdm_ttt_desc &e = d[(size_t)ixg * R.nzg + izg];
      e.idx = (int64_t)(ixg * R.nzg + izg) * R.per_cell;
      e.nys = (uint16_t)g->nny;
      e.nxs = (uint16_t)g->nnx;
      e.dys = (uint16_t)g->ndy;
      e.dxs = (uint16_t)g->ndx;
    }
  QPSDM_HIP(hipDeviceSynchronize());  // the tables' producers (any stream) are done
  float *dev = nullptr;
  QPSDM_HIP(hipMalloc(&dev, sizeof(float) * R.per_line));
  std::vector<float> host(R.per_line);
  char path[4096];
  for (int comp = 0; comp < 2 && !rc; ++comp) {
    snprintf(path, sizeof(path), "%s/Lxxx.%s", dir, kIdxSuffix[comp]);
    rc = qpsdm_write_tidx(path, &h, d.data());
    for (int l = 0; l < L->nlines && !rc; ++l) {
      const int grid = (int)std::min<long long>((R.per_line + 255) / 256, 65536);
      hipLaunchKernelGGL(ref_pack_kernel, dim3(grid), dim3(256), 0, 0, R, (const float2 *)tab, l, comp, dev);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpy(host.data(), dev, sizeof(float) * R.per_line, hipMemcpyDeviceToHost) != hipSuccess) {
        rc = qpsdm::set_error(QPSDM_EHIP, "tables_write_ref: copy of line %d failed", L->line1 + l);
        break;
      }
      snprintf(path, sizeof(path), "%s/L%d.%s", dir, L->line1 + l, kDatSuffix[comp]);
      FILE *fp = fopen(path, "wb");
      if (!fp) {
        rc = qpsdm::set_error(QPSDM_EIO, "cannot create %s", path);
        break;
      }
      const bool ok = fwrite(host.data(), sizeof(float), R.per_line, fp) == (size_t)R.per_line;
      fclose(fp);
      if (!ok) rc = qpsdm::set_error(QPSDM_EIO, "short write %s", path);
    }
  }
  (void)hipFree(dev);
  return rc;
}

extern "C" int qpsdm_tables_read_ref(const qpsdm_tgeom *g, const qpsdm_ref_layout *L, const char *dir, float *tab) {
  RefGeo R;
  int rc = ref_geo(g, L, &R);
  if (rc) return rc;
  QPSDM_REQUIRE(tab && dir, "tables_read_ref: null args");
  char path[4096];
  std::vector<qpsdm_ttt_desc> d((size_t)R.nxg * R.nzg);
  // both index files must describe this layout with the node grid as every cell's source grid
  for (int comp = 0; comp < 2; ++comp) {
    snprintf(path, sizeof(path), "%s/Lxxx.%s", dir, kIdxSuffix[comp]);
    qpsdm_tidx_header h;
    rc = qpsdm_read_tidx(path, &h, nullptr, 0);
    if (rc) return rc;
    if (h.cdp1 != L->cdp1 || h.cdp2 != L->cdp1 + L->ncol - 1 || h.nxi0 != L->nxi0 || h.nzi0 != L->nzi0 ||
        h.nxg != R.nxg || h.nzg != R.nzg || h.dx != L->dx || h.dz != L->dz)
      return qpsdm::set_error(QPSDM_EINVAL, "%s: header (cdp %d-%d, cells %d x %d of %d x %d) is not the layout's",
                              path, h.cdp1, h.cdp2, h.nxg, h.nzg, h.nxi0, h.nzi0);
    rc = qpsdm_read_tidx(path, &h, d.data(), (int)d.size());
    if (rc) return rc;
    for (size_t i = 0; i < d.size(); ++i)
      if (d[i].nys != g->nny || d[i].nxs != g->nnx || d[i].dys != (uint16_t)g->ndy || d[i].dxs != (uint16_t)g->ndx ||
          d[i].idx != (int64_t)i * R.per_cell)
        return qpsdm::set_error(QPSDM_EINVAL,
                                "%s: cell %zu's source grid %u x %u at %u x %u m (offset %lld) is not the surface-node "
                                "grid %d x %d at %g x %g m",
                                path, i, d[i].nys, d[i].nxs, d[i].dys, d[i].dxs, (long long)d[i].idx, g->nny, g->nnx,
                                g->ndy, g->ndx);
  }
  const size_t nfl = (size_t)L->nlines * R.per_line;
  QPSDM_HIP(hipDeviceSynchronize());  // no earlier work on any stream still uses tab
  float *dev = nullptr;
  QPSDM_HIP(hipMalloc(&dev, sizeof(float) * nfl));
  std::vector<float> host(R.per_line);
  for (int comp = 0; comp < 2 && !rc; ++comp) {
    for (int l = 0; l < L->nlines && !rc; ++l) {
      const long n = qpsdm_load_line_table(dir, kDatSuffix[comp], L->line1 + l, L->line1,
                                           L->line1 + L->nlines - 1, host.data(), host.size(), nullptr, nullptr);
      if (n < 0) {
        rc = (int)n;
        break;
      }
      if (n != R.per_line) {
        rc = qpsdm::set_error(QPSDM_EIO, "L%d.%s holds %ld floats, the layout %lld", L->line1 + l, kDatSuffix[comp],
                              n, R.per_line);
        break;
      }
      if (hipMemcpy(dev + (size_t)l * R.per_line, host.data(), sizeof(float) * R.per_line, hipMemcpyHostToDevice) !=
          hipSuccess)
        rc = qpsdm::set_error(QPSDM_EHIP, "tables_read_ref: upload of line %d failed", L->line1 + l);
    }
    if (rc) break;
    const long long total = (long long)g->nv * g->nu * g->nw * g->nnx * g->nny;
    const int grid = (int)std::min<long long>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(ref_unpack_kernel, dim3(grid), dim3(256), 0, 0, R, dev, comp, (float2 *)tab);
    if (hipDeviceSynchronize() != hipSuccess) rc = qpsdm::set_error(QPSDM_EHIP, "tables_read_ref: unpack failed");
  }
  (void)hipFree(dev);
  return rc;
}

// ---- ray integrals at arbitrary points --------------------------------------
namespace {

__global__ void rays_kernel(qpsdm_dev::ModelDev m, const double *__restrict__ pts, int n, double *__restrict__ out) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double *p = pts + 5 * (size_t)i;
    double t, ts, sg;
    qpsdm_dev::ray_integrals(m, p[0], p[1], p[2], p[3], p[4], &t, &ts, &sg);
    out[3 * (size_t)i] = t;
    out[3 * (size_t)i + 1] = ts;
    out[3 * (size_t)i + 2] = sg;
  }
}

}  // namespace

extern "C" int qpsdm_rays(const qpsdm_model *m, const double *pts, int n, double *out, void *stream) {
  QPSDM_REQUIRE(m && (pts || n == 0) && (out || n == 0) && n >= 0, "rays: bad args");
  if (n == 0) return QPSDM_OK;
  hipLaunchKernelGGL(rays_kernel, dim3(std::min((n + 255) / 256, 65536)), dim3(256), 0, (hipStream_t)stream,
                     qpsdm_dev::make_model(*m), pts, n, out);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}
