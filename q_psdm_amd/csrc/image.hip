// Kernel (b): Kirchhoff imaging sum for gfx950.
//
// Output-stationary: a one-wave work-group owns an image tile of 8 x 8
// columns x ZP depth points (one column per lane, ZP accumulators in
// registers) and streams a run of traces.  The source leg (t, t*) of every
// tile point is interpolated into registers and reused while consecutive
// traces share their source (once per shot gather).  Per
// trace and lane the receiver table column is interpolated bilinearly in
// (x, y), walked down the table depths the tile spans (two nodes live in
// registers) and lerped in z; each point then applies the dip-bound taper and
// the (tr + 1/tr) weight of the reference (src/q-psdm.cu:2303-2378), picks
// the bank level by t* and interpolates the filtered trace linearly in time.  Tiles whose whole box
// lies beyond the dip aperture of a trace skip it (a conservative bound).
// The image is float [ny][nx][nz], z fastest (the reference's img[ix][iz]
// per line, src/q-psdm.cu:2376-2378).
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "common.h"
#include "image_dev.h"

namespace qk = qpsdm_dev;

namespace {

// One wave per tile of tile_x x tile_y = 64 columns x ZP depth points; the
// depth of a point index is the same in every lane, so the table-depth walk
// and the tile skip are wave-uniform branches.
template <int ZP, typename BankT, typename Tab>
__global__ __launch_bounds__(64) void image_kernel(qk::ImgDev P, Tab tab, float *__restrict__ img,
                                                   const float4 *__restrict__ geom,
                                                   const BankT *__restrict__ bank) {
  const int lane = threadIdx.x;
  const int ntx = (P.nx + P.tile_x - 1) / P.tile_x;
  const int tx0 = (blockIdx.x % ntx) * P.tile_x;
  const int ty0 = (blockIdx.x / ntx) * P.tile_y;
  const int z0 = blockIdx.y * ZP;
  const int nzv = min(ZP, P.nz - z0);
  const int cx = tx0 + lane % P.tile_x, cy = ty0 + lane / P.tile_x;
  const bool valid = cx < P.nx && cy < P.ny;
  const float x = (float)cx * P.dx, y = (float)cy * P.dy;
  const size_t col = ((size_t)(valid ? cy : 0) * P.nx + (valid ? cx : 0)) * P.nz + z0;

  // per depth point (wave-uniform): z, z^2, table depth node and weight
  float zz[ZP], z2[ZP], fz[ZP];
  int wz[ZP];
#pragma unroll
  for (int i = 0; i < ZP; ++i) {
    zz[i] = (float)(z0 + i) * P.dz;
    z2[i] = zz[i] * zz[i];
    wz[i] = qk::depth_node(P, zz[i], &fz[i]);
  }

  float acc[ZP];
#pragma unroll
  for (int i = 0; i < ZP; ++i) acc[i] = (valid && i < nzv) ? img[col + i] : 0.f;

  // Source leg (t, t*, 1/t, 1/ds, ds) of every tile point: computed for the
  // first trace and again only when the source moves (a shot gather pays it
  // once; a common-offset gather per trace).  A trace whose geometry is NaN
  // (rejected outside the model extent) is skipped.
  float ts[ZP], tss[ZP], rts[ZP], rds[ZP], ds[ZP];
  float sx = NAN, sy = NAN, dsx = 0.f, dsy = 0.f;
  const qk::TileBox box = qk::tile_box(P, tx0, ty0, z0, ZP);
  const unsigned bank_bytes = (unsigned)(P.K * P.nt * sizeof(BankT));

  for (int t = P.tr0; t < P.tr0 + P.ntr; ++t) {
    const float4 g = geom[t];
    if (!(g.x == g.x)) continue;
    if (qk::tile_culled(P, box, g.x, g.y, g.z, g.w)) continue;
    if (g.x != sx || g.y != sy) {
      sx = g.x;
      sy = g.y;
      dsx = sx - x;
      dsy = sy - y;
      const float hs2 = dsx * dsx + dsy * dsy;
      const auto c = qk::column_setup(P, tab, sx, sy, x, y);
      int cur = wz[0];
      float2 lo = tab.at(c, cur), hi = tab.at(c, cur + 1);
#pragma unroll
      for (int i = 0; i < ZP; ++i) {
        while (cur < wz[i]) {
          ++cur;
          lo = hi;
          hi = tab.at(c, cur + 1);
        }
        ts[i] = fmaxf(lo.x + fz[i] * (hi.x - lo.x), P.tmin);
        tss[i] = lo.y + fz[i] * (hi.y - lo.y);
        rts[i] = __builtin_amdgcn_rcpf(ts[i]);
        const float d2 = hs2 + z2[i];
        rds[i] = __builtin_amdgcn_rsqf(d2);
        ds[i] = d2 * rds[i];
      }
    }
    const float dgx = g.z - x, dgy = g.w - y;
    const float hg2 = dgx * dgx + dgy * dgy;
    const auto c = qk::column_setup(P, tab, g.z, g.w, x, y);
    const __amdgpu_buffer_rsrc_t br =
        qk::make_rsrc(bank + (size_t)(t - P.tr0) * P.K * P.nt, bank_bytes);
    int cur = wz[0];
    float2 lo = tab.at(c, cur), hi = tab.at(c, cur + 1);
#pragma unroll
    for (int i = 0; i < ZP; ++i) {
      if (i >= nzv) continue;
      while (cur < wz[i]) {
        ++cur;
        lo = hi;
        hi = tab.at(c, cur + 1);
      }
      if (!valid || z0 + i == 0) continue;
      const float rdg = __builtin_amdgcn_rsqf(hg2 + z2[i]);
      float wt;
      if (!qk::dip_weight(P, dsx * rds[i] + dgx * rdg, dsy * rds[i] + dgy * rdg,
                          zz[i] * (rds[i] + rdg), &wt))
        continue;
      const float tg = fmaxf(lo.x + fz[i] * (hi.x - lo.x), P.tmin);
      const float tsg = lo.y + fz[i] * (hi.y - lo.y);
      int j = (int)((tss[i] + tsg) * P.inv_dtstar + 0.5f);
      j = j < 0 ? 0 : (j > P.K - 1 ? P.K - 1 : j);
      const float f = (ts[i] + tg) * P.inv_dt;
      const int i0 = (int)f;
      if (!(i0 > 0 && i0 + 1 < P.nt)) continue;
      const float frt = f - (float)i0;
      const float2 b = qk::bank_pair(br, j * P.nt + i0, bank);
      const float rr = ds[i] * rdg;
      const float tr = rr * rr * rr * tg * rts[i];
      acc[i] += wt * (tr + __builtin_amdgcn_rcpf(tr)) * (b.x + frt * (b.y - b.x));
    }
  }
  if (valid) {
#pragma unroll
    for (int i = 0; i < ZP; ++i)
      if (i < nzv) img[col + i] = acc[i];
  }
}

// Update count: pairs (trace, point) inside the dip aperture with z > 0,
// same geometry and the same conservative tile skip as the imaging kernel.
template <int ZP>
__global__ __launch_bounds__(64) void count_kernel(qk::ImgDev P, const float4 *__restrict__ geom,
                                                   unsigned long long *count) {
  const int lane = threadIdx.x;
  const int ntx = (P.nx + P.tile_x - 1) / P.tile_x;
  const int tx0 = (blockIdx.x % ntx) * P.tile_x;
  const int ty0 = (blockIdx.x / ntx) * P.tile_y;
  const int z0 = blockIdx.y * ZP;
  const int nzv = min(ZP, P.nz - z0);
  const int cx = tx0 + lane % P.tile_x, cy = ty0 + lane / P.tile_x;
  const bool valid = cx < P.nx && cy < P.ny;
  const float x = (float)cx * P.dx, y = (float)cy * P.dy;
  const qk::TileBox box = qk::tile_box(P, tx0, ty0, z0, ZP);
  unsigned int n = 0;
  for (int t = P.tr0; t < P.tr0 + P.ntr; ++t) {
    const float4 g = geom[t];
    if (!(g.x == g.x)) continue;
    if (qk::tile_culled(P, box, g.x, g.y, g.z, g.w)) continue;
    const float dsx = g.x - x, dsy = g.y - y, hs2 = dsx * dsx + dsy * dsy;
    const float dgx = g.z - x, dgy = g.w - y, hg2 = dgx * dgx + dgy * dgy;
#pragma unroll
    for (int i = 0; i < ZP; ++i) {
      const int iz = z0 + i;
      if (i >= nzv || !valid || iz == 0) continue;
      const float z = (float)iz * P.dz, z2 = z * z;
      const float rds = __builtin_amdgcn_rsqf(hs2 + z2), rdg = __builtin_amdgcn_rsqf(hg2 + z2);
      float w;
      n += qk::dip_weight(P, dsx * rds + dgx * rdg, dsy * rds + dgy * rdg, z * (rds + rdg), &w);
    }
  }
  unsigned long long v = n;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (lane == 0 && v) atomicAdd(count, v);
}

__global__ void scale_kernel(float *p, size_t n, float s) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] *= s;
}

template <int ZP, typename BankT, typename Tab>
int launch_t(const qk::ImgDev &P, const Tab &tab, float *img, const float4 *geom, const BankT *bank,
             hipStream_t s) {
  const int nxt = (P.nx + P.tile_x - 1) / P.tile_x, nyt = (P.ny + P.tile_y - 1) / P.tile_y;
  const dim3 grid(nxt * nyt, (P.nz + ZP - 1) / ZP);
  hipLaunchKernelGGL((image_kernel<ZP, BankT, Tab>), grid, dim3(64), 0, s, P, tab, img, geom, bank);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

template <typename BankT, typename Tab>
int launch_tz(const qk::ImgDev &P, const Tab &tab, float *img, const float4 *geom,
              const BankT *bank, hipStream_t s) {
  switch (P.tile_z) {
    case 4: return launch_t<4>(P, tab, img, geom, bank, s);
    case 6: return launch_t<6>(P, tab, img, geom, bank, s);
    case 8: return launch_t<8>(P, tab, img, geom, bank, s);
    case 12: return launch_t<12>(P, tab, img, geom, bank, s);
    case 16: return launch_t<16>(P, tab, img, geom, bank, s);
    default:
      return qpsdm::set_error(QPSDM_EINVAL, "image: unsupported tile_z=%d (4, 6, 8, 12, 16)",
                              P.tile_z);
  }
}

}  // namespace

extern "C" int qpsdm_image_shot(const qpsdm_image_params *p, const qpsdm_launch *shape, float *img,
                                const float *geom, int tr0, int ntr, const void *bank,
                                int bank_dtype, const void *tab, int table_dtype, void *stream) {
  qk::ImgDev P;
  int rc = qk::make_dev(p, shape, tr0, ntr, &P);
  if (rc) return rc;
  QPSDM_REQUIRE(img && geom && bank && tab, "image: null buffer");
  if (ntr == 0) return QPSDM_OK;
  hipStream_t s = (hipStream_t)stream;
  const float4 *g4 = (const float4 *)geom;
  if (table_dtype == QPSDM_F32) {
    qk::TabF32 T = qk::make_tab_f32(p->tg, tab);
    if (bank_dtype == QPSDM_F32)
      return launch_tz(P, T, img, g4, (const float *)bank, s);
    if (bank_dtype == QPSDM_BF16)
      return launch_tz(P, T, img, g4, (const __hip_bfloat16 *)bank, s);
  } else if (table_dtype == QPSDM_F16B) {
    auto T = qk::make_tab_b16<false>(p->tg, tab);
    if (bank_dtype == QPSDM_F32)
      return launch_tz(P, T, img, g4, (const float *)bank, s);
    if (bank_dtype == QPSDM_BF16)
      return launch_tz(P, T, img, g4, (const __hip_bfloat16 *)bank, s);
  } else if (table_dtype == QPSDM_U16B) {
    auto T = qk::make_tab_b16<true>(p->tg, tab);
    if (bank_dtype == QPSDM_F32)
      return launch_tz(P, T, img, g4, (const float *)bank, s);
    if (bank_dtype == QPSDM_BF16)
      return launch_tz(P, T, img, g4, (const __hip_bfloat16 *)bank, s);
  }
  return qpsdm::set_error(QPSDM_EINVAL, "image: bad dtypes bank=%d table=%d", bank_dtype,
                          table_dtype);
}

extern "C" int qpsdm_count_shot(const qpsdm_image_params *p, const float *geom, int tr0, int ntr,
                                unsigned long long *count, void *stream) {
  qpsdm_launch sh = {8, 8, 8, 1};  // count kernel shape
  qk::ImgDev P;
  int rc = qk::make_dev(p, &sh, tr0, ntr, &P);
  if (rc) return rc;
  QPSDM_REQUIRE(geom && count, "count: null buffer");
  if (ntr == 0) return QPSDM_OK;
  const int nxt = (P.nx + 7) / 8, nyt = (P.ny + 7) / 8;
  hipLaunchKernelGGL(count_kernel<8>, dim3(nxt * nyt, (P.nz + 7) / 8), dim3(64), 0,
                     (hipStream_t)stream, P, (const float4 *)geom, count);
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
