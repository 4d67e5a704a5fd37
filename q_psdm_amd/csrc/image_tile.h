// The imaging work of one image tile: one wave owns tile_x x tile_y = 64
// columns x ZP depth points (one column per lane, ZP accumulators in
// registers) and streams runs of traces into it.
//
// The source leg (t, t*, 1/t, 1/ds, ds) of every tile point is interpolated
// into registers and reused while consecutive traces share their source
// (a shot gather pays it once, a common-offset gather per trace).  Per trace
// and lane the receiver table column is interpolated bilinearly in (x, y),
// walked down the table depths the tile spans (two nodes in registers) and
// lerped in z; each point applies the run's dip-bound taper and the
// (tr + 1/tr) weight of the reference (src/q-psdm.cu:2303-2378), picks the
// bank level by t* and interpolates the filtered trace linearly in time.  A
// run whose dip aperture misses the whole tile box skips the trace.  Depth
// quantities are wave-uniform, so the depth walk and the skips are scalar
// branches.  The image is float [ny][nx][nz], z fastest (the reference's
// img[ix][iz] per line, src/q-psdm.cu:2376-2378).
#ifndef QPSDM_IMAGE_TILE_H_
#define QPSDM_IMAGE_TILE_H_

#include "image_dev.h"

namespace qpsdm_dev {

// A run of traces [tr0, tr0 + ntr) of geom; bank row of trace t is
// t - tr0 + brow0; dip bounds of the run.  When src is set, the run's
// source leg comes from its own table (one node at the source, relative
// origin (su0, sv0), snv x snu x nw nodes at the node tables' spacing)
// instead of the nearest node's table.
struct Run {
  int tr0, ntr, brow0;
  Dip dip;
  const float2 *src;
  float su0, sv0;
  int snu, snv;
};

// Source leg of the tile points: walk the column c of table tab down the
// depth nodes; t (floored at tmin), t*, 1/t, 1/ds, ds per point.
template <int ZP, typename Tab>
__device__ inline void source_leg(const ImgDev &P, const Tab &tab, const typename Tab::Col &c, const int (&wz)[ZP],
                                  const float (&fz)[ZP], const float (&z2)[ZP], float hs2, float (&ts)[ZP],
                                  float (&tss)[ZP], float (&rts)[ZP], float (&rds)[ZP], float (&ds)[ZP]) {
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

template <int ZP, typename BankT, typename Tab, typename Runs>
__device__ inline void image_tile(const ImgDev &P, const Tab &tab, float *__restrict__ img,
                                  const float4 *__restrict__ geom, const BankT *__restrict__ bank,
                                  int tile_col, int zgroup, const Runs &runs, int nrun) {
  const int lane = threadIdx.x;
  const int ntx = (P.nx + P.tile_x - 1) / P.tile_x;
  const int tx0 = (tile_col % ntx) * P.tile_x;
  const int ty0 = (tile_col / ntx) * P.tile_y;
  const int z0 = zgroup * ZP;
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
    wz[i] = depth_node(P, zz[i], &fz[i]);
  }

  float acc[ZP];
#pragma unroll
  for (int i = 0; i < ZP; ++i) acc[i] = (valid && i < nzv) ? img[col + i] : 0.f;

  float ts[ZP], tss[ZP], rts[ZP], rds[ZP], ds[ZP];
  float sx = NAN, sy = NAN, dsx = 0.f, dsy = 0.f;
  int cache_run = -1;
  const unsigned bank_bytes = (unsigned)(P.K * P.nt * sizeof(BankT));

  for (int r = 0; r < nrun; ++r) {
    const Run run = runs(r);
    const Dip &D = run.dip;
    const TileBox box = tile_box(P, D, tx0, ty0, z0, ZP);
    for (int t = run.tr0; t < run.tr0 + run.ntr; ++t) {
      const float4 g = geom[t];
      if (!(g.x == g.x)) continue;  // trace rejected upstream (NaN geometry)
      if (tile_culled(box, g.x, g.y, g.z, g.w)) continue;
      if (g.x != sx || g.y != sy || r != cache_run) {
        cache_run = r;
        sx = g.x;
        sy = g.y;
        dsx = sx - x;
        dsy = sy - y;
        const float hs2 = dsx * dsx + dsy * dsy;
        if (run.src) {
          TabF32 st;
          st.p = run.src;
          st.nstride = (size_t)run.snv * run.snu * P.nw;
          st.nu = run.snu;
          st.nw = P.nw;
          int iu, iv;
          float a, b;
          frac_index((x - sx - run.su0) * P.inv_du, run.snu, &iu, &a);
          frac_index((y - sy - run.sv0) * P.inv_dv, run.snv, &iv, &b);
          source_leg<ZP>(P, st, st.column(0, iv, iu, a, b), wz, fz, z2, hs2, ts, tss, rts, rds, ds);
        } else {
          source_leg<ZP>(P, tab, column_setup(P, tab, sx, sy, x, y), wz, fz, z2, hs2, ts, tss, rts, rds, ds);
        }
      }
      const float dgx = g.z - x, dgy = g.w - y;
      const float hg2 = dgx * dgx + dgy * dgy;
      const auto c = column_setup(P, tab, g.z, g.w, x, y);
      const __amdgpu_buffer_rsrc_t br =
          make_rsrc(bank + (size_t)(t - run.tr0 + run.brow0) * P.K * P.nt, bank_bytes);
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
        if (!dip_weight(D, dsx * rds[i] + dgx * rdg, dsy * rds[i] + dgy * rdg, zz[i] * (rds[i] + rdg),
                        &wt))
          continue;
        const float tg = fmaxf(lo.x + fz[i] * (hi.x - lo.x), P.tmin);
        const float tsg = lo.y + fz[i] * (hi.y - lo.y);
        int j = (int)((tss[i] + tsg) * P.inv_dtstar + 0.5f);
        j = j < 0 ? 0 : (j > P.K - 1 ? P.K - 1 : j);
        const float f = (ts[i] + tg) * P.inv_dt;
        const int i0 = (int)f;
        if (!(i0 > 0 && i0 + 1 < P.nt)) continue;
        const float frt = f - (float)i0;
        const float2 b = bank_pair(br, j * P.nt + i0, bank);
        const float rr = ds[i] * rdg;
        const float tr = rr * rr * rr * tg * rts[i];
        acc[i] += wt * (tr + __builtin_amdgcn_rcpf(tr)) * (b.x + frt * (b.y - b.x));
      }
    }
  }
  if (valid) {
#pragma unroll
    for (int i = 0; i < ZP; ++i)
      if (i < nzv) img[col + i] = acc[i];
  }
}

// Update count of one tile: pairs (trace, point) with z > 0 inside the run's
// dip aperture, same geometry and tile skip as image_tile.
template <int ZP, typename Runs>
__device__ inline unsigned count_tile(const ImgDev &P, const float4 *__restrict__ geom, int tile_col,
                                      int zgroup, const Runs &runs, int nrun) {
  const int lane = threadIdx.x;
  const int ntx = (P.nx + P.tile_x - 1) / P.tile_x;
  const int tx0 = (tile_col % ntx) * P.tile_x;
  const int ty0 = (tile_col / ntx) * P.tile_y;
  const int z0 = zgroup * ZP;
  const int nzv = min(ZP, P.nz - z0);
  const int cx = tx0 + lane % P.tile_x, cy = ty0 + lane / P.tile_x;
  const bool valid = cx < P.nx && cy < P.ny;
  const float x = (float)cx * P.dx, y = (float)cy * P.dy;
  unsigned n = 0;
  for (int r = 0; r < nrun; ++r) {
    const Run run = runs(r);
    const TileBox box = tile_box(P, run.dip, tx0, ty0, z0, ZP);
    for (int t = run.tr0; t < run.tr0 + run.ntr; ++t) {
      const float4 g = geom[t];
      if (!(g.x == g.x)) continue;
      if (tile_culled(box, g.x, g.y, g.z, g.w)) continue;
      const float dsx = g.x - x, dsy = g.y - y, hs2 = dsx * dsx + dsy * dsy;
      const float dgx = g.z - x, dgy = g.w - y, hg2 = dgx * dgx + dgy * dgy;
#pragma unroll
      for (int i = 0; i < ZP; ++i) {
        const int iz = z0 + i;
        if (i >= nzv || !valid || iz == 0) continue;
        const float z = (float)iz * P.dz, z2 = z * z;
        const float rds = __builtin_amdgcn_rsqf(hs2 + z2), rdg = __builtin_amdgcn_rsqf(hg2 + z2);
        float w;
        n += dip_weight(run.dip, dsx * rds + dgx * rdg, dsy * rds + dgy * rdg, z * (rds + rdg), &w);
      }
    }
  }
  return n;
}

}  // namespace qpsdm_dev

#endif
