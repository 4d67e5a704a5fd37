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

#include <type_traits>

#include "image_dev.h"

namespace qpsdm_dev {

// A run of traces [tr0, tr0 + ntr) of geom; bank row of trace t is
// t - tr0 + brow0; dip bounds of the run.  When src is set, the run's
// source leg comes from its own table (built for the source (ssx, ssy),
// relative origin (su0, sv0), snv x snu x nw nodes at the node tables'
// spacing) instead of the nearest node's table, and only the run's traces
// whose source is (ssx, ssy), bit for bit, are imaged.
struct Run {
  int tr0, ntr, brow0;
  Dip dip;
  const float2 *src;
  float ssx, ssy, su0, sv0;
  int snu, snv;
};

// A trace of a source-table run whose source is not the table's own
// (compared bitwise, on the scalar unit).
__device__ inline bool foreign_source(const Run &run, const float4 &g) {
  return (__float_as_int(g.x) != __float_as_int(run.ssx)) | (__float_as_int(g.y) != __float_as_int(run.ssy));
}

// Source-leg cache of the tile points, per source position: t in samples,
// t* in bank levels plus the rounding half, the distance ds and ds^3 / t
// (the source part of the (ds/dg)^3 tg/ts weight) per point, and the
// column's horizontal offsets to the source.  The bisector
// (dsx/ds + dgx/dg, dsy/ds + dgy/dg, z/ds + z/dg) is used scaled by ds,
// (dsx + dgx q, dsy + dgy q, z + z q) with q = ds/dg, which leaves its
// direction (the only thing the taper reads) unchanged and needs one value
// per point instead of three.
template <int ZP>
struct SrcCache {
  float tsn[ZP], tsl[ZP], ds[ZP], A[ZP];
  float dsx, dsy;
};

// Source leg of the tile points: walk the column c of table tab down the
// depth nodes (t floored at tmin).
template <int ZP, typename Tab>
__device__ inline void source_leg(const ImgDev &P, const Tab &tab, typename Tab::Col c, const int (&wz)[ZP],
                                  const float (&fz)[ZP], const float (&zz)[ZP], const float (&z2)[ZP], float dsx,
                                  float dsy, SrcCache<ZP> &S) {
  const float hs2 = dsx * dsx + dsy * dsy;
  S.dsx = dsx;
  S.dsy = dsy;
  int cur = wz[0];
  float2 lo = tab.at(c, cur), hi = tab.at(c, cur + 1);
#pragma unroll
  for (int i = 0; i < ZP; ++i) {
    while (cur < wz[i]) {
      ++cur;
      lo = hi;
      hi = tab.at(c, cur + 1);
    }
    const float ts = fmaxf(fmaf(fz[i], hi.x - lo.x, lo.x), P.tmin);
    const float tss = fmaf(fz[i], hi.y - lo.y, lo.y);
    const float d2 = hs2 + z2[i];
    const float rds = __builtin_amdgcn_rsqf(d2);
    const float ds = d2 * rds;
    S.tsn[i] = ts * P.inv_dt;
    S.tsl[i] = fmaf(tss, P.inv_dtstar, 0.5f);
    S.ds[i] = ds;
    S.A[i] = ds * ds * ds * __builtin_amdgcn_rcpf(ts);
  }
}

__device__ inline float sat(float v) { return __builtin_amdgcn_fmed3f(v, 0.0f, 1.0f); }

// Dip-bound taper weight of the reference (src/q-psdm.cu:2322-2340) as
// clamped ramps on tan = X/Z: 1 inside [tan1, tan2], linear to 0 at the
// widened bounds, 0 outside (the reference rejects those points).
// The ramp offsets come in vector registers (Ramps): a vector ALU
// instruction reads at most one scalar operand, so with slope and offset
// both scalar every ramp would first copy one of them.
struct Ramps {
  float cx1, cx2, cy1, cy2;
};

__device__ inline Ramps vector_ramps(const Dip &D) {
  Ramps R{D.cx1, D.cx2, D.cy1, D.cy2};
  asm volatile("" : "+v"(R.cx1), "+v"(R.cx2), "+v"(R.cy1), "+v"(R.cy2));
  return R;
}

__device__ inline float dip_taper(const Dip &D, const Ramps &R, float X, float Y, float Z) {
  const float rz = __builtin_amdgcn_rcpf(Z);
  const float tx = X * rz, ty = Y * rz;
  const float wx = fminf(sat(fmaf(tx, D.rx1, R.cx1)), sat(fmaf(tx, -D.rx2, R.cx2)));
  const float wy = fminf(sat(fmaf(ty, D.ry1, R.cy1)), sat(fmaf(ty, -D.ry2, R.cy2)));
  return wx * wy;
}

// SRC: every run carries its own source table (Run::src), and traces of
// another source are skipped; without it the source leg comes from the
// node tables.  RANGE: the tile's depth groups start at ImgDev::zb and stop
// at ze (the node form's exact parts); without it they cover [0, nz).  Both
// are compile-time so that the common kernels carry neither the registers
// nor the per-trace tests (they cost the exact body 4-14 % in time).
// CULL: traces are tested 64 at a time, one per lane, against the tile box
// and the dip aperture over the tile's points (live_traces); without it one
// at a time against the tile box on the scalar unit (next_live).
template <int ZP, bool SRC, bool RANGE, bool CULL, typename BankT, typename Tab, typename Runs>
__device__ inline void image_tile(const ImgDev &P, const Tab &tab, float *__restrict__ img,
                                  const float4 *__restrict__ geom, const BankT *__restrict__ bank,
                                  int tile_col, int zgroup, const Runs &runs, int nrun) {
  const int lane = threadIdx.x;
  const int ntx = (P.nx + P.tile_x - 1) / P.tile_x;
  const int tx0 = (tile_col % ntx) * P.tile_x;
  const int ty0 = (tile_col / ntx) * P.tile_y;
  const int z0 = (RANGE ? P.zb : 0) + zgroup * ZP;
  const int nzv = min(ZP, (RANGE ? P.ze : P.nz) - z0);
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

  SrcCache<ZP> S;
  float sx = NAN, sy = NAN;
  int cache_run = -1;
  const unsigned nt = (unsigned)P.nt;
  const float tmin = __builtin_canonicalizef(P.tmin);  // lets fmaxf skip re-quieting it per point
  const float jmax = (float)(P.K - 1);

  for (int r = 0; r < nrun; ++r) {
    const Run run = runs(r);
    const Dip &D = run.dip;
    const TileBox box = tile_box(P, D, tx0, ty0, z0, ZP);
    const Ramps R = vector_ramps(D);
    const int tend = run.tr0 + run.ntr;
    // source leg of the tile points for trace geometry gs, unless cached
    auto set_source = [&](const float4 &gs) {
      // same source as the cached leg: compared bitwise, on the scalar unit
      if (__float_as_int(gs.x) == __float_as_int(sx) && __float_as_int(gs.y) == __float_as_int(sy) &&
          r == cache_run)
        return;
      cache_run = r;
      sx = gs.x;
      sy = gs.y;
      if constexpr (SRC) {
        TabF32 st;
        st.p = run.src;
        st.nstride = (size_t)run.snv * run.snu * P.nw;
        st.nu = run.snu;
        st.nw = P.nw;
        int iu, iv;
        float a, b;
        frac_index((x - sx - run.su0) * P.inv_du, run.snu, &iu, &a);
        frac_index((y - sy - run.sv0) * P.inv_dv, run.snv, &iv, &b);
        source_leg<ZP>(P, st, st.column(0, iv, iu, a, b), wz, fz, zz, z2, sx - x, sy - y, S);
      } else {
        source_leg<ZP>(P, tab, column_setup(P, tab, sx, sy, x, y), wz, fz, zz, z2, sx - x, sy - y, S);
      }
    };
    auto image_trace = [&](int t, const float4 &g, int rnode) {
      if constexpr (SRC) {
        if (foreign_source(run, g)) return;
      }
      set_source(g);
      const float dgx = g.z - x, dgy = g.w - y;
      const float hg2 = dgx * dgx + dgy * dgy;
      auto c = column_at_node(P, tab, rnode, g.z, g.w, x, y);
      float unit;  // stored sample -> filtered-trace units (FP16 banks carry a per-trace scale)
      const __amdgpu_buffer_rsrc_t br = bank_row(bank, (size_t)(t - run.tr0 + run.brow0), P.K, P.nt, &unit);
      int cur = wz[0];
      float2 lo = tab.at(c, cur), hi = tab.at(c, cur + 1);
      float2 dl = make_float2(hi.x - lo.x, hi.y - lo.y);
      // Three phases: the weights, time fractions and sample indices of all
      // depth points; then every sample gather in one branch-free block (so
      // they are in flight together); then decode and accumulate.  A lane
      // without an update keeps weight 0 and gathers at an index that is
      // either inside this trace's bank or past the buffer's end (where the
      // buffer load returns 0), with a time fraction in [0, 1): its term is 0.
      float cw[ZP], cf[ZP];
      int ci[ZP];
#pragma unroll
      for (int i = 0; i < ZP; ++i) {
        cw[i] = 0.0f;
        cf[i] = 0.0f;
        ci[i] = 0;
        if (i >= nzv) continue;
        while (cur < wz[i]) {
          ++cur;
          lo = hi;
          hi = tab.at(c, cur + 1);
          dl = make_float2(hi.x - lo.x, hi.y - lo.y);
        }
        if (z0 + i == 0) continue;
        const float rdg = __builtin_amdgcn_rsqf(hg2 + z2[i]);
        const float q = rdg * S.ds[i];
        const float w = dip_taper(D, R, fmaf(dgx, q, S.dsx), fmaf(dgy, q, S.dsy), fmaf(zz[i], q, zz[i]));
        const float tg = fmaxf(fmaf(fz[i], dl.x, lo.x), tmin);
        const float f = fmaf(tg, P.inv_dt, S.tsn[i]);
        const int i0 = (int)f;
        cf[i] = __builtin_amdgcn_fractf(f);  // f >= 1 where the point updates: f - i0
        ci[i] = i0;
        // an update: inside the taper, and i0 > 0 && i0 + 1 < nt
        if (valid && w > 0.0f && (unsigned)(i0 - 1) < nt - 2u) {
          // level: round(t*/dt*) clamped to [0, K-1] (the half is in tsl)
          const float tsg = fmaf(fz[i], dl.y, lo.y);
          const int j = (int)__builtin_amdgcn_fmed3f(fmaf(tsg, P.inv_dtstar, S.tsl[i]), 0.0f, jmax);
          ci[i] += (int)__umul24((unsigned)j, nt);
          const float tr = S.A[i] * (rdg * rdg * rdg) * tg;
          if constexpr (bank_scaled<BankT>())
            cw[i] = (w * unit) * (tr + __builtin_amdgcn_rcpf(tr));
          else
            cw[i] = w * (tr + __builtin_amdgcn_rcpf(tr));
        }
      }
      typename BankRaw<BankT>::type cb[ZP];
#pragma unroll
      for (int i = 0; i < ZP; ++i) cb[i] = bank_raw(br, ci[i], bank);
#pragma unroll
      for (int i = 0; i < ZP; ++i) acc[i] = fmaf(cw[i], bank_interp(cb[i], cf[i], bank), acc[i]);
    };
    if constexpr (CULL) {
      // the traces of each 64 that may update the tile, tested one per lane
      // against the tile box and the aperture over the tile's points
      const CullBox cbox{(float)tx0 * P.dx, (float)(min(tx0 + P.tile_x, P.nx) - 1) * P.dx, (float)ty0 * P.dy,
                         (float)(min(ty0 + P.tile_y, P.ny) - 1) * P.dy, zz[0], (float)(z0 + nzv - 1) * P.dz};
      for (int t0 = run.tr0; t0 < tend; t0 += 64) {
        int lnode;
        for (unsigned long long live = live_traces(P, geom, t0, tend, box, cbox, D, lane, &lnode); live;
             live &= live - 1) {
          const int b = (int)__builtin_ctzll(live);
          image_trace(t0 + b, geom_at(geom, t0 + b), __builtin_amdgcn_readlane(lnode, b));
        }
      }
    } else {
      float4 g;
      for (int t = next_live(geom, box, run.tr0, tend, g); t < tend; t = next_live(geom, box, t + 1, tend, g))
        image_trace(t, g, surface_node(P, g.z, g.w));
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
      if (rejected(g.x)) continue;
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
