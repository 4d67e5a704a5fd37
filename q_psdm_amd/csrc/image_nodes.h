// Node form of the imaging sum: the tile body that evaluates the dip taper
// ramps and the amplitude weight at the tables' depth nodes and interpolates
// them linearly to the image points between.
//
// Along an image column both legs' (t, t*) are read from tables that are
// interpolated linearly in z between the same depth nodes (every dw metres),
// so the total time in samples F = (ts + tg)/dt and the bank level
// L = (t*s + t*g)/dt* + 1/2 are linear in z between two nodes: a point's
// values are the linear interpolation of the nodes' values, as in the
// exact form.  The dip taper (src/q-psdm.cu:2322-2340) is four ramps
// u = tan * r + c on the bisector's tangents, and the amplitude weight
// w * (tr + 1/tr), tr = (ds/dg)^3 tg/ts (src/q-psdm.cu:2372-2374), is
// smooth in z; the node form evaluates the four ramps and
// a = unit * (tr + 1/tr) at the nodes, interpolates them linearly, and
// clamps the ramps to [0, 1] per point, so the taper's corners stay at the
// points where the interpolated ramps cross 0 and 1.  The transcendental
// work (1/dg, 1/Z, 1/tr) and the receiver-table read happen once per node
// (every dw / dz = 2.5 image points at 25 m tables and a 10 m image)
// instead of once per point.
//
// Tiles are node-aligned: a tile owns NI consecutive node intervals of a
// column group (one column per lane) and every image point whose interval
// is one of them, at most PM points per interval (a register slot each).
// Points above the first node interval of the node range (the z_exact
// depth) and beyond the tables' last node are imaged by the exact body
// (image_tile); the node range's geometry is fixed by the host (ImgDev
// nf_*) with the same float arithmetic as the device, so the two bodies
// cover every point once.
#ifndef QPSDM_IMAGE_NODES_H_
#define QPSDM_IMAGE_NODES_H_

#include "image_tile.h"

namespace qpsdm_dev {

// Interval (upper depth node) of image point iz and its fraction in it, the
// rule depth_node applies: fw = min(z / dw, nw - 1), w = min(floor(fw), nw - 2).
__host__ __device__ inline int point_interval(int iz, float dz, float inv_dw, int nw, float *frac) {
  const float z = (float)iz * dz;
  float fw = z * inv_dw;
  if (fw > (float)(nw - 1)) fw = (float)(nw - 1);
  int w = (int)fw;
  if (w > nw - 2) w = nw - 2;
  *frac = fw - (float)w;
  return w;
}

// Smallest point index in [0, nz] whose interval is >= W (nz if none).
__host__ __device__ inline int first_point(int W, int nz, float dz, float inv_dw, int nw) {
  float f;
  int iz = (int)((float)W / (dz * inv_dw));
  iz = iz < 0 ? 0 : (iz > nz ? nz : iz);
  while (iz < nz && point_interval(iz, dz, inv_dw, nw, &f) < W) ++iz;
  while (iz > 0 && point_interval(iz - 1, dz, inv_dw, nw, &f) >= W) --iz;
  return iz;
}

// The ramps and amplitude of the node form at one node, and the time and
// level there.
struct NodeVal {
  float F, L, u1, u2, u3, u4, a;
};

__device__ inline NodeVal node_sub(const NodeVal &b, const NodeVal &a) {
  return NodeVal{b.F - a.F, b.L - a.L, b.u1 - a.u1, b.u2 - a.u2, b.u3 - a.u3, b.u4 - a.u4, b.a - a.a};
}

// Source leg of the tile's nodes: t in samples, t* in levels plus the
// rounding half, ds and 1/ts per node, and the column's offsets to the
// source.
template <int NI>
struct NodeSrc {
  float tsn[NI + 1], tsl[NI + 1], ds[NI + 1], its[NI + 1];
  float dsx, dsy;
};

template <int NI, typename Tab>
__device__ inline void node_source(const ImgDev &P, const Tab &tab, typename Tab::Col c, const int (&wn)[NI + 1],
                                   const float (&zn)[NI + 1], float dsx, float dsy, NodeSrc<NI> &S) {
  const float hs2 = dsx * dsx + dsy * dsy;
  S.dsx = dsx;
  S.dsy = dsy;
#pragma unroll
  for (int m = 0; m <= NI; ++m) {
    const float2 v = tab.at(c, wn[m]);
    const float ts = fmaxf(v.x, P.tmin);
    const float d2 = fmaf(zn[m], zn[m], hs2);
    const float ds = d2 * __builtin_amdgcn_rsqf(d2);
    S.tsn[m] = ts * P.inv_dt;
    S.tsl[m] = fmaf(v.y, P.inv_dtstar, 0.5f);
    S.ds[m] = ds;
    S.its[m] = __builtin_amdgcn_rcpf(ts);
  }
}

// Receiver leg, ramps and amplitude at node slot m.
template <int NI>
__device__ inline NodeVal node_value(const ImgDev &P, const Dip &D, const Ramps &R, const NodeSrc<NI> &S, int m,
                                     float2 tg, float zn, float dgx, float dgy, float hg2, float unit,
                                     float tmin) {
  const float rdg = __builtin_amdgcn_rsqf(fmaf(zn, zn, hg2));
  const float q = S.ds[m] * rdg;
  const float X = fmaf(dgx, q, S.dsx), Y = fmaf(dgy, q, S.dsy), Z = fmaf(zn, q, zn);
  const float rz = __builtin_amdgcn_rcpf(Z);
  const float tx = X * rz, ty = Y * rz;
  const float t = fmaxf(tg.x, tmin);
  const float tr = (q * q) * q * (t * S.its[m]);
  NodeVal v;
  v.u1 = fmaf(tx, D.rx1, R.cx1);
  v.u2 = fmaf(tx, -D.rx2, R.cx2);
  v.u3 = fmaf(ty, D.ry1, R.cy1);
  v.u4 = fmaf(ty, -D.ry2, R.cy2);
  v.a = unit * (tr + __builtin_amdgcn_rcpf(tr));
  v.F = fmaf(t, P.inv_dt, S.tsn[m]);
  v.L = fmaf(tg.y, P.inv_dtstar, S.tsl[m]);
  return v;
}

__device__ inline float uniform_float(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// Node-form tile body: tile_col as in image_tile; group g owns intervals
// [nf_ws + g NI, nf_ws + (g + 1) NI) of the node range.
// SRC as in image_tile.
template <int NI, int PM, bool SRC, typename BankT, typename Tab, typename Runs>
__device__ inline void image_tile_nodes(const ImgDev &P, const Tab &tab, float *__restrict__ img,
                                        const float4 *__restrict__ geom, const BankT *__restrict__ bank,
                                        int tile_col, int group, const Runs &runs, int nrun) {
  const int lane = threadIdx.x;
  const int ntx = (P.nx + P.tile_x - 1) / P.tile_x;
  const int tx0 = (tile_col % ntx) * P.tile_x;
  const int ty0 = (tile_col / ntx) * P.tile_y;
  const int cx = tx0 + lane % P.tile_x, cy = ty0 + lane / P.tile_x;
  const bool valid = cx < P.nx && cy < P.ny;
  const float x = (float)cx * P.dx, y = (float)cy * P.dy;
  const size_t col = ((size_t)(valid ? cy : 0) * P.nx + (valid ? cx : 0)) * P.nz;

  // wave-uniform tile geometry: nodes, their depths, each interval's first
  // point and each point slot's fraction in its interval
  const int wa = P.nf_ws + group * NI;
  int wn[NI + 1], pa[NI + 1];
  float zn[NI + 1];
#pragma unroll
  for (int m = 0; m <= NI; ++m) {
    wn[m] = min(wa + m, P.nw - 1);  // slots past the range read a valid node and own no points
    zn[m] = uniform_float((float)wn[m] * P.dw);
    const int w = wa + m;
    int p = (w >= P.nf_we) ? P.nf_pe : first_point(w, P.nz, P.dz, P.inv_dw, P.nw);
    p = min(max(p, P.nf_ps), P.nf_pe);
    pa[m] = uniform_int(p);
  }
  // slot (m, p) holds a point when bit m PM + p of the wave-uniform word
  // `on` is set.  Inside the trace loop the word passes through an empty
  // asm that claims to change it at the top of every phase: each slot's
  // test is then a scalar bit test in its own block, instead of a per-slot
  // lane mask that the compiler would hoist out of the loop, keep live in
  // SGPR pairs and, short of SGPRs, spill to VGPR lanes and read back with
  // vector instructions in every iteration.
  unsigned on = 0;
#pragma unroll
  for (int m = 0; m < NI; ++m)
#pragma unroll
    for (int p = 0; p < PM; ++p)
      if (pa[m] + p < pa[m + 1]) on |= 1u << (m * PM + p);
  on = (unsigned)uniform_int((int)on);
  unsigned onl = on;
#define QPSDM_ON(m, p) ((onl >> ((m) * PM + (p))) & 1u)
#define QPSDM_PHASE() asm volatile("" : "+s"(onl))
  float cf_pt[NI][PM];
#pragma unroll
  for (int m = 0; m < NI; ++m)
#pragma unroll
    for (int p = 0; p < PM; ++p) {
      float f = 0.0f;
      point_interval(min(pa[m] + p, P.nf_pe - 1), P.dz, P.inv_dw, P.nw, &f);
      cf_pt[m][p] = uniform_float(f);
    }
  const int pfirst = pa[0], plast = pa[NI];
  if (pfirst >= plast) return;

  float acc[NI][PM];
#pragma unroll
  for (int m = 0; m < NI; ++m)
#pragma unroll
    for (int p = 0; p < PM; ++p) acc[m][p] = (valid && QPSDM_ON(m, p)) ? img[col + pa[m] + p] : 0.f;

  NodeSrc<NI> S;
  float sx = NAN, sy = NAN;
  int cache_run = -1;
  const unsigned nt = (unsigned)P.nt;
  const float tmin = __builtin_canonicalizef(P.tmin);
  const float jmax = (float)(P.K - 1);

  for (int r = 0; r < nrun; ++r) {
    const Run run = runs(r);
    const Dip &D = run.dip;
    const TileBox box = tile_box(P, D, tx0, ty0, pfirst, plast - pfirst);
    const Ramps R = vector_ramps(D);
    const int tend = run.tr0 + run.ntr;
    auto set_source = [&](const float4 &gs) {
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
        node_source<NI>(P, st, st.column(0, iv, iu, a, b), wn, zn, sx - x, sy - y, S);
      } else {
        node_source<NI>(P, tab, column_setup(P, tab, sx, sy, x, y), wn, zn, sx - x, sy - y, S);
      }
    };
    // the aperture cull's box: the tile's columns from its first node to
    // its last (they bound its points)
    const CullBox cbox{(float)tx0 * P.dx, (float)(min(tx0 + P.tile_x, P.nx) - 1) * P.dx, (float)ty0 * P.dy,
                       (float)(min(ty0 + P.tile_y, P.ny) - 1) * P.dy, zn[0], zn[NI]};
    for (int t0 = run.tr0; t0 < tend; t0 += 64) {
      // the live traces of the next 64, tested one per lane
      int lnode;
      for (unsigned long long live = live_traces(P, geom, t0, tend, box, cbox, D, lane, &lnode); live;
           live &= live - 1) {
        const int b = (int)__builtin_ctzll(live), t = t0 + b;
        const float4 g = geom_at(geom, t);
        if constexpr (SRC) {
          if (foreign_source(run, g)) continue;
        }
        set_source(g);
        const float dgx = g.z - x, dgy = g.w - y;
        const float hg2 = dgx * dgx + dgy * dgy;
        auto c = column_at_node(P, tab, __builtin_amdgcn_readlane(lnode, b), g.z, g.w, x, y);
        float unit;
        const __amdgpu_buffer_rsrc_t br = bank_row(bank, (size_t)(t - run.tr0 + run.brow0), P.K, P.nt, &unit);
        // Three phases as in image_tile: the weights, fractions and sample
        // indices of every point (node values computed once per node), then
        // every gather in one branch-free block, then decode and accumulate.
        // slots that hold no point are left unset: every later read of a
        // slot is under the same test
        float cw[NI][PM], cf[NI][PM];
        int ci[NI][PM];
        QPSDM_PHASE();
        NodeVal n0 = node_value<NI>(P, D, R, S, 0, tab.at(c, wn[0]), zn[0], dgx, dgy, hg2, unit, tmin);
#pragma unroll
        for (int m = 0; m < NI; ++m) {
          const NodeVal n1 = node_value<NI>(P, D, R, S, m + 1, tab.at(c, wn[m + 1]), zn[m + 1], dgx, dgy, hg2,
                                            unit, tmin);
          const NodeVal d = node_sub(n1, n0);
#pragma unroll
          for (int p = 0; p < PM; ++p) {
            if (!QPSDM_ON(m, p)) continue;
            const float s = cf_pt[m][p];
            const float f = fmaf(s, d.F, n0.F);
            // min(sat(a), sat(b)) = sat(min(a, b)): one clamp per axis
            const float wx = sat(fminf(fmaf(s, d.u1, n0.u1), fmaf(s, d.u2, n0.u2)));
            const float wy = sat(fminf(fmaf(s, d.u3, n0.u3), fmaf(s, d.u4, n0.u4)));
            const float w = (wx * wy) * fmaf(s, d.a, n0.a);
            const int i0 = (int)f;
            const int j = (int)__builtin_amdgcn_fmed3f(fmaf(s, d.L, n0.L), 0.0f, jmax);
            // a lane outside the trace keeps weight 0 and reads at an index
            // inside the bank or past its end (which reads 0): no branch
            const bool ok = valid && (unsigned)(i0 - 1) < nt - 2u;
            cf[m][p] = __builtin_amdgcn_fractf(f);
            ci[m][p] = i0 + (int)__umul24((unsigned)j, nt);
            cw[m][p] = ok ? w : 0.0f;
          }
          n0 = n1;
        }
        typename BankRaw<BankT>::type cb[NI][PM];
        QPSDM_PHASE();
#pragma unroll
        for (int m = 0; m < NI; ++m)
#pragma unroll
          for (int p = 0; p < PM; ++p)
            if (QPSDM_ON(m, p)) cb[m][p] = bank_raw(br, ci[m][p], bank);
        QPSDM_PHASE();
#pragma unroll
        for (int m = 0; m < NI; ++m)
#pragma unroll
          for (int p = 0; p < PM; ++p) {
            if (!QPSDM_ON(m, p)) continue;
            const float2 b = bank_decode(cb[m][p], bank);
            acc[m][p] = fmaf(cw[m][p], fmaf(cf[m][p], b.y - b.x, b.x), acc[m][p]);
          }
      }
    }
  }
  if (valid) {
#pragma unroll
    for (int m = 0; m < NI; ++m)
#pragma unroll
      for (int p = 0; p < PM; ++p)
        if (QPSDM_ON(m, p)) img[col + pa[m] + p] = acc[m][p];
  }
#undef QPSDM_PHASE
#undef QPSDM_ON
}

}  // namespace qpsdm_dev

#endif
