// Span form of the imaging sum (weight_nodes = 2): the node form's times and
// levels, with the dip-taper ramps and the amplitude evaluated at coarser
// weight nodes and interpolated between them in 1/z.
//
// As in the node form, a point's total time in samples F and bank level L
// are interpolated linearly between the two table depth nodes of its
// interval (the tables are linear in z there), with L clamped to [0, K - 1]
// at the nodes.  The four taper ramps u = tan * r + c (src/q-psdm.cu:
// 2322-2340) and the amplitude a = unit * (tr + 1/tr), tr = (ds/dg)^3 tg/ts
// (src/q-psdm.cu:2372-2374), are evaluated only at the two ends of a span
// of S table intervals and interpolated at the fraction
//   s = ((z - z0) z1) / ((z1 - z0) z)
// of 1/z between the span's end depths z0 and z1: the bisector's tangents
// vary with depth as (horizontal offset) / z, so they are close to linear in
// 1/z, and the interpolation error at a given span length is several times
// smaller than linear in z.  The ramps are clamped per point as in the node
// form.  Spans come from a depth schedule: below the node range's first
// interval ws, band k starts at the first interval at or below span_z[k]
// rounded up to a multiple of 8 intervals past ws and has spans of span_s[k]
// intervals (1, 2, 4 or 8) from its start; the last span of a band ends at
// the band's end.
//
// A tile is NI consecutive intervals of a band (one work-group of 64
// columns, one column per lane); S and NI are compile-time.  With S <= NI
// the tile holds NI / S whole spans whose weight nodes are tile nodes; with
// S > NI it lies inside one span whose two weight nodes can lie outside it
// (their receiver times are read separately).  Per trace and span the wave
// classifies the span from the ramps at its two ends: every valid lane has
// some ramp <= 0 at both ends (the taper is 0 over the span: its points are
// skipped, no gathers), or every ramp >= 1 at both ends (the taper is 1: the
// weight is the interpolated amplitude), or neither (the full clamped
// ramps).  The shortcuts give the full formula's values exactly: the fused
// interpolation of two values that are both <= 0 (both >= 1) is <= 0
// (>= 1) in floating point too, since the rounded difference is bounded by
// the exactly representable 0 - u0 (1 - u0) and rounding is monotone.
//
// The bank is a pair bank (QPSDM_F32P / QPSDM_F16P): entries 0 and nt - 1
// of each row are zero, so the sample index is clamped to nt - 1 instead of
// tested, and the interpolation in time is one fused multiply-add.
#ifndef QPSDM_IMAGE_SPANS_H_
#define QPSDM_IMAGE_SPANS_H_

#include <float.h>

#include "image_nodes.h"

namespace qpsdm_dev {

typedef const __attribute__((address_space(3))) unsigned long long *lds_u64;

// max(x, y) and min(x, y) of finite values as one v_med3_f32 against
// +-FLT_MAX.  fmaxf / fminf first quiet each operand the compiler cannot
// prove canonical (any value from another basic block: loop-carried ramps,
// times after a conditional reload), one more instruction per operand.
__device__ inline float fmax_fin(float x, float y) { return __builtin_amdgcn_fmed3f(x, y, FLT_MAX); }
__device__ inline float fmin_fin(float x, float y) { return __builtin_amdgcn_fmed3f(x, y, -FLT_MAX); }

// true when pred holds for no lane of the wave
__device__ inline bool wave_none(bool pred) { return __builtin_amdgcn_ballot_w64(pred) == 0; }

// The ramps and amplitude at a weight node.
struct WVal {
  float u1, u2, u3, u4, a;
};

// the smallest of a node's four ramps
__device__ inline float wval_lo(const WVal &v) { return fminf(fminf(v.u1, v.u2), fminf(v.u3, v.u4)); }

__device__ inline WVal wval_sub(const WVal &b, const WVal &a) {
  return WVal{b.u1 - a.u1, b.u2 - a.u2, b.u3 - a.u3, b.u4 - a.u4, b.a - a.a};
}

// Ramps and amplitude at depth zn for a receiver leg (dgx, dgy, hg2, time
// tg) and the source leg's distance ds and 1/ts at that depth; the node
// form's arithmetic (node_value).
__device__ inline WVal weight_value(const Dip &D, const Ramps &R, float ds, float its, float tg, float zn, float dsx,
                                    float dsy, float dgx, float dgy, float hg2, float unit, float tmin) {
  const float rdg = __builtin_amdgcn_rsqf(fmaf(zn, zn, hg2));
  const float q = ds * rdg;
  const float X = fmaf(dgx, q, dsx), Y = fmaf(dgy, q, dsy), Z = fmaf(zn, q, zn);
  const float rz = __builtin_amdgcn_rcpf(Z);
  const float tx = X * rz, ty = Y * rz;
  const float t = fmax_fin(tg, tmin);
  const float tr = (q * q) * q * (t * its);
  WVal v;
  v.u1 = fmaf(tx, D.rx1, R.cx1);
  v.u2 = fmaf(tx, -D.rx2, R.cx2);
  v.u3 = fmaf(ty, D.ry1, R.cy1);
  v.u4 = fmaf(ty, -D.ry2, R.cy2);
  v.a = unit * (tr + __builtin_amdgcn_rcpf(tr));
  return v;
}

// Source leg of a span tile: t in samples and t* in levels (+ the rounding
// half) at the NI + 1 table nodes, ds and 1/ts at the NW weight nodes.
template <int NI, int NW>
struct SpanSrc {
  float tsn[NI + 1], tsl[NI + 1], ds[NW], its[NW];
  float dsx, dsy;
};

template <int NI, int NW, typename Tab>
__device__ inline void span_source(const ImgDev &P, const Tab &tab, typename Tab::Col c, const int (&wn)[NI + 1],
                                   const int (&wk)[NW], float dsx, float dsy, SpanSrc<NI, NW> &S) {
  const float hs2 = dsx * dsx + dsy * dsy;
  S.dsx = dsx;
  S.dsy = dsy;
#pragma unroll
  for (int m = 0; m <= NI; ++m) {
    const float2 v = tab.at(c, wn[m]);
    S.tsn[m] = fmaxf(v.x, P.tmin) * P.inv_dt;
    S.tsl[m] = fmaf(v.y, P.inv_dtstar, 0.5f);
  }
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    const float zn = (float)wk[k] * P.dw;
    const float2 v = tab.at(c, wk[k]);
    const float d2 = fmaf(zn, zn, hs2);
    S.ds[k] = d2 * __builtin_amdgcn_rsqf(d2);
    S.its[k] = __builtin_amdgcn_rcpf(fmaxf(v.x, P.tmin));
  }
}

// The table words of a tile's nodes wn[m0..m1).  U16Q: nodes 0..NI-1 lie in
// the depth brick of wn[0] (the span form's tiles start at multiples of 4 or
// 8 intervals from a multiple of 8), node NI in its own; their two headers
// are loaded once per trace instead of checked per node.  Nodes past the
// tables' last node are clamped to it and own no points.
template <int NI>
__device__ inline void tile_words(const TabF32 &tab, TabF32::Col &c, const int (&wn)[NI + 1], int m0, int m1,
                                  TabF32::Raw *raw, const float4 &, const float4 &) {
#pragma unroll
  for (int m = m0; m < m1; ++m) raw[m - m0] = tab.fetch(c, wn[m]);
}
template <int NI>
__device__ inline void tile_words(const TabU16Q &tab, TabU16Q::Col &c, const int (&wn)[NI + 1], int m0, int m1,
                                  TabU16Q::Raw *raw, const float4 &h0, const float4 &h1) {
#pragma unroll
  for (int m = m0; m < m1; ++m) raw[m - m0] = tab.word(c, wn[m], m == NI ? h1 : h0);
}
__device__ inline float4 tile_header(const TabF32 &, const TabF32::Col &, int) { return float4{}; }
__device__ inline float4 tile_header(const TabU16Q &, const TabU16Q::Col &c, int w) {
  return TabU16Q::header(c, (w >> 3) * kQBrickBytes);
}

// The span tile body: tile_col as in image_tile; group g owns intervals
// [sp_b + (sp_g0 + g) NI, sp_b + (sp_g0 + g + 1) NI) of the band
// [sp_b, sp_e) (ImgDev), cut at the node range's points [nf_ps, nf_pe).
// SRC as in image_tile.
// PAT: the node range alternates 3 and 2 points per interval from the
// band's start (a 10 m image under 25 m tables, or any geometry with at
// most that many), so the slots are 3, 2, 3, 2, ... instead of PM per
// interval; PAT = 2 gives the tile's last interval 3 slots too (the
// range's last interval, which holds the points down to the tables' last
// node).
template <int NI, int S, int PM, int PAT, bool SRC, typename BankT, typename Tab, typename Runs>
__device__ inline void image_tile_spans(const ImgDev &P, const Tab &tab, float *__restrict__ img,
                                        const float4 *__restrict__ geom, const BankT *__restrict__ bank,
                                        int tile_col, int group, const Runs &runs, int nrun) {
  static_assert(S == 1 || S == 2 || S == 4 || S == 8, "span of 1, 2, 4 or 8 intervals");
  static_assert(NI == 2 || NI == 4 || NI == 8, "tiles of 2, 4 or 8 intervals");
  constexpr bool INSIDE = S <= NI;           // whole spans inside the tile
  constexpr int NSP = INSIDE ? NI / S : 1;   // spans per tile
  constexpr int SI = INSIDE ? S : NI;        // intervals per span inside the tile
  constexpr int NW = NSP + 1;                // weight nodes
  constexpr int CI = SI < 2 ? SI : 2;        // intervals per gather chunk
  static_assert(!PAT || (PM == 3 && NI % 2 == 0), "the 3, 2 slot pattern takes PM = 3 and an even NI");
  // point slots of interval m
  auto nslot = [](int m) { return PAT ? (((m & 1) && !(PAT == 2 && m == NI - 1)) ? 2 : 3) : PM; };
  __shared__ float2 cst[NI * PM];            // per point slot: (fraction in its interval, in its span)

  const int lane = threadIdx.x;
  const int ntx = (P.nx + P.tile_x - 1) / P.tile_x;
  const int tx0 = (tile_col % ntx) * P.tile_x;
  const int ty0 = (tile_col / ntx) * P.tile_y;
  const int cx = tx0 + lane % P.tile_x, cy = ty0 + lane / P.tile_x;
  const bool valid = cx < P.nx && cy < P.ny;
  const float x = (float)cx * P.dx, y = (float)cy * P.dy;
  const size_t col = ((size_t)(valid ? cy : 0) * P.nx + (valid ? cx : 0)) * P.nz;

  // wave-uniform tile geometry: table nodes, weight nodes, each interval's
  // first point, and per point slot its two fractions (in LDS)
  const int wa = P.sp_b + (P.sp_g0 + group) * NI;
  int wn[NI + 1], pa[NI + 1], wk[NW];
#pragma unroll
  for (int m = 0; m <= NI; ++m) {
    wn[m] = min(wa + m, P.nw - 1);  // nodes past the range read a valid node and own no points
    // points of the band only: past its end the next band's first point
    int p = first_point(min(wa + m, P.sp_e), P.nz, P.dz, P.inv_dw, P.nw);
    p = min(max(p, P.nf_ps), P.nf_pe);
    pa[m] = uniform_int(p);
  }
  if constexpr (INSIDE) {
#pragma unroll
    for (int k = 0; k < NW; ++k) wk[k] = min(wa + k * S, P.sp_e);
  } else {
    const int w0 = P.sp_b + (wa - P.sp_b) / S * S;
    wk[0] = w0;
    wk[1] = min(w0 + S, P.sp_e);
  }
  unsigned on = 0;
#pragma unroll
  for (int m = 0; m < NI; ++m)
#pragma unroll
    for (int p = 0; p < PM; ++p)
      if (p < nslot(m) && pa[m] + p < pa[m + 1]) on |= 1u << (m * PM + p);
  on = (unsigned)uniform_int((int)on);
  if (lane < NI * PM) {
    // a slot that holds no point gets (0, 0): it reads its interval's first
    // node, inside the tables and the bank
    const int m = lane / PM, p = lane % PM;
    const int k = INSIDE ? m / S : 0;
    float2 v = make_float2(0.0f, 0.0f);
    if (pa[m] + p < pa[m + 1]) {
      const int iz = pa[m] + p;
      point_interval(iz, P.dz, P.inv_dw, P.nw, &v.x);
      const float z = (float)iz * P.dz;
      const float z0 = (float)wk[k] * P.dw, z1 = (float)wk[k + 1] * P.dw;
      v.y = ((z - z0) * z1) / ((z1 - z0) * z);
    }
    cst[lane] = v;
  }
  __syncthreads();
  const int pfirst = pa[0], plast = pa[NI];
  if (pfirst >= plast) return;
#define QPSDM_ON(m, p) ((on >> ((m) * PM + (p))) & 1u)

  float acc[NI][PM];
#pragma unroll
  for (int m = 0; m < NI; ++m)
#pragma unroll
    for (int p = 0; p < PM; ++p) acc[m][p] = (valid && QPSDM_ON(m, p)) ? img[col + pa[m] + p] : 0.f;

  SpanSrc<NI, NW> SC;
  float sx = NAN, sy = NAN;
  int cache_run = -1;
  const unsigned ntm1 = (unsigned)P.nt - 1u;
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
        span_source<NI, NW>(P, st, st.column(0, iv, iu, a, b), wn, wk, sx - x, sy - y, SC);
      } else {
        span_source<NI, NW>(P, tab, column_setup(P, tab, sx, sy, x, y), wn, wk, sx - x, sy - y, SC);
      }
    };
    // the aperture cull's box: the tile's columns from its first weight node
    // to its last (they bound its points)
    const CullBox cbox{(float)tx0 * P.dx, (float)(min(tx0 + P.tile_x, P.nx) - 1) * P.dx, (float)ty0 * P.dy,
                       (float)(min(ty0 + P.tile_y, P.ny) - 1) * P.dy, (float)wk[0] * P.dw, (float)wk[NW - 1] * P.dw};
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
        // time (samples) and level at the table nodes; the receiver time kept
        // at the weight nodes that are table nodes
        float Fn[NI + 1], Ln[NI + 1], tw[NW];
        // (in two halves of the nodes: the raw words of all nine nodes of an
        // eight-interval tile would hold 36 registers at once)
        constexpr int NH = NI / 2 + 1;
        const float4 h0 = tile_header(tab, c, wn[0]), h1 = tile_header(tab, c, wn[NI]);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int m0 = h ? NH : 0, m1 = h ? NI + 1 : NH;
          typename Tab::Raw raw[NH];
          tile_words<NI>(tab, c, wn, m0, m1, raw, h0, h1);
#pragma unroll
          for (int m = m0; m < m1; ++m) {
            const float2 v = tab.reduce(c, raw[m - m0]);
            Fn[m] = fmaf(fmax_fin(v.x, tmin), P.inv_dt, SC.tsn[m]);
            Ln[m] = __builtin_amdgcn_fmed3f(fmaf(v.y, P.inv_dtstar, SC.tsl[m]), 0.0f, jmax);
            if constexpr (INSIDE) {
              if (m % S == 0) tw[m / S] = v.x;
            }
          }
        }
        if constexpr (!INSIDE) {
#pragma unroll
          for (int k = 0; k < NW; ++k) tw[k] = tab.at(c, wk[k]).x;
        } else {
          // a span cut at the band's end has its end node inside the tile's
          // last interval, not at a tile node
#pragma unroll
          for (int k = 1; k < NW; ++k)
            if (wk[k] != wa + k * S) tw[k] = tab.at(c, wk[k]).x;
        }
        // the ramps and amplitude at the weight nodes, and per span: bit k of
        // out_m when the taper is 0 over span k for every valid lane, of in_m
        // when it is 1
        WVal W[NW];
        float lo[NW];
#pragma unroll
        for (int k = 0; k < NW; ++k) {
          W[k] = weight_value(D, R, SC.ds[k], SC.its[k], tw[k], (float)wk[k] * P.dw, SC.dsx, SC.dsy, dgx, dgy, hg2,
                              unit, tmin);
          lo[k] = wval_lo(W[k]);
        }
        unsigned out_m = 0, in_m = 0;
#pragma unroll
        for (int k = 0; k < NSP; ++k) {
          const WVal &A = W[k], &B = W[k + 1];
          const float mlo = fmin_fin(lo[k], lo[k + 1]);
          const float mout = fminf(fminf(fmax_fin(A.u1, B.u1), fmax_fin(A.u2, B.u2)),
                                   fminf(fmax_fin(A.u3, B.u3), fmax_fin(A.u4, B.u4)));
          if (wave_none(valid && !(mout <= 0.0f))) out_m |= 1u << k;
          if (wave_none(valid && !(mlo >= 1.0f))) in_m |= 1u << k;
        }
        // the per-slot constants are read from LDS in every trace (ds_read;
        // the asm keeps the compiler from hoisting them into registers)
        lds_u64 cs = (lds_u64)cst;
        asm volatile("" : "+v"(cs));
#pragma unroll
        for (int k = 0; k < NSP; ++k) {
          if (!((out_m >> k) & 1u)) {
            const WVal &W0 = W[k];
            const WVal dW = wval_sub(W[k + 1], W0);
            const bool all_in = (in_m >> k) & 1u;
            // every slot of the span's intervals, branch-free (a slot that
            // holds no point reads its interval's first node and accumulates
            // into a register that is never stored), in chunks of CI
            // intervals: time and level, the gathers, the weights, the sums
#pragma unroll
            for (int c0 = k * SI; c0 < (k + 1) * SI; c0 += CI) {
              float cw[CI][PM], cf[CI][PM], sf[CI][PM];
              int ci[CI][PM];
              typename BankRaw<BankT>::type cb[CI][PM];
#pragma unroll
              for (int i = 0; i < CI; ++i) {
                const int m = c0 + i;
                const float dF = Fn[m + 1] - Fn[m], dL = Ln[m + 1] - Ln[m];
#pragma unroll
                for (int p = 0; p < PM; ++p) {
                  if (p >= nslot(m)) continue;
                  const float2 q = __builtin_bit_cast(float2, cs[m * PM + p]);
                  const float f = fmaf(q.x, dF, Fn[m]);
                  const unsigned j = (unsigned)fmaf(q.x, dL, Ln[m]);
                  const unsigned i0 = min((unsigned)f, ntm1);
                  cf[i][p] = __builtin_amdgcn_fractf(f);
                  ci[i][p] = (int)(__umul24(j, nt) + i0);
                  sf[i][p] = q.y;
                }
              }
#pragma unroll
              for (int i = 0; i < CI; ++i)
#pragma unroll
                for (int p = 0; p < PM; ++p)
                  if (p < nslot(c0 + i)) cb[i][p] = bank_raw(br, ci[i][p], bank);
              // the weights: the interpolated amplitude, times the clamped
              // ramps unless the span's taper is 1 (one branch per chunk)
#pragma unroll
              for (int i = 0; i < CI; ++i)
#pragma unroll
                for (int p = 0; p < PM; ++p)
                  if (p < nslot(c0 + i)) cw[i][p] = fmaf(sf[i][p], dW.a, W0.a);
              if (!all_in) {
#pragma unroll
                for (int i = 0; i < CI; ++i)
#pragma unroll
                  for (int p = 0; p < PM; ++p) {
                    if (p >= nslot(c0 + i)) continue;
                    const float s = sf[i][p];
                    const float wx = sat(fminf(fmaf(s, dW.u1, W0.u1), fmaf(s, dW.u2, W0.u2)));
                    const float wy = sat(fminf(fmaf(s, dW.u3, W0.u3), fmaf(s, dW.u4, W0.u4)));
                    cw[i][p] *= wx * wy;
                  }
              }
#pragma unroll
              for (int i = 0; i < CI; ++i)
#pragma unroll
                for (int p = 0; p < PM; ++p)
                  if (p < nslot(c0 + i))
                    acc[c0 + i][p] = fmaf(cw[i][p], bank_interp(cb[i][p], cf[i][p], bank), acc[c0 + i][p]);
            }
          }
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
#undef QPSDM_ON
}

}  // namespace qpsdm_dev

#endif
