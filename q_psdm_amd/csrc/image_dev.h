// Device-side pieces shared by the imaging kernels: parameter block, table
// readers (FP32 and FP16 offset+scale bricks), the bilinear table column of
// a surface position, the dip-bound taper and the conservative tile skip.
#ifndef QPSDM_IMAGE_DEV_H_
#define QPSDM_IMAGE_DEV_H_

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <type_traits>

#include "common.h"
#include "pair_bank.h"

namespace qpsdm_dev {

// Dip-bound taper of one run of traces: bounds tan1 <= tan2 per axis, the
// bounds widened by taperzone*(1 + tan^2) (src/q-psdm.cu:2324-2333), the
// reciprocal ramp widths and the |tan| beyond which every point is rejected.
struct Dip {
  float tx1, tx2, ty1, ty2, tx1c, tx2c, ty1c, ty2c;
  float rx1, rx2, ry1, ry2;
  float cx1, cx2, cy1, cy2;  // ramp offsets: ramp = tan * r + c
  float Tx, Ty;
};

// Computed on the host for single-run launches and on the device for
// batches: no contraction, so both give the same IEEE-rounded constants and
// the two paths stay bit-identical.
__host__ __device__ inline Dip make_dip(float tanx1, float tanx2, float tany1, float tany2, float tz) {
#pragma clang fp contract(off)
  Dip d;
  d.tx1 = tanx1;
  d.tx2 = tanx2;
  d.ty1 = tany1;
  d.ty2 = tany2;
  d.tx1c = tanx1 - tz * (1 + tanx1 * tanx1);
  d.tx2c = tanx2 + tz * (1 + tanx2 * tanx2);
  d.ty1c = tany1 - tz * (1 + tany1 * tany1);
  d.ty2c = tany2 + tz * (1 + tany2 * tany2);
  d.rx1 = 1.0f / (d.tx1 - d.tx1c);
  d.rx2 = 1.0f / (d.tx2c - d.tx2);
  d.ry1 = 1.0f / (d.ty1 - d.ty1c);
  d.ry2 = 1.0f / (d.ty2c - d.ty2);
  d.cx1 = -d.tx1c * d.rx1;
  d.cx2 = d.tx2c * d.rx2;
  d.cy1 = -d.ty1c * d.ry1;
  d.cy2 = d.ty2c * d.ry2;
  d.Tx = fmaxf(fabsf(d.tx1c), fabsf(d.tx2c));
  d.Ty = fmaxf(fabsf(d.ty1c), fabsf(d.ty2c));
  return d;
}

// A wave-uniform Dip computed in vector registers, moved to scalar ones.
__device__ inline Dip uniform_dip(const Dip &d) {
  auto u = [](float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); };
  Dip o;
  o.tx1 = u(d.tx1), o.tx2 = u(d.tx2), o.ty1 = u(d.ty1), o.ty2 = u(d.ty2);
  o.tx1c = u(d.tx1c), o.tx2c = u(d.tx2c), o.ty1c = u(d.ty1c), o.ty2c = u(d.ty2c);
  o.rx1 = u(d.rx1), o.rx2 = u(d.rx2), o.ry1 = u(d.ry1), o.ry2 = u(d.ry2);
  o.cx1 = u(d.cx1), o.cx2 = u(d.cx2), o.cy1 = u(d.cy1), o.cy2 = u(d.cy2);
  o.Tx = u(d.Tx), o.Ty = u(d.Ty);
  return o;
}

__host__ __device__ inline bool dip_ok(const Dip &d) {
  return d.tx1c < d.tx1 && d.tx1 <= d.tx2 && d.tx2 < d.tx2c && d.ty1c < d.ty1 && d.ty1 <= d.ty2 &&
         d.ty2 < d.ty2c;
}

struct ImgDev {
  int nx, ny, nz;
  float dx, dy, dz;
  int nt;
  float inv_dt;
  int K;
  float inv_dtstar;
  float taperzone;
  Dip dip;  // of single-run launches
  float tmin;
  int nnx, nny;
  float n0x, n0y, inv_ndx, inv_ndy;
  int nu, nv, nw;
  float u0, v0, inv_du, inv_dv, inv_dw;
  int tile_x, tile_y, tile_z;
  int tr0, ntr;
  float dw;
  // the exact body's point range [zb, ze) (all of [0, nz) unless the node
  // form splits it)
  int zb, ze;
  // node form (weight_nodes = 1): node intervals [nf_ws, nf_we) and their
  // points [nf_ps, nf_pe), nf_ngroups tiles of tile_z intervals per column
  // group, at most nf_pm points per interval; nf_ngroups = 0 without it
  int nf_ws, nf_we, nf_ps, nf_pe, nf_ngroups, nf_pm;
  // span form (weight_nodes = 2): the node range's bands (sp_n of them;
  // band k covers intervals [sp_w[k], sp_w[k + 1]) in spans of sp_s[k]),
  // and the band [sp_b, sp_e) of one launch
  int sp_n, sp_w[6], sp_s[5];
  int sp_b, sp_e, sp_g0;  // a launch's band and its first tile group
  // the range's intervals (sp_pat: all but the last; sp_pat_last = 1: the
  // last too) hold at most 3, 2, 3, 2, ... points from its start;
  // sp_pat_last = 2: the last holds 3 and is the last of its tile
  int sp_pat, sp_pat_last;
};

int node_range(const qpsdm_image_params *p, int ni, ImgDev *o);
int span_range(const qpsdm_image_params *p, int ni, ImgDev *o);

// The parameters with the exact form selected: for the kernels that evaluate
// single points (update count, table lookup, dip-angle gathers), which have
// no node form.
inline qpsdm_image_params exact_form(const qpsdm_image_params *p) {
  qpsdm_image_params q = *p;
  q.weight_nodes = 0;
  return q;
}

inline int make_dev(const qpsdm_image_params *p, const qpsdm_launch *sh, int tr0, int ntr,
                    ImgDev *o) {
  QPSDM_REQUIRE(p, "image: null params");
  QPSDM_REQUIRE(p->nx > 0 && p->ny > 0 && p->nz > 1, "image: empty grid");
  QPSDM_REQUIRE(p->dx > 0 && p->dy > 0 && p->dz > 0, "image: bad spacing");
  QPSDM_REQUIRE(p->nt > 2 && p->dt > 0 && p->K >= 1 && p->dtstar > 0, "image: bad nt/dt/K/dtstar");
  QPSDM_REQUIRE(tr0 >= 0 && ntr >= 0, "image: bad trace range");
  const qpsdm_tgeom &g = p->tg;
  QPSDM_REQUIRE(g.nnx >= 1 && g.nny >= 1 && g.nu >= 2 && g.nv >= 2 && g.nw >= 2,
                "image: table dims must be >= 2 (nodes >= 1)");
  QPSDM_REQUIRE(g.du > 0 && g.dv > 0 && g.dw > 0 && g.ndx > 0 && g.ndy > 0, "image: bad table spacing");
  QPSDM_REQUIRE((size_t)g.nv * g.nu * g.nw * 8 < (1ull << 31),
                "image: one node's table must stay below 2 GiB (32-bit buffer offsets)");
  QPSDM_REQUIRE((size_t)p->K * p->nt * 4 < (1ull << 31), "image: one trace's bank must stay below 2 GiB");
  QPSDM_REQUIRE((size_t)g.nv * g.nu < (1u << 24) && p->nt < (1 << 24),
                "image: table columns per node and trace length must stay below 2^24 (24-bit index products)");
  o->nx = p->nx;
  o->ny = p->ny;
  o->nz = p->nz;
  o->dx = p->dx;
  o->dy = p->dy;
  o->dz = p->dz;
  o->nt = p->nt;
  o->inv_dt = 1.0f / p->dt;
  o->K = p->K;
  o->inv_dtstar = 1.0f / p->dtstar;
  o->taperzone = p->taperzone;
  o->dip = make_dip(p->tanx1, p->tanx2, p->tany1, p->tany2, p->taperzone);
  QPSDM_REQUIRE(dip_ok(o->dip), "image: dip bounds must satisfy tan1c < tan1 <= tan2 < tan2c");
  o->tmin = p->tmin;
  o->nnx = g.nnx;
  o->nny = g.nny;
  o->n0x = g.n0x;
  o->n0y = g.n0y;
  o->inv_ndx = 1.0f / g.ndx;
  o->inv_ndy = 1.0f / g.ndy;
  o->nu = g.nu;
  o->nv = g.nv;
  o->nw = g.nw;
  o->u0 = g.u0;
  o->v0 = g.v0;
  o->inv_du = 1.0f / g.du;
  o->inv_dv = 1.0f / g.dv;
  o->inv_dw = 1.0f / g.dw;
  int tx = 8, ty = 8, tzp = 6;
  if (sh) {
    if (sh->tile_x > 0) tx = sh->tile_x;
    if (sh->tile_y > 0) ty = sh->tile_y;
    if (sh->tile_z > 0) tzp = sh->tile_z;
  }
  QPSDM_REQUIRE(tx * ty == 64, "image: tile_x*tile_y must be 64 (one wave of columns)");
  o->tile_x = tx;
  o->tile_y = ty;
  o->tile_z = tzp;
  o->tr0 = tr0;
  o->ntr = ntr;
  o->dw = g.dw;
  o->zb = 0;
  o->ze = p->nz;
  o->nf_ws = o->nf_we = o->nf_ngroups = o->nf_pm = 0;
  o->nf_ps = o->nf_pe = p->nz;
  o->sp_n = 0;
  o->sp_b = o->sp_e = o->sp_g0 = 0;
  o->sp_pat = o->sp_pat_last = 0;
  QPSDM_REQUIRE(p->weight_nodes >= 0 && p->weight_nodes <= 2,
                "image: weight_nodes must be 0 (exact form), 1 (node form) or 2 (span form)");
  if (p->weight_nodes == 1) return node_range(p, tzp, o);
  if (p->weight_nodes == 2) return span_range(p, tzp, o);
  return QPSDM_OK;
}


// Buffer resource (base, size) for raw buffer loads: per-lane 32-bit byte
// offsets plus a wave-uniform scalar offset; offsets past `bytes` read 0.
__device__ inline __amdgpu_buffer_rsrc_t make_rsrc(const void *base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}

__device__ inline float2 ld_f2(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

// FP32 tables: float2 (t, t*) [node][nv][nu][nw].
struct TabF32 {
  const float2 *p;
  size_t nstride;  // float2 per node
  int nu, nw;
  // The four corner columns of one image column under one surface position:
  // the (iv, iu) corner's byte offset per lane; the other three sit at fixed
  // distances (nw and nu nw nodes), added as scalar offsets.
  struct Col {
    __amdgpu_buffer_rsrc_t r;
    int o;
    float a, b;
  };
  __device__ inline Col column(int node, int iv, int iu, float a, float b) const {
    Col c;
    c.r = make_rsrc(p + (size_t)node * nstride, (unsigned)(nstride * sizeof(float2)));
    c.o = (int)__umul24(__umul24((unsigned)iv, (unsigned)nu) + (unsigned)iu, (unsigned)nw * 8u);
    c.a = a;
    c.b = b;
    return c;
  }
  // The four corners of depth node w as loaded (fetch) and their bilinear
  // combination (reduce); at = reduce(fetch).  Split so that a caller can
  // issue the loads of one trace's column well before it needs the values.
  struct Raw {
    float2 v00, v01, v10, v11;
  };
  __device__ inline Raw fetch(Col &c, int w) const {
    const int so = w * 8, du = nw * 8, dv = nu * nw * 8;
    return Raw{ld_f2(c.r, c.o, so), ld_f2(c.r, c.o, so + du), ld_f2(c.r, c.o, so + dv), ld_f2(c.r, c.o, so + dv + du)};
  }
  __device__ inline float2 reduce(const Col &c, const Raw &q) const {
    const float lx = fmaf(c.a, q.v01.x - q.v00.x, q.v00.x), ly = fmaf(c.a, q.v01.y - q.v00.y, q.v00.y);
    const float hx = fmaf(c.a, q.v11.x - q.v10.x, q.v10.x), hy = fmaf(c.a, q.v11.y - q.v10.y, q.v10.y);
    return make_float2(fmaf(c.b, hx - lx, lx), fmaf(c.b, hy - ly, ly));
  }
  __device__ inline float2 at(Col &c, int w) const { return reduce(c, fetch(c, w)); }
};

inline TabF32 make_tab_f32(const qpsdm_tgeom &g, const void *p) {
  TabF32 t;
  t.p = (const float2 *)p;
  t.nstride = (size_t)g.nv * g.nu * g.nw;
  t.nu = g.nu;
  t.nw = g.nw;
  return t;
}

// 16-bit bricks: per node [bv][bu][bw] bricks of 8x8x8; each brick a float4
// header {t_off, t_scale, s_off, s_scale} then 512 (t, t*) code pairs
// [v][u][w], value = off + scale * h with h in [0, 1] stored as __half
// (QPSDM_F16B) or as unsigned 16-bit fixed point h = code / 65535
// (QPSDM_U16B).
constexpr int kBrickBytes = 16 + 512 * 4;

template <bool UNORM>
struct TabB16 {
  const unsigned char *p;
  size_t nstride;  // bytes per node
  int nbu, nbw;
  struct Col {
    __amdgpu_buffer_rsrc_t r;
    int hdr[4], val[4];  // per corner: header and code offsets in depth brick 0
    float a, b;
  };
  __device__ inline Col column(int node, int iv, int iu, float a, float b) const {
    Col c;
    c.r = make_rsrc(p + (size_t)node * nstride, (unsigned)nstride);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int v = iv + (k >> 1), u = iu + (k & 1);
      c.hdr[k] = ((v >> 3) * nbu + (u >> 3)) * nbw * kBrickBytes;
      c.val[k] = c.hdr[k] + 16 + (((v & 7) << 3) + (u & 7)) * 32;
    }
    c.a = a;
    c.b = b;
    return c;
  }
  __device__ inline float2 corner(const Col &c, int k, int w) const {
    const int sb = (w >> 3) * kBrickBytes;
    const float4 h = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(c.r, c.hdr[k], sb, 0));
    const int code = __builtin_amdgcn_raw_buffer_load_b32(c.r, c.val[k], sb + (w & 7) * 4, 0);
    float2 f;
    if (UNORM) {
      f = make_float2((float)(code & 0xffff) * (1.0f / 65535.0f),
                      (float)((unsigned)code >> 16) * (1.0f / 65535.0f));
    } else {
      f = __half22float2(__builtin_bit_cast(__half2, code));
    }
    return make_float2(h.x + h.y * f.x, h.z + h.w * f.y);
  }
  __device__ inline float2 at(Col &c, int w) const {
    const float2 v00 = corner(c, 0, w), v01 = corner(c, 1, w);
    const float2 v10 = corner(c, 2, w), v11 = corner(c, 3, w);
    const float lx = fmaf(c.a, v01.x - v00.x, v00.x), ly = fmaf(c.a, v01.y - v00.y, v00.y);
    const float hx = fmaf(c.a, v11.x - v10.x, v10.x), hy = fmaf(c.a, v11.y - v10.y, v10.y);
    return make_float2(fmaf(c.b, hx - lx, lx), fmaf(c.b, hy - ly, ly));
  }
  // No split: the fetch already combines (each corner's header is needed to
  // decode it).
  struct Raw {
    float2 v;
  };
  __device__ inline Raw fetch(Col &c, int w) const { return Raw{at(c, w)}; }
  __device__ inline float2 reduce(const Col &, const Raw &q) const { return q.v; }
};

// 16-bit fixed-point bricks sharing their v and u edges (QPSDM_U16X): brick
// (bv, bu, bw) holds nodes v in [8 bv, 8 bv + 8], u in [8 bu, 8 bu + 8]
// (edges clamped to the table) and w in [8 bw, 8 bw + 7], as a float4 header
// {t_off, t_step, s_off, s_step} then 9 x 9 x 8 codes (t | t* << 16) in
// [v][u][w] order; value = off + step * code.  The four bilinear corners of
// any table cell lie in one brick, so a column reads one header per depth
// brick and interpolates the raw codes before decoding once.
constexpr int kXCodes = 9 * 9 * 8;
constexpr int kXBrickBytes = 16 + kXCodes * 4;

struct TabU16X {
  const unsigned char *p;
  size_t nstride;  // bytes per node
  int nbu, nbw;
  // The other three corners' codes sit 32, 9 x 32 and 10 x 32 bytes after
  // the (iv, iu) corner's (immediate offsets).
  struct Col {
    __amdgpu_buffer_rsrc_t r;
    int base;  // header of the column's brick at depth brick 0
    int o;     // code offset of the (iv, iu) corner (w = 0 of the brick)
    float a, b;
    int hb;  // depth brick of the cached header
    float4 h;
  };
  __device__ inline Col column(int node, int iv, int iu, float a, float b) const {
    Col c;
    c.r = make_rsrc(p + (size_t)node * nstride, (unsigned)nstride);
    c.base = (int)__umul24(__umul24((unsigned)(iv >> 3), (unsigned)nbu) + (unsigned)(iu >> 3),
                           (unsigned)(nbw * kXBrickBytes));
    c.o = c.base + 16 + ((iv & 7) * 9 + (iu & 7)) * 32;
    c.a = a;
    c.b = b;
    c.hb = -1;
    return c;
  }
  // Codes of the four corners at depth node w with their brick's header
  // (fetch; loads the header when the depth brick changes), and the bilinear
  // combination of the codes decoded once (reduce); at = reduce(fetch).
  struct Raw {
    unsigned q00, q01, q10, q11;
    float4 h;
  };
  __device__ inline Raw fetch(Col &c, int w) const {
    const int bw = w >> 3;
    const int sb = bw * kXBrickBytes;
    if (bw != c.hb) {
      c.h = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(c.r, c.base, sb, 0));
      c.hb = bw;
    }
    const int so = sb + (w & 7) * 4;
    return Raw{__builtin_amdgcn_raw_buffer_load_b32(c.r, c.o, so, 0),
               __builtin_amdgcn_raw_buffer_load_b32(c.r, c.o + 32, so, 0),
               __builtin_amdgcn_raw_buffer_load_b32(c.r, c.o + 9 * 32, so, 0),
               __builtin_amdgcn_raw_buffer_load_b32(c.r, c.o + 10 * 32, so, 0), c.h};
  }
  __device__ inline float2 reduce(const Col &c, const Raw &q) const {
    auto lo = [](unsigned v) { return (float)(v & 0xffffu); };
    auto hi = [](unsigned v) { return (float)(v >> 16); };
    const float t0 = fmaf(c.a, lo(q.q01) - lo(q.q00), lo(q.q00)), t1 = fmaf(c.a, lo(q.q11) - lo(q.q10), lo(q.q10));
    const float s0 = fmaf(c.a, hi(q.q01) - hi(q.q00), hi(q.q00)), s1 = fmaf(c.a, hi(q.q11) - hi(q.q10), hi(q.q10));
    return make_float2(fmaf(q.h.y, fmaf(c.b, t1 - t0, t0), q.h.x), fmaf(q.h.w, fmaf(c.b, s1 - s0, s0), q.h.z));
  }
  __device__ inline float2 at(Col &c, int w) const { return reduce(c, fetch(c, w)); }
};

// Corner-quad bricks (QPSDM_U16Q): the U16X bricks, same codes (header
// steps stored divided by 65535), with the codes stored per cell: brick
// (bv, bu, bw) is the float4 header, then for cells v in [8 bv, 8 bv + 8),
// u in [8 bu, 8 bu + 8) and depth nodes w in [8 bw, 8 bw + 8), in [v][u][w]
// order, the codes of the cell's corners 00 = (v, u), 01 = (v, u + 1),
// 10 = (v + 1, u), 11 = (v + 1, u + 1) as one 16-byte word
// {t00 | t01 << 16, t10 | t11 << 16, t*00 | t*01 << 16, t*10 | t*11 << 16}.  A column reads one word per depth node (the corners share
// a cache line with the next seven nodes) instead of four: the imaging
// kernel's vector-memory pipeline, not its arithmetic, sets its pace.
constexpr int kQBrickBytes = 16 + 8 * 8 * 8 * 16;

// Bilinear weights of a cell in 16-bit fixed point: wa = round(65535 a),
// wb = round(65535 b) for the u and v weights, then w11 = floor(wa wb /
// 65535), w01 = wa - w11, w10 = wb - w11 and w00 = 65535 - wa - wb + w11,
// all in [0, 65535] and summing to exactly 65535, so a weighted sum of
// 16-bit codes fits 32 bits and, divided by 65535, is a convex combination
// of them.  floor(x / 65535) = (x + (x >> 16) + 1) >> 16 holds for every
// x <= 65535^2 without overflow (checked over that whole range).
// Packed per row: (w00 | w01 << 16) and (w10 | w11 << 16).
__host__ __device__ inline void cell_weights(float a, float b, unsigned *row0, unsigned *row1) {
  const unsigned wa = (unsigned)fmaf(a, 65535.0f, 0.5f), wb = (unsigned)fmaf(b, 65535.0f, 0.5f);
  // a, b in [0, 1]: 16-bit weights, whose product a 24-bit multiply forms
  __builtin_assume(wa <= 65535u && wb <= 65535u);
  const unsigned x = wa * wb;
  const unsigned w11 = (x + (x >> 16) + 1u) >> 16;
  *row0 = (65535u - wa - wb + w11) | ((wa - w11) << 16);
  *row1 = (wb - w11) | (w11 << 16);
}

typedef unsigned short u16pair_t __attribute__((ext_vector_type(2)));

// x.lo w.lo + x.hi w.hi + acc over unsigned 16-bit halves (v_dot2_u32_u16)
__device__ inline unsigned dot2_u16(unsigned x, unsigned w, unsigned acc) {
  return __builtin_amdgcn_udot2(__builtin_bit_cast(u16pair_t, x), __builtin_bit_cast(u16pair_t, w), acc, false);
}

struct TabU16Q {
  const unsigned char *p;
  size_t nstride;  // bytes per node
  int nbu, nbw;
  struct Col {
    __amdgpu_buffer_rsrc_t r;
    int base;  // header of the column's brick at depth brick 0
    int o;     // the cell's corner word at w = 0 of the brick
    unsigned row0, row1;  // the cell's fixed-point bilinear weights
    int hb;    // depth brick of the cached header
    float4 h;  // its header (steps divided by 65535, the weights' sum)
  };
  __device__ inline Col column(int node, int iv, int iu, float a, float b) const {
    Col c;
    c.r = make_rsrc(p + (size_t)node * nstride, (unsigned)nstride);
    c.base = (int)__umul24(__umul24((unsigned)(iv >> 3), (unsigned)nbu) + (unsigned)(iu >> 3),
                           (unsigned)(nbw * kQBrickBytes));
    c.o = c.base + 16 + ((iv & 7) * 8 + (iu & 7)) * 128;
    cell_weights(a, b, &c.row0, &c.row1);
    c.hb = -1;
    return c;
  }
  typedef TabU16X::Raw Raw;
  // header of the column's brick sb bytes below depth brick 0 (its steps
  // stored divided by 65535, as reduce takes them)
  __device__ static inline float4 header(const Col &c, int sb) {
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(c.r, c.base, sb, 0));
  }
  __device__ inline Raw fetch(Col &c, int w) const {
    const int bw = w >> 3;
    const int sb = bw * kQBrickBytes;
    if (bw != c.hb) {
      c.h = header(c, sb);
      c.hb = bw;
    }
    const uint4 q = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(c.r, c.o, sb + (w & 7) * 16, 0));
    return Raw{q.x, q.y, q.z, q.w, c.h};
  }
  // The word holds each row's two t codes and two t* codes already paired
  // ({t00 | t01, t10 | t11, s00 | s01, s10 | s11}), each pair weighted by
  // one 16-bit dot product; the weighted code sum decodes as
  // off + (step / 65535) * sum.
  __device__ inline float2 reduce(const Col &c, const Raw &q) const {
    const float t = (float)dot2_u16(q.q01, c.row1, dot2_u16(q.q00, c.row0, 0u));
    const float s = (float)dot2_u16(q.q11, c.row1, dot2_u16(q.q10, c.row0, 0u));
    return make_float2(fmaf(q.h.y, t, q.h.x), fmaf(q.h.w, s, q.h.z));
  }
  __device__ inline float2 at(Col &c, int w) const { return reduce(c, fetch(c, w)); }
  // The corner word of depth node w with a given header (a caller that knows
  // which depth brick the node lies in loads the header once).
  __device__ inline Raw word(const Col &c, int w, const float4 &h) const {
    const uint4 q = __builtin_bit_cast(
        uint4, __builtin_amdgcn_raw_buffer_load_b128(c.r, c.o, (w >> 3) * kQBrickBytes + (w & 7) * 16, 0));
    return Raw{q.x, q.y, q.z, q.w, h};
  }
};

inline void u16x_bricks(const qpsdm_tgeom &g, int *nbv, int *nbu, int *nbw) {
  *nbv = (g.nv - 2) / 8 + 1;
  *nbu = (g.nu - 2) / 8 + 1;
  *nbw = (g.nw + 7) / 8;
}

inline size_t u16x_node_bytes(const qpsdm_tgeom &g) {
  int nbv, nbu, nbw;
  u16x_bricks(g, &nbv, &nbu, &nbw);
  return (size_t)nbv * nbu * nbw * kXBrickBytes;
}

inline size_t u16q_node_bytes(const qpsdm_tgeom &g) {
  int nbv, nbu, nbw;
  u16x_bricks(g, &nbv, &nbu, &nbw);
  return (size_t)nbv * nbu * nbw * kQBrickBytes;
}

// U16Q nodes are addressed with 32-bit byte offsets.
inline bool u16q_fits(const qpsdm_tgeom &g) { return u16q_node_bytes(g) < (1ull << 31); }

inline TabU16Q make_tab_u16q(const qpsdm_tgeom &g, const void *p) {
  TabU16Q t;
  int nbv;
  u16x_bricks(g, &nbv, &t.nbu, &t.nbw);
  t.p = (const unsigned char *)p;
  t.nstride = u16q_node_bytes(g);
  return t;
}

inline TabU16X make_tab_u16x(const qpsdm_tgeom &g, const void *p) {
  TabU16X t;
  int nbv;
  u16x_bricks(g, &nbv, &t.nbu, &t.nbw);
  t.p = (const unsigned char *)p;
  t.nstride = u16x_node_bytes(g);
  return t;
}

inline size_t b16_node_bytes(const qpsdm_tgeom &g) {
  return (size_t)((g.nv + 7) / 8) * ((g.nu + 7) / 8) * ((g.nw + 7) / 8) * kBrickBytes;
}

template <bool UNORM>
inline TabB16<UNORM> make_tab_b16(const qpsdm_tgeom &g, const void *p) {
  TabB16<UNORM> t;
  t.p = (const unsigned char *)p;
  t.nstride = b16_node_bytes(g);
  t.nbu = (g.nu + 7) / 8;
  t.nbw = (g.nw + 7) / 8;
  return t;
}

__device__ inline void frac_index(float q, int n, int *i0, float *a) {
  // clamp to [0, n - 1] as one v_med3_f32 (q is finite)
  q = __builtin_amdgcn_fmed3f(q, 0.0f, (float)(n - 1));
  int i = min((int)q, n - 2);
  *i0 = i;
  *a = q - (float)i;
}

// Table column of the surface position (px, py) under the image column
// (x, y): nearest node, then bilinear weights at the relative position
// (x - px, y - py).
// The nearest surface node of (px, py).
__device__ inline int surface_node(const ImgDev &P, float px, float py) {
  const int in = min(max((int)floorf((px - P.n0x) * P.inv_ndx + 0.5f), 0), P.nnx - 1);
  const int jn = min(max((int)floorf((py - P.n0y) * P.inv_ndy + 0.5f), 0), P.nny - 1);
  return jn * P.nnx + in;
}
// (the node given: a caller that has it from a lane-parallel pass)
template <typename Tab>
__device__ inline typename Tab::Col column_at_node(const ImgDev &P, const Tab &tab, int node, float px, float py,
                                                   float x, float y) {
  int iu, iv;
  float a, b;
  frac_index((x - px - P.u0) * P.inv_du, P.nu, &iu, &a);
  frac_index((y - py - P.v0) * P.inv_dv, P.nv, &iv, &b);
  return tab.column(node, iv, iu, a, b);
}
template <typename Tab>
__device__ inline typename Tab::Col column_setup(const ImgDev &P, const Tab &tab, float px, float py,
                                                 float x, float y) {
  return column_at_node(P, tab, surface_node(P, px, py), px, py, x, y);
}

// Depth node and weight of depth z: w = min(floor(z/dw), nw-2), frac clamped.
__device__ inline int depth_node(const ImgDev &P, float z, float *frac) {
  const float fw = fminf(z * P.inv_dw, (float)(P.nw - 1));
  const int w = min((int)fw, P.nw - 2);
  *frac = fw - (float)w;
  return w;
}

// Per-shot source table sub-grid: the relative nodes that can reach the
// image from any source position (nodes at the node tables' spacing).
inline void src_subgrid(const qpsdm_image_params *p, int *nus, int *nvs) {
  const qpsdm_tgeom &g = p->tg;
  *nus = std::min(g.nu, (int)ceil((p->nx - 1) * (double)p->dx / g.du) + 2);
  *nvs = std::min(g.nv, (int)ceil((p->ny - 1) * (double)p->dy / g.dv) + 2);
}

// Order-preserving map of a float's bits to a signed integer: a < b as
// floats (neither NaN) implies key(a) <= key(b).  Trace geometry is loaded
// into scalar registers, and the scalar unit has no float compares, so the
// per-trace tests below compare keys there instead of spending vector
// instructions on wave-uniform values.
__device__ inline int ord_key(float v) {
  const int b = __float_as_int(v);
  return b ^ ((b >> 31) & 0x7fffffff);
}

__device__ inline int uniform_int(int v) { return __builtin_amdgcn_readfirstlane(v); }

// A tile's cull bounds: the tile's horizontal extent widened by Tx*zmax
// (Ty*zmax), a 1e-4 relative and 0.01 m absolute margin, as keys.
struct TileBox {
  int xlo, xhi, ylo, yhi;
};

__device__ inline TileBox tile_box(const ImgDev &P, const Dip &D, int tx0, int ty0, int z0, int tz) {
  const float xa = (float)tx0 * P.dx;
  const float xb = (float)(min(tx0 + P.tile_x, P.nx) - 1) * P.dx;
  const float ya = (float)ty0 * P.dy;
  const float yb = (float)(min(ty0 + P.tile_y, P.ny) - 1) * P.dy;
  const float zmax = (float)(min(z0 + tz, P.nz) - 1) * P.dz;
  const float limx = D.Tx * zmax * 1.0001f + 0.01f;
  const float limy = D.Ty * zmax * 1.0001f + 0.01f;
  TileBox b;
  b.xlo = uniform_int(ord_key(xa - limx));
  b.xhi = uniform_int(ord_key(xb + limx));
  b.ylo = uniform_int(ord_key(ya - limy));
  b.yhi = uniform_int(ord_key(yb + limy));
  return b;
}

// True when every point of the tile lies on one side of both the source and
// the receiver by more than Tx*z (or Ty*z): then the bisector's |tan| exceeds
// Tx (Ty) at every point and the dip test rejects all of them.
__device__ inline bool tile_culled(const TileBox &b, float sx, float sy, float gx, float gy) {
  const int ksx = ord_key(sx), ksy = ord_key(sy), kgx = ord_key(gx), kgy = ord_key(gy);
  // (bitwise or: all four coordinates are needed, so one 16-byte scalar load
  // serves the test instead of dependent loads of single fields)
  return (max(ksx, kgx) < b.xlo) | (min(ksx, kgx) > b.xhi) | (max(ksy, kgy) < b.ylo) | (min(ksy, kgy) > b.yhi);
}

// A trace rejected upstream carries NaN coordinates.
__device__ inline bool rejected(float sx) { return (__float_as_int(sx) & 0x7fffffff) > 0x7f800000; }

// First trace in [t, tend) that is neither rejected nor culled for the tile
// box (tend if none), with its geometry in g; loads and tests on the scalar
// unit.  The geometry is read through the constant address space: it is not
// written during a launch, and that lets the compiler use a scalar load even
// where the kernel also writes memory (the persistent kernel's atomics).
typedef const __attribute__((address_space(4))) float *const_float_ptr;
__device__ inline int next_live(const float4 *__restrict__ geom, const TileBox &box, int t, int tend, float4 &g) {
  for (; t < tend; ++t) {
    const const_float_ptr cg = (const_float_ptr)(geom + t);
    g = make_float4(cg[0], cg[1], cg[2], cg[3]);
    if (!(rejected(g.x) | tile_culled(box, g.x, g.y, g.z, g.w))) break;
  }
  return t;
}

// The geometry of trace t as a scalar load (see next_live).
__device__ inline float4 geom_at(const float4 *__restrict__ geom, int t) {
  const const_float_ptr cg = (const_float_ptr)(geom + t);
  return make_float4(cg[0], cg[1], cg[2], cg[3]);
}

// A box of image points: columns x in [xa, xb], y in [ya, yb], depths z in
// [z0, z1].
struct CullBox {
  float xa, xb, ya, yb, z0, z1;
};

// Bounds over a box of the unit vector from a point toward the surface
// position (px, py): its x and y components (lo, hi) and its z component
// (hi).  u_x = d / sqrt(d^2 + r^2), d = px - x, r^2 = (py - y)^2 + z^2, is
// increasing in d and, at fixed d, falls with r^2 for d >= 0 and rises for
// d < 0, and d and r^2 vary independently over the box; u_z = z / sqrt(h^2
// + z^2) is largest at the smallest horizontal distance h and the deepest z.
// 0 / 0 (a point on the surface position) gives NaN, which no test below
// passes.
struct LegBounds {
  float xlo, xhi, ylo, yhi, zhi;
};
__device__ inline float unit_part(float d, float r2) { return d * __builtin_amdgcn_rsqf(fmaf(d, d, r2)); }
__device__ inline LegBounds leg_bounds(float px, float py, const CullBox &B) {
  const float dxl = px - B.xb, dxh = px - B.xa, dyl = py - B.yb, dyh = py - B.ya;
  const float hxmin = fmaxf(fmaxf(dxl, -dxh), 0.0f), hxmax = fmaxf(fabsf(dxl), fabsf(dxh));
  const float hymin = fmaxf(fmaxf(dyl, -dyh), 0.0f), hymax = fmaxf(fabsf(dyl), fabsf(dyh));
  const float z02 = B.z0 * B.z0, z12 = B.z1 * B.z1;
  const float rxmin = fmaf(hymin, hymin, z02), rxmax = fmaf(hymax, hymax, z12);  // r^2 for u_x
  const float rymin = fmaf(hxmin, hxmin, z02), rymax = fmaf(hxmax, hxmax, z12);  // r^2 for u_y
  LegBounds b;
  b.xlo = unit_part(dxl, dxl >= 0.0f ? rxmax : rxmin);
  b.xhi = unit_part(dxh, dxh >= 0.0f ? rxmin : rxmax);
  b.ylo = unit_part(dyl, dyl >= 0.0f ? rymax : rymin);
  b.yhi = unit_part(dyh, dyh >= 0.0f ? rymin : rymax);
  b.zhi = B.z1 * __builtin_amdgcn_rsqf(fmaf(hxmin, hxmin, fmaf(hymin, hymin, z12)));
  return b;
}

// True when the trace (sx, sy, gx, gy) is out of the dip aperture at every
// point of the box, by one side of one axis: the bisector's tangent
// tan = X / Z = (u_x,s + u_x,g) / (u_z,s + u_z,g) (the unit vectors toward
// the source and the receiver) stays beyond one cut-off tangent, with a
// margin of 1e-4 (1 + |cut-off|) that covers the rounding of these bounds
// and of the kernels' own per-point and per-node tests.  Then the exact and
// node forms find every point outside the taper, and the span form finds
// the same ramp <= 0 at both ends of every span in the box (the span is
// out).  A numerator bound of the wrong sign, or NaN, culls nothing.
__device__ inline bool aperture_out(const CullBox &B, const Dip &D, float sx, float sy, float gx, float gy) {
  const LegBounds s = leg_bounds(sx, sy, B), g = leg_bounds(gx, gy, B);
  const float den = s.zhi + g.zhi;
  const float nxlo = s.xlo + g.xlo, nxhi = s.xhi + g.xhi, nylo = s.ylo + g.ylo, nyhi = s.yhi + g.yhi;
  auto above = [&](float n, float c) { return n >= 0.0f && n >= (c + 1e-4f * (1.0f + fabsf(c))) * den; };
  auto below = [&](float n, float c) { return n <= 0.0f && n <= (c - 1e-4f * (1.0f + fabsf(c))) * den; };
  return above(nxlo, D.tx2c) || below(nxhi, D.tx1c) || above(nylo, D.ty2c) || below(nyhi, D.ty1c);
}

// The traces of [t0, t0 + 64) that may update the tile, one per lane (bit l
// for trace t0 + l): not rejected, not culled by the tile box, and not out
// of the aperture over the cull box; *node: lane l's receiver's surface
// node (read back per trace with readlane).  Sixty-four traces cost one
// pass of the bounds above.
__device__ inline unsigned long long live_traces(const ImgDev &P, const float4 *__restrict__ geom, int t0, int tend,
                                                 const TileBox &box, const CullBox &cb, const Dip &D, int lane,
                                                 int *node) {
  const int t = t0 + lane;
  bool live = false;
  *node = 0;
  if (t < tend) {
    const float4 g = geom[t];
    live = !(rejected(g.x) | tile_culled(box, g.x, g.y, g.z, g.w)) && !aperture_out(cb, D, g.x, g.y, g.z, g.w);
    if (live) *node = surface_node(P, g.z, g.w);
  }
  return __builtin_amdgcn_ballot_w64(live);
}

// Dip-bound taper of the reference (src/q-psdm.cu:2322-2340) on the bisector
// (X, Y, Z), Z > 0: tan = X/Z is compared as X against bound * Z, and 1/Z
// is formed only for points inside a taper ramp.  false when outside.
__device__ inline bool dip_weight(const Dip &P, float X, float Y, float Z, float *w) {
  if (X <= P.tx1c * Z || P.tx2c * Z <= X || Y <= P.ty1c * Z || P.ty2c * Z <= Y) return false;
  float v = 1.0f;
  if (X < P.tx1 * Z || X > P.tx2 * Z || Y < P.ty1 * Z || Y > P.ty2 * Z) {
    const float rz = __builtin_amdgcn_rcpf(Z);
    const float tanx = X * rz, tany = Y * rz;
    if (tanx < P.tx1) v *= (tanx - P.tx1c) * P.rx1;
    if (tanx > P.tx2) v *= (P.tx2c - tanx) * P.rx2;
    if (tany < P.ty1) v *= (tany - P.ty1c) * P.ry1;
    if (tany > P.ty2) v *= (P.ty2c - tany) * P.ry2;
  }
  *w = v;
  return true;
}

// Samples b[idx], b[idx + 1] of a trace's bank row (resource over K*nt),
// as raw bits (bank_raw: nothing waits on the loads) and decoded to FP32
// (bank_decode, where the samples are consumed).  A BF16 pair is one 32-bit
// load at a 2-byte aligned offset (the device runs with unaligned access
// enabled, as the compiler itself assumes for such pairs), so it takes one
// load and one register: low half b[idx], high half b[idx + 1].
template <typename BankT>
struct BankRaw;
template <>
struct BankRaw<float> {
  typedef uint2 type;
};
template <>
struct BankRaw<__hip_bfloat16> {
  typedef unsigned type;
};
template <>
struct BankRaw<__half> {
  typedef unsigned type;
};

__device__ inline uint2 bank_raw(__amdgpu_buffer_rsrc_t r, int idx, const float *) {
  return make_uint2(__builtin_amdgcn_raw_buffer_load_b32(r, idx * 4, 0, 0),
                    __builtin_amdgcn_raw_buffer_load_b32(r, idx * 4 + 4, 0, 0));
}
__device__ inline unsigned bank_raw(__amdgpu_buffer_rsrc_t r, int idx, const __hip_bfloat16 *) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, idx * 2, 0, 0);
}
__device__ inline unsigned bank_raw(__amdgpu_buffer_rsrc_t r, int idx, const __half *) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, idx * 2, 0, 0);
}
__device__ inline float2 bank_decode(uint2 v, const float *) {
  return make_float2(__uint_as_float(v.x), __uint_as_float(v.y));
}
__device__ inline float2 bank_decode(unsigned v, const __hip_bfloat16 *) {
  return make_float2(__uint_as_float(v << 16), __uint_as_float(v & 0xffff0000u));
}
__device__ inline float2 bank_decode(unsigned v, const __half *) {
  return make_float2(__half2float(__ushort_as_half((unsigned short)(v & 0xffffu))),
                     __half2float(__ushort_as_half((unsigned short)(v >> 16))));
}

// Row `row` of a bank of K x nt samples per trace: a buffer resource over
// its samples, and in *unit the factor from stored samples to filtered-trace
// units (1 for FP32 and BF16; for FP16 the 1/scale of the row's header,
// read on the scalar unit).
__device__ inline __amdgpu_buffer_rsrc_t bank_row(const float *bank, size_t row, int K, int nt, float *unit) {
  *unit = 1.0f;
  return make_rsrc(bank + row * K * nt, (unsigned)(K * nt * 4));
}
__device__ inline __amdgpu_buffer_rsrc_t bank_row(const __hip_bfloat16 *bank, size_t row, int K, int nt,
                                                  float *unit) {
  *unit = 1.0f;
  return make_rsrc(bank + row * K * nt, (unsigned)(K * nt * 2));
}
__device__ inline __amdgpu_buffer_rsrc_t bank_row(const __half *bank, size_t row, int K, int nt, float *unit) {
  typedef const __attribute__((address_space(4))) float *cfp;
  const size_t halves = (16 + ((size_t)K * nt * 2 + 15) / 16 * 16) / 2;  // qpsdm_bank_row_bytes / 2
  const __half *r = bank + row * halves;
  *unit = *(cfp)r;
  return make_rsrc(r + 8, (unsigned)(K * nt * 2));
}

template <typename BankT>
__device__ inline float2 bank_pair(__amdgpu_buffer_rsrc_t r, int idx, const BankT *b) {
  return bank_decode(bank_raw(r, idx, b), b);
}

// Pair banks (QPSDM_F32P, QPSDM_F16P): sample i of a level row is stored
// with its forward difference, (b[i], b[i + 1] - b[i]), so the linear
// interpolation at fraction f is one fused multiply-add, b[i] + f d[i]
// (for FP16 one v_fma_mix_f32 reading both halves of the 32-bit word).
// Entries i = 0 and i = nt - 1 of every row hold (0, 0): the reference
// images a sample pair only for 0 < i and i + 1 < nt (src/q-psdm.cu:
// 2354-2357), and those two entries are exactly the indices in [0, nt - 1]
// outside that range, so a kernel that clamps i to nt - 1 needs no range
// test per point.  F32P rows are K * nt float2; F16P rows carry the FP16
// bank's 16-byte {1/scale, scale, 0, 0} header, then K * nt half2.
template <>
struct BankRaw<PairF32> {
  typedef uint2 type;
};
template <>
struct BankRaw<PairF16> {
  typedef unsigned type;
};
__device__ inline uint2 bank_raw(__amdgpu_buffer_rsrc_t r, int idx, const PairF32 *) {
  return __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, idx * 8, 0, 0));
}
__device__ inline unsigned bank_raw(__amdgpu_buffer_rsrc_t r, int idx, const PairF16 *) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, idx * 4, 0, 0);
}

// Linear interpolation at fraction f of the raw pair (b[i], b[i + 1]) or
// (b[i], d[i]) as loaded.
__device__ inline float bank_interp(uint2 v, float f, const float *) {
  const float b0 = __uint_as_float(v.x), b1 = __uint_as_float(v.y);
  return fmaf(f, b1 - b0, b0);
}
__device__ inline float bank_interp(unsigned v, float f, const __hip_bfloat16 *b) {
  const float2 p = bank_decode(v, b);
  return fmaf(f, p.y - p.x, p.x);
}
__device__ inline float bank_interp(unsigned v, float f, const __half *b) {
  const float2 p = bank_decode(v, b);
  return fmaf(f, p.y - p.x, p.x);
}
__device__ inline float bank_interp(uint2 v, float f, const PairF32 *) {
  return fmaf(f, __uint_as_float(v.y), __uint_as_float(v.x));
}
__device__ inline float bank_interp(unsigned v, float f, const PairF16 *) {
  const __half2 h = __builtin_bit_cast(__half2, v);
  return fmaf(f, __high2float(h), __low2float(h));
}

// Banks stored under a per-trace scale (the row header's 1/scale folds into
// the weight).
template <typename BankT>
__host__ __device__ constexpr bool bank_scaled() {
  return std::is_same<BankT, __half>::value || std::is_same<BankT, PairF16>::value;
}

__device__ inline __amdgpu_buffer_rsrc_t bank_row(const PairF32 *bank, size_t row, int K, int nt, float *unit) {
  *unit = 1.0f;
  return make_rsrc(bank + row * K * nt, (unsigned)(K * nt * 8));
}
__device__ inline __amdgpu_buffer_rsrc_t bank_row(const PairF16 *bank, size_t row, int K, int nt, float *unit) {
  typedef const __attribute__((address_space(4))) float *cfp;
  const size_t words = (16 + ((size_t)K * nt * 4 + 15) / 16 * 16) / 4;  // qpsdm_bank_row_bytes / 4
  const PairF16 *r = bank + row * words;
  *unit = *(cfp)r;
  return make_rsrc(r + 4, (unsigned)(K * nt * 4));
}

}  // namespace qpsdm_dev

#endif
