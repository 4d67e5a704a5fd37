// Device-side pieces shared by the imaging kernels: parameter block, table
// readers (FP32 and FP16 offset+scale bricks), the bilinear table column of
// a surface position, the dip-bound taper and the conservative tile skip.
#ifndef QPSDM_IMAGE_DEV_H_
#define QPSDM_IMAGE_DEV_H_

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <math.h>

#include "common.h"

namespace qpsdm_dev {

struct ImgDev {
  int nx, ny, nz;
  float dx, dy, dz;
  int nt;
  float inv_dt;
  int K;
  float inv_dtstar;
  float tx1, tx2, ty1, ty2, tx1c, tx2c, ty1c, ty2c;
  float rx1, rx2, ry1, ry2;  // reciprocal ramp widths
  float Tx, Ty;              // |tan| beyond which a point is rejected
  float tmin;
  int nnx, nny;
  float n0x, n0y, inv_ndx, inv_ndy;
  int nu, nv, nw;
  float u0, v0, inv_du, inv_dv, inv_dw;
  int tile_x, tile_y, tile_z;
  int tr0, ntr;
};

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
  const float tz = p->taperzone;
  o->tx1 = p->tanx1;
  o->tx2 = p->tanx2;
  o->ty1 = p->tany1;
  o->ty2 = p->tany2;
  o->tx1c = p->tanx1 - tz * (1 + p->tanx1 * p->tanx1);
  o->tx2c = p->tanx2 + tz * (1 + p->tanx2 * p->tanx2);
  o->ty1c = p->tany1 - tz * (1 + p->tany1 * p->tany1);
  o->ty2c = p->tany2 + tz * (1 + p->tany2 * p->tany2);
  QPSDM_REQUIRE(o->tx1c < o->tx1 && o->tx1 <= o->tx2 && o->tx2 < o->tx2c && o->ty1c < o->ty1 &&
                    o->ty1 <= o->ty2 && o->ty2 < o->ty2c,
                "image: dip bounds must satisfy tan1c < tan1 <= tan2 < tan2c");
  o->rx1 = 1.0f / (o->tx1 - o->tx1c);
  o->rx2 = 1.0f / (o->tx2c - o->tx2);
  o->ry1 = 1.0f / (o->ty1 - o->ty1c);
  o->ry2 = 1.0f / (o->ty2c - o->ty2);
  o->Tx = fmaxf(fabsf(o->tx1c), fabsf(o->tx2c));
  o->Ty = fmaxf(fabsf(o->ty1c), fabsf(o->ty2c));
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
  int tx = 8, ty = 8, tzp = 8;
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
  return QPSDM_OK;
}

struct TabF32 {
  const float2 *p;
  size_t nstride;
  int nu, nw;
  __device__ inline float2 fetch(int node, int iv, int iu, int iw) const {
    return p[(size_t)node * nstride + ((size_t)iv * nu + iu) * nw + iw];
  }
};

inline TabF32 make_tab_f32(const qpsdm_tgeom &g, const void *p) {
  TabF32 t;
  t.p = (const float2 *)p;
  t.nstride = (size_t)g.nv * g.nu * g.nw;
  t.nu = g.nu;
  t.nw = g.nw;
  return t;
}

// FP16 bricks: per node [bv][bu][bw] bricks of 8x8x8; each brick a float4
// header {t_off, t_scale, s_off, s_scale} then 512 half2 (t, t*) [v][u][w].
constexpr int kBrickBytes = 16 + 512 * 4;

struct TabF16B {
  const unsigned char *p;
  size_t nstride;  // bytes per node
  int nbu, nbw;
  __device__ inline float2 fetch(int node, int iv, int iu, int iw) const {
    const size_t brick = ((size_t)(iv >> 3) * nbu + (iu >> 3)) * nbw + (iw >> 3);
    const unsigned char *b = p + (size_t)node * nstride + brick * kBrickBytes;
    const float4 h = *(const float4 *)b;
    const int inner = (((iv & 7) << 3) + (iu & 7)) * 8 + (iw & 7);
    const __half2 v = *(const __half2 *)(b + 16 + inner * 4);
    const float2 f = __half22float2(v);
    return make_float2(h.x + h.y * f.x, h.z + h.w * f.y);
  }
};

inline size_t f16b_node_bytes(const qpsdm_tgeom &g) {
  return (size_t)((g.nv + 7) / 8) * ((g.nu + 7) / 8) * ((g.nw + 7) / 8) * kBrickBytes;
}

inline TabF16B make_tab_f16b(const qpsdm_tgeom &g, const void *p) {
  TabF16B t;
  t.p = (const unsigned char *)p;
  t.nstride = f16b_node_bytes(g);
  t.nbu = (g.nu + 7) / 8;
  t.nbw = (g.nw + 7) / 8;
  return t;
}

// Bilinear (x, y) interpolation weights of the nearest node's table at the
// relative position (x - px, y - py): the table column of one surface
// position under one image column.
struct Column {
  int node, iv, iu;
  float a, b;
};

__device__ inline void frac_index(float q, int n, int *i0, float *a) {
  q = fminf(fmaxf(q, 0.0f), (float)(n - 1));
  int i = min((int)q, n - 2);
  *i0 = i;
  *a = q - (float)i;
}

__device__ inline Column column_setup(const ImgDev &P, float px, float py, float x, float y) {
  Column c;
  const int in = min(max((int)floorf((px - P.n0x) * P.inv_ndx + 0.5f), 0), P.nnx - 1);
  const int jn = min(max((int)floorf((py - P.n0y) * P.inv_ndy + 0.5f), 0), P.nny - 1);
  c.node = jn * P.nnx + in;
  frac_index((x - px - P.u0) * P.inv_du, P.nu, &c.iu, &c.a);
  frac_index((y - py - P.v0) * P.inv_dv, P.nv, &c.iv, &c.b);
  return c;
}

// (t, t*) of the column at table depth node w.
template <typename Tab>
__device__ inline float2 column_at(const Tab &tab, const Column &c, int w) {
  const float2 v00 = tab.fetch(c.node, c.iv, c.iu, w), v01 = tab.fetch(c.node, c.iv, c.iu + 1, w);
  const float2 v10 = tab.fetch(c.node, c.iv + 1, c.iu, w);
  const float2 v11 = tab.fetch(c.node, c.iv + 1, c.iu + 1, w);
  const float lx = v00.x + c.a * (v01.x - v00.x), ly = v00.y + c.a * (v01.y - v00.y);
  const float hx = v10.x + c.a * (v11.x - v10.x), hy = v10.y + c.a * (v11.y - v10.y);
  return make_float2(lx + c.b * (hx - lx), ly + c.b * (hy - ly));
}

// Depth node and weight of depth z: w = min(floor(z/dw), nw-2), frac clamped.
__device__ inline int depth_node(const ImgDev &P, float z, float *frac) {
  const float fw = fminf(z * P.inv_dw, (float)(P.nw - 1));
  const int w = min((int)fw, P.nw - 2);
  *frac = fw - (float)w;
  return w;
}

struct TileBox {
  float xa, xb, ya, yb, limx, limy;
};

__device__ inline TileBox tile_box(const ImgDev &P, int tx0, int ty0, int z0, int tz) {
  TileBox b;
  b.xa = (float)tx0 * P.dx;
  b.xb = (float)(min(tx0 + P.tile_x, P.nx) - 1) * P.dx;
  b.ya = (float)ty0 * P.dy;
  b.yb = (float)(min(ty0 + P.tile_y, P.ny) - 1) * P.dy;
  const float zmax = (float)(min(z0 + tz, P.nz) - 1) * P.dz;
  b.limx = P.Tx * zmax * 1.0001f + 0.01f;
  b.limy = P.Ty * zmax * 1.0001f + 0.01f;
  return b;
}

// True when every point of the tile lies on one side of both the source and
// the receiver by more than Tx*z (or Ty*z): then the bisector's |tan| exceeds
// Tx (Ty) at every point and the dip test rejects all of them.
__device__ inline bool tile_culled(const ImgDev &P, const TileBox &b, float sx, float sy, float gx,
                                   float gy) {
  return (b.xa - fmaxf(sx, gx) > b.limx) || (fminf(sx, gx) - b.xb > b.limx) ||
         (b.ya - fmaxf(sy, gy) > b.limy) || (fminf(sy, gy) - b.yb > b.limy);
}

// Dip-bound taper of the reference (src/q-psdm.cu:2322-2340) on the bisector
// (X, Y, Z); false when the point is outside the widened bounds.
__device__ inline bool dip_weight(const ImgDev &P, float X, float Y, float Z, float *w) {
  const float rz = __builtin_amdgcn_rcpf(Z);
  const float tanx = X * rz, tany = Y * rz;
  if (tanx <= P.tx1c || P.tx2c <= tanx || tany <= P.ty1c || P.ty2c <= tany) return false;
  float v = 1.0f;
  if (tanx < P.tx1) v *= (tanx - P.tx1c) * P.rx1;
  if (tanx > P.tx2) v *= (P.tx2c - tanx) * P.rx2;
  if (tany < P.ty1) v *= (tany - P.ty1c) * P.ry1;
  if (tany > P.ty2) v *= (P.ty2c - tany) * P.ry2;
  *w = v;
  return true;
}

__device__ inline float ld(const float *p) { return *p; }
__device__ inline float ld(const __hip_bfloat16 *p) { return __bfloat162float(*p); }

}  // namespace qpsdm_dev

#endif
