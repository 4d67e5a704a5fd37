// Straight-ray integrals through the synthetic earth models, in FP64 on the
// device: traveltime t = int ds / v, t* = int ds / (v Q) and the spreading
// integral sigma = int v ds, on the segment from the surface point (sx, sy, 0)
// to (x, y, z).  Linear v(z): closed forms.  Layers + Gaussian anomaly:
// 8-point Gauss-Legendre on each layer's piece of the segment.  Shared by the
// table builder and the surrogate's training-sample generator.
#ifndef QPSDM_RAY_DEV_H_
#define QPSDM_RAY_DEV_H_

#include <hip/hip_runtime.h>
#include <math.h>

#include "qpsdm.h"

namespace qpsdm_dev {

struct ModelDev {
  int kind;
  double v0, kgrad, q0;
  int nlayer;
  double layer_dz;
  double vl[16], ql[16];
  double ax, ay, az, is2, av, aq;
};

inline ModelDev make_model(const qpsdm_model &m) {
  ModelDev d;
  d.kind = m.kind;
  d.v0 = m.v0;
  d.kgrad = m.kgrad;
  d.q0 = m.q0;
  d.nlayer = m.nlayer;
  d.layer_dz = m.layer_dz;
  for (int i = 0; i < 16; ++i) d.vl[i] = m.vl[i], d.ql[i] = m.ql[i];
  d.ax = m.ax;
  d.ay = m.ay;
  d.az = m.az;
  d.is2 = m.asig > 0 ? 1.0 / (2.0 * m.asig * m.asig) : 0.0;
  d.av = m.av;
  d.aq = m.aq;
  return d;
}

inline bool model_ok(const qpsdm_model *m) {
  return m && (m->kind == 0 || (m->kind == 1 && m->nlayer >= 1 && m->nlayer <= 16 && m->layer_dz > 0));
}

// Gauss-Legendre nodes and weights on [-1, 1], 8 points (Abramowitz & Stegun 25.4.30).
__device__ inline double gl_x(int q) {
  constexpr double x[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290,
                           -0.1834346424956498, 0.1834346424956498,  0.5255324099163290,
                           0.7966664774136267,  0.9602898564975363};
  return x[q];
}
__device__ inline double gl_w(int q) {
  constexpr double w[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873,
                           0.3626837833783620, 0.3626837833783620, 0.3137066458778873,
                           0.2223810344533745, 0.1012285362903763};
  return w[q];
}

__device__ inline void ray_integrals(const ModelDev &m, double sx, double sy, double x, double y, double z,
                                     double *t, double *tstar, double *sigma) {
  const double hx = x - sx, hy = y - sy;
  const double L = sqrt(hx * hx + hy * hy + z * z);
  if (m.kind == 0) {
    const double kz = m.kgrad * z;
    double tt, sg;
    if (kz > 1e-9 * m.v0) {
      tt = L / kz * log((m.v0 + kz) / m.v0);
      sg = L * (m.v0 + 0.5 * kz);
    } else {
      tt = L / m.v0;
      sg = L * m.v0;
    }
    *t = tt;
    *tstar = tt / m.q0;
    if (sigma) *sigma = sg;
    return;
  }
  double at = 0, aq = 0, av = 0;
  double prev = 0.0;
  for (int l = 0; l < m.nlayer; ++l) {
    double nxt = 1.0;
    if (l + 1 < m.nlayer) {
      const double zb = (l + 1) * m.layer_dz;
      if (zb < z) nxt = zb / z;
    }
    if (nxt > prev) {
      const double zm = 0.5 * (prev + nxt) * z;
      int li = (int)(zm / m.layer_dz);
      li = li >= m.nlayer ? m.nlayer - 1 : (li < 0 ? 0 : li);
      const double half = 0.5 * (nxt - prev), mid = 0.5 * (prev + nxt);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const double la = mid + half * gl_x(q);
        const double dx = sx + la * hx - m.ax, dy = sy + la * hy - m.ay, dz = la * z - m.az;
        const double G = exp(-(dx * dx + dy * dy + dz * dz) * m.is2);
        const double v = m.vl[li] * (1.0 - m.av * G);
        const double qq = m.ql[li] * (1.0 - m.aq * G);
        const double w = gl_w(q) * half;
        at += w / v;
        aq += w / (v * qq);
        av += w * v;
      }
    }
    prev = nxt;
    if (nxt >= 1.0) break;
  }
  *t = at * L;
  *tstar = aq * L;
  if (sigma) *sigma = av * L;
}

}  // namespace qpsdm_dev

#endif
