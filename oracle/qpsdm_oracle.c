/*
 * CPU oracle for the Q-compensated Kirchhoff PSDM imaging path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libqpsdm.so and the timed CPU
 * baseline.  Nothing in the product path links or loads this file's library.
 *
 * Each function cites the reference text it restates.  The reference's
 * imaging kernel is a stub (src/q-psdm.cu:2350-2352 fixes both traveltimes
 * at 1.0 and the trace never reaches the device), so the imaging formula and
 * the Q bank are this build's own, written in the reference's conventions
 * (windows, taper, gain, -i*omega filter, gain limiter, dip-taper weight,
 * (tr + 1/tr) obliquity weight, z-fastest image, img /= nsb).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Threads of the OpenMP loops that follow (n >= 1). */
void or_set_num_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n < 1 ? 1 : n);
#else
  (void)n;
#endif
}

/* ---------------------------------------------------------------------- */
/* FFT lengths.  The reference's valid prime-factor lengths are the 240     */
/* divisors of 720720 = 16*9*5*7*11*13 (src/fft.cc:274-349, NTAB 240);      */
/* npfa returns the smallest one >= nmin, or 720720 (src/fft.cc:351-372);   */
/* npfar = 2*npfa((nmin+1)/2) (src/fft.cc:412-439); get_opt_n_fftw = npfar */
/* (src/fft.cc:3721).  Restated as a divisor search.                       */
/* ---------------------------------------------------------------------- */
int or_npfa(int nmin) {
  const int big = 720720;
  if (nmin >= big) return big;
  for (int n = (nmin < 1 ? 1 : nmin); n < big; ++n)
    if (big % n == 0) return n;
  return big;
}

int or_npfar(int nmin) { return 2 * or_npfa((nmin + 1) / 2); }

/* Two-window set-up of one offset class, src/q-psdm.cu:1593-1645, with the */
/* reference's own float/int types so truncations land on the same values.  */
void or_window_params(int ntl, float dt, float f1, float f3, float f4, or_window *w) {
  float t00, t11, tdmid, tdstart1, tdstart2, df, dw;
  int it00pre = 0.48 / dt;
  int it00suf = 0.36 / dt;
  int itb1, itb2, nsb;
  t00 = 0;
  t11 = (ntl - 1) * dt;
  nsb = (t11 - t00 - it00pre * dt - it00suf * dt) / 2.0 / dt;
  itb1 = t00 / dt;
  tdstart1 = itb1 * dt;
  tdmid = t00 + it00pre * dt + nsb * dt;
  nsb = nsb + it00pre + it00suf;
  nsb = or_npfar(nsb);
  itb2 = t11 / dt - nsb + 1;
  tdstart2 = itb2 * dt;
  df = 1.0 / nsb / dt;
  dw = 2.0 * M_PI * df;
  w->nsb = nsb;
  w->itb1 = itb1;
  w->itb2 = itb2;
  w->tdstart1 = tdstart1;
  w->tdstart2 = tdstart2;
  w->tdmid = tdmid;
  w->df = df;
  w->dw = dw;
  w->nf1 = (int)(f1 / df);
  w->nf4 = (int)(f4 / df);
  w->nww = w->nf4 - w->nf1 + 1;
  w->ncoef1 = (int)(10.0 / df);
  /* per-point high-cut bin, src/q-psdm.cu:1640-1644 (f3 -> nf3) */
  w->nf3 = (int)(f3 / df + 0.5);
}

/* Gain-limiter coefficients, src/q-psdm.cu:1277-1283. */
void or_gain_coefs(float threshold, float *lnG, float *c1, float *c2, float *c3, float *c4) {
  float l = logf(threshold);
  *lnG = l;
  *c1 = threshold * (1.0 - l - 2.5 * l * l);
  *c2 = threshold * (1.0 + 5.0 * l);
  *c3 = -2.5 * threshold;
  *c4 = 1.1 * threshold;
}

/* Gain-limited exp(x): e^x up to ln G, then the quadratic c1 + c2 x + c3 x^2
 * (C1-continuous at ln G, reaching its maximum 1.1 G at ln G + 0.2), then
 * the cap 1.1 G.  The coefficients are the reference's (src/q-psdm.cu:
 * 1277-1283); evaluated here in FP64. */
double or_gain_limit(double x, double G) {
  double L = log(G);
  if (x <= L) return exp(x);
  if (x <= L + 0.2) {
    double c1 = G * (1.0 - L - 2.5 * L * L), c2 = G * (1.0 + 5.0 * L), c3 = -2.5 * G;
    return c1 + c2 * x + c3 * x * x;
  }
  return 1.1 * G;
}

/* Azimuth-sector table, src/q-psdm.cu:1167-1193 (same float/double mix). */
void or_iaztab(int *iaztab, int naztab) {
  float eps = 0.001f, c1, c2, c3;
  int i, i1, i2, iaz;
  c2 = (1.0 - cosf(10.0 * 0.0174532925199)) * M_PI * 0.5;
  c2 = 1.0 - cosf(c2 * 0.5);
  i2 = (int)(c2 / eps);
  for (i = 0; i < i2 && i < naztab; ++i) iaztab[i] = 0;
  for (iaz = 1; iaz < 18; ++iaz) {
    c2 = iaz * 10.0;
    c1 = c2 - 10.0;
    c3 = c2 + 10.0;
    c1 = (1.0 - cosf(c1 * 0.0174532925199)) * M_PI * 0.5;
    c2 = (1.0 - cosf(c2 * 0.0174532925199)) * M_PI * 0.5;
    c3 = (1.0 - cosf(c3 * 0.0174532925199)) * M_PI * 0.5;
    c1 = 1.0 - cosf(c2 - (c2 - c1) * 0.5);
    c2 = 1.0 - cosf(c2 + (c3 - c2) * 0.5);
    i1 = (int)(c1 / eps);
    i2 = (int)(c2 / eps);
    for (i = i1; i < i2 && i < naztab; ++i) iaztab[i] = iaz;
  }
  for (i = i2; i < naztab; ++i) iaztab[i] = 0;
}

/* Azimuth index of a trace, src/q-psdm.cu:2125-2144. */
int or_azimuth_index(const int *iaztab, double sx, double sy, double gx, double gy) {
  double ddx = gx - sx, ddy = gy - sy;
  double offset = sqrtf(ddx * ddx + ddy * ddy);
  float cosa, cosx, c1;
  if (offset > 1.0) {
    cosa = ddx / offset;
    cosx = fabs(cosa);
    if (ddx * ddy < 0.0) cosx = -cosx;
    c1 = 1.0 - cosx;
    return iaztab[(int)(c1 / 0.001f)];
  }
  return 0;
}

/* Window copy with zero fill and 200-sample raised-cosine end tapers,
 * src/psdmpkg.cc:208-231. */
int or_fetch_and_taper(const float *data, int ns, int it1, int it2, float *out) {
  int it, ns1;
  float c;
  for (it = it1; it <= it2; ++it) out[it - it1] = (it < 0 || it >= ns) ? 0.0f : data[it];
  ns1 = it2 - it1 + 1;
  for (it = 0; it < 200 && it < ns1; ++it) {
    c = 0.5 - 0.5 * cosf(M_PI * it / 200.0);
    out[it] *= c;
    out[ns1 - 1 - it] *= c;
  }
  return 1;
}

/* Forward real DFT (sign -1, unnormalised, the PFA convention documented at
 * src/fft.cc:1838-1850) of bins [m0, m0+nb), in FP64 with exact index
 * reduction of the twiddles. */
static void dft_bins(const float *x, int n, int m0, int nb, double *re, double *im) {
  double *cs = (double *)malloc(sizeof(double) * 2 * n);
  for (int a = 0; a < n; ++a) {
    cs[2 * a] = cos(2.0 * M_PI * a / n);
    cs[2 * a + 1] = sin(2.0 * M_PI * a / n);
  }
  for (int k = 0; k < nb; ++k) {
    long m = m0 + k;
    double sr = 0, si = 0;
    for (int j = 0; j < n; ++j) {
      long a = (m * j) % n;
      sr += x[j] * cs[2 * a];
      si -= x[j] * cs[2 * a + 1];
    }
    re[k] = sr;
    im[k] = si;
  }
  free(cs);
}

/* Per-trace Q pre-processing of the reference, src/q-psdm.cu:2173-2216:
 * t^2 gain (x0.01 before 0.1 s), full-trace taper, two windowed and tapered
 * segments, forward FFT of each, bins nf1.. multiplied by -i*omega
 * (re = w*Im, im = -w*Re, with the reference's float accumulation of w).
 * Output re[2*nww], im[2*nww]: segment 1 then segment 2. */
void or_trace_spectra(const float *trace, int nt, float dt, const or_window *w,
                      float *re, float *im) {
  float *g = (float *)malloc(sizeof(float) * nt);
  float *d = (float *)malloc(sizeof(float) * nt);
  float *seg = (float *)malloc(sizeof(float) * w->nsb);
  double *xr = (double *)malloc(sizeof(double) * w->nww);
  double *xi = (double *)malloc(sizeof(double) * w->nww);
  for (int it = 0; it < nt; ++it) {
    g[it] = trace[it];
    if (it * dt < 0.1)
      g[it] *= 0.1 * 0.1;
    else
      g[it] *= (it * dt) * (it * dt);
  }
  or_fetch_and_taper(g, nt, 0, nt - 1, d);
  for (int s = 0; s < 2; ++s) {
    int itb = s == 0 ? w->itb1 : w->itb2;
    or_fetch_and_taper(d, nt, itb, itb + w->nsb - 1, seg);
    dft_bins(seg, w->nsb, w->nf1, w->nww, xr, xi);
    float wv = w->nf1 * w->dw;
    for (int iw = 0; iw < w->nww; ++iw, wv += w->dw) {
      re[s * w->nww + iw] = +(float)xi[iw] * wv;
      im[s * w->nww + iw] = -(float)xr[iw] * wv;
    }
  }
  free(g);
  free(d);
  free(seg);
  free(xr);
  free(xi);
}

/* High-cut taper of bin m: the reference's derivative-filter shape
 * (src/q-psdm.cu:1236-1247): 1 below nf3, cos^2 ramp from nf3 to nf4,
 * 0 above nf4. */
static double hicut(int m, int nf3, int nf4) {
  if (m > nf4) return 0.0;
  if (m < nf3 || nf4 <= nf3) return 1.0;
  double c = cos((double)(m - nf3) / (nf4 - nf3) * M_PI * 0.5);
  return c * c;
}

/* Q filter bank of one trace (this build's formula, in the reference's
 * conventions).  For t* level j (t*_j = j*dtstar) and segment s the
 * filtered spectrum is C_s(w) * H_j(w) with
 *     H_j(w) = hicut(w) * g(w t*_j / 2) * exp(-i w t*_j ln(w/w_r) / pi),
 * g the gain limiter and w_r = 2 pi fc; the level's time trace at sample it
 * is the unnormalised one-sided inverse transform of segment s at
 * n = it - itb_s, where s = 2 when it*dt > tdmid (the kernel's segment
 * select, src/q-psdm.cu:2361-2365):
 *     bank_j[it] = 2 Re sum_k C_s[k] H_j(w_k) exp(+2 pi i (nf1+k) n / nsb).
 * The image is divided by nsb on output (src/q-psdm.cu:1760-1766). */
#define QBANK_BODY(REAL)                                                          \
  const or_window *w = &p->win;                                                   \
  int nww = w->nww, nsb = w->nsb, K = p->K, nt = p->nt;                           \
  double *cs = (double *)malloc(sizeof(double) * 2 * nsb);                        \
  for (int a = 0; a < nsb; ++a) {                                                 \
    cs[2 * a] = cos(2.0 * M_PI * a / nsb);                                        \
    cs[2 * a + 1] = sin(2.0 * M_PI * a / nsb);                                    \
  }                                                                               \
  double dwx = 2.0 * M_PI / ((double)nsb * p->dt);                                \
  double wr = 2.0 * M_PI * p->fc;                                                 \
  double *hr = (double *)malloc(sizeof(double) * K * nww);                        \
  double *hi = (double *)malloc(sizeof(double) * K * nww);                        \
  for (int j = 0; j < K; ++j)                                                     \
    for (int k = 0; k < nww; ++k) {                                               \
      int m = w->nf1 + k;                                                         \
      double om = m * dwx, ts = j * (double)p->dtstar;                            \
      double amp = hicut(m, w->nf3, w->nf4) * or_gain_limit(om * ts * 0.5, p->threshold); \
      double ph = om > 0 ? -om * ts * log(om / wr) / M_PI : 0.0;                   \
      hr[j * nww + k] = amp * cos(ph);                                            \
      hi[j * nww + k] = amp * sin(ph);                                            \
    }                                                                             \
  _Pragma("omp parallel") {                                                       \
    float *re = (float *)malloc(sizeof(float) * 2 * nww);                         \
    float *im = (float *)malloc(sizeof(float) * 2 * nww);                         \
    REAL *fr = (REAL *)malloc(sizeof(REAL) * 2 * nww);                            \
    REAL *fi = (REAL *)malloc(sizeof(REAL) * 2 * nww);                            \
    _Pragma("omp for schedule(dynamic)")                                          \
    for (int itr = 0; itr < ntr; ++itr) {                                         \
      or_trace_spectra(traces + (size_t)itr * nt, nt, p->dt, w, re, im);          \
      for (int j = 0; j < K; ++j) {                                               \
        for (int q = 0; q < 2 * nww; ++q) {                                       \
          int k = q % nww;                                                        \
          double a = re[q], b = im[q], c = hr[j * nww + k], d = hi[j * nww + k];  \
          fr[q] = (REAL)(a * c - b * d);                                          \
          fi[q] = (REAL)(a * d + b * c);                                          \
        }                                                                         \
        for (int it = 0; it < nt; ++it) {                                         \
          int s = ((float)it * p->dt > w->tdmid) ? 1 : 0;                         \
          int n = it - (s ? w->itb2 : w->itb1);                                   \
          REAL acc = 0;                                                           \
          if (n >= 0 && n < nsb) {                                                \
            for (int k = 0; k < nww; ++k) {                                       \
              long a = ((long)(w->nf1 + k) * n) % nsb;                            \
              acc += fr[s * nww + k] * (REAL)cs[2 * a] - fi[s * nww + k] * (REAL)cs[2 * a + 1]; \
            }                                                                     \
          }                                                                       \
          bank[((size_t)itr * K + j) * nt + it] = 2 * acc;                        \
        }                                                                         \
      }                                                                           \
    }                                                                             \
    free(re);                                                                     \
    free(im);                                                                     \
    free(fr);                                                                     \
    free(fi);                                                                     \
  }                                                                               \
  free(cs);                                                                       \
  free(hr);                                                                       \
  free(hi);

void or_qbank_f64(const float *traces, int ntr, const or_bparams *p, double *bank) {
  QBANK_BODY(double)
}

void or_qbank_f32(const float *traces, int ntr, const or_bparams *p, float *bank) {
  QBANK_BODY(float)
}

/* ---------------------------------------------------------------------- */
/* Straight-ray traveltime and t* (this build's table builder; the         */
/* reference ships tables but no builder).  t = int slowness dl and         */
/* t* = int slowness / Q dl along the straight segment from the surface     */
/* point (sx, sy, 0) to (x, y, z).                                          */
/* ---------------------------------------------------------------------- */
static const double GLX[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290,
                              -0.1834346424956498, 0.1834346424956498,  0.5255324099163290,
                              0.7966664774136267,  0.9602898564975363};
static const double GLW[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873,
                              0.3626837833783620, 0.3626837833783620, 0.3137066458778873,
                              0.2223810344533745, 0.1012285362903763};

void or_ray_times(const or_model *m, double sx, double sy, double x, double y, double z,
                  double *t, double *tstar) {
  double hx = x - sx, hy = y - sy;
  double L = sqrt(hx * hx + hy * hy + z * z);
  if (m->kind == 0) {
    /* linear v(z): t = L/(k z) ln(v(z)/v0), L/v0 when k z -> 0 */
    double tt;
    if (m->kgrad * z > 1e-9 * m->v0)
      tt = L / (m->kgrad * z) * log((m->v0 + m->kgrad * z) / m->v0);
    else
      tt = L / m->v0;
    *t = tt;
    *tstar = tt / m->q0;
    return;
  }
  /* layers + Gaussian anomaly: split at layer boundaries, 8-point
     Gauss-Legendre on each piece */
  double acc_t = 0, acc_q = 0;
  int nb = 0;
  double lam[20];
  lam[nb++] = 0.0;
  for (int l = 1; l < m->nlayer; ++l) {
    double zb = l * m->layer_dz;
    if (zb < z) lam[nb++] = zb / z;
  }
  lam[nb++] = 1.0;
  double is2 = 1.0 / (2.0 * m->asig * m->asig);
  for (int sgi = 0; sgi + 1 < nb; ++sgi) {
    double a = lam[sgi], b = lam[sgi + 1];
    if (b <= a) continue;
    double zm = 0.5 * (a + b) * z;
    int l = (int)(zm / m->layer_dz);
    if (l >= m->nlayer) l = m->nlayer - 1;
    if (l < 0) l = 0;
    double half = 0.5 * (b - a), mid = 0.5 * (a + b);
    for (int q = 0; q < 8; ++q) {
      double la = mid + half * GLX[q];
      double px = sx + la * hx, py = sy + la * hy, pz = la * z;
      double dx = px - m->ax, dy = py - m->ay, dz = pz - m->az;
      double G = exp(-(dx * dx + dy * dy + dz * dz) * is2);
      double v = m->vl[l] * (1.0 - m->av * G);
      double qq = m->ql[l] * (1.0 - m->aq * G);
      acc_t += GLW[q] * half / v;
      acc_q += GLW[q] * half / (v * qq);
    }
  }
  *t = acc_t * L;
  *tstar = acc_q * L;
}

/* Spreading integral sigma = int v ds on the same straight segment (the
 * amplitude target A = 1/sigma of the learned surrogate); closed form
 * L (v0 + k z / 2) for linear v(z), the same Gauss-Legendre pieces otherwise. */
double or_ray_sigma(const or_model *m, double sx, double sy, double x, double y, double z) {
  double hx = x - sx, hy = y - sy;
  double L = sqrt(hx * hx + hy * hy + z * z);
  if (m->kind == 0) return L * (m->v0 + 0.5 * m->kgrad * z);
  double acc = 0;
  int nb = 0;
  double lam[20];
  lam[nb++] = 0.0;
  for (int l = 1; l < m->nlayer; ++l) {
    double zb = l * m->layer_dz;
    if (zb < z) lam[nb++] = zb / z;
  }
  lam[nb++] = 1.0;
  double is2 = 1.0 / (2.0 * m->asig * m->asig);
  for (int sgi = 0; sgi + 1 < nb; ++sgi) {
    double a = lam[sgi], b = lam[sgi + 1];
    if (b <= a) continue;
    double zm = 0.5 * (a + b) * z;
    int l = (int)(zm / m->layer_dz);
    if (l >= m->nlayer) l = m->nlayer - 1;
    if (l < 0) l = 0;
    double half = 0.5 * (b - a), mid = 0.5 * (a + b);
    for (int q = 0; q < 8; ++q) {
      double la = mid + half * GLX[q];
      double dx = sx + la * hx - m->ax, dy = sy + la * hy - m->ay, dz = la * z - m->az;
      double G = exp(-(dx * dx + dy * dy + dz * dz) * is2);
      acc += GLW[q] * half * m->vl[l] * (1.0 - m->av * G);
    }
  }
  return acc * L;
}

/* Table set: node (jn, in) sits at (n0x + in*ndx, n0y + jn*ndy); its table
 * holds (t, t*) at relative offsets (u, v) = (u0 + iu*du, v0 + iv*dv) and
 * depth w = iw*dw, layout [node][iv][iu][iw][2] (z fastest, as the
 * reference's "z contiguous" tables, src/q-psdm.cu:1158-1159). */
void or_build_tables(const or_model *m, const or_tgeom *g, float *tab) {
  int nn = g->nnx * g->nny;
  size_t per = (size_t)g->nv * g->nu * g->nw;
#pragma omp parallel for collapse(2) schedule(dynamic)
  for (int n = 0; n < nn; ++n)
    for (int iv = 0; iv < g->nv; ++iv) {
      double px = g->n0x + (double)(n % g->nnx) * g->ndx;
      double py = g->n0y + (double)(n / g->nnx) * g->ndy;
      for (int iu = 0; iu < g->nu; ++iu)
        for (int iw = 0; iw < g->nw; ++iw) {
          double t, ts;
          or_ray_times(m, px, py, px + g->u0 + (double)iu * g->du,
                       py + g->v0 + (double)iv * g->dv, (double)iw * g->dw, &t, &ts);
          size_t o = (size_t)n * per + ((size_t)iv * g->nu + iu) * g->nw + iw;
          tab[2 * o] = (float)t;
          tab[2 * o + 1] = (float)ts;
        }
    }
}

static int nearest_node(double p, float o, float d, int n) {
  int i = (int)floor((p - o) / d + 0.5);
  return i < 0 ? 0 : (i > n - 1 ? n - 1 : i);
}

static void frac_index(double q, int n, int *i0, double *a) {
  if (q < 0) q = 0;
  if (q > n - 1) q = n - 1;
  int i = (int)q;
  if (i > n - 2) i = n - 2;
  *i0 = i;
  *a = q - i;
}

/* Nearest-node, shifted-table lookup: (t, t*) from position (px, py, 0) to
 * (x, y, z) = trilinear interpolation in the nearest node's table at the
 * relative position (x - px, y - py, z). */
/* Trilinear (t, t*) of one relative-coordinate grid T [nv][nu][nw][2]
 * (origin (u0, v0), spacing (du, dv, dw)) at the relative position
 * (rx, ry, z), indices clamped to the grid. */
static void lookup_grid(const float *T, int nu, int nv, int nw, double u0, double v0, double du,
                        double dv, double dw, double rx, double ry, double z, double *t,
                        double *tstar) {
  int iu, iv, iw;
  double a, b, c;
  frac_index((rx - u0) / du, nu, &iu, &a);
  frac_index((ry - v0) / dv, nv, &iv, &b);
  frac_index(z / dw, nw, &iw, &c);
  double r[2] = {0, 0};
  for (int k = 0; k < 8; ++k) {
    int ku = k & 1, kv = (k >> 1) & 1, kz = (k >> 2) & 1;
    double wgt = (ku ? a : 1 - a) * (kv ? b : 1 - b) * (kz ? c : 1 - c);
    size_t o = ((size_t)(iv + kv) * nu + (iu + ku)) * nw + (iw + kz);
    r[0] += wgt * T[2 * o];
    r[1] += wgt * T[2 * o + 1];
  }
  *t = r[0];
  *tstar = r[1];
}

/* Nearest-node, shifted-table lookup: (t, t*) from position (px, py, 0) to
 * (x, y, z) = trilinear interpolation in the nearest node's table at the
 * relative position (x - px, y - py, z). */
void or_lookup(const or_tgeom *g, const float *tab, double px, double py, double x,
               double y, double z, double *t, double *tstar) {
  int in = nearest_node(px, g->n0x, g->ndx, g->nnx);
  int jn = nearest_node(py, g->n0y, g->ndy, g->nny);
  const float *T = tab + 2 * (size_t)(jn * g->nnx + in) * g->nv * g->nu * g->nw;
  lookup_grid(T, g->nu, g->nv, g->nw, g->u0, g->v0, g->du, g->dv, g->dw, x - px, y - py, z, t, tstar);
}

/* A trace belongs to a per-shot source table when its source is the
 * table's own, bit for bit (the table was built for that one position). */
static int same_source(const float *geom_row, const or_src *src) {
  return memcmp(geom_row, src->hdr, 2 * sizeof(float)) == 0;
}

/* ---------------------------------------------------------------------- */
/* Kirchhoff sum.  Per (trace, point): distances and bisector as in        */
/* src/q-psdm.cu:2303-2320; dip-bound taper widened by taperzone*(1+tan^2) */
/* with linear ramps, src/q-psdm.cu:2322-2340; traveltimes and t* from the */
/* tables (the reference stubs them at 1.0); segment/level pick of the     */
/* filtered trace by t*, linear interpolation in time with the reference's */
/* bounds it1 > 0 && it2 < nt (src/q-psdm.cu:2354-2357); weight            */
/* w*(tr + 1/tr), tr = (ds/dg)^3 tg/ts (src/q-psdm.cu:2372-2374); image    */
/* z fastest, traces accumulated in order.                                  */
/* geom: per trace (sx, sy, gx, gy) in image-relative metres; a NaN sx     */
/* marks a trace the extent check rejected (src/q-psdm.cu:1379-1383).       */
/* Returns the update count (pairs inside the dip aperture).                */
/* ---------------------------------------------------------------------- */
/* Bank level of a two-leg t* (in units of the level spacing): the nearest
 * level, round(t* / dt*) (this build's pick; the reference's kernel takes
 * the t* table qqq and the spacing arguments but is a stub,
 * src/q-psdm.cu:2246-2259).  The other rules exist only so that tests can
 * plant a wrong pick in the FP64 sum and show that a known answer fails. */
static int g_level_rule = 0; /* 0 round (the product's), 1 floor, 2 level 0 */
void or_set_level_rule(int rule) { g_level_rule = rule; }
static inline int level_pick(double x) {
  if (g_level_rule == 1) return (int)floor(x);
  if (g_level_rule == 2) return 0;
  return (int)(x + 0.5);
}

#define IMAGE_BODY(REAL, SQRT)                                                             \
  int64_t count = 0;                                                                       \
  const or_tgeom *g = &p->tg;                                                              \
  REAL tz = p->taperzone;                                                                  \
  REAL tx1 = p->tanx1, tx2 = p->tanx2, ty1 = p->tany1, ty2 = p->tany2;                     \
  REAL tx1c = tx1 - tz * (1 + tx1 * tx1), tx2c = tx2 + tz * (1 + tx2 * tx2);               \
  REAL ty1c = ty1 - tz * (1 + ty1 * ty1), ty2c = ty2 + tz * (1 + ty2 * ty2);               \
  _Pragma("omp parallel for schedule(dynamic) reduction(+:count)")                         \
  for (int iy = 0; iy < p->ny; ++iy) {                                                     \
    for (int itr = 0; itr < ntr; ++itr) {                                                  \
      REAL sx = geom[4 * itr], sy = geom[4 * itr + 1];                                     \
      REAL gx = geom[4 * itr + 2], gy = geom[4 * itr + 3];                                 \
      if (sx != sx) continue; /* rejected by the extent check */                           \
      if (src && !same_source(geom + 4 * itr, src)) continue;                              \
      const BANK *b = bank + (size_t)itr * p->K * p->nt;                                   \
      REAL y = (REAL)iy * p->dy;                                                           \
      for (int ix = 0; ix < p->nx; ++ix) {                                                 \
        REAL x = (REAL)ix * p->dx;                                                         \
        REAL dsx = sx - x, dsy = sy - y, dgx = gx - x, dgy = gy - y;                       \
        for (int iz = 1; iz < p->nz; ++iz) {                                               \
          REAL z = (REAL)iz * p->dz;                                                       \
          REAL ds = SQRT(dsx * dsx + dsy * dsy + z * z);                                   \
          REAL dg = SQRT(dgx * dgx + dgy * dgy + z * z);                                   \
          REAL X = dsx / ds + dgx / dg, Y = dsy / ds + dgy / dg, Z = z / ds + z / dg;      \
          REAL tanx = X / Z, tany = Y / Z;                                                 \
          if (tanx <= tx1c || tx2c <= tanx || tany <= ty1c || ty2c <= tany) continue;      \
          ++count;                                                                         \
          REAL wt = 1;                                                                     \
          if (tanx < tx1) wt *= (tanx - tx1c) / (tx1 - tx1c);                              \
          if (tanx > tx2) wt *= (tx2c - tanx) / (tx2c - tx2);                              \
          if (tany < ty1) wt *= (tany - ty1c) / (ty1 - ty1c);                              \
          if (tany > ty2) wt *= (ty2c - tany) / (ty2c - ty2);                              \
          double a1, a2, b1, b2;                                                           \
          if (src)                                                                         \
            lookup_grid(src->tab, src->nus, src->nvs, g->nw, src->hdr[2], src->hdr[3], g->du,  \
                        g->dv, g->dw, x - sx, y - sy, z, &a1, &a2);                          \
          else                                                                             \
            or_lookup(g, tab, sx, sy, x, y, z, &a1, &a2);                                  \
          or_lookup(g, tab, gx, gy, x, y, z, &b1, &b2);                                    \
          REAL ts = (REAL)a1, tg = (REAL)b1;                                               \
          if (ts < p->tmin) ts = p->tmin;                                                  \
          if (tg < p->tmin) tg = p->tmin;                                                  \
          REAL t = ts + tg, tst = (REAL)a2 + (REAL)b2;                                     \
          int j = level_pick((double)tst / (double)p->dtstar);                             \
          if (j < 0) j = 0;                                                                \
          if (j > p->K - 1) j = p->K - 1;                                                  \
          REAL f = t / p->dt;                                                              \
          int i0 = (int)f;                                                                 \
          if (!(i0 > 0 && i0 + 1 < p->nt)) continue;                                       \
          REAL fr = f - (REAL)i0;                                                          \
          const BANK *bj = b + (size_t)j * p->nt;                                          \
          REAL sv = ((REAL)1 - fr) * (REAL)bj[i0] + fr * (REAL)bj[i0 + 1];                 \
          REAL r = ds / dg;                                                                \
          REAL tr = r * r * r * tg / ts;                                                   \
          REAL W = wt * (tr + (REAL)1 / tr);                                               \
          img[((size_t)iy * p->nx + ix) * p->nz + iz] += W * sv;                           \
        }                                                                                  \
      }                                                                                    \
    }                                                                                      \
  }                                                                                        \
  return count;

int64_t or_image_src_f64(const double *bank, const float *geom, int ntr, const float *tab,
                         const or_src *src, const or_iparams *p, double *img) {
  typedef double BANK;
  IMAGE_BODY(double, sqrt)
}

int64_t or_image_f64(const double *bank, const float *geom, int ntr, const float *tab,
                     const or_iparams *p, double *img) {
  return or_image_src_f64(bank, geom, ntr, tab, NULL, p, img);
}


/* FP32 CPU path of the sum (the timed CPU baseline and the FP32 parity
 * partner of the GPU form): the same formula evaluated in float with the
 * GPU's order of operations -- reciprocal spacings, bilinear (x, y) table
 * interpolation at two depth nodes then a lerp in z, 1/sqrt distances,
 * reciprocal weights. */
static inline void lookup_f32_grid(const float *T, int nu, int nv, int nw, float u0, float v0, float idu,
                                   float idv, float idw, float px, float py, float x, float y, float z,
                                   float *t, float *ts) {
  float q = fminf(fmaxf((x - px - u0) * idu, 0.0f), (float)(nu - 1));
  int iu = (int)q < nu - 2 ? (int)q : nu - 2;
  const float a = q - (float)iu;
  q = fminf(fmaxf((y - py - v0) * idv, 0.0f), (float)(nv - 1));
  int iv = (int)q < nv - 2 ? (int)q : nv - 2;
  const float b = q - (float)iv;
  const float fw = fminf(z * idw, (float)(nw - 1));
  int w = (int)fw < nw - 2 ? (int)fw : nw - 2;
  const float c = fw - (float)w;
  float col[2][2];
  for (int k = 0; k < 2; ++k) {
#define TV(dv_, du_, comp_) T[2 * (((size_t)(iv + dv_) * nu + (iu + du_)) * nw + (w + k)) + comp_]
    for (int comp = 0; comp < 2; ++comp) {
      const float lx = fmaf(a, TV(0, 1, comp) - TV(0, 0, comp), TV(0, 0, comp));
      const float hx = fmaf(a, TV(1, 1, comp) - TV(1, 0, comp), TV(1, 0, comp));
      col[k][comp] = fmaf(b, hx - lx, lx);
    }
#undef TV
  }
  *t = fmaf(c, col[1][0] - col[0][0], col[0][0]);
  *ts = fmaf(c, col[1][1] - col[0][1], col[0][1]);
}

static inline void lookup_f32(const or_tgeom *g, const float *tab, float px, float py, float x,
                              float y, float z, float *t, float *ts) {
  const float inx = 1.0f / g->ndx, iny = 1.0f / g->ndy;
  int in = (int)floorf((px - g->n0x) * inx + 0.5f);
  int jn = (int)floorf((py - g->n0y) * iny + 0.5f);
  in = in < 0 ? 0 : (in > g->nnx - 1 ? g->nnx - 1 : in);
  jn = jn < 0 ? 0 : (jn > g->nny - 1 ? g->nny - 1 : jn);
  const float *T = tab + 2 * (size_t)(jn * g->nnx + in) * g->nv * g->nu * g->nw;
  lookup_f32_grid(T, g->nu, g->nv, g->nw, g->u0, g->v0, 1.0f / g->du, 1.0f / g->dv, 1.0f / g->dw, px, py, x, y,
                  z, t, ts);
}

/* Source leg of the FP32 path: the per-shot source table when given (read
 * at the table's own source position), else the nearest node's table. */
static inline void source_f32(const or_tgeom *g, const float *tab, const or_src *src, float sx, float sy,
                              float x, float y, float z, float *t, float *ts) {
  if (src)
    lookup_f32_grid(src->tab, src->nus, src->nvs, g->nw, src->hdr[2], src->hdr[3], 1.0f / g->du, 1.0f / g->dv,
                    1.0f / g->dw, sx, sy, x, y, z, t, ts);
  else
    lookup_f32(g, tab, sx, sy, x, y, z, t, ts);
}

static inline float sat01(float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); }

int64_t or_image_f32(const float *bank, const float *geom, int ntr, const float *tab,
                     const or_iparams *p, float *img) {
  return or_image_src_f32(bank, geom, ntr, tab, NULL, p, img);
}

static int64_t image_f32_impl(const float *bank, const float *geom, int ntr, const float *tab,
                              const or_src *src, const or_iparams *p, int skip_lo, int skip_hi, float *img);

int64_t or_image_src_f32(const float *bank, const float *geom, int ntr, const float *tab,
                         const or_src *src, const or_iparams *p, float *img) {
  return image_f32_impl(bank, geom, ntr, tab, src, p, 0, 0, img);
}

/* FP32 path over the points outside [skip_lo, skip_hi). */
static int64_t image_f32_impl(const float *bank, const float *geom, int ntr, const float *tab,
                              const or_src *src, const or_iparams *p, int skip_lo, int skip_hi, float *img) {
  int64_t count = 0;
  const or_tgeom *g = &p->tg;
  const float tz = p->taperzone;
  const float tx1 = p->tanx1, tx2 = p->tanx2, ty1 = p->tany1, ty2 = p->tany2;
  const float tx1c = tx1 - tz * (1 + tx1 * tx1), tx2c = tx2 + tz * (1 + tx2 * tx2);
  const float ty1c = ty1 - tz * (1 + ty1 * ty1), ty2c = ty2 + tz * (1 + ty2 * ty2);
  const float rx1 = 1.0f / (tx1 - tx1c), rx2 = 1.0f / (tx2c - tx2);
  const float ry1 = 1.0f / (ty1 - ty1c), ry2 = 1.0f / (ty2c - ty2);
  const float cx1 = -tx1c * rx1, cx2 = tx2c * rx2, cy1 = -ty1c * ry1, cy2 = ty2c * ry2;
  const float inv_dt = 1.0f / p->dt, inv_dts = 1.0f / p->dtstar;
#pragma omp parallel for schedule(dynamic) reduction(+ : count)
  for (int iy = 0; iy < p->ny; ++iy) {
    for (int itr = 0; itr < ntr; ++itr) {
      const float sx = geom[4 * itr], sy = geom[4 * itr + 1];
      const float gx = geom[4 * itr + 2], gy = geom[4 * itr + 3];
      if (sx != sx) continue;
      if (src && !same_source(geom + 4 * itr, src)) continue;
      const float *bt = bank + (size_t)itr * p->K * p->nt;
      const float y = (float)iy * p->dy;
      for (int ix = 0; ix < p->nx; ++ix) {
        const float x = (float)ix * p->dx;
        const float dsx = sx - x, dsy = sy - y, dgx = gx - x, dgy = gy - y;
        const float hs2 = dsx * dsx + dsy * dsy, hg2 = dgx * dgx + dgy * dgy;
        for (int iz = 1; iz < p->nz; ++iz) {
          if (iz >= skip_lo && iz < skip_hi) continue;
          const float z = (float)iz * p->dz;
          /* source leg: t in samples, t* in levels + 1/2, bisector terms, ds^3/t */
          float ts, tss;
          source_f32(g, tab, src, sx, sy, x, y, z, &ts, &tss);
          ts = fmaxf(ts, p->tmin);
          const float d2s = hs2 + z * z;
          const float rds = 1.0f / sqrtf(d2s), ds = d2s * rds;
          const float tsn = ts * inv_dt, tsl = fmaf(tss, inv_dts, 0.5f);
          const float A = ds * ds * ds * (1.0f / ts);
          /* receiver leg and bisector (scaled by ds: (ds/dg) dg + ds); dip
           * taper as clamped ramps */
          const float rdg = 1.0f / sqrtf(hg2 + z * z);
          const float q = rdg * ds;
          const float X = fmaf(dgx, q, dsx), Y = fmaf(dgy, q, dsy), Z = fmaf(z, q, z);
          const float rz = 1.0f / Z, tanx = X * rz, tany = Y * rz;
          const float wx = fminf(sat01(fmaf(tanx, rx1, cx1)), sat01(fmaf(tanx, -rx2, cx2)));
          const float wy = fminf(sat01(fmaf(tany, ry1, cy1)), sat01(fmaf(tany, -ry2, cy2)));
          const float wt = wx * wy;
          if (!(wt > 0.0f)) continue;
          ++count;
          float tg, tsg;
          lookup_f32(g, tab, gx, gy, x, y, z, &tg, &tsg);
          tg = fmaxf(tg, p->tmin);
          const float f = fmaf(tg, inv_dt, tsn);
          const int i0 = (int)f;
          if (!(i0 > 0 && i0 + 1 < p->nt)) continue;
          int j = (int)fmaf(tsg, inv_dts, tsl);
          j = j < 0 ? 0 : (j > p->K - 1 ? p->K - 1 : j);
          const float fr = f - (float)i0;
          const float *b = bt + (size_t)j * p->nt + i0;
          const float tr = A * (rdg * rdg * rdg) * tg;
          float *o = &img[((size_t)iy * p->nx + ix) * p->nz + iz];
          *o = fmaf(wt * (tr + 1.0f / tr), fmaf(fr, b[1] - b[0], b[0]), *o);
        }
      }
    }
  }
  return count;
}

/* ---------------------------------------------------------------------- */
/* Node form of the FP32 path (the library's weight_nodes = 1): below      */
/* z_exact and within the tables' depth range, the four dip-taper ramps,   */
/* the amplitude (tr + 1/tr), the time in samples and the level are        */
/* evaluated at the tables' depth nodes and interpolated linearly to the   */
/* points between, the ramps clamped per point; elsewhere the exact FP32   */
/* path.  The FP32 restatement of the device's node body, in its order of  */
/* operations (this build's own method: the reference's kernel evaluates   */
/* every point, src/q-psdm.cu:2303-2378).                                   */
/* ---------------------------------------------------------------------- */
static int nf_interval(int iz, float dz, float inv_dw, int nw, float *frac) {
  const float z = (float)iz * dz;
  float fw = z * inv_dw;
  if (fw > (float)(nw - 1)) fw = (float)(nw - 1);
  int w = (int)fw;
  if (w > nw - 2) w = nw - 2;
  *frac = fw - (float)w;
  return w;
}

static int nf_first_point(int W, int nz, float dz, float inv_dw, int nw) {
  float f;
  int iz = (int)((float)W / (dz * inv_dw));
  iz = iz < 0 ? 0 : (iz > nz ? nz : iz);
  while (iz < nz && nf_interval(iz, dz, inv_dw, nw, &f) < W) ++iz;
  while (iz > 0 && nf_interval(iz - 1, dz, inv_dw, nw, &f) >= W) --iz;
  return iz;
}

void or_node_range(const or_iparams *p, float z_exact, int *ps_out, int *pe_out) {
  or_node_range_aligned(p, z_exact, 1, ps_out, pe_out);
}

/* The node range with its first interval rounded up to a multiple of align
 * (the span form's: 8). */
void or_node_range_aligned(const or_iparams *p, float z_exact, int align, int *ps_out, int *pe_out) {
  const int nz = p->nz, nw = p->tg.nw;
  const float dz = p->dz, inv_dw = 1.0f / p->tg.dw;
  int ws = (int)ceil((double)z_exact / (double)p->tg.dw);
  if (ws < 1) ws = 1;
  ws = (ws + align - 1) / align * align;
  int pe = nz;
  for (int iz = 0; iz < nz; ++iz) {
    const float z = (float)iz * dz;
    if (z * inv_dw > (float)(nw - 1)) {
      pe = iz;
      break;
    }
  }
  int ps = pe;
  if (ws <= nw - 2) {
    ps = nf_first_point(ws, nz, dz, inv_dw, nw);
    if (ps > pe) ps = pe;
  }
  *ps_out = ps;
  *pe_out = pe;
}

/* Bilinear (t, t*) at depth node w of grid T (no depth interpolation). */
static void lookup_f32_node(const float *T, int nu, int nv, int nw, float u0, float v0, float idu, float idv,
                            float px, float py, float x, float y, int w, float *t, float *ts) {
  float q = fminf(fmaxf((x - px - u0) * idu, 0.0f), (float)(nu - 1));
  int iu = (int)q < nu - 2 ? (int)q : nu - 2;
  const float a = q - (float)iu;
  q = fminf(fmaxf((y - py - v0) * idv, 0.0f), (float)(nv - 1));
  int iv = (int)q < nv - 2 ? (int)q : nv - 2;
  const float b = q - (float)iv;
  float r[2];
#define TV(dv_, du_, comp_) T[2 * (((size_t)(iv + dv_) * nu + (iu + du_)) * nw + w) + comp_]
  for (int comp = 0; comp < 2; ++comp) {
    const float lx = fmaf(a, TV(0, 1, comp) - TV(0, 0, comp), TV(0, 0, comp));
    const float hx = fmaf(a, TV(1, 1, comp) - TV(1, 0, comp), TV(1, 0, comp));
    r[comp] = fmaf(b, hx - lx, lx);
  }
#undef TV
  *t = r[0];
  *ts = r[1];
}

static void node_table_f32(const or_tgeom *g, const float *tab, const or_src *src, int is_source, float px,
                           float py, float x, float y, int w, float *t, float *ts) {
  if (is_source && src) {
    lookup_f32_node(src->tab, src->nus, src->nvs, g->nw, src->hdr[2], src->hdr[3], 1.0f / g->du, 1.0f / g->dv, px,
                    py, x, y, w, t, ts);
    return;
  }
  int in = (int)floorf((px - g->n0x) * (1.0f / g->ndx) + 0.5f);
  int jn = (int)floorf((py - g->n0y) * (1.0f / g->ndy) + 0.5f);
  in = in < 0 ? 0 : (in > g->nnx - 1 ? g->nnx - 1 : in);
  jn = jn < 0 ? 0 : (jn > g->nny - 1 ? g->nny - 1 : jn);
  const float *T = tab + 2 * (size_t)(jn * g->nnx + in) * g->nv * g->nu * g->nw;
  lookup_f32_node(T, g->nu, g->nv, g->nw, g->u0, g->v0, 1.0f / g->du, 1.0f / g->dv, px, py, x, y, w, t, ts);
}

typedef struct {
  float F, L, u[4], a;
} nf_node;

/* Ramps, amplitude, time and level of a column at depth node wn (one
 * node of the node form). */
typedef struct {
  const or_tgeom *g;
  const float *tab;
  const or_src *src;
  float sx, sy, gx, gy, x, y, dsx, dsy, dgx, dgy, hs2, hg2;
  float tmin, inv_dt, inv_dts, rx1, rx2, ry1, ry2, cx1, cx2, cy1, cy2;
} nf_col;

static nf_node nf_value(const nf_col *C, int wn) {
  const float zn = (float)wn * C->g->dw;
  float ts, tss, tg, tsg;
  node_table_f32(C->g, C->tab, C->src, 1, C->sx, C->sy, C->x, C->y, wn, &ts, &tss);
  node_table_f32(C->g, C->tab, C->src, 0, C->gx, C->gy, C->x, C->y, wn, &tg, &tsg);
  ts = fmaxf(ts, C->tmin);
  const float d2 = fmaf(zn, zn, C->hs2);
  const float ds = d2 * (1.0f / sqrtf(d2));
  const float tsn = ts * C->inv_dt, tsl = fmaf(tss, C->inv_dts, 0.5f), its = 1.0f / ts;
  const float rdg = 1.0f / sqrtf(fmaf(zn, zn, C->hg2));
  const float q = ds * rdg;
  const float X = fmaf(C->dgx, q, C->dsx), Y = fmaf(C->dgy, q, C->dsy), Z = fmaf(zn, q, zn);
  const float rz = 1.0f / Z, tanx = X * rz, tany = Y * rz;
  const float t = fmaxf(tg, C->tmin);
  const float tr = (q * q) * q * (t * its);
  nf_node v;
  v.u[0] = fmaf(tanx, C->rx1, C->cx1);
  v.u[1] = fmaf(tanx, -C->rx2, C->cx2);
  v.u[2] = fmaf(tany, C->ry1, C->cy1);
  v.u[3] = fmaf(tany, -C->ry2, C->cy2);
  v.a = tr + 1.0f / tr;
  v.F = fmaf(t, C->inv_dt, tsn);
  v.L = fmaf(tsg, C->inv_dts, tsl);
  return v;
}

/* The node form (spans == NULL) and the span form (this build's own
 * method, like the node form).  Span form: the time and the level are
 * interpolated between the interval's two depth nodes as in the node form,
 * with the level clamped to [0, K-1] at the nodes; the four ramps and the
 * amplitude are interpolated between the weight nodes w0 = spans[2w] and
 * w1 = spans[2w + 1] of the interval's span (w0 <= w < w1) linearly in 1/z:
 * at the fraction ((z - z0) z1) / ((z1 - z0) z) of a point at depth z,
 * z0 = w0 dw, z1 = w1 dw; the ramps are clamped per point. */
static int64_t nodes_impl(const float *bank, const float *geom, int ntr, const float *tab, const or_src *src,
                          const or_iparams *p, float z_exact, const int *spans, float *img) {
  int ps, pe;
  or_node_range_aligned(p, z_exact, spans ? 8 : 1, &ps, &pe);
  int64_t count = image_f32_impl(bank, geom, ntr, tab, src, p, ps, pe, img);
  if (ps >= pe) return count;
  const or_tgeom *g = &p->tg;
  const float tz = p->taperzone;
  const float tx1 = p->tanx1, tx2 = p->tanx2, ty1 = p->tany1, ty2 = p->tany2;
  const float tx1c = tx1 - tz * (1 + tx1 * tx1), tx2c = tx2 + tz * (1 + tx2 * tx2);
  const float ty1c = ty1 - tz * (1 + ty1 * ty1), ty2c = ty2 + tz * (1 + ty2 * ty2);
  const float inv_dw = 1.0f / g->dw;
  const float jmax = (float)(p->K - 1);
  nf_col C0;
  C0.g = g;
  C0.tab = tab;
  C0.src = src;
  C0.tmin = p->tmin;
  C0.inv_dt = 1.0f / p->dt;
  C0.inv_dts = 1.0f / p->dtstar;
  C0.rx1 = 1.0f / (tx1 - tx1c);
  C0.rx2 = 1.0f / (tx2c - tx2);
  C0.ry1 = 1.0f / (ty1 - ty1c);
  C0.ry2 = 1.0f / (ty2c - ty2);
  C0.cx1 = -tx1c * C0.rx1;
  C0.cx2 = tx2c * C0.rx2;
  C0.cy1 = -ty1c * C0.ry1;
  C0.cy2 = ty2c * C0.ry2;
  int64_t ncount = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : ncount)
  for (int iy = 0; iy < p->ny; ++iy) {
    for (int itr = 0; itr < ntr; ++itr) {
      nf_col C = C0;
      C.sx = geom[4 * itr];
      C.sy = geom[4 * itr + 1];
      C.gx = geom[4 * itr + 2];
      C.gy = geom[4 * itr + 3];
      if (C.sx != C.sx) continue;
      if (src && !same_source(geom + 4 * itr, src)) continue;
      const float *bt = bank + (size_t)itr * p->K * p->nt;
      C.y = (float)iy * p->dy;
      for (int ix = 0; ix < p->nx; ++ix) {
        C.x = (float)ix * p->dx;
        C.dsx = C.sx - C.x;
        C.dsy = C.sy - C.y;
        C.dgx = C.gx - C.x;
        C.dgy = C.gy - C.y;
        C.hs2 = C.dsx * C.dsx + C.dsy * C.dsy;
        C.hg2 = C.dgx * C.dgx + C.dgy * C.dgy;
        int wcur = -1, scur = -1;
        nf_node nv[2], sv[2];
        for (int iz = ps; iz < pe; ++iz) {
          float c;
          const int w = nf_interval(iz, p->dz, inv_dw, g->nw, &c);
          if (w != wcur) {
            nv[0] = nf_value(&C, w);
            nv[1] = nf_value(&C, w + 1);
            if (spans) {
              nv[0].L = nv[0].L < 0.0f ? 0.0f : (nv[0].L > jmax ? jmax : nv[0].L);
              nv[1].L = nv[1].L < 0.0f ? 0.0f : (nv[1].L > jmax ? jmax : nv[1].L);
            }
            wcur = w;
          }
          const nf_node *n0 = &nv[0], *n1 = &nv[1], *m0 = n0, *m1 = n1;
          float s = c;
          if (spans) {
            const int w0 = spans[2 * w], w1 = spans[2 * w + 1];
            if (w0 != scur) {
              sv[0] = nf_value(&C, w0);
              sv[1] = nf_value(&C, w1);
              scur = w0;
            }
            m0 = &sv[0];
            m1 = &sv[1];
            const float z = (float)iz * p->dz, z0 = (float)w0 * g->dw, z1 = (float)w1 * g->dw;
            s = ((z - z0) * z1) / ((z1 - z0) * z);
          }
          const float f = fmaf(c, n1->F - n0->F, n0->F);
          const float wx = fminf(sat01(fmaf(s, m1->u[0] - m0->u[0], m0->u[0])),
                                 sat01(fmaf(s, m1->u[1] - m0->u[1], m0->u[1])));
          const float wy = fminf(sat01(fmaf(s, m1->u[2] - m0->u[2], m0->u[2])),
                                 sat01(fmaf(s, m1->u[3] - m0->u[3], m0->u[3])));
          const float wt = (wx * wy) * fmaf(s, m1->a - m0->a, m0->a);
          const int i0 = (int)f;
          float lv = fmaf(c, n1->L - n0->L, n0->L);
          if (!spans) lv = lv < 0.0f ? 0.0f : (lv > jmax ? jmax : lv);
          const int j = (int)lv;
          if (!((unsigned)(i0 - 1) < (unsigned)(p->nt - 2))) continue;
          if (wt > 0.0f) ++ncount;
          const float fr = f - floorf(f);
          const float *b = bt + (size_t)j * p->nt + i0;
          float *o = &img[((size_t)iy * p->nx + ix) * p->nz + iz];
          *o = fmaf(wt, fmaf(fr, b[1] - b[0], b[0]), *o);
        }
      }
    }
  }
  return count + ncount;
}

int64_t or_image_nodes_f32(const float *bank, const float *geom, int ntr, const float *tab, const or_src *src,
                           const or_iparams *p, float z_exact, float *img) {
  return nodes_impl(bank, geom, ntr, tab, src, p, z_exact, NULL, img);
}

int64_t or_image_spans_f32(const float *bank, const float *geom, int ntr, const float *tab, const or_src *src,
                           const or_iparams *p, float z_exact, const int *spans, float *img) {
  return nodes_impl(bank, geom, ntr, tab, src, p, z_exact, spans, img);
}

/* Update count alone: pairs (trace, point) with z > 0 inside the widened
 * dip bounds, evaluated in FP32 with the FP32 path's order of operations. */
int64_t or_count(const float *geom, int ntr, const or_iparams *p) {
  int64_t count = 0;
  float tz = p->taperzone;
  float tx1 = p->tanx1, tx2 = p->tanx2, ty1 = p->tany1, ty2 = p->tany2;
  float tx1c = tx1 - tz * (1 + tx1 * tx1), tx2c = tx2 + tz * (1 + tx2 * tx2);
  float ty1c = ty1 - tz * (1 + ty1 * ty1), ty2c = ty2 + tz * (1 + ty2 * ty2);
#pragma omp parallel for schedule(dynamic) reduction(+ : count)
  for (int iy = 0; iy < p->ny; ++iy)
    for (int itr = 0; itr < ntr; ++itr) {
      float sx = geom[4 * itr], sy = geom[4 * itr + 1];
      float gx = geom[4 * itr + 2], gy = geom[4 * itr + 3];
      if (sx != sx) continue;
      float y = (float)iy * p->dy;
      for (int ix = 0; ix < p->nx; ++ix) {
        float x = (float)ix * p->dx;
        float dsx = sx - x, dsy = sy - y, dgx = gx - x, dgy = gy - y;
        for (int iz = 1; iz < p->nz; ++iz) {
          float z = (float)iz * p->dz;
          const float rds = 1.0f / sqrtf(dsx * dsx + dsy * dsy + z * z);
          const float rdg = 1.0f / sqrtf(dgx * dgx + dgy * dgy + z * z);
          const float X = dsx * rds + dgx * rdg, Y = dsy * rds + dgy * rdg, Z = z * (rds + rdg);
          if (X <= tx1c * Z || tx2c * Z <= X || Y <= ty1c * Z || ty2c * Z <= Y) continue;
          ++count;
        }
      }
    }
  return count;
}

/* ---------------------------------------------------------------------- */
/* SEG-Y rev0 trace decode (src/segy.cc:350-1547, the format branches at
 * 1496-1545).  ibm2ieee as the reference states it: zero fraction -> 0;
 * else shift the 24-bit fraction left one bit at a time until bit 23 is set,
 * decrementing the binary exponent 4*(E-64); drop the implicit bit, bias by
 * 126, replace an exponent outside [0, 255] by 0. */
uint32_t or_ibm_to_ieee(uint32_t w) {
  if ((w & 0xffffffu) == 0) return 0u;
  uint32_t mant = w & 0xffffffu;
  int ex = ((int)((w & 0x7f000000u) >> 24) - 64) * 4;
  const uint32_t sign = w & 0x80000000u;
  while ((mant & 0x800000u) == 0) {
    mant <<= 1;
    ex--;
  }
  mant &= 0x7fffffu;
  ex += 126;
  if (ex < 0 || ex > 255) ex = 0;
  return sign | ((uint32_t)ex << 23) | mant;
}

/* One header key as the reader stores it (src/segy.cc:369-1490): kind 0 an
 * int field (2-byte key sign-extended), 1 a short field, 2 an unsigned short
 * field (4-byte key truncated to its low 16 bits); -> the field's value. */
long long or_segy_key(const unsigned char *hdr240, int start, int bytes, int kind) {
  const unsigned char *p = hdr240 + start - 1;
  if (bytes == 4) {
    const int32_t v = (int32_t)(((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]);
    if (kind == 0) return v;
    if (kind == 1) return (int16_t)(v & 0xffff);
    return (uint16_t)(v & 0xffff);
  }
  const uint16_t u = (uint16_t)((p[0] << 8) | p[1]);
  if (kind == 2) return u;
  return (int16_t)u;
}

/* ns big-endian samples of one record (after its 240-byte header): format
 * 1 IBM, 5 IEEE; the result as FP32 bits.  Returns 0, or -1 for another
 * format. */
int or_segy_samples(const unsigned char *rec, int ns, int format, uint32_t *out) {
  const unsigned char *s = rec + 240;
  for (int i = 0; i < ns; ++i) {
    const unsigned char *p = s + 4 * (size_t)i;
    const uint32_t w = ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
    if (format == 1)
      out[i] = or_ibm_to_ieee(w);
    else if (format == 5)
      out[i] = w;
    else
      return -1;
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Dip-angle gathers (options at src/q-psdm.cu:651-702, bin count          */
/* ndip = (int)(dip / ddip + 0.5) at src/q-psdm.cu:828-829; the reference  */
/* leaves gthrxd/gthryd, src/q-psdm.cu:211-212, unfilled): the FP32 path's */
/* update of each (trace, point) at the selected columns, binned by the    */
/* bisector's dip tangent against bin boundaries e[0..n] (bin b covers     */
/* [e[b], e[b+1])) and accumulated in trace order.  gx [ncol][nbx][nz],    */
/* gy [ncol][nby][nz]; nbx or nby 0 skips that axis.                        */
/* ---------------------------------------------------------------------- */
static int bin_linear(float v, const float *e, int n) {
  if (!(v >= e[0] && v < e[n])) return -1;
  int b = 0;
  while (b + 1 < n && v >= e[b + 1]) ++b;
  return b;
}

void or_gather_edges(int ndip, double ddip, float *e) {
  const double rad = M_PI / 180.0;
  for (int b = 0; b <= 2 * ndip + 1; ++b) e[b] = (float)tan(((double)(b - ndip) - 0.5) * ddip * rad);
}

void or_gathers_f32(const float *bank, const float *geom, int ntr, const float *tab, const or_iparams *p,
                    const int *cols, int ncol, const float *ex, int nbx, const float *ey, int nby, float *gx,
                    float *gy) {
  const or_tgeom *g = &p->tg;
  const float tz = p->taperzone;
  const float tx1 = p->tanx1, tx2 = p->tanx2, ty1 = p->tany1, ty2 = p->tany2;
  const float tx1c = tx1 - tz * (1 + tx1 * tx1), tx2c = tx2 + tz * (1 + tx2 * tx2);
  const float ty1c = ty1 - tz * (1 + ty1 * ty1), ty2c = ty2 + tz * (1 + ty2 * ty2);
  const float rx1 = 1.0f / (tx1 - tx1c), rx2 = 1.0f / (tx2c - tx2);
  const float ry1 = 1.0f / (ty1 - ty1c), ry2 = 1.0f / (ty2c - ty2);
  const float cx1 = -tx1c * rx1, cx2 = tx2c * rx2, cy1 = -ty1c * ry1, cy2 = ty2c * ry2;
  const float inv_dt = 1.0f / p->dt, inv_dts = 1.0f / p->dtstar;
#pragma omp parallel for schedule(dynamic)
  for (int c = 0; c < ncol; ++c) {
    const float x = (float)cols[2 * c] * p->dx, y = (float)cols[2 * c + 1] * p->dy;
    for (int itr = 0; itr < ntr; ++itr) {
      const float sx = geom[4 * itr], sy = geom[4 * itr + 1];
      const float gxx = geom[4 * itr + 2], gyy = geom[4 * itr + 3];
      if (sx != sx) continue;
      const float *bt = bank + (size_t)itr * p->K * p->nt;
      const float dsx = sx - x, dsy = sy - y, dgx = gxx - x, dgy = gyy - y;
      const float hs2 = dsx * dsx + dsy * dsy, hg2 = dgx * dgx + dgy * dgy;
      for (int iz = 1; iz < p->nz; ++iz) {
        const float z = (float)iz * p->dz;
        float ts, tss;
        lookup_f32(g, tab, sx, sy, x, y, z, &ts, &tss);
        ts = fmaxf(ts, p->tmin);
        const float d2s = hs2 + z * z;
        const float rds = 1.0f / sqrtf(d2s), ds = d2s * rds;
        const float tsn = ts * inv_dt, tsl = fmaf(tss, inv_dts, 0.5f);
        const float A = ds * ds * ds * (1.0f / ts);
        const float rdg = 1.0f / sqrtf(hg2 + z * z);
        const float q = rdg * ds;
        const float X = fmaf(dgx, q, dsx), Y = fmaf(dgy, q, dsy), Z = fmaf(z, q, z);
        const float rz = 1.0f / Z, tanx = X * rz, tany = Y * rz;
        const float wx = fminf(sat01(fmaf(tanx, rx1, cx1)), sat01(fmaf(tanx, -rx2, cx2)));
        const float wy = fminf(sat01(fmaf(tany, ry1, cy1)), sat01(fmaf(tany, -ry2, cy2)));
        const float wt = wx * wy;
        if (!(wt > 0.0f)) continue;
        float tg, tsg;
        lookup_f32(g, tab, gxx, gyy, x, y, z, &tg, &tsg);
        tg = fmaxf(tg, p->tmin);
        const float f = fmaf(tg, inv_dt, tsn);
        const int i0 = (int)f;
        if (!(i0 > 0 && i0 + 1 < p->nt)) continue;
        int j = (int)fmaf(tsg, inv_dts, tsl);
        j = j < 0 ? 0 : (j > p->K - 1 ? p->K - 1 : j);
        const float fr = f - (float)i0;
        const float *b = bt + (size_t)j * p->nt + i0;
        const float tr = A * (rdg * rdg * rdg) * tg;
        const float W = wt * (tr + 1.0f / tr), sv = fmaf(fr, b[1] - b[0], b[0]);
        if (nbx) {
          const int k = bin_linear(tanx, ex, nbx);
          if (k >= 0) {
            float *o = &gx[((size_t)c * nbx + k) * p->nz + iz];
            *o = fmaf(W, sv, *o);
          }
        }
        if (nby) {
          const int k = bin_linear(tany, ey, nby);
          if (k >= 0) {
            float *o = &gy[((size_t)c * nby + k) * p->nz + iz];
            *o = fmaf(W, sv, *o);
          }
        }
      }
    }
  }
}
