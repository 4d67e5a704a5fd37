// Host-side parts of the imaging path: window/filter set-up, FFT length,
// gain limiter, azimuth table, per-trace geometry and the table-file layout.
// Written for this build; each function cites the reference block whose
// behaviour it keeps.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include "common.h"
#include "segy.h"

namespace qpsdm {

static thread_local char g_err[512];

int set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace qpsdm

extern "C" const char *qpsdm_last_error(void) { return qpsdm::g_err; }

extern "C" const char *qpsdm_version(void) { return "qpsdm-mi355x 0.1 (gfx950)"; }

// Valid prime-factor lengths are the divisors of 720720 = 16*9*5*7*11*13
// (the reference's 240-entry table, src/fft.cc:274-349).
static int npfa_len(int nmin) {
  const int big = 720720;
  if (nmin >= big) return big;
  for (int n = nmin < 1 ? 1 : nmin; n < big; ++n)
    if (big % n == 0) return n;
  return big;
}

extern "C" int qpsdm_npfar(int nmin) { return 2 * npfa_len((nmin + 1) / 2); }

// The window block of qpsdm() (src/q-psdm.cu:1593-1645) keeps float dt and
// int truncations; the same expressions with the same types give the same
// integers (e.g. 0.48/0.002f truncates to 239).
extern "C" int qpsdm_window_params(int ntl, float dt, float f1, float f3, float f4,
                                   qpsdm_window *o) {
  QPSDM_REQUIRE(o && ntl > 1 && dt > 0, "window: bad ntl/dt");
  const int pre = 0.48 / dt;
  const int suf = 0.36 / dt;
  const float t0 = 0;
  const float t1 = (ntl - 1) * dt;
  int half = (t1 - t0 - pre * dt - suf * dt) / 2.0 / dt;
  QPSDM_REQUIRE(half > 0, "window: trace too short for the two-segment windows");
  const int itb1 = t0 / dt;
  const float tmid = t0 + pre * dt + half * dt;
  const int nsb = qpsdm_npfar(half + pre + suf);
  const int itb2 = t1 / dt - nsb + 1;
  const float df = 1.0 / nsb / dt;
  o->nsb = nsb;
  o->itb1 = itb1;
  o->itb2 = itb2;
  o->tdstart1 = itb1 * dt;
  o->tdstart2 = itb2 * dt;
  o->tdmid = tmid;
  o->df = df;
  o->dw = 2.0 * M_PI * df;
  o->nf1 = (int)(f1 / df);
  o->nf4 = (int)(f4 / df);
  o->nww = o->nf4 - o->nf1 + 1;
  o->ncoef1 = (int)(10.0 / df);
  o->nf3 = (int)(f3 / df + 0.5);
  QPSDM_REQUIRE(o->nww > 0 && o->nf4 < nsb / 2, "window: band [f1,f4] empty or above Nyquist");
  return QPSDM_OK;
}

// main()'s band sanity (src/q-psdm.cu:711-723).
extern "C" int qpsdm_clamp_band(float dt, float *f3, float *f4) {
  QPSDM_REQUIRE(dt > 0 && f3 && f4, "clamp_band: bad args");
  if (*f3 > 0.5 / dt - 10.0) *f3 = 0.5 / dt - 10.0;
  if (*f4 > 0.5 / dt - 5.0) *f4 = 0.5 / dt - 5.0;
  if (*f4 - *f3 <= 5.0) *f3 = *f4 - 5.0;
  return QPSDM_OK;
}

// lnG and c1..c4 of the gain limiter (src/q-psdm.cu:1277-1283).
extern "C" int qpsdm_gain_coefs(float G, float out[5]) {
  QPSDM_REQUIRE(out && G > 1, "gain: threshold must exceed 1");
  const float l = logf(G);
  out[0] = l;
  out[1] = G * (1.0 - l - 2.5 * l * l);
  out[2] = G * (1.0 + 5.0 * l);
  out[3] = -2.5 * G;
  out[4] = 1.1 * G;
  return QPSDM_OK;
}

// 18 azimuth sectors of 10 degrees on a (1 - cos) axis warped through a
// second cosine, sampled every 0.001 (src/q-psdm.cu:1167-1193).
extern "C" int qpsdm_iaztab(int *tab, int n) {
  QPSDM_REQUIRE(tab && n > 0, "iaztab: bad args");
  const float eps = 0.001f;
  const double deg = 0.0174532925199;
  auto warp = [&](float a_deg) -> float {  // (1 - cos a) * pi/2, float-rounded
    float v = (1.0 - cosf(a_deg * deg)) * M_PI * 0.5;
    return v;
  };
  float lo = warp(10.0f);
  lo = 1.0 - cosf(lo * 0.5);
  int end = (int)(lo / eps);
  for (int i = 0; i < end && i < n; ++i) tab[i] = 0;
  for (int s = 1; s < 18; ++s) {
    float a = warp(s * 10.0f - 10.0f), b = warp(s * 10.0f), c = warp(s * 10.0f + 10.0f);
    float edge0 = 1.0 - cosf(b - (b - a) * 0.5);
    float edge1 = 1.0 - cosf(b + (c - b) * 0.5);
    int i0 = (int)(edge0 / eps);
    end = (int)(edge1 / eps);
    for (int i = i0; i < end && i < n; ++i) tab[i] = s;
  }
  for (int i = end; i < n; ++i) tab[i] = 0;
  return QPSDM_OK;
}

// Per-trace geometry of psdm_kernel (src/q-psdm.cu:2107-2144, 2221-2224).
extern "C" int qpsdm_trace_geometry(const qpsdm_geom_params *p, const int32_t *xy, int ntr,
                                    const int *iaztab, float *geom, int32_t *info) {
  QPSDM_REQUIRE(p && xy && geom && info && ntr >= 0, "geometry: bad args");
  for (int i = 0; i < ntr; ++i) {
    const double sx = xy[4 * i] * p->scaled, sy = xy[4 * i + 1] * p->scaled;
    const double gx = xy[4 * i + 2] * p->scaled, gy = xy[4 * i + 3] * p->scaled;
    int keep = !(sx < p->fxv1 || sx > p->exv1 || gx < p->fxv1 || gx > p->exv1 ||
                 sy < p->fyv1 || sy > p->eyv1 || gy < p->fyv1 || gy > p->eyv1);
    const double ddx = gx - sx, ddy = gy - sy;
    const double off = sqrtf(ddx * ddx + ddy * ddy);
    const int cdpx = (int)(((sx + gx) * 0.5 - p->fxi) / p->dxi + 0.5);
    const int cdpy = (int)(((sy + gy) * 0.5 - p->cyi) / p->dyi + 0.5);
    int iaz = 0;
    if (iaztab && off > 1.0) {
      float cx = fabs((float)(ddx / off));
      if (ddx * ddy < 0.0) cx = -cx;
      const float c1 = 1.0 - cx;
      iaz = iaztab[(int)(c1 / 0.001f)];
    }
    geom[4 * i] = (float)(sx - p->ox);
    geom[4 * i + 1] = (float)(sy - p->oy);
    geom[4 * i + 2] = (float)(gx - p->ox);
    geom[4 * i + 3] = (float)(gy - p->oy);
    info[4 * i] = keep;
    info[4 * i + 1] = cdpx;
    info[4 * i + 2] = cdpy;
    info[4 * i + 3] = iaz;
  }
  return QPSDM_OK;
}

// ---- table-file layout (src/q-psdm.cu:103-109, 2382-2622) -------------
extern "C" int qpsdm_read_tidx(const char *path, qpsdm_tidx_header *h, qpsdm_ttt_desc *d,
                               int maxdesc) {
  QPSDM_REQUIRE(path && h, "read_tidx: bad args");
  static_assert(sizeof(qpsdm_ttt_desc) == 16, "descriptor is 16 bytes on disk");
  FILE *fp = fopen(path, "rb");
  if (!fp) return qpsdm::set_error(QPSDM_EIO, "cannot open index file %s", path);
  int32_t raw[8];
  if (fread(raw, 4, 8, fp) != 8) {
    fclose(fp);
    return qpsdm::set_error(QPSDM_EIO, "short index header in %s", path);
  }
  h->cdp1 = raw[0];
  h->cdp2 = raw[1];
  memcpy(&h->dx, &raw[2], 4);
  memcpy(&h->dz, &raw[3], 4);
  h->nzg = raw[4];
  h->nxg = raw[5];
  h->nzi0 = raw[6];
  h->nxi0 = raw[7];
  int rc = QPSDM_OK;
  if (h->nxg < 0 || h->nzg < 0 || h->nzi0 < 0 || h->nxi0 < 0) {
    rc = qpsdm::set_error(QPSDM_EIO, "negative grid size in index file %s", path);
  } else if (d) {
    const size_t n = (size_t)h->nxg * (size_t)h->nzg;
    if (maxdesc < 0 || n > (size_t)maxdesc) {
      rc = qpsdm::set_error(QPSDM_ERANGE, "index has %zu descriptors, room for %d", n, maxdesc);
    } else if (fread(d, sizeof(qpsdm_ttt_desc), n, fp) != n) {
      rc = qpsdm::set_error(QPSDM_EIO, "travel time table or qfield index size error %s", path);
    }
  }
  fclose(fp);
  return rc;
}

extern "C" int qpsdm_write_tidx(const char *path, const qpsdm_tidx_header *h,
                                const qpsdm_ttt_desc *d) {
  QPSDM_REQUIRE(path && h && d, "write_tidx: bad args");
  QPSDM_REQUIRE(h->nxg >= 0 && h->nzg >= 0, "write_tidx: negative grid size");
  FILE *fp = fopen(path, "wb");
  if (!fp) return qpsdm::set_error(QPSDM_EIO, "cannot create %s", path);
  int32_t raw[8] = {h->cdp1, h->cdp2, 0, 0, h->nzg, h->nxg, h->nzi0, h->nxi0};
  memcpy(&raw[2], &h->dx, 4);
  memcpy(&raw[3], &h->dz, 4);
  const size_t n = (size_t)h->nxg * h->nzg;
  const bool ok = fwrite(raw, 4, 8, fp) == 8 && fwrite(d, sizeof(qpsdm_ttt_desc), n, fp) == n;
  fclose(fp);
  return ok ? QPSDM_OK : qpsdm::set_error(QPSDM_EIO, "short write %s", path);
}

static long file_size(const char *p) {
  struct stat st;
  if (stat(p, &st) != 0) return -1;
  return (long)st.st_size;
}

static long read_floats(const char *path, float *out, size_t maxfloat) {
  const long sz = file_size(path);
  if (sz < 0) return qpsdm::set_error(QPSDM_EIO, "missing %s", path);
  if (sz % 4 != 0) return qpsdm::set_error(QPSDM_EIO, "%s: size %ld is not a whole number of floats", path, sz);
  if ((size_t)sz / 4 > maxfloat) return qpsdm::set_error(QPSDM_ERANGE, "%s larger than buffer", path);
  FILE *fp = fopen(path, "rb");
  if (!fp) return qpsdm::set_error(QPSDM_EIO, "cannot open %s", path);
  const size_t got = fread(out, 1, (size_t)sz, fp);
  fclose(fp);
  if ((long)got != sz) return qpsdm::set_error(QPSDM_EIO, "travle time table size error %s", path);
  return sz / 4;
}

// get_ttt_all's search and blend (src/q-psdm.cu:2486-2600): the nearest
// existing L<n>.<suffix> at or below `line` (down to line1) and at or above
// it (up to line2); one file if both are `line`, else the linear blend
// t1 (1-c) + t2 c with c = (line - lo)/(hi - lo); error if either is absent.
extern "C" long qpsdm_load_line_table(const char *ttpath, const char *suffix, int line,
                                      int line1, int line2, float *out, size_t maxfloat,
                                      int *lo_out, int *hi_out) {
  QPSDM_REQUIRE(ttpath && suffix && out, "load_line_table: bad args");
  char path[4096];
  auto exists = [&](int l) {
    snprintf(path, sizeof(path), "%s/L%d.%s", ttpath, l, suffix);
    return access(path, R_OK) == 0;
  };
  int lo = -1, hi = -1;
  for (int l = line; l >= line1; --l)
    if (exists(l)) { lo = l; break; }
  for (int l = line; l <= line2; ++l)
    if (exists(l)) { hi = l; break; }
  if (lo_out) *lo_out = lo;
  if (hi_out) *hi_out = hi;
  if (lo == -1 || hi == -1) return qpsdm::set_error(QPSDM_EIO, "no travel time table exists");
  snprintf(path, sizeof(path), "%s/L%d.%s", ttpath, lo, suffix);
  const long n = read_floats(path, out, maxfloat);
  if (n < 0 || (lo == line && hi == line)) return n;
  if (lo == hi) return n;
  float *b = (float *)malloc(sizeof(float) * (size_t)n);
  if (!b) return qpsdm::set_error(QPSDM_EINVAL, "host allocation failed");
  snprintf(path, sizeof(path), "%s/L%d.%s", ttpath, hi, suffix);
  const long n2 = read_floats(path, b, (size_t)n);
  if (n2 != n) {
    free(b);
    return n2 < 0 ? n2 : qpsdm::set_error(QPSDM_EIO, "travel time table size not confirm");
  }
  const float c2 = 1.0 * (line - lo) / (hi - lo);
  const float c1 = 1.0 - c2;
  for (long i = 0; i < n; ++i) out[i] = out[i] * c1 + b[i] * c2;
  free(b);
  return n;
}

// The offset-class image writer of qpsdm() (src/q-psdm.cu:1769-1800).
extern "C" int qpsdm_write_image_su(const char *path, int append, const float *img, int nx, int nz,
                                    float dzi, int line, float offset, int fci, const int *cdps,
                                    int ncdps, const int32_t *vel_xy, short vel_scalco) {
  QPSDM_REQUIRE(path && img && cdps && vel_xy && nx > 0 && nz > 0 && nz <= SU_NFLTS,
                "write_image_su: bad args");
  FILE *fp = fopen(path, append ? "ab" : "wb");
  if (!fp) return qpsdm::set_error(QPSDM_EIO, "Error: File [%s] Open Error", path);
  segy *tr = (segy *)calloc(1, sizeof(segy));
  tr->ns = nz;
  tr->dt = (int)(dzi * 1.0E3);
  tr->fldr = line;
  tr->offset = (int)offset;
  tr->trid = 1;
  tr->counit = 1;
  tr->nvs = 1;
  int k = 0;
  for (int i = 0; i < ncdps; ++i) {
    const int ixi = cdps[i] - fci;
    if (ixi < 0 || ixi >= nx) continue;
    tr->tracl = ++k;
    tr->cdp = fci + ixi;
    tr->sx = tr->gx = vel_xy[2 * ixi];
    tr->sy = tr->gy = vel_xy[2 * ixi + 1];
    tr->scalco = vel_scalco;
    memcpy(tr->data, img + (size_t)ixi * nz, sizeof(float) * nz);
    fputtr(fp, tr);
  }
  free(tr);
  fclose(fp);
  return QPSDM_OK;
}
