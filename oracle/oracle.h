/*
 * CPU oracle for the Q-compensated Kirchhoff PSDM imaging path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path (libqpsdm.so) never links,
 * loads or calls anything here.
 *
 * Every function restates the reference algorithm from its text; the comment
 * on each names the reference file:line it follows.  Where the reference is a
 * stub (its imaging kernel fixes traveltimes at 1.0, src/q-psdm.cu:2350-2352)
 * the oracle states this build's own formula and says so.
 */
#ifndef QPSDM_ORACLE_H_
#define QPSDM_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Segment windows and filter bins of the reference's two-window spectral
 * pre-processing (src/q-psdm.cu:1593-1645). */
typedef struct {
  int nsb;       /* per-segment FFT length, npfar(...)                  */
  int itb1;      /* first sample of segment 1                           */
  int itb2;      /* first sample of segment 2                           */
  int nf1;       /* first kept bin                                      */
  int nf4;       /* last kept bin                                       */
  int nww;       /* kept bins, nf4 - nf1 + 1                            */
  int ncoef1;    /* (int)(10/df), kernel argument of the reference      */
  int nf3;       /* high-cut taper start bin, (int)(f3/df + 0.5)        */
  float tdstart1, tdstart2, tdmid;
  float df, dw;
} or_window;

/* Velocity / Q model used to build traveltime and t* tables. */
typedef struct {
  int kind;            /* 0: v(z) = v0 + k z, constant Q;  1: layers + Gaussian anomaly */
  double v0, kgrad, q0;
  int nlayer;
  double layer_dz;
  double vl[16], ql[16];
  double ax, ay, az, asig; /* anomaly centre and sigma (image-relative metres) */
  double av, aq;           /* fractional velocity drop and Q drop at the centre */
} or_model;

/* Geometry of a relative-coordinate table set (one 3-D table per surface
 * node; a position p looks up its nearest node's table at x - p). */
typedef struct {
  int nnx, nny;
  float n0x, n0y, ndx, ndy;
  int nu, nv, nw;
  float u0, v0, du, dv, dw;
} or_tgeom;

/* Parameters of the imaging sum. */
typedef struct {
  int nx, ny, nz;
  float dx, dy, dz;
  int nt;
  float dt;
  int K;
  float dtstar;                        /* t* spacing of the bank levels       */
  float tanx1, tanx2, tany1, tany2;    /* dip bounds (tan)                     */
  float taperzone;                     /* radians                              */
  float tmin;                          /* floor on each leg's traveltime (s)   */
  or_tgeom tg;
} or_iparams;

/* Parameters of the Q filter bank. */
typedef struct {
  int nt;
  float dt;
  or_window win;
  int K;
  float dtstar;
  float fc;          /* dispersion reference frequency (Hz) */
  float threshold;   /* gain cap G of the gain limiter      */
} or_bparams;

int or_npfa(int nmin);
int or_npfar(int nmin);
void or_window_params(int ntl, float dt, float f1, float f3, float f4, or_window *w);
void or_gain_coefs(float threshold, float *lnG, float *c1, float *c2, float *c3, float *c4);
double or_gain_limit(double x, double G);
void or_iaztab(int *tab, int n);
int or_azimuth_index(const int *tab, double sx, double sy, double gx, double gy);
int or_fetch_and_taper(const float *data, int ns, int it1, int it2, float *out);
void or_trace_spectra(const float *trace, int nt, float dt, const or_window *w,
                      float *re, float *im);
void or_qbank_f64(const float *traces, int ntr, const or_bparams *p, double *bank);
void or_qbank_f32(const float *traces, int ntr, const or_bparams *p, float *bank);
double or_ray_sigma(const or_model *m, double sx, double sy, double x, double y, double z);
void or_ray_times(const or_model *m, double sx, double sy, double x, double y,
                  double z, double *t, double *tstar);
void or_build_tables(const or_model *m, const or_tgeom *g, float *tab);
void or_lookup(const or_tgeom *g, const float *tab, double px, double py,
               double x, double y, double z, double *t, double *tstar);
/* Per-shot source table, in the layout of the library's per-shot source
 * tables: hdr = {sx, sy, su0, sv0}; tab = (t, t*) [nvs][nus][nw][2] on the
 * node tables' (du, dv, dw) spacing from the relative origin (su0, sv0).
 * The source leg of a trace is read at (x - sx - su0, y - sy - sv0, z) with
 * (sx, sy) from hdr; a trace whose source differs from hdr (bitwise) is not
 * imaged. */
typedef struct {
  const float *hdr;
  const float *tab;
  int nus, nvs;
} or_src;

int64_t or_image_f64(const double *bank, const float *geom, int ntr,
                     const float *tab, const or_iparams *p, double *img);
int64_t or_image_f32(const float *bank, const float *geom, int ntr,
                     const float *tab, const or_iparams *p, float *img);
int64_t or_image_src_f64(const double *bank, const float *geom, int ntr, const float *tab,
                         const or_src *src, const or_iparams *p, double *img);
int64_t or_image_src_f32(const float *bank, const float *geom, int ntr, const float *tab,
                         const or_src *src, const or_iparams *p, float *img);
/* Node form (the library's weight_nodes = 1) of the FP32 path; the points
 * [ps, pe) it interpolates from depth nodes come from or_node_range. */
void or_node_range(const or_iparams *p, float z_exact, int *ps, int *pe);
/* The same with the first interval rounded up to a multiple of align (the
 * span form's range: align 8). */
void or_node_range_aligned(const or_iparams *p, float z_exact, int align, int *ps, int *pe);
int64_t or_image_nodes_f32(const float *bank, const float *geom, int ntr, const float *tab, const or_src *src,
                           const or_iparams *p, float z_exact, float *img);
/* Span form (the library's weight_nodes = 2): the node form over the node
 * range aligned to 8 intervals, with the level clamped at the nodes and the
 * ramps and amplitude interpolated, linearly in 1/z, between the weight
 * nodes spans[2w] <= w < spans[2w + 1] of each interval w. */
int64_t or_image_spans_f32(const float *bank, const float *geom, int ntr, const float *tab, const or_src *src,
                           const or_iparams *p, float z_exact, const int *spans, float *img);
/* Test-only: the FP64 sum's level rule (0 round, the product's; 1 floor;
 * 2 always level 0), for planting a wrong pick in known-answer tests. */
void or_set_level_rule(int rule);
int64_t or_count(const float *geom, int ntr, const or_iparams *p);
int or_num_threads(void);
void or_set_num_threads(int n);
uint32_t or_ibm_to_ieee(uint32_t w);
void or_gather_edges(int ndip, double ddip, float *e);
void or_gathers_f32(const float *bank, const float *geom, int ntr, const float *tab, const or_iparams *p,
                    const int *cols, int ncol, const float *ex, int nbx, const float *ey, int nby, float *gx,
                    float *gy);
long long or_segy_key(const unsigned char *hdr240, int start, int bytes, int kind);
int or_segy_samples(const unsigned char *rec, int ns, int format, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
