/*
 * qpsdm.h -- C ABI of the MI355X Q-compensated Kirchhoff PSDM imaging path.
 *
 * Plain C types only: device buffers are passed as raw device pointers
 * (void* / float*), HIP streams as void* (NULL = default stream).  Every
 * entry point returns 0 on success or a negative QPSDM_E* code; the text of
 * the last error is available from qpsdm_last_error().  Nothing here
 * allocates device memory except the functions whose name says so
 * (qpsdm_ctx_create, qpsdm_load_tables_files).
 *
 * The reference is one C++/CUDA program whose imaging operator is
 * psdm_kernel() (src/q-psdm.cu:252-283, 2063-2244) driving image_depth_gpu
 * (src/q-psdm.cu:2246-2380); each entry point below names the part of it
 * that it replaces.
 */
#ifndef QPSDM_H_
#define QPSDM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define QPSDM_OK 0
#define QPSDM_EINVAL -1   /* bad argument / shape                     */
#define QPSDM_EHIP -2     /* HIP runtime error                        */
#define QPSDM_EIO -3      /* file missing, short read, size mismatch  */
#define QPSDM_ERANGE -4   /* size beyond a documented maximum         */

/* Bank / table storage types. */
#define QPSDM_F32 0
#define QPSDM_BF16 1 /* bank only                         */
#define QPSDM_F16B 2 /* tables: FP16 codes, offset+scale per 8^3 brick            */
#define QPSDM_U16B 3 /* tables: unsigned 16-bit codes, offset+scale per 8^3 brick */
#define QPSDM_U16X 4 /* tables: unsigned 16-bit codes, offset+step per 9x9x8 brick whose
                        v and u edges repeat the neighbour's first nodes (one brick holds
                        every bilinear corner of its cells) */
#define QPSDM_F16 5  /* bank only: FP16 samples under a power-of-two scale per trace; a
                        trace's row is a 16-byte header {float 1/scale, float scale, 0, 0}
                        then [K][nt] f16, rows qpsdm_bank_row_bytes() apart */
#define QPSDM_U16Q 6 /* tables: the U16X bricks' codes stored per cell: at each depth node
                        the four bilinear corners' codes of a cell as one 16-byte word,
                        u16 halves t00 t01 t10 t11 t*00 t*01 t*10 t*11 (3.1x the U16X
                        bytes; one vector load per table node instead of four); the brick
                        header's steps are stored divided by 65535 */
#define QPSDM_F32P 7 /* bank only: pair bank, sample i of a level row stored with its forward
                        difference as float2 (b[i], b[i+1] - b[i]); entries 0 and nt-1 of
                        every row are (0, 0) (outside the samples the sum may pick);
                        [ntr][K][nt] float2 */
#define QPSDM_F16P 8 /* bank only: the pair bank in FP16 under a power-of-two scale per
                        trace: a 16-byte header {1/scale, scale, 0, 0} then [K][nt] half2,
                        rows qpsdm_bank_row_bytes() apart */

/* Two-window spectral set-up of one offset class; replaces the window block
 * of qpsdm() (src/q-psdm.cu:1593-1645) and the f3 -> nf3 conversion
 * (src/q-psdm.cu:1640-1644).  Same float/int semantics as the reference. */
typedef struct {
  int nsb, itb1, itb2, nf1, nf4, nww, ncoef1, nf3;
  float tdstart1, tdstart2, tdmid, df, dw;
} qpsdm_window;

int qpsdm_window_params(int ntl, float dt, float f1, float f3, float f4, qpsdm_window *out);

/* Band sanity of main() (src/q-psdm.cu:711-723): f3 <= Nyquist - 10 Hz,
 * f4 <= Nyquist - 5 Hz, and f3 <= f4 - 5 Hz. */
int qpsdm_clamp_band(float dt, float *f3, float *f4);

/* Reference FFT length choice: npfar (src/fft.cc:412-439) via
 * get_opt_n_fftw (src/fft.cc:3721). */
int qpsdm_npfar(int nmin);

/* Gain-limiter coefficients lnG, c1..c4 (src/q-psdm.cu:1277-1283). */
int qpsdm_gain_coefs(float threshold, float out5[5]);

/* Azimuth-sector table (src/q-psdm.cu:1167-1193); n = (int)(2/0.001 + 1.5). */
int qpsdm_iaztab(int *tab, int n);

/* Per-trace geometry of psdm_kernel (src/q-psdm.cu:2107-2144): scaled
 * coordinates in FP64, extent reject, offset, midpoint CDP bins, azimuth
 * sector; then the image-origin shift done in FP64 before the FP32 cast
 * (src/q-psdm.cu:2221-2224).  hdr_xy = {sx, sy, gx, gy} as read from the
 * SU header (int), scalco applied through `scaled`.  Outputs per trace:
 * geom[4] = image-relative (sx, sy, gx, gy) float, info[4] = {keep, cdpx,
 * cdpy, iaz}. */
typedef struct {
  double scaled;               /* coordinate scale (getscale(scalco))          */
  double fxv1, exv1, fyv1, eyv1; /* model extent with margins (src/q-psdm.cu:1379-1383) */
  double fxi, cyi;             /* image origin x and line y (src/q-psdm.cu:1430-1431) */
  double ox, oy;               /* origin subtracted before the FP32 cast         */
  float dxi, dyi;              /* CDP and line spacing                          */
} qpsdm_geom_params;

int qpsdm_trace_geometry(const qpsdm_geom_params *p, const int32_t *hdr_xy, int ntr,
                         const int *iaztab, float *geom, int32_t *info);

/* ---------------------------------------------------------------------- */
/* Tables.                                                                 */
/* ---------------------------------------------------------------------- */
/* Geometry of a relative-coordinate table set: one 3-D (t, t*) table per
 * surface node; a surface position p uses its nearest node's table at the
 * relative position (x - px, y - py, z).  FP32 layout:
 * float2 [node][nv][nu][nw] (z fastest).  FP16-brick layout: per node a
 * grid of 8x8x8 bricks [bv][bu][bw], each a float4 header {t_off, t_scale,
 * s_off, s_scale} followed by 512 half2 values [v][u][w]. */
typedef struct {
  int nnx, nny;
  float n0x, n0y, ndx, ndy;
  int nu, nv, nw;
  float u0, v0, du, dv, dw;
} qpsdm_tgeom;

/* Velocity/Q model for the table builder. kind 0: v(z)=v0+k z, constant Q
 * (closed form); kind 1: layers of thickness layer_dz with a Gaussian
 * anomaly (8-point Gauss-Legendre per layer piece). */
typedef struct {
  int kind;
  double v0, kgrad, q0;
  int nlayer;
  double layer_dz;
  double vl[16], ql[16];
  double ax, ay, az, asig, av, aq;
} qpsdm_model;

/* Table builder (no counterpart in the reference, which ships tables but no
 * generator): straight-ray t and t* into the FP32 layout, on the device. */
int qpsdm_tables_build(const qpsdm_model *m, const qpsdm_tgeom *g, float *tab_dev, void *stream);

/* The table builder's straight-ray integrals at arbitrary points, in FP64
 * on the device: pts [n][5] = (sx, sy, x, y, z) (source at the surface),
 * out [n][3] = (t, t*, sigma = int v ds).  For modelling data and checking
 * tables off their nodes. */
int qpsdm_rays(const qpsdm_model *m, const double *pts_dev, int n, double *out_dev, void *stream);

/* Size in bytes of the 16-bit brick form of a table set (QPSDM_F16B and
 * QPSDM_U16B have the same size). */
size_t qpsdm_tables_f16b_bytes(const qpsdm_tgeom *g);

/* Size in bytes of a table set in storage type dtype (F32, F16B, U16B, U16X, U16Q). */
size_t qpsdm_tables_bytes(const qpsdm_tgeom *g, int dtype);

/* Quantise an FP32 table set into 16-bit bricks: per 8^3 brick of (t, t*)
 * an offset and a scale per quantity, codes h = (v - off)/scale in [0, 1]
 * stored as FP16 (QPSDM_F16B) or as round(h * 65535) (QPSDM_U16B); or per
 * 9x9x8 edge-sharing brick an offset and a step, codes round((v - off)/step)
 * (QPSDM_U16X), or the same codes as corner quads (QPSDM_U16Q, header
 * {off, step / 65535, off*, step* / 65535}). */
int qpsdm_tables_quantize(const qpsdm_tgeom *g, const float *tab_dev, void *out_dev, int code,
                          void *stream);

/* Line interpolation of two table files, t = t1 (1-c) + t2 c, as
 * get_ttt_all does on the host (src/q-psdm.cu:2582-2594); on the device. */
int qpsdm_tables_blend(const float *t1_dev, const float *t2_dev, size_t n, float c,
                       float *out_dev, void *stream);

/* Reference table-file layout (src/q-psdm.cu:103-109, 2382-2622):
 * Lxxx.tidx|qidx header + descriptors, L<line>.ttt|qqq payloads. */
typedef struct {
  int32_t cdp1, cdp2;
  float dx, dz;
  int32_t nzg, nxg, nzi0, nxi0;
} qpsdm_tidx_header;

typedef struct {
  int64_t idx;
  uint16_t nys, nxs, dys, dxs;
} qpsdm_ttt_desc;

/* A reference-layout table set over the image, as this build writes and
 * reads it (the layout of src/q-psdm.cu:103-109, 1158-1159, 2382-2622).  The
 * reference leaves the order inside a cell and the origin of a cell's
 * source grid unstated (its kernel computes a cell index and then uses
 * fixed traveltimes, src/q-psdm.cu:2343-2352); this build's convention:
 *   - columns cdp1 .. cdp1 + ncol - 1 sit at image-relative
 *     x = x0 + (c - cdp1) dx, line file L<line1 + l> at y = y0 + l dy,
 *     depth node k at z = k dz (the index header's dx and dz);
 *   - cell (ixg, izg), descriptor ixg nzg + izg, holds columns
 *     cdp1 + ixg nxi0 + [0, nxi0) and depth nodes izg nzi0 + [0, nzi0)
 *     (past the last column or node: the last one's values);
 *   - its payload at float offset idx is [iys][ixs][column][depth], depth
 *     fastest ("z contiguous", src/q-psdm.cu:1158-1159);
 *   - every cell's source grid is the surface-node grid of the node tables:
 *     nys = nny, nxs = nnx, dys = ndy, dxs = ndx, source node (iys, ixs) at
 *     (n0x + ixs ndx, n0y + iys ndy).
 * L<n>.ttt and Lxxx.tidx hold t, L<n>.qqq and Lxxx.qidx hold t*. */
typedef struct {
  int line1, nlines;  /* first line number and number of lines   */
  int cdp1, ncol;     /* first column index and number of columns */
  int nz;             /* depth nodes                              */
  int nxi0, nzi0;     /* columns and depth nodes per cell         */
  float x0, y0;       /* image-relative x of column cdp1, y of line line1 */
  float dx, dy, dz;   /* column, line and depth-node spacing      */
} qpsdm_ref_layout;

/* Table builder output in the reference layout: writes Lxxx.tidx,
 * Lxxx.qidx and L<n>.ttt / L<n>.qqq for every line of the layout into dir,
 * from FP32 node tables on the device.  The layout's lattice must be the node
 * tables' (dx = du, dy = dv, dz = dw, aligned to the relative grid): each
 * payload value is a node table entry. */
int qpsdm_tables_write_ref(const qpsdm_tgeom *g, const float *tab_dev, const qpsdm_ref_layout *L, const char *dir);
/* The inverse: reads the index files and each line's payload (a missing
 * line blended from the nearest existing lines around it, as get_ttt_all
 * does) and fills the FP32 node tables on the device; a node-table entry
 * whose point lies outside the layout's columns, lines or depths takes the
 * nearest covered point's value.  Allocates a staging buffer for the
 * payload. */
int qpsdm_tables_read_ref(const qpsdm_tgeom *g, const qpsdm_ref_layout *L, const char *dir, float *tab_dev);

/* Read an index file; desc may be NULL to query the header only. */
int qpsdm_read_tidx(const char *path, qpsdm_tidx_header *hdr, qpsdm_ttt_desc *desc, int maxdesc);
int qpsdm_write_tidx(const char *path, const qpsdm_tidx_header *hdr, const qpsdm_ttt_desc *desc);
/* get_ttt_all's host part: find the bracketing existing L<n>.<suffix> in
 * [line1, line2] around `line`, load one or blend two (linear in line),
 * into out (nfloat floats).  Returns the number of floats, or < 0. */
long qpsdm_load_line_table(const char *ttpath, const char *suffix, int line, int line1,
                           int line2, float *out, size_t maxfloat, int *line_lo, int *line_hi);

/* ---------------------------------------------------------------------- */
/* Q filter bank (kernel a).  Replaces psdm_kernel's host pre-processing   */
/* and the per-trace H2D copies (src/q-psdm.cu:2173-2226) plus the PFA FFT */
/* (src/fft.cc), and tabulates the per-point Q-compensated inverse         */
/* transform that image_depth_gpu's argument list implies                  */
/* (src/q-psdm.cu:2246-2259: red/imd, nf1, fc, dw, nww, lnG, coef1-4).     */
/* ---------------------------------------------------------------------- */
typedef struct {
  int nt;
  float dt;
  qpsdm_window win;
  int K;          /* t* levels                                  */
  float dtstar;   /* t* level spacing (s)                       */
  float fc;       /* dispersion reference frequency (Hz)        */
  float threshold;/* gain cap G                                 */
} qpsdm_bank_params;

/* Reference pre-processing only: re/im spectra [ntr][2*nww] of psdm_kernel
 * (src/q-psdm.cu:2173-2216) for a batch of decoded traces. */
int qpsdm_trace_spectra(const qpsdm_bank_params *p, const float *traces_dev, int ntr,
                        float *re_dev, float *im_dev, void *stream);

/* Q filter bank: traces [ntr][nt] FP32 -> bank of dtype QPSDM_F32 or
 * QPSDM_BF16 ([ntr][K][nt]) or QPSDM_F16 (per trace: a 16-byte header
 * {1/scale, scale, 0, 0} then [K][nt] samples times scale = 2^(15 - ex), ex
 * the binary exponent of a bound on the trace's largest |sample| over all
 * levels taken from its spectrum, so every stored sample is below 2^15), or
 * the pair banks QPSDM_F32P ([ntr][K][nt] float2) and QPSDM_F16P (header,
 * then [K][nt] half2 scaled by 2^(14 - ex), so that the differences too stay
 * below 2^15). */
/* Bytes between consecutive traces of a bank (K * nt * sample size for
 * QPSDM_F32, QPSDM_BF16 and QPSDM_F32P; 16 + K * nt * 2 (QPSDM_F16) or
 * 16 + K * nt * 4 (QPSDM_F16P) rounded up to 16); 0 for another dtype. */
size_t qpsdm_bank_row_bytes(int K, int nt, int dtype);
int qpsdm_qbank(const qpsdm_bank_params *p, const float *traces_dev, int ntr, void *bank_dev,
                int bank_dtype, void *stream);

/* qpsdm_qbank for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same arguments and output as
 * qpsdm_qbank; instead of the per-parameter-set coefficient buffers that
 * qpsdm_qbank allocates on a set's first call, the FP64 filter coefficients
 * are generated on the device, on `stream`, into a fixed library-owned
 * device array (1.1 MB, one per GPU, held until the process ends) that the
 * bank kernel then reads.  Limits: nsb <= 8192, K * nww <= 131072,
 * nww <= 4096 (QPSDM_ERANGE beyond).  Calls no hipMalloc, hipFree or
 * allocating library routine and copies nothing from the host.  Every call
 * on a GPU rewrites that GPU's array: calls that overlap on one GPU (two
 * streams, two threads) race. */
int qpsdm_qbank_noalloc(const qpsdm_bank_params *p, const float *traces_dev, int ntr, void *bank_dev,
                        int bank_dtype, void *stream);

/* ---------------------------------------------------------------------- */
/* Imaging (kernel b).  Replaces image_depth_gpu (src/q-psdm.cu:2246-2380) */
/* and its per-trace launch (src/q-psdm.cu:2228-2243).                    */
/* ---------------------------------------------------------------------- */
typedef struct {
  int nx, ny, nz;
  float dx, dy, dz;
  int nt;
  float dt;
  int K;
  float dtstar;
  float tanx1, tanx2, tany1, tany2; /* dip bounds (tan), src/q-psdm.cu:1464-1472 */
  float taperzone;                  /* radians, src/q-psdm.cu:1287                */
  float tmin;                       /* floor on each leg's traveltime (s)         */
  qpsdm_tgeom tg;
  /* Form of the sum.  0: the exact form -- every image point evaluates its
   * own dip taper and amplitude weight w (tr + 1/tr).  1: the node form --
   * below z_exact (and above the tables' last depth node) the four taper
   * ramps and the amplitude are evaluated at the tables' depth nodes and
   * interpolated linearly to the points between them, the ramps clamped per
   * point; times and t* levels are exact in both forms (both legs are linear
   * in z between the nodes).  The node form's difference from the exact
   * form is an approximation error, not rounding: stated against the FP64
   * oracle where it is used. */
  int weight_nodes;
  float z_exact;                    /* node form: depth (m) imaged exactly above  */
  /* 2: the span form -- the node form's times and levels (the level clamped
   * to [0, K-1] at the table nodes), with the four ramps and the amplitude
   * evaluated at the ends of spans of span_s[k] table intervals and
   * interpolated between them linearly in 1/z (at the fraction
   * ((z - z0) z1) / ((z1 - z0) z) of a point at depth z in the span
   * [z0, z1]), the ramps clamped per point.  Band k of the node range starts
   * at the first table interval at or below span_z[k] metres, rounded up to a
   * multiple of 8 intervals past the node range's first interval, and ends
   * where the next band starts (the last at the node range's end); its spans
   * start at the band's start, the last cut at the band's end.  span_n bands
   * (1..4), span_z ascending, span_s 1, 2, 4 or 8; intervals above the first
   * band are spans of one interval.  Batched imaging only
   * (qpsdm_image_batch*), with QPSDM_F32P or QPSDM_F16P banks and QPSDM_F32 or
   * QPSDM_U16Q tables. */
  int span_n;
  float span_z[4];
  int span_s[4];
} qpsdm_image_params;

/* Launch shape picked by the autotuner; a NULL shape or 0 fields take the
 * default 8 x 8 columns x 6 depth points.  In the node form tile_z counts
 * table depth intervals per tile (2, 3 or 4), in the span form too (4 or 8;
 * a band of spans of 8 intervals runs tiles of 8); the exact body's share of
 * those forms runs 8 depth points per tile. */
typedef struct {
  int tile_x, tile_y, tile_z; /* image tile per work-group (columns x, y; depth points) */
} qpsdm_launch;

/* Table lookup as the imaging kernel does it (nearest node of the surface
 * position (px, py), bilinear in the relative (x, y), linear in z): out
 * [n][2] = (t, t*) at xyz [n][3], for any table storage type.  Replaces the
 * table indexing of image_depth_gpu (src/q-psdm.cu:2303-2318). */
int qpsdm_tables_lookup(const qpsdm_image_params *p, const void *tab_dev, int table_dtype, float px, float py,
                        const float *xyz_dev, int n, float *out_dev, void *stream);

/* A run of traces: geom_dev holds (sx, sy, gx, gy) per trace, image-relative,
 * and the run is geom rows [tr0, tr0+ntr); bank_dev holds the run's ntr
 * filtered traces [ntr][K][nt] (its first row is trace tr0).  Consecutive
 * traces with the same source share one source-leg evaluation (a shot
 * gather pays it once).  A trace whose sx is NaN is skipped.  Image float
 * [ny][nx][nz] accumulated in place.  table_dtype QPSDM_F32 or QPSDM_F16B;
 * bank_dtype QPSDM_F32, QPSDM_BF16 or QPSDM_F16. */
int qpsdm_image_shot(const qpsdm_image_params *p, const qpsdm_launch *shape, float *img_dev,
                     const float *geom_dev, int tr0, int ntr, const void *bank_dev,
                     int bank_dtype, const void *tab_dev, int table_dtype, void *stream);

/* Dip-angle gathers at selected image columns.  The reference declares
 * them -- options opcrp, cdps, aspect, dipx, dipy, ddipx, ddipy
 * (src/q-psdm.cu:651-702), ndipx = (int)(dipx / ddipx + 0.5)
 * (src/q-psdm.cu:828-829), buffers gthrxd/gthryd (src/q-psdm.cu:211-212) --
 * but does not fill them.  Every update the imaging sum makes at column
 * cols[c] = (ix, iy) (int pairs, device memory) is also added, in trace
 * order, to bin k of gx [ncol][nbx][nz] and gy [ncol][nby][nz]: k - ndip is
 * the nearest multiple of ddip to the bisector's dip angle atan(X / Z) in x
 * (atan(Y / Z) in y), k in [0, 2 ndip]; updates beyond the outer bins are not
 * binned.  aspect: 1 both axes (the reference's "All"), 2 x only, 3 y only,
 * 4 none; an axis not asked for may be NULL.  Gathers accumulate (+=) as the
 * image does.  Tables QPSDM_F32, QPSDM_U16X or QPSDM_U16Q; any bank dtype. */
typedef struct {
  float dipx, dipy;   /* largest dip binned (degrees)  */
  float ddipx, ddipy; /* bin width (degrees)            */
  int aspect;         /* 1 all, 2 x, 3 y, 4 none        */
} qpsdm_gather_params;
/* Bins per axis for g: 2 ndip + 1, or 0 for an axis aspect leaves out. */
int qpsdm_gather_bins(const qpsdm_gather_params *g, int *nbx, int *nby);
int qpsdm_gathers_shot(const qpsdm_image_params *p, const qpsdm_gather_params *g, const int *cols_dev, int ncol,
                       float *gx_dev, float *gy_dev, const float *geom_dev, int tr0, int ntr, const void *bank_dev,
                       int bank_dtype, const void *tab_dev, int table_dtype, void *stream);

/* Update count of a run (pairs (trace, point) inside the dip aperture, z > 0),
 * added to *count_dev (uint64). */
int qpsdm_count_shot(const qpsdm_image_params *p, const float *geom_dev, int tr0, int ntr,
                     unsigned long long *count_dev, void *stream);

/* ---------------------------------------------------------------------- */
/* Batched imaging (no counterpart in the reference, which launches once   */
/* per trace): shot gathers of uneven sizes with per-shot dip apertures    */
/* through one call, one dispatch per shot in shot order; a tile is owned  */
/* by one work-group per shot (fixed summation order, no image atomics).   */
/* ---------------------------------------------------------------------- */
typedef struct {
  int tr0, ntr;                      /* geom/bank rows of the shot          */
  float tanx1, tanx2, tany1, tany2;  /* the shot's dip bounds (tan)         */
} qpsdm_shot;

/* shots_dev: device array of nshot descriptors; geom/bank hold every trace
 * of the batch (bank row = geom row); work_dev: a device uint32 work counter
 * (reset by the call before each shot); nwg persistent work-groups (0 =
 * resident capacity), or nwg = -1: one work-group per tile, scheduled by the
 * hardware instead of the counter.  Either grid is dispatched once per shot,
 * in shot order (same tiles and summation order, so the same image bit for
 * bit).  The image params' own dip bounds are not used; taperzone is.
 * Traces that provably miss a tile's dip aperture are skipped (exactly:
 * they would add nothing).  The span form (weight_nodes = 2) spreads its
 * exact part and its bands over library streams of the current device,
 * forked from and joined back to `stream` within the call (captures as one
 * fork and join); calls from several host threads on one device are
 * serialised while they enqueue. */
int qpsdm_image_batch(const qpsdm_image_params *p, const qpsdm_launch *shape, float *img_dev,
                      const float *geom_dev, const void *bank_dev, int bank_dtype, const void *tab_dev,
                      int table_dtype, const qpsdm_shot *shots_dev, int nshot, unsigned *work_dev, int nwg,
                      void *stream);
/* Persistent grid (work-groups) that qpsdm_image_batch uses with nwg = 0
 * for this launch shape and these storage types: the resident capacity. */
int qpsdm_image_batch_grid(const qpsdm_image_params *p, const qpsdm_launch *shape, int bank_dtype, int table_dtype,
                           int *nwg);
int qpsdm_count_batch(const qpsdm_image_params *p, const float *geom_dev, const qpsdm_shot *shots_dev,
                      int nshot, unsigned long long *count_dev, void *stream);

/* As qpsdm_image_batch, with each shot's source leg read from its own
 * source table (src_dev as written by qpsdm_surrogate_src_tables or
 * qpsdm_src_tables_nodes for the same shots); the receiver legs use tab_dev.
 * A shot has one source: its table is built for its first valid trace's
 * (sx, sy), and traces of the shot with another source are not imaged.
 * src_dev == NULL is qpsdm_image_batch. */
int qpsdm_image_batch_src(const qpsdm_image_params *p, const qpsdm_launch *shape, float *img_dev,
                          const float *geom_dev, const void *bank_dev, int bank_dtype, const void *tab_dev,
                          int table_dtype, const qpsdm_shot *shots_dev, int nshot, const void *src_dev,
                          unsigned *work_dev, int nwg, void *stream);

/* ---------------------------------------------------------------------- */
/* Learned source-leg surrogate (no counterpart in the reference, whose    */
/* traveltimes come only from the table files, src/q-psdm.cu:2382-2622).  */
/* MLP 5-128-128-128-128-3 with SiLU, BF16 MFMA operands, FP32 accumulate; */
/* inputs (sx, sy, x, y, z) mapped from box [lo, hi] to [-1, 1]; outputs    */
/* (t, t*, log A), A = 1 / int v ds, as residuals (scaled by out_scale)     */
/* over the model without its anomaly, evaluated in closed form.  Trained  */
/* on the device by its own forward/backward/Adam kernels.                 */
/* ---------------------------------------------------------------------- */
typedef struct qpsdm_surrogate qpsdm_surrogate;

int qpsdm_surrogate_nparams(void);
int qpsdm_surrogate_create(const qpsdm_model *m, const float box_lo[5], const float box_hi[5],
                           const float out_scale[3], unsigned long long seed, qpsdm_surrogate **out);
int qpsdm_surrogate_destroy(qpsdm_surrogate *s);
/* FP32 master parameters, flattened W1[128][5], b1, W2, b2, W3, b3, W4, b4,
 * W5[3][128], b5 (host memory). */
int qpsdm_surrogate_get_params(const qpsdm_surrogate *s, float *host);
int qpsdm_surrogate_set_params(qpsdm_surrogate *s, const float *host);
/* n training points uniform in the input box, X [n][5], with exact targets
 * T [n][3] from the FP64 straight-ray integrals of the full model. */
int qpsdm_surrogate_samples(const qpsdm_surrogate *s, unsigned long long seed, long long offset, int n,
                            float *X_dev, float *T_dev, void *stream);
/* Outputs [n][3] (t, t*, log A) at X [n][5]. */
int qpsdm_surrogate_eval(const qpsdm_surrogate *s, const float *X_dev, int n, float *out_dev, void *stream);
/* Loss (mean weighted squared residual error) and its gradient [nparams]
 * on n points (n a multiple of 64); no update. */
int qpsdm_surrogate_grad(qpsdm_surrogate *s, const float *X_dev, const float *T_dev, int n, float *grad_dev,
                         float *loss_host, void *stream);
/* Adam over `steps` mini-batches of `batch` points taken in order from the
 * nsamples points (X, T), learning rate cosine-decayed lr0 -> lr1; the loss
 * of every step into loss_host[steps] when given. */
int qpsdm_surrogate_train(qpsdm_surrogate *s, const float *X_dev, const float *T_dev, long long nsamples,
                          int steps, int batch, float lr0, float lr1, float *loss_host, void *stream);
/* Judge: the surrogate at each node's own source position against that
 * node's FP32 table, at the table points inside the input box (the only
 * points source tables query), and at off-node sources inside the node's
 * cell (displaced by 0.45 of the node spacing in eight directions) against
 * the FP64 ray integrals at the brick's corners; mask[brick] = 1 where
 * max |dt| <= tol over all of them (node-major, [node][bv][bu][bw]);
 * err_dev [bricks] gets the max. */
long long qpsdm_surrogate_mask_bricks(const qpsdm_tgeom *g);
int qpsdm_surrogate_mask(const qpsdm_surrogate *s, const qpsdm_tgeom *g, const float *tab_dev, float tol,
                         unsigned char *mask_dev, float *err_dev, void *stream);
/* Per-shot source tables on the relative sub-grid that reaches the image:
 * header float4 {sx, sy, u0, v0} per shot (the shot's first valid trace's
 * source), then [nshot][nvs][nus][nw] float2 (t, t*) from the surrogate,
 * or, where the mask (optional) marks the nearest node's brick invalid,
 * from the table builder at the shot's own source (the FP64 straight-ray
 * integrals the node tables and the training samples come from).  tab_dev
 * is not read (may be NULL): the node tables judge the network in
 * qpsdm_surrogate_mask. */
size_t qpsdm_surrogate_src_bytes(const qpsdm_image_params *ip, int nshot);
int qpsdm_surrogate_src_tables(const qpsdm_surrogate *s, const qpsdm_image_params *ip, const float *geom_dev,
                               const qpsdm_shot *shots_dev, int nshot, const float *tab_dev,
                               const unsigned char *mask_dev, void *src_dev, void *stream);
/* The same per-shot source tables from the table path alone (no network):
 * every value the bilinear blend of the four surrounding node tables at
 * the shot's source.  Replaces the nearest-node source leg of the node
 * tables (the reference's tables carry per-cell source grids for the same
 * purpose, src/q-psdm.cu:103-109); qpsdm_surrogate_src_bytes sizes it. */
int qpsdm_src_tables_nodes(const qpsdm_image_params *ip, const float *geom_dev, const qpsdm_shot *shots_dev,
                           int nshot, const float *tab_dev, void *src_dev, void *stream);

/* A captured hipGraph of one batch: Q bank of traces_dev [ntr][nt] into
 * bank_dev, then (with a surrogate) the shots' source tables into src_dev,
 * then the persistent imaging.  Replays read the buffers' contents at
 * launch time, so the caller refills traces/geom/shots between replays. */
typedef struct qpsdm_graph qpsdm_graph;
int qpsdm_batch_graph_create(const qpsdm_bank_params *bp, const float *traces_dev, int ntr, void *bank_dev,
                             int bank_dtype, const qpsdm_image_params *ip, const qpsdm_launch *shape,
                             float *img_dev, const float *geom_dev, const void *tab_dev, int table_dtype,
                             const qpsdm_shot *shots_dev, int nshot, const qpsdm_surrogate *sur,
                             const float *tab_f32_dev, const unsigned char *mask_dev, void *src_dev,
                             unsigned *work_dev, int nwg, int stages, qpsdm_graph **out);
/* stages: bit mask of what the graph holds, 1 = Q bank, 2 = source tables,
 * 4 = imaging (7 = all); a caller that times the imaging alone captures
 * 1|2 and 4 as two graphs and launches them in order on one stream. */
int qpsdm_graph_launch(qpsdm_graph *g, void *stream);
int qpsdm_graph_destroy(qpsdm_graph *g);

/* Plain device copy of n floats (n % 4 == 0, 16-byte aligned): a kernel of
 * exactly known traffic, for calibrating hardware byte counters. */
int qpsdm_copy_f32(const float *src_dev, float *dst_dev, size_t n, void *stream);

/* Output normalisation img /= nsb (src/q-psdm.cu:1760-1766). */
int qpsdm_image_scale(float *img_dev, size_t n, float s, void *stream);

/* ---------------------------------------------------------------------- */
/* Pipeline context.  Replaces qpsdm()'s producer/consumer threads and      */
/* queue (src/q-psdm.cu:1693-1738, 1862-2061; src/queue.h:11-72) and       */
/* psdm_kernel's per-trace synchronous copies and launch                   */
/* (src/q-psdm.cu:2218-2243): two pinned staging slots, a copy stream and  */
/* a compute stream (SU decode on the GPU, Q bank, imaging).  Owns the     */
/* device image, the tables and every staging buffer.                      */
/* ---------------------------------------------------------------------- */
typedef struct qpsdm_ctx qpsdm_ctx;

typedef struct {
  qpsdm_image_params image;  /* grid, traces, dip taper, table geometry       */
  qpsdm_bank_params bank;    /* windows and Q bank (nt, K, dtstar as image)   */
  qpsdm_geom_params geom;    /* SU coordinate scale, model extent, origin     */
  int bank_dtype;            /* QPSDM_F32, QPSDM_BF16 or QPSDM_F16; the span  */
                             /* form (image.weight_nodes = 2): QPSDM_F16P   */
                             /* with QPSDM_U16Q or QPSDM_F32P with QPSDM_F32 */
                             /* tables, each slot imaged as one batched shot,*/
                             /* the two slots' shots in one call (a banked  */
                             /* slot waits for the other; flush, image,     */
                             /* reset and set_tables image it alone first)  */
  int table_dtype;           /* QPSDM_F32, QPSDM_F16B, QPSDM_U16B, QPSDM_U16X or QPSDM_U16Q */
  int batch_traces;          /* traces per staging slot                       */
  qpsdm_launch shape;        /* imaging launch shape (zeros = defaults)       */
} qpsdm_ctx_params;

int qpsdm_ctx_create(const qpsdm_ctx_params *p, qpsdm_ctx **out);
int qpsdm_ctx_destroy(qpsdm_ctx *c);
/* Tables in the FP32 layout, host (on_device = 0) or device memory;
 * quantised on the device for 16-bit table dtypes. */
int qpsdm_ctx_set_tables(qpsdm_ctx *c, const float *tab, int on_device);
int qpsdm_ctx_build_tables(qpsdm_ctx *c, const qpsdm_model *m);
/* Big-endian SU records with ns = nt (whole records only). */
int qpsdm_ctx_push_su(qpsdm_ctx *c, const void *bytes, size_t nbytes);
/* SEG-Y rev0 trace records (no 3600-byte reel header) with ns = nt, format
 * 1 (IBM float) or 5 (IEEE float), header keys at keyidx (struct keyindex
 * [75] of segy.h; NULL: standard positions), decoded on the device as
 * fsgettr decodes them (src/segy.cc:350-1547). */
struct keyindex;
int qpsdm_ctx_push_segy(qpsdm_ctx *c, const void *bytes, size_t nbytes, int format, const struct keyindex *keyidx);
/* Decoded host traces [ntr][nt] with image-relative geometry [ntr][4]. */
int qpsdm_ctx_push_traces(qpsdm_ctx *c, const float *samples, const float *geom, int ntr);
/* One decoded trace with its header, as psdm_kernel(segy tr, ...) takes it
 * (src/q-psdm.cu:252-283); traces outside the model extent are dropped. */
struct segy;
int qpsdm_ctx_psdm_kernel(qpsdm_ctx *c, const struct segy *tr);
int qpsdm_ctx_flush(qpsdm_ctx *c);
/* Image [ny][nx][nz] to host memory; divided by nsb when normalize != 0
 * (src/q-psdm.cu:1760-1766). */
int qpsdm_ctx_image(qpsdm_ctx *c, float *host_img, int normalize);
/* Zeroes the image and the gathers. */
int qpsdm_ctx_reset_image(qpsdm_ctx *c);
/* Dip-angle gathers (qpsdm_gathers_shot) at the columns cols (host int
 * pairs (ix, iy) inside the image), accumulated beside the image by every
 * later batch: the reference's opcrp/cdps/aspect/dipx/dipy/ddipx/ddipy
 * options (src/q-psdm.cu:651-702, 820-835) and gthrxd/gthryd buffers
 * (src/q-psdm.cu:211-212).  Replaces an earlier set; ncol = 0 stops
 * gathering.  Needs QPSDM_F32, QPSDM_U16X or QPSDM_U16Q tables. */
int qpsdm_ctx_set_gathers(qpsdm_ctx *c, const qpsdm_gather_params *g, const int *cols, int ncol);
/* The gathers to host memory, gx [ncol][nbx][nz] and gy [ncol][nby][nz] (an
 * axis not gathered may be NULL), divided by nsb when normalize != 0. */
int qpsdm_ctx_gathers(qpsdm_ctx *c, float *gx, float *gy, int normalize);
long long qpsdm_ctx_traces_pushed(const qpsdm_ctx *c);

/* Image writer of qpsdm() (src/q-psdm.cu:1769-1800): for each CDP in cdps
 * that lies in [fci, fci+nx), append one big-endian SU trace of nz samples
 * (img_line [nx][nz], one inline) with ns = nz, dt = (int)(dzi*1e3),
 * fldr = line, offset = (int)offset, trid = 1, counit = 1, nvs = 1,
 * tracl = running index, cdp, and sx = gx, sy = gy, scalco from the
 * velocity file's coordinates vel_xy [nx][2] of that CDP. */
int qpsdm_write_image_su(const char *path, int append, const float *img_line, int nx, int nz, float dzi,
                         int line, float offset, int fci, const int *cdps, int ncdps,
                         const int32_t *vel_xy, short vel_scalco);

const char *qpsdm_last_error(void);
const char *qpsdm_version(void);

#ifdef __cplusplus
}
#endif
#endif
