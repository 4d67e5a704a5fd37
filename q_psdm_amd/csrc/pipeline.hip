// Single-GPU imaging pipeline behind the qpsdm_ctx_* entry points.
//
// Replaces the reference's producer/consumer threads and bounded queue
// (src/q-psdm.cu:1862-2061, src/queue.h:11-72) and psdm_kernel's per-trace
// synchronous copies and launch (src/q-psdm.cu:2218-2243): traces are
// batched into two pinned host slots; a copy stream moves a slot to the
// device while the compute stream decodes the previous one (big-endian SU
// -> FP32 samples + image-relative geometry, on the GPU), runs the Q bank
// and the imaging kernel (in the span form, the two slots' shots in one
// batched call).  Events order slot reuse.  The device image is owned by
// the context and copied back (divided by nsb) on request.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "ibm_float.h"
#include "segy.h"

namespace {

constexpr int kSlots = 2;

// A trace-header key in a record: 0-based byte offset and width (2 or 4;
// any other width reads as 0, a key the reader does not decode).
struct KeyPos {
  int off, bytes;
};
struct GeomKeys {
  KeyPos sx, sy, gx, gy;
};
// sx, sy, gx, gy at their SU / standard SEG-Y positions
constexpr GeomKeys kStdGeom{{72, 4}, {76, 4}, {80, 4}, {84, 4}};

struct Slot {
  unsigned char *host = nullptr;  // pinned staging: SU records or decoded samples+geom
  unsigned char *raw = nullptr;   // device copy of staged SU records
  float *samples = nullptr;       // [batch][nt]
  float *geom = nullptr;          // [batch][4], slot i's rows of the context's geometry
  void *bank = nullptr;           // [batch][K][nt], slot i's rows of the context's bank
  int ntr = 0;                    // traces staged
  bool su = false;                // staging holds SU / SEG-Y records (else samples + geom)
  int fmt = 5;                    // records' sample format: 1 IBM float, 5 IEEE float
  GeomKeys keys = kStdGeom;       // records' sx, sy, gx, gy positions
  hipEvent_t copied = nullptr, done = nullptr;
  bool used = false;
};

struct GeomDev {
  double scaled, fxv1, exv1, fyv1, eyv1, ox, oy;
};

// Big-endian key as a 4-byte integer field: a 2-byte key sign-extends.
__device__ inline int key_i32(const unsigned char *r, KeyPos k) {
  const unsigned char *p = r + k.off;
  if (k.bytes == 4) return (int)(((unsigned)p[0] << 24) | ((unsigned)p[1] << 16) | ((unsigned)p[2] << 8) | p[3]);
  if (k.bytes == 2) return (int)(short)((p[0] << 8) | p[1]);
  return 0;
}
// Big-endian key as a 2-byte field: a 4-byte key keeps its low 16 bits.
__device__ inline unsigned short key_u16(const unsigned char *r, KeyPos k) {
  const unsigned char *p = r + k.off;
  if (k.bytes == 2) return (unsigned short)((p[0] << 8) | p[1]);
  if (k.bytes == 4) return (unsigned short)((p[2] << 8) | p[3]);
  return 0;
}

// Sample word of a record (big-endian) -> FP32 bits.
template <int FMT>
__device__ inline float sample_in(unsigned w) {
  w = __builtin_bswap32(w);
  if constexpr (FMT == 1) {
    bool range = false;
    w = qpsdm_ibm::to_ieee_bits(w, &range);
  }
  return __uint_as_float(w);
}

// The samples of one record by one work-group: each thread loads all of its
// words before converting and storing any, so every load of the record is
// in flight at once (one pass for ns <= 8 * blockDim).
template <int FMT>
__device__ inline void decode_samples(const unsigned *__restrict__ src, float *__restrict__ dst, int ns) {
  constexpr int U = 8;
  const int step = blockDim.x;
  for (int base = threadIdx.x; base < ns; base += U * step) {
    unsigned w[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * step;
      w[u] = i < ns ? src[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * step;
      if (i < ns) dst[i] = sample_in<FMT>(w[u]);
    }
  }
}

// One work-group per SU / SEG-Y record: convert the samples (FMT 5:
// big-endian IEEE, the SU case; FMT 1: big-endian IBM), read sx, sy, gx, gy
// at their key positions (and scalco-independent `scaled`), reject traces
// outside the model extent by writing NaN geometry
// (src/q-psdm.cu:2115-2123), else image-relative FP32 coordinates with the
// origin removed in FP64.
template <int FMT>
__global__ void record_decode_kernel(const unsigned char *__restrict__ raw, int ns, GeomKeys keys, GeomDev g,
                                     float *__restrict__ samples, float4 *__restrict__ geom) {
  const size_t rec = 240 + 4 * (size_t)ns;
  const unsigned char *r = raw + (size_t)blockIdx.x * rec;
  const unsigned *src = (const unsigned *)(r + 240);
  float *dst = samples + (size_t)blockIdx.x * ns;
  decode_samples<FMT>(src, dst, ns);
  if (threadIdx.x == 0) {
    const double sx = key_i32(r, keys.sx) * g.scaled;
    const double sy = key_i32(r, keys.sy) * g.scaled;
    const double gx = key_i32(r, keys.gx) * g.scaled;
    const double gy = key_i32(r, keys.gy) * g.scaled;
    const bool keep = !(sx < g.fxv1 || sx > g.exv1 || gx < g.fxv1 || gx > g.exv1 || sy < g.fyv1 ||
                        sy > g.eyv1 || gy < g.fyv1 || gy > g.eyv1);
    geom[blockIdx.x] = keep ? make_float4((float)(sx - g.ox), (float)(sy - g.oy), (float)(gx - g.ox),
                                          (float)(gy - g.oy))
                            : make_float4(NAN, NAN, NAN, NAN);
  }
}

struct HdrKeys {
  KeyPos sx, sy, gx, gy, fldr, cdp, offset, tracl, scalco, ns, dt;
};

// One work-group per SEG-Y record: samples as record_decode_kernel, and the
// header subset of qpsdm_trhdr read at its key positions.
template <int FMT>
__global__ void segy_decode_kernel(const unsigned char *__restrict__ raw, int ns, HdrKeys k,
                                   float *__restrict__ samples, qpsdm_trhdr *__restrict__ hdrs) {
  const size_t rec = 240 + 4 * (size_t)ns;
  const unsigned char *r = raw + (size_t)blockIdx.x * rec;
  const unsigned *src = (const unsigned *)(r + 240);
  float *dst = samples + (size_t)blockIdx.x * ns;
  decode_samples<FMT>(src, dst, ns);
  if (threadIdx.x == 0 && hdrs) {
    qpsdm_trhdr h;
    h.sx = key_i32(r, k.sx);
    h.sy = key_i32(r, k.sy);
    h.gx = key_i32(r, k.gx);
    h.gy = key_i32(r, k.gy);
    h.fldr = key_i32(r, k.fldr);
    h.cdp = key_i32(r, k.cdp);
    h.offset = key_i32(r, k.offset);
    h.tracl = key_i32(r, k.tracl);
    h.scalco = (short)key_u16(r, k.scalco);
    h.pad = 0;
    h.ns = key_u16(r, k.ns);
    h.dt = key_u16(r, k.dt);
    hdrs[blockIdx.x] = h;
  }
}

}  // namespace

extern "C" int qpsdm_segy_decode_dev(const void *raw_dev, int ntr, int ns, int format, const keyindex *keyidx,
                                     float *samples_dev, qpsdm_trhdr *hdrs_dev, void *stream) {
  QPSDM_REQUIRE(ntr >= 0 && ns > 0 && ns <= SU_NFLTS, "segy_decode_dev: bad sizes ntr=%d ns=%d", ntr, ns);
  QPSDM_REQUIRE(format == 1 || format == 5, "segy_decode_dev: format %d (1 = IBM float, 5 = IEEE float)", format);
  QPSDM_REQUIRE(ntr == 0 || (raw_dev && samples_dev), "segy_decode_dev: null buffer");
  QPSDM_REQUIRE(((uintptr_t)raw_dev & 3) == 0, "segy_decode_dev: records must start on a 4-byte boundary");
  if (ntr == 0) return QPSDM_OK;
  // key numbers in struct segy order, standard positions when keyidx is NULL
  auto pos = [&](int k, int std_off, int std_bytes) {
    if (!keyidx) return KeyPos{std_off, std_bytes};
    const int off = keyidx[k].start - 1, w = keyidx[k].bytes;
    if ((w != 2 && w != 4) || off < 0 || off + w > 240) return KeyPos{0, 0};  // not read: 0
    return KeyPos{off, w};
  };
  HdrKeys k{pos(21, 72, 4), pos(22, 76, 4), pos(23, 80, 4), pos(24, 84, 4), pos(2, 8, 4), pos(5, 20, 4),
            pos(11, 36, 4), pos(0, 0, 4),   pos(20, 70, 2), pos(38, 114, 2), pos(39, 116, 2)};
  const hipStream_t s = (hipStream_t)stream;
  if (format == 1)
    hipLaunchKernelGGL(segy_decode_kernel<1>, dim3(ntr), dim3(256), 0, s, (const unsigned char *)raw_dev, ns, k,
                       samples_dev, hdrs_dev);
  else
    hipLaunchKernelGGL(segy_decode_kernel<5>, dim3(ntr), dim3(256), 0, s, (const unsigned char *)raw_dev, ns, k,
                       samples_dev, hdrs_dev);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

struct qpsdm_ctx {
  qpsdm_ctx_params p;
  int device = 0;
  hipStream_t copy = nullptr, compute = nullptr;
  float *img = nullptr;
  void *tab = nullptr;
  bool have_tab = false;
  unsigned *work = nullptr;  // the batched entry point's work counter (span form)
  Slot slot[kSlots];
  // the slots' geometry and bank rows in one allocation each (slot i from
  // row i * batch_traces), so that the span form images both slots' shots
  // in one batched call
  float *geom_all = nullptr;
  void *bank_all = nullptr;
  // span form: slot i's traces as one shot (pinned host copy per slot; on the
  // device slot i's descriptor at entry i, slot 0's also at entry kSlots, so
  // that entries [i, i + 1] hold slot i's shot then the other slot's)
  qpsdm_shot *shot_host = nullptr, *shots_dev = nullptr;
  int pending = -1;  // span form: a slot banked and not yet imaged (or -1)
  int cur = 0;
  size_t rec_bytes = 0, slot_bytes = 0;
  long long pushed = 0, rejected_host = 0;
  // dip-angle gathers accumulated beside the image (ngcol = 0: none)
  qpsdm_gather_params gp{};
  int ngcol = 0, nbx = 0, nby = 0;
  int *gcols = nullptr;
  float *gx = nullptr, *gy = nullptr;
};

namespace {

// Span form: the shots of slots a, a + 1, ... (n of them, in that order:
// the order their traces came in) as one batched call, each slot's traces
// one shot under the context's dip bounds (each trace's source leg from the
// node tables at its own source); the slots' buffers are free again when
// their done events fire.
int image_slots(qpsdm_ctx *c, int a, int n) {
  const qpsdm_image_params &ip = c->p.image;
  int rc = qpsdm_image_batch(&ip, &c->p.shape, c->img, c->geom_all, c->bank_all, c->p.bank_dtype, c->tab,
                             c->p.table_dtype, c->shots_dev + a, n, c->work, -1, c->compute);
  for (int k = 0; k < n; ++k) {
    const hipError_t e = hipEventRecord(c->slot[(a + k) % kSlots].done, c->compute);
    if (!rc && e != hipSuccess) rc = qpsdm::set_error(QPSDM_EHIP, "ctx: recording a slot's event failed");
  }
  c->pending = -1;
  return rc;
}

// A held slot imaged alone (before anything that changes what imaging reads).
int image_pending(qpsdm_ctx *c) { return c->pending < 0 ? QPSDM_OK : image_slots(c, c->pending, 1); }

int submit(qpsdm_ctx *c) {
  Slot &s = c->slot[c->cur];
  if (s.ntr == 0) return QPSDM_OK;
  const int nt = c->p.image.nt;
  // the device buffers of this slot may still be read by its last batch
  if (s.used) QPSDM_HIP(hipStreamWaitEvent(c->copy, s.done, 0));
  if (s.su) {
    QPSDM_HIP(hipMemcpyAsync(s.raw, s.host, s.ntr * c->rec_bytes, hipMemcpyHostToDevice, c->copy));
  } else {
    // staging layout: samples [B][nt] then geometry [B][4]
    QPSDM_HIP(hipMemcpyAsync(s.samples, s.host, (size_t)s.ntr * nt * 4, hipMemcpyHostToDevice, c->copy));
    QPSDM_HIP(hipMemcpyAsync(s.geom, s.host + (size_t)c->p.batch_traces * nt * 4, (size_t)s.ntr * 16,
                             hipMemcpyHostToDevice, c->copy));
  }
  if (c->p.image.weight_nodes == 2) {
    // the slot's shot descriptor, with the staged data (the host rewrites it
    // only after this slot's copies are done; the device entries only after
    // the slot's last imaging, which the copy stream waited for above)
    const qpsdm_image_params &ip = c->p.image;
    const int si = c->cur;
    c->shot_host[si] = qpsdm_shot{si * c->p.batch_traces, s.ntr, ip.tanx1, ip.tanx2, ip.tany1, ip.tany2};
    QPSDM_HIP(hipMemcpyAsync(c->shots_dev + si, c->shot_host + si, sizeof(qpsdm_shot), hipMemcpyHostToDevice,
                             c->copy));
    if (si == 0)
      QPSDM_HIP(hipMemcpyAsync(c->shots_dev + kSlots, c->shot_host, sizeof(qpsdm_shot), hipMemcpyHostToDevice,
                               c->copy));
  }
  QPSDM_HIP(hipEventRecord(s.copied, c->copy));
  QPSDM_HIP(hipStreamWaitEvent(c->compute, s.copied, 0));
  if (s.su) {
    GeomDev g{c->p.geom.scaled, c->p.geom.fxv1, c->p.geom.exv1, c->p.geom.fyv1, c->p.geom.eyv1,
              c->p.geom.ox, c->p.geom.oy};
    if (s.fmt == 1)
      hipLaunchKernelGGL(record_decode_kernel<1>, dim3(s.ntr), dim3(256), 0, c->compute, s.raw, nt, s.keys, g,
                         s.samples, (float4 *)s.geom);
    else
      hipLaunchKernelGGL(record_decode_kernel<5>, dim3(s.ntr), dim3(256), 0, c->compute, s.raw, nt, s.keys, g,
                         s.samples, (float4 *)s.geom);
    QPSDM_HIP(hipGetLastError());
  }
  int rc = qpsdm_qbank(&c->p.bank, s.samples, s.ntr, s.bank, c->p.bank_dtype, c->compute);
  if (rc) return rc;
  if (c->p.image.weight_nodes == 2) {
    // the span form images through the batched entry point, two slots per
    // call (the bands of the second shot then follow the first's on each
    // band stream instead of waiting for all of them): this slot is held
    // until the other slot is banked, or imaged alone at a flush
    if (c->pending < 0) {
      c->pending = c->cur;
    } else {
      rc = image_slots(c, c->pending, 2);
      if (rc) return rc;
    }
  } else {
    rc = qpsdm_image_shot(&c->p.image, &c->p.shape, c->img, s.geom, 0, s.ntr, s.bank, c->p.bank_dtype, c->tab,
                          c->p.table_dtype, c->compute);
    if (rc) return rc;
    if (c->ngcol) {
      rc = qpsdm_gathers_shot(&c->p.image, &c->gp, c->gcols, c->ngcol, c->gx, c->gy, s.geom, 0, s.ntr, s.bank,
                              c->p.bank_dtype, c->tab, c->p.table_dtype, c->compute);
      if (rc) return rc;
    }
    QPSDM_HIP(hipEventRecord(s.done, c->compute));
  }
  s.used = true;
  s.ntr = 0;
  c->cur = (c->cur + 1) % kSlots;
  // the next slot's pinned buffer is refilled by the host: its previous copy
  // must have finished
  Slot &n = c->slot[c->cur];
  if (n.used) QPSDM_HIP(hipEventSynchronize(n.copied));
  return QPSDM_OK;
}

inline bool same_keys(const GeomKeys &a, const GeomKeys &b) {
  return a.sx.off == b.sx.off && a.sx.bytes == b.sx.bytes && a.sy.off == b.sy.off && a.sy.bytes == b.sy.bytes &&
         a.gx.off == b.gx.off && a.gx.bytes == b.gx.bytes && a.gy.off == b.gy.off && a.gy.bytes == b.gy.bytes;
}

// The current slot takes what comes next: records of one format and key
// layout, or decoded traces; a slot holding something else is submitted.
int ensure_mode(qpsdm_ctx *c, bool su, int fmt = 5, const GeomKeys &keys = kStdGeom) {
  Slot &s = c->slot[c->cur];
  if (s.ntr > 0 && (s.su != su || (su && (s.fmt != fmt || !same_keys(s.keys, keys))))) {
    int rc = submit(c);
    if (rc) return rc;
  }
  Slot &n = c->slot[c->cur];
  n.su = su;
  n.fmt = fmt;
  n.keys = keys;
  return QPSDM_OK;
}

// Records (240-byte header + nt big-endian samples each) into the staging
// slots; every record's ns (read at ns_key) must equal nt.
int push_records(qpsdm_ctx *c, const unsigned char *b, size_t nbytes, int fmt, const GeomKeys &keys,
                 KeyPos ns_key, const char *who) {
  QPSDM_REQUIRE(c->have_tab, "%s: tables not loaded", who);
  QPSDM_REQUIRE(nbytes % c->rec_bytes == 0, "%s: %zu bytes is not a whole number of %zu-byte records", who, nbytes,
                c->rec_bytes);
  const size_t n = nbytes / c->rec_bytes;
  for (size_t i = 0; i < n; ++i) {
    const unsigned char *p = b + i * c->rec_bytes + ns_key.off;
    const int ns = ns_key.bytes == 2 ? (p[0] << 8) | p[1] : (p[2] << 8) | p[3];
    QPSDM_REQUIRE(ns == c->p.image.nt, "%s: record %zu has ns=%d, expected %d", who, i, ns, c->p.image.nt);
  }
  int rc = ensure_mode(c, true, fmt, keys);
  if (rc) return rc;
  size_t done = 0;
  while (done < n) {
    Slot &s = c->slot[c->cur];
    const size_t take = std::min(n - done, (size_t)(c->p.batch_traces - s.ntr));
    memcpy(s.host + (size_t)s.ntr * c->rec_bytes, b + done * c->rec_bytes, take * c->rec_bytes);
    s.ntr += (int)take;
    done += take;
    c->pushed += (long long)take;
    if (s.ntr == c->p.batch_traces) {
      rc = submit(c);
      if (rc) return rc;
      Slot &t = c->slot[c->cur];
      t.su = true;
      t.fmt = fmt;
      t.keys = keys;
    }
  }
  return QPSDM_OK;
}

}  // namespace

extern "C" int qpsdm_ctx_create(const qpsdm_ctx_params *p, qpsdm_ctx **out) {
  QPSDM_REQUIRE(p && out, "ctx_create: null args");
  QPSDM_REQUIRE(p->batch_traces >= 1 && p->batch_traces <= (1 << 20), "ctx_create: batch_traces in [1, 2^20]");
  QPSDM_REQUIRE(p->image.nt == p->bank.nt && p->image.K == p->bank.K && p->image.dtstar == p->bank.dtstar,
                "ctx_create: image and bank parameters disagree on nt/K/dtstar");
  if (p->image.weight_nodes == 2)
    QPSDM_REQUIRE((p->bank_dtype == QPSDM_F16P && p->table_dtype == QPSDM_U16Q) ||
                      (p->bank_dtype == QPSDM_F32P && p->table_dtype == QPSDM_F32),
                  "ctx_create: the span form takes a QPSDM_F16P bank with QPSDM_U16Q tables or a QPSDM_F32P bank "
                  "with QPSDM_F32 tables");
  else
    QPSDM_REQUIRE(p->bank_dtype == QPSDM_F32 || p->bank_dtype == QPSDM_BF16 || p->bank_dtype == QPSDM_F16,
                  "ctx_create: bank dtype");
  QPSDM_REQUIRE(p->table_dtype == QPSDM_F32 || p->table_dtype == QPSDM_F16B || p->table_dtype == QPSDM_U16B ||
                    p->table_dtype == QPSDM_U16X || p->table_dtype == QPSDM_U16Q,
                "ctx_create: table dtype");
  qpsdm_ctx *c = new qpsdm_ctx();
  c->p = *p;
  auto fail = [&](int rc) {
    qpsdm_ctx_destroy(c);
    return rc;
  };
  if (hipGetDevice(&c->device) != hipSuccess) return fail(qpsdm::set_error(QPSDM_EHIP, "no HIP device"));
  const int nt = p->image.nt, B = p->batch_traces;
  c->rec_bytes = 240 + 4 * (size_t)nt;
  c->slot_bytes = (size_t)B * (c->rec_bytes > (size_t)(nt + 4) * 4 ? c->rec_bytes : (size_t)(nt + 4) * 4);
  const size_t bank_bytes = (size_t)B * qpsdm_bank_row_bytes(p->bank.K, nt, p->bank_dtype);
  const size_t img_bytes = (size_t)p->image.nx * p->image.ny * p->image.nz * 4;
  if (hipStreamCreateWithFlags(&c->copy, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->compute, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->img, img_bytes) != hipSuccess || hipMemset(c->img, 0, img_bytes) != hipSuccess)
    return fail(qpsdm::set_error(QPSDM_EHIP, "ctx_create: stream/image allocation failed"));
  if (hipMalloc((void **)&c->geom_all, (size_t)kSlots * B * 16) != hipSuccess ||
      hipMalloc(&c->bank_all, kSlots * bank_bytes) != hipSuccess ||
      hipHostMalloc((void **)&c->shot_host, kSlots * sizeof(qpsdm_shot), hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void **)&c->shots_dev, (kSlots + 1) * sizeof(qpsdm_shot)) != hipSuccess)
    return fail(qpsdm::set_error(QPSDM_EHIP, "ctx_create: slot allocation failed"));
  for (int i = 0; i < kSlots; ++i) {
    Slot &s = c->slot[i];
    s.geom = c->geom_all + (size_t)i * B * 4;
    s.bank = (char *)c->bank_all + i * bank_bytes;
    if (hipHostMalloc(&s.host, c->slot_bytes, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&s.raw, c->slot_bytes) != hipSuccess ||
        hipMalloc(&s.samples, (size_t)B * nt * 4) != hipSuccess ||
        hipEventCreateWithFlags(&s.copied, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess)
      return fail(qpsdm::set_error(QPSDM_EHIP, "ctx_create: slot allocation failed"));
  }
  if (hipMalloc((void **)&c->work, sizeof(unsigned)) != hipSuccess)
    return fail(qpsdm::set_error(QPSDM_EHIP, "ctx_create: work counter allocation failed"));
  *out = c;
  return QPSDM_OK;
}

extern "C" int qpsdm_ctx_destroy(qpsdm_ctx *c) {
  if (!c) return QPSDM_OK;
  // teardown: errors here have nowhere useful to go, release everything
  if (c->compute) (void)hipStreamSynchronize(c->compute);
  if (c->copy) (void)hipStreamSynchronize(c->copy);
  for (Slot &s : c->slot) {
    if (s.host) (void)hipHostFree(s.host);
    if (s.raw) (void)hipFree(s.raw);
    if (s.samples) (void)hipFree(s.samples);
    if (s.copied) (void)hipEventDestroy(s.copied);
    if (s.done) (void)hipEventDestroy(s.done);
  }
  if (c->geom_all) (void)hipFree(c->geom_all);
  if (c->bank_all) (void)hipFree(c->bank_all);
  if (c->shot_host) (void)hipHostFree(c->shot_host);
  if (c->shots_dev) (void)hipFree(c->shots_dev);
  if (c->work) (void)hipFree(c->work);
  if (c->img) (void)hipFree(c->img);
  if (c->gcols) (void)hipFree(c->gcols);
  if (c->gx) (void)hipFree(c->gx);
  if (c->gy) (void)hipFree(c->gy);
  if (c->tab) (void)hipFree(c->tab);
  if (c->copy) (void)hipStreamDestroy(c->copy);
  if (c->compute) (void)hipStreamDestroy(c->compute);
  delete c;
  return QPSDM_OK;
}

// Tables in the FP32 layout (host or device); quantised on the device when
// the context stores 16-bit bricks.  The new set is complete before the old
// one is released: on any error the context keeps its previous tables (or
// none) and nothing allocated here is left behind.
extern "C" int qpsdm_ctx_set_tables(qpsdm_ctx *c, const float *tab, int on_device) {
  QPSDM_REQUIRE(c && tab, "ctx_set_tables: null args");
  {
    // a held slot is imaged under the tables current when it was submitted
    const int rc = image_pending(c);
    if (rc) return rc;
  }
  const qpsdm_tgeom &g = c->p.image.tg;
  const size_t n32 = (size_t)g.nnx * g.nny * g.nv * g.nu * g.nw * 8;
  float *t32 = nullptr;
  void *q = nullptr;
  auto fail = [&](int rc) {
    (void)hipStreamSynchronize(c->compute);
    if (t32) (void)hipFree(t32);
    if (q) (void)hipFree(q);
    return rc;
  };
  if (hipMalloc(&t32, n32) != hipSuccess) {
    t32 = nullptr;
    return fail(qpsdm::set_error(QPSDM_EHIP, "ctx_set_tables: hipMalloc of %zu bytes failed", n32));
  }
  if (hipMemcpyAsync(t32, tab, n32, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->compute) !=
      hipSuccess)
    return fail(qpsdm::set_error(QPSDM_EHIP, "ctx_set_tables: table copy failed"));
  void *fresh = t32;
  if (c->p.table_dtype != QPSDM_F32) {
    const size_t nq = qpsdm_tables_bytes(&g, c->p.table_dtype);
    if (hipMalloc(&q, nq) != hipSuccess) {
      q = nullptr;
      return fail(qpsdm::set_error(QPSDM_EHIP, "ctx_set_tables: hipMalloc of %zu bytes failed", nq));
    }
    const int rc = qpsdm_tables_quantize(&g, t32, q, c->p.table_dtype, c->compute);
    if (rc) return fail(rc);
    if (hipStreamSynchronize(c->compute) != hipSuccess)
      return fail(qpsdm::set_error(QPSDM_EHIP, "ctx_set_tables: quantize failed"));
    (void)hipFree(t32);
    t32 = nullptr;
    fresh = q;
  }
  if (hipStreamSynchronize(c->compute) != hipSuccess)
    return fail(qpsdm::set_error(QPSDM_EHIP, "ctx_set_tables: table copy failed"));
  if (c->tab) (void)hipFree(c->tab);
  c->tab = fresh;
  c->have_tab = true;
  return QPSDM_OK;
}

extern "C" int qpsdm_ctx_build_tables(qpsdm_ctx *c, const qpsdm_model *m) {
  QPSDM_REQUIRE(c && m, "ctx_build_tables: null args");
  const qpsdm_tgeom &g = c->p.image.tg;
  float *t32 = nullptr;
  QPSDM_HIP(hipMalloc(&t32, (size_t)g.nnx * g.nny * g.nv * g.nu * g.nw * 8));
  int rc = qpsdm_tables_build(m, &g, t32, c->compute);
  if (!rc) rc = qpsdm_ctx_set_tables(c, t32, 1);
  QPSDM_HIP(hipStreamSynchronize(c->compute));
  QPSDM_HIP(hipFree(t32));
  return rc;
}

// Big-endian SU records of ns = nt samples (a line file's bytes, as the
// producer reads them, src/q-psdm.cu:1887-1916).
extern "C" int qpsdm_ctx_push_su(qpsdm_ctx *c, const void *bytes, size_t nbytes) {
  QPSDM_REQUIRE(c && (bytes || nbytes == 0), "ctx_push_su: null args");
  return push_records(c, (const unsigned char *)bytes, nbytes, 5, kStdGeom, KeyPos{114, 2}, "ctx_push_su");
}

// SEG-Y rev0 trace records (no reel header) of ns = nt samples in format 1
// (IBM float) or 5 (IEEE float), sx/sy/gx/gy and ns at their keyindex
// positions (NULL: standard), decoded on the device as fsgettr decodes them
// (src/segy.cc:350-1547).
extern "C" int qpsdm_ctx_push_segy(qpsdm_ctx *c, const void *bytes, size_t nbytes, int format,
                                   const keyindex *keyidx) {
  QPSDM_REQUIRE(c && (bytes || nbytes == 0), "ctx_push_segy: null args");
  QPSDM_REQUIRE(format == 1 || format == 5, "ctx_push_segy: format %d (1 = IBM float, 5 = IEEE float)", format);
  GeomKeys keys = kStdGeom;
  KeyPos ns_key{114, 2};
  if (keyidx) {
    auto pos = [&](int k, KeyPos *out) {
      const int off = keyidx[k].start - 1, w = keyidx[k].bytes;
      if ((w != 2 && w != 4) || off < 0 || off + w > 240)
        return qpsdm::set_error(QPSDM_EINVAL, "ctx_push_segy: key %d at byte %d width %d does not fit the header", k,
                                keyidx[k].start, w);
      *out = KeyPos{off, w};
      return QPSDM_OK;
    };
    // key numbers in struct segy order: sx 21, sy 22, gx 23, gy 24, ns 38
    int rc = pos(21, &keys.sx);
    if (!rc) rc = pos(22, &keys.sy);
    if (!rc) rc = pos(23, &keys.gx);
    if (!rc) rc = pos(24, &keys.gy);
    if (!rc) rc = pos(38, &ns_key);
    if (rc) return rc;
  }
  return push_records(c, (const unsigned char *)bytes, nbytes, format, keys, ns_key, "ctx_push_segy");
}

// Decoded traces from host memory: samples [ntr][nt], geometry [ntr][4]
// image-relative (NaN sx rejects a trace).
extern "C" int qpsdm_ctx_push_traces(qpsdm_ctx *c, const float *samples, const float *geom, int ntr) {
  QPSDM_REQUIRE(c && ((samples && geom) || ntr == 0) && ntr >= 0, "ctx_push_traces: bad args");
  QPSDM_REQUIRE(c->have_tab, "ctx_push_traces: tables not loaded");
  int rc = ensure_mode(c, false);
  if (rc) return rc;
  const int nt = c->p.image.nt, B = c->p.batch_traces;
  int done = 0;
  while (done < ntr) {
    Slot &s = c->slot[c->cur];
    if (s.ntr == 0) s.su = false;
    const int take = std::min(ntr - done, B - s.ntr);
    // staging layout: samples of the whole slot first, then its geometry
    float *hs = (float *)s.host;
    memcpy(hs + (size_t)s.ntr * nt, samples + (size_t)done * nt, (size_t)take * nt * 4);
    float *hg = hs + (size_t)B * nt;
    memcpy(hg + (size_t)s.ntr * 4, geom + (size_t)done * 4, (size_t)take * 16);
    s.ntr += take;
    done += take;
    c->pushed += take;
    if (s.ntr == B) {
      rc = submit(c);
      if (rc) return rc;
    }
  }
  return QPSDM_OK;
}

// The per-trace entry point of the reference, psdm_kernel(segy tr, ...)
// (src/q-psdm.cu:2063-2244): one host trace with its SU header; geometry
// from sx, sy, gx, gy with the context's scale and model extent.
extern "C" int qpsdm_ctx_psdm_kernel(qpsdm_ctx *c, const segy *tr) {
  QPSDM_REQUIRE(c && tr, "ctx_psdm_kernel: null args");
  QPSDM_REQUIRE(tr->ns == c->p.image.nt, "ctx_psdm_kernel: ns=%d, expected %d", tr->ns, c->p.image.nt);
  const qpsdm_geom_params &g = c->p.geom;
  const double sx = tr->sx * g.scaled, sy = tr->sy * g.scaled, gx = tr->gx * g.scaled, gy = tr->gy * g.scaled;
  if (sx < g.fxv1 || sx > g.exv1 || gx < g.fxv1 || gx > g.exv1 || sy < g.fyv1 || sy > g.eyv1 ||
      gy < g.fyv1 || gy > g.eyv1) {
    ++c->rejected_host;
    return QPSDM_OK;  // outside the model extent: no contribution (src/q-psdm.cu:2120-2123)
  }
  const float geom[4] = {(float)(sx - g.ox), (float)(sy - g.oy), (float)(gx - g.ox), (float)(gy - g.oy)};
  return qpsdm_ctx_push_traces(c, tr->data, geom, 1);
}

extern "C" int qpsdm_ctx_flush(qpsdm_ctx *c) {
  QPSDM_REQUIRE(c, "ctx_flush: null ctx");
  int rc = submit(c);
  if (!rc) rc = image_pending(c);
  if (rc) return rc;
  QPSDM_HIP(hipStreamSynchronize(c->compute));
  return QPSDM_OK;
}

// Image [ny][nx][nz] to host memory, divided by nsb when normalize != 0
// (src/q-psdm.cu:1760-1766).
extern "C" int qpsdm_ctx_image(qpsdm_ctx *c, float *host, int normalize) {
  QPSDM_REQUIRE(c && host, "ctx_image: null args");
  int rc = qpsdm_ctx_flush(c);
  if (rc) return rc;
  const size_t n = (size_t)c->p.image.nx * c->p.image.ny * c->p.image.nz;
  QPSDM_HIP(hipMemcpy(host, c->img, n * 4, hipMemcpyDeviceToHost));
  if (normalize) {
    const float s = 1.0f / (float)c->p.bank.win.nsb;
    for (size_t i = 0; i < n; ++i) host[i] *= s;
  }
  return QPSDM_OK;
}

extern "C" int qpsdm_ctx_reset_image(qpsdm_ctx *c) {
  QPSDM_REQUIRE(c, "ctx_reset_image: null ctx");
  int rc = qpsdm_ctx_flush(c);
  if (rc) return rc;
  QPSDM_HIP(hipMemsetAsync(c->img, 0, (size_t)c->p.image.nx * c->p.image.ny * c->p.image.nz * 4, c->compute));
  const size_t nz = (size_t)c->p.image.nz;
  if (c->gx) QPSDM_HIP(hipMemsetAsync(c->gx, 0, (size_t)c->ngcol * c->nbx * nz * 4, c->compute));
  if (c->gy) QPSDM_HIP(hipMemsetAsync(c->gy, 0, (size_t)c->ngcol * c->nby * nz * 4, c->compute));
  QPSDM_HIP(hipStreamSynchronize(c->compute));
  return QPSDM_OK;
}

// Dip-angle gathers at the columns cols (host (ix, iy) pairs), accumulated
// by every later batch beside the image (the reference's opcrp/cdps/aspect/
// dip options, src/q-psdm.cu:651-702, 820-835, and its gthrxd/gthryd
// buffers, src/q-psdm.cu:211-212).  Replaces an earlier set; ncol = 0 stops
// gathering.
extern "C" int qpsdm_ctx_set_gathers(qpsdm_ctx *c, const qpsdm_gather_params *g, const int *cols, int ncol) {
  QPSDM_REQUIRE(c && ncol >= 0 && (cols || ncol == 0) && (g || ncol == 0), "ctx_set_gathers: bad args");
  int nbx = 0, nby = 0;
  if (ncol) {
    int rc = qpsdm_gather_bins(g, &nbx, &nby);
    if (rc) return rc;
    QPSDM_REQUIRE(c->p.table_dtype == QPSDM_F32 || c->p.table_dtype == QPSDM_U16X || c->p.table_dtype == QPSDM_U16Q,
                  "ctx_set_gathers: gathers need QPSDM_F32, QPSDM_U16X or QPSDM_U16Q tables");
    QPSDM_REQUIRE(c->p.bank_dtype != QPSDM_F16P && c->p.bank_dtype != QPSDM_F32P,
                  "ctx_set_gathers: gathers read a plain bank (the span form's pair bank is not one)");
    for (int i = 0; i < ncol; ++i)
      QPSDM_REQUIRE(cols[2 * i] >= 0 && cols[2 * i] < c->p.image.nx && cols[2 * i + 1] >= 0 &&
                        cols[2 * i + 1] < c->p.image.ny,
                    "ctx_set_gathers: column %d (%d, %d) outside the %d x %d image", i, cols[2 * i],
                    cols[2 * i + 1], c->p.image.nx, c->p.image.ny);
  }
  int rc = qpsdm_ctx_flush(c);
  if (rc) return rc;
  QPSDM_HIP(hipStreamSynchronize(c->compute));
  if (c->gcols) QPSDM_HIP(hipFree(c->gcols));
  if (c->gx) QPSDM_HIP(hipFree(c->gx));
  if (c->gy) QPSDM_HIP(hipFree(c->gy));
  c->gcols = nullptr;
  c->gx = c->gy = nullptr;
  c->ngcol = c->nbx = c->nby = 0;
  if (!ncol) return QPSDM_OK;
  const size_t nz = (size_t)c->p.image.nz;
  QPSDM_HIP(hipMalloc(&c->gcols, (size_t)ncol * 8));
  QPSDM_HIP(hipMemcpy(c->gcols, cols, (size_t)ncol * 8, hipMemcpyHostToDevice));
  if (nbx) {
    QPSDM_HIP(hipMalloc(&c->gx, (size_t)ncol * nbx * nz * 4));
    QPSDM_HIP(hipMemset(c->gx, 0, (size_t)ncol * nbx * nz * 4));
  }
  if (nby) {
    QPSDM_HIP(hipMalloc(&c->gy, (size_t)ncol * nby * nz * 4));
    QPSDM_HIP(hipMemset(c->gy, 0, (size_t)ncol * nby * nz * 4));
  }
  c->gp = *g;
  c->ngcol = ncol;
  c->nbx = nbx;
  c->nby = nby;
  return QPSDM_OK;
}

// The gathers to host memory, divided by nsb when normalize != 0 (as the
// image, src/q-psdm.cu:1760-1766).
extern "C" int qpsdm_ctx_gathers(qpsdm_ctx *c, float *gx, float *gy, int normalize) {
  QPSDM_REQUIRE(c, "ctx_gathers: null ctx");
  QPSDM_REQUIRE(c->ngcol, "ctx_gathers: no gathers set (qpsdm_ctx_set_gathers)");
  QPSDM_REQUIRE((!c->nbx || gx) && (!c->nby || gy), "ctx_gathers: null output for a gathered axis");
  int rc = qpsdm_ctx_flush(c);
  if (rc) return rc;
  QPSDM_HIP(hipStreamSynchronize(c->compute));
  const float sc = 1.0f / (float)c->p.bank.win.nsb;
  const size_t nz = (size_t)c->p.image.nz;
  float *dev[2] = {c->gx, c->gy};
  float *host[2] = {gx, gy};
  const int nb[2] = {c->nbx, c->nby};
  for (int a = 0; a < 2; ++a) {
    if (!nb[a]) continue;
    const size_t n = (size_t)c->ngcol * nb[a] * nz;
    QPSDM_HIP(hipMemcpy(host[a], dev[a], n * 4, hipMemcpyDeviceToHost));
    if (normalize)
      for (size_t i = 0; i < n; ++i) host[a][i] *= sc;
  }
  return QPSDM_OK;
}

extern "C" long long qpsdm_ctx_traces_pushed(const qpsdm_ctx *c) { return c ? c->pushed : -1; }
