// Single-GPU imaging pipeline behind the qpsdm_ctx_* entry points.
//
// Replaces the reference's producer/consumer threads and bounded queue
// (src/q-psdm.cu:1862-2061, src/queue.h:11-72) and psdm_kernel's per-trace
// synchronous copies and launch (src/q-psdm.cu:2218-2243): traces are
// batched into two pinned host slots; a copy stream moves a slot to the
// device while the compute stream decodes the previous one (big-endian SU
// -> FP32 samples + image-relative geometry, on the GPU), runs the Q bank
// and the imaging kernel.  Events order slot reuse.  The device image is
// owned by the context and copied back (divided by nsb) on request.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "segy.h"

namespace {

constexpr int kSlots = 2;

struct Slot {
  unsigned char *host = nullptr;  // pinned staging: SU records or decoded samples+geom
  unsigned char *raw = nullptr;   // device copy of staged SU records
  float *samples = nullptr;       // [batch][nt]
  float *geom = nullptr;          // [batch][4]
  void *bank = nullptr;           // [batch][K][nt]
  int ntr = 0;                    // traces staged
  bool su = false;                // staging holds SU records (else samples + geom)
  hipEvent_t copied = nullptr, done = nullptr;
  bool used = false;
};

struct GeomDev {
  double scaled, fxv1, exv1, fyv1, eyv1, ox, oy;
};

// One work-group per SU record: swap the samples, read sx, sy, gx, gy (and
// scalco-independent `scaled`), reject traces outside the model extent by
// writing NaN geometry (src/q-psdm.cu:2115-2123), else image-relative FP32
// coordinates with the origin removed in FP64.
__global__ void su_decode_kernel(const unsigned char *__restrict__ raw, int ns, GeomDev g,
                                 float *__restrict__ samples, float4 *__restrict__ geom) {
  const size_t rec = 240 + 4 * (size_t)ns;
  const unsigned char *r = raw + (size_t)blockIdx.x * rec;
  const unsigned *src = (const unsigned *)(r + 240);
  float *dst = samples + (size_t)blockIdx.x * ns;
  for (int i = threadIdx.x; i < ns; i += blockDim.x) dst[i] = __uint_as_float(__builtin_bswap32(src[i]));
  if (threadIdx.x == 0) {
    const unsigned *h = (const unsigned *)r;
    const double sx = (int)__builtin_bswap32(h[18]) * g.scaled;
    const double sy = (int)__builtin_bswap32(h[19]) * g.scaled;
    const double gx = (int)__builtin_bswap32(h[20]) * g.scaled;
    const double gy = (int)__builtin_bswap32(h[21]) * g.scaled;
    const bool keep = !(sx < g.fxv1 || sx > g.exv1 || gx < g.fxv1 || gx > g.exv1 || sy < g.fyv1 ||
                        sy > g.eyv1 || gy < g.fyv1 || gy > g.eyv1);
    geom[blockIdx.x] = keep ? make_float4((float)(sx - g.ox), (float)(sy - g.oy), (float)(gx - g.ox),
                                          (float)(gy - g.oy))
                            : make_float4(NAN, NAN, NAN, NAN);
  }
}

}  // namespace

struct qpsdm_ctx {
  qpsdm_ctx_params p;
  int device = 0;
  hipStream_t copy = nullptr, compute = nullptr;
  float *img = nullptr;
  void *tab = nullptr;
  bool have_tab = false;
  Slot slot[kSlots];
  int cur = 0;
  size_t rec_bytes = 0, slot_bytes = 0;
  long long pushed = 0, rejected_host = 0;
};

namespace {

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
  QPSDM_HIP(hipEventRecord(s.copied, c->copy));
  QPSDM_HIP(hipStreamWaitEvent(c->compute, s.copied, 0));
  if (s.su) {
    GeomDev g{c->p.geom.scaled, c->p.geom.fxv1, c->p.geom.exv1, c->p.geom.fyv1, c->p.geom.eyv1,
              c->p.geom.ox, c->p.geom.oy};
    hipLaunchKernelGGL(su_decode_kernel, dim3(s.ntr), dim3(256), 0, c->compute, s.raw, nt, g,
                       s.samples, (float4 *)s.geom);
    QPSDM_HIP(hipGetLastError());
  }
  int rc = qpsdm_qbank(&c->p.bank, s.samples, s.ntr, s.bank, c->p.bank_dtype, c->compute);
  if (rc) return rc;
  rc = qpsdm_image_shot(&c->p.image, &c->p.shape, c->img, s.geom, 0, s.ntr, s.bank, c->p.bank_dtype,
                        c->tab, c->p.table_dtype, c->compute);
  if (rc) return rc;
  QPSDM_HIP(hipEventRecord(s.done, c->compute));
  s.used = true;
  s.ntr = 0;
  c->cur = (c->cur + 1) % kSlots;
  // the next slot's pinned buffer is refilled by the host: its previous copy
  // must have finished
  Slot &n = c->slot[c->cur];
  if (n.used) QPSDM_HIP(hipEventSynchronize(n.copied));
  return QPSDM_OK;
}

int ensure_mode(qpsdm_ctx *c, bool su) {
  Slot &s = c->slot[c->cur];
  if (s.ntr > 0 && s.su != su) {
    int rc = submit(c);
    if (rc) return rc;
  }
  c->slot[c->cur].su = su;
  return QPSDM_OK;
}

}  // namespace

extern "C" int qpsdm_ctx_create(const qpsdm_ctx_params *p, qpsdm_ctx **out) {
  QPSDM_REQUIRE(p && out, "ctx_create: null args");
  QPSDM_REQUIRE(p->batch_traces >= 1 && p->batch_traces <= (1 << 20), "ctx_create: batch_traces in [1, 2^20]");
  QPSDM_REQUIRE(p->image.nt == p->bank.nt && p->image.K == p->bank.K && p->image.dtstar == p->bank.dtstar,
                "ctx_create: image and bank parameters disagree on nt/K/dtstar");
  QPSDM_REQUIRE(p->bank_dtype == QPSDM_F32 || p->bank_dtype == QPSDM_BF16 || p->bank_dtype == QPSDM_F16,
                "ctx_create: bank dtype");
  QPSDM_REQUIRE(p->table_dtype == QPSDM_F32 || p->table_dtype == QPSDM_F16B || p->table_dtype == QPSDM_U16B ||
                    p->table_dtype == QPSDM_U16X,
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
  for (Slot &s : c->slot) {
    if (hipHostMalloc(&s.host, c->slot_bytes, hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&s.raw, c->slot_bytes) != hipSuccess ||
        hipMalloc(&s.samples, (size_t)B * nt * 4) != hipSuccess ||
        hipMalloc(&s.geom, (size_t)B * 16) != hipSuccess ||
        hipMalloc(&s.bank, bank_bytes) != hipSuccess ||
        hipEventCreateWithFlags(&s.copied, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess)
      return fail(qpsdm::set_error(QPSDM_EHIP, "ctx_create: slot allocation failed"));
  }
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
    if (s.geom) (void)hipFree(s.geom);
    if (s.bank) (void)hipFree(s.bank);
    if (s.copied) (void)hipEventDestroy(s.copied);
    if (s.done) (void)hipEventDestroy(s.done);
  }
  if (c->img) (void)hipFree(c->img);
  if (c->tab) (void)hipFree(c->tab);
  if (c->copy) (void)hipStreamDestroy(c->copy);
  if (c->compute) (void)hipStreamDestroy(c->compute);
  delete c;
  return QPSDM_OK;
}

// Tables in the FP32 layout (host or device); quantised on the device when
// the context stores 16-bit bricks.
extern "C" int qpsdm_ctx_set_tables(qpsdm_ctx *c, const float *tab, int on_device) {
  QPSDM_REQUIRE(c && tab, "ctx_set_tables: null args");
  const qpsdm_tgeom &g = c->p.image.tg;
  const size_t n32 = (size_t)g.nnx * g.nny * g.nv * g.nu * g.nw * 8;
  float *t32 = nullptr;
  QPSDM_HIP(hipMalloc(&t32, n32));
  QPSDM_HIP(hipMemcpyAsync(t32, tab, n32, on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                           c->compute));
  if (c->tab) {
    QPSDM_HIP(hipStreamSynchronize(c->compute));
    QPSDM_HIP(hipFree(c->tab));
    c->tab = nullptr;
  }
  if (c->p.table_dtype == QPSDM_F32) {
    c->tab = t32;
  } else {
    void *q = nullptr;
    QPSDM_HIP(hipMalloc(&q, qpsdm_tables_bytes(&g, c->p.table_dtype)));
    int rc = qpsdm_tables_quantize(&g, t32, q, c->p.table_dtype, c->compute);
    QPSDM_HIP(hipStreamSynchronize(c->compute));
    QPSDM_HIP(hipFree(t32));
    if (rc) return rc;
    c->tab = q;
  }
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
  QPSDM_REQUIRE(c->have_tab, "ctx_push_su: tables not loaded");
  QPSDM_REQUIRE(nbytes % c->rec_bytes == 0, "ctx_push_su: %zu bytes is not a whole number of %zu-byte records",
                nbytes, c->rec_bytes);
  const unsigned char *b = (const unsigned char *)bytes;
  const size_t n = nbytes / c->rec_bytes;
  for (size_t i = 0; i < n; ++i) {
    const unsigned char *r = b + i * c->rec_bytes;
    const int ns = (r[114] << 8) | r[115];
    QPSDM_REQUIRE(ns == c->p.image.nt, "ctx_push_su: record %zu has ns=%d, expected %d", i, ns, c->p.image.nt);
  }
  int rc = ensure_mode(c, true);
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
      c->slot[c->cur].su = true;
    }
  }
  return QPSDM_OK;
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
  QPSDM_HIP(hipStreamSynchronize(c->compute));
  return QPSDM_OK;
}

extern "C" long long qpsdm_ctx_traces_pushed(const qpsdm_ctx *c) { return c ? c->pushed : -1; }
