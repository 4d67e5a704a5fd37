// Big-endian SU codec: selective header byte swap driven by a field table
// (the enabled set of the reference's src/segy.h:22-65), samples swapped as
// 32-bit words.  Written for this build.
#include "segy.h"

#include <stddef.h>
#include <string.h>

#include "common.h"

static_assert(offsetof(segy, data) == 240, "SU header is 240 bytes");
static_assert(offsetof(segy, sx) == 72 && offsetof(segy, ns) == 114 && offsetof(segy, trwf) == 168,
              "SU header offsets");

namespace {

struct Field {
  unsigned short off, size;
};

// Byte offset and width of every header field the codec converts.
constexpr Field kFields[] = {
    {0, 4},   {4, 4},   {8, 4},   {12, 4},  {16, 4},  {20, 4},  {24, 4},   // tracl..cdpt
    {28, 2},  {30, 2},  {32, 2},                                         // trid nvs nhs
    {36, 4},  {40, 4},  {44, 4},  {48, 4},  {52, 4},  {56, 4},            // offset..sdel
    {68, 2},  {70, 2},                                                   // scalel scalco
    {72, 4},  {76, 4},  {80, 4},  {84, 4},                               // sx sy gx gy
    {88, 2},  {90, 2},                                                   // counit wevel
    {102, 2}, {108, 2}, {110, 2}, {112, 2},                              // tstat delrt muts mute
    {114, 2}, {116, 2},                                                  // ns dt
    {162, 2}, {168, 2},                                                  // minute trwf
    {180, 4}, {184, 4}, {188, 4}, {192, 4},                              // cdpx cdpy iline xline
};

inline void swap_field(unsigned char *p, int size) {
  if (size == 2) {
    unsigned char t = p[0];
    p[0] = p[1];
    p[1] = t;
  } else {
    unsigned char a = p[0], b = p[1];
    p[0] = p[3];
    p[1] = p[2];
    p[2] = b;
    p[3] = a;
  }
}

inline uint32_t bswap32(uint32_t v) { return __builtin_bswap32(v); }

// Raw big-endian header -> host header (non-enabled fields stay raw).
void header_in(const unsigned char *raw, segy *tr) {
  unsigned char *h = reinterpret_cast<unsigned char *>(tr);
  memcpy(h, raw, 240);
  for (const Field &f : kFields) swap_field(h + f.off, f.size);
}

// Host header -> big-endian header with only the enabled fields set.
void header_out(const segy *tr, unsigned char *raw) {
  const unsigned char *h = reinterpret_cast<const unsigned char *>(tr);
  memset(raw, 0, 240);
  for (const Field &f : kFields) {
    memcpy(raw + f.off, h + f.off, f.size);
    swap_field(raw + f.off, f.size);
  }
}

void samples_swap(const void *in, void *out, int ns) {
  const uint32_t *a = static_cast<const uint32_t *>(in);
  uint32_t *b = static_cast<uint32_t *>(out);
  for (int i = 0; i < ns; ++i) b[i] = bswap32(a[i]);
}

}  // namespace

int fgettr(FILE *fp, segy *tr) {
  unsigned char raw[240];
  if (fread(raw, 240, 1, fp) != 1) return 0;
  header_in(raw, tr);
  const int ns = tr->ns;
  if ((int)fread(tr->data, sizeof(float), ns, fp) != ns) return 0;
  samples_swap(tr->data, tr->data, ns);
  return ns;
}

void fputtr(FILE *fp, const segy *tr) {
  static thread_local unsigned char buf[240 + 4 * SU_NFLTS];
  header_out(tr, buf);
  samples_swap(tr->data, buf + 240, tr->ns);
  fwrite(buf, 240 + 4 * (size_t)tr->ns, 1, fp);
}

int sgettr(unsigned char *buf, segy *tr) {
  header_in(buf, tr);
  const int ns = tr->ns;
  samples_swap(buf + 240, tr->data, ns);
  return ns;
}

void sputtr(unsigned char *buf, const segy *tr) {
  header_out(tr, buf);
  samples_swap(tr->data, buf + 240, tr->ns);
}

extern "C" int qpsdm_su_fgettr(FILE *fp, segy *tr) { return fgettr(fp, tr); }
extern "C" int qpsdm_su_fputtr(FILE *fp, const segy *tr) {
  fputtr(fp, tr);
  return QPSDM_OK;
}
extern "C" int qpsdm_su_sgettr(const unsigned char *buf, segy *tr) {
  return sgettr(const_cast<unsigned char *>(buf), tr);
}
extern "C" int qpsdm_su_sputtr(unsigned char *buf, const segy *tr) {
  sputtr(buf, tr);
  return QPSDM_OK;
}

extern "C" long qpsdm_su_decode(const unsigned char *bytes, size_t nbytes, int ns, float *samples,
                                qpsdm_trhdr *hdrs) {
  QPSDM_REQUIRE(bytes && ns > 0 && ns <= SU_NFLTS, "su_decode: bad args");
  const size_t rec = 240 + 4 * (size_t)ns;
  const long n = (long)(nbytes / rec);
  for (long i = 0; i < n; ++i) {
    const unsigned char *r = bytes + (size_t)i * rec;
    segy h;
    header_in(r, &h);
    if (h.ns != ns) return qpsdm::set_error(QPSDM_EINVAL, "record %ld has ns=%d, expected %d", i, h.ns, ns);
    if (samples) samples_swap(r + 240, samples + (size_t)i * ns, ns);
    if (hdrs) {
      qpsdm_trhdr &o = hdrs[i];
      o.sx = h.sx;
      o.sy = h.sy;
      o.gx = h.gx;
      o.gy = h.gy;
      o.fldr = h.fldr;
      o.cdp = h.cdp;
      o.offset = h.offset;
      o.tracl = h.tracl;
      o.scalco = h.scalco;
      o.pad = 0;
      o.ns = h.ns;
      o.dt = h.dt;
    }
  }
  return n;
}
