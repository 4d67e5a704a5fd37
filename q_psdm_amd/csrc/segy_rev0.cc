// SEG-Y rev0 reader: reel header (EBCDIC text + big-endian binary header),
// trace headers with keys at caller-chosen positions, IBM or IEEE samples.
// Behaviour follows the reference's fsgethd/initkeyindex/setkeyindex/
// fsgettr/ssgettr (src/segy.cc:83-2752); written for this build.
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include "common.h"
#include "ibm_float.h"
#include "segy.h"

static_assert(sizeof(hd3600) == 3600, "reel header is 3600 bytes");
static_assert(offsetof(hd3600, jobid) == 3200 && offsetof(hd3600, ntrpr) == 3212 &&
                  offsetof(hd3600, vpol) == 3258 && offsetof(hd3600, hunass) == 3260,
              "binary header offsets");

int keyindex_inited = 0;

namespace {

enum Kind : unsigned char { I32, I16, U16 };

struct Key {
  const char *name;
  unsigned short off;  // byte offset in the 240-byte header / struct segy
  Kind kind;
  bool read;           // decoded by the reader (the SU codec's header set)
};

// The 75 keys of struct segy from tracl to xline, in struct order, then the
// five the name lookup also knows (sp .. shortpad, never read).
constexpr Key kKeys[] = {
    {"tracl", 0, I32, true},     {"tracr", 4, I32, true},     {"fldr", 8, I32, true},
    {"tracf", 12, I32, true},    {"ep", 16, I32, true},       {"cdp", 20, I32, true},
    {"cdpt", 24, I32, true},     {"trid", 28, I16, true},     {"nvs", 30, I16, true},
    {"nhs", 32, I16, true},      {"duse", 34, I16, false},    {"offset", 36, I32, true},
    {"gelev", 40, I32, true},    {"selev", 44, I32, true},    {"sdepth", 48, I32, true},
    {"gdel", 52, I32, true},     {"sdel", 56, I32, true},     {"swdep", 60, I32, false},
    {"gwdep", 64, I32, false},   {"scalel", 68, I16, true},   {"scalco", 70, I16, true},
    {"sx", 72, I32, true},       {"sy", 76, I32, true},       {"gx", 80, I32, true},
    {"gy", 84, I32, true},       {"counit", 88, I16, true},   {"wevel", 90, I16, true},
    {"swevel", 92, I16, false},  {"sut", 94, I16, false},     {"gut", 96, I16, false},
    {"sstat", 98, I16, false},   {"gstat", 100, I16, false},  {"tstat", 102, I16, true},
    {"laga", 104, I16, false},   {"lagb", 106, I16, false},   {"delrt", 108, I16, true},
    {"muts", 110, I16, true},    {"mute", 112, I16, true},    {"ns", 114, U16, true},
    {"dt", 116, U16, true},      {"gain", 118, I16, false},   {"igc", 120, I16, false},
    {"igi", 122, I16, false},    {"corr", 124, I16, false},   {"sfs", 126, I16, false},
    {"sfe", 128, I16, false},    {"slen", 130, I16, false},   {"styp", 132, I16, false},
    {"stas", 134, I16, false},   {"stae", 136, I16, false},   {"tatyp", 138, I16, false},
    {"afilf", 140, I16, false},  {"afils", 142, I16, false},  {"nofilf", 144, I16, false},
    {"nofils", 146, I16, false}, {"lcf", 148, I16, false},    {"hcf", 150, I16, false},
    {"lcs", 152, I16, false},    {"hcs", 154, I16, false},    {"year", 156, I16, false},
    {"day", 158, I16, false},    {"hour", 160, I16, false},   {"minute", 162, I16, true},
    {"sec", 164, I16, false},    {"timbas", 166, I16, false}, {"trwf", 168, I16, true},
    {"grnors", 170, I16, false}, {"grnofr", 172, I16, false}, {"grnlof", 174, I16, false},
    {"gaps", 176, I16, false},   {"otrav", 178, I16, false},  {"cdpx", 180, I32, true},
    {"cdpy", 184, I32, true},    {"iline", 188, I32, true},   {"xline", 192, I32, true},
    {"sp", 196, I32, false},     {"unscale", 200, I32, false}, {"ntr", 204, I32, false},
    {"mark", 208, I16, false},   {"shortpad", 210, I16, false},
};
constexpr int kNamed = sizeof(kKeys) / sizeof(kKeys[0]);
static_assert(kNamed == SEGY_TOTAL_KEY_NUM + 5, "key table");
static_assert(offsetof(segy, xline) == 192 && offsetof(segy, sp) == 196 && offsetof(segy, trwf) == 168,
              "struct segy offsets");

// EBCDIC -> the reference's ASCII (fsgethd): printable ASCII 0x21..0x7e
// and three runs of 8-bit codes, each listed as the EBCDIC code it comes
// from; every other EBCDIC code (0x40, the EBCDIC space, included) becomes
// an ASCII space.
constexpr unsigned char kFromPrintable[94] = {
    0x4f, 0x7f, 0x7b, 0x5b, 0x6c, 0x50, 0x7d, 0x4d, 0x5d, 0x5c, 0x4e, 0x6b, 0x60, 0x4b, 0x61,  // !"#$%&'()*+,-./
    0xf0, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9,                                // 0-9
    0x7a, 0x5e, 0x4c, 0x7e, 0x6e, 0x6f, 0x7c,                                                  // :;<=>?@
    0xc1, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9,                                      // A-I
    0xd1, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9,                                      // J-R
    0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9,                                            // S-Z
    0x4a, 0xe0, 0x5a, 0x5f, 0x6d, 0x79,                                                        // [\]^_`
    0x81, 0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89,                                      // a-i
    0x91, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99,                                      // j-r
    0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9,                                            // s-z
    0xc0, 0x6a, 0xd0, 0xa1,                                                                    // {|}~
};
struct Run {
  unsigned char ascii0;
  unsigned char n;
  unsigned char from[32];
};
constexpr Run kRuns[] = {
    {0x80, 32, {0xb9, 0xba, 0xed, 0xbf, 0xbc, 0xbd, 0xec, 0xfa, 0xcb, 0xcc, 0xcd, 0xce, 0xcf, 0xda, 0xdb, 0xdc,
                0xde, 0xdf, 0xea, 0xeb, 0xbe, 0xca, 0xbb, 0xfe, 0xfb, 0xfd, 0x00, 0xef, 0xee, 0xfc, 0xb8, 0xdd}},
    {0xa0, 16, {0x77, 0x78, 0xaf, 0x8d, 0x8a, 0x8b, 0xae, 0xb2, 0x8f, 0x90, 0x9a, 0x9b, 0x9c, 0x9d, 0x9e, 0x9f}},
    {0xe0, 16, {0xaa, 0xab, 0xac, 0xad, 0x8c, 0x8e, 0x80, 0xb6, 0xb3, 0xb5, 0xb7, 0xb1, 0xb0, 0xb4, 0x76, 0xa0}},
};

struct Ebcdic {
  unsigned char to_ascii[256];
  Ebcdic() {
    memset(to_ascii, 0x20, sizeof to_ascii);
    for (int a = 0x21; a <= 0x7e; ++a) to_ascii[kFromPrintable[a - 0x21]] = (unsigned char)a;
    for (const Run &r : kRuns)
      for (int i = 0; i < r.n; ++i)
        if (r.from[i]) to_ascii[r.from[i]] = (unsigned char)(r.ascii0 + i);  // 0: ASCII 0x9a has no source
  }
};
const Ebcdic &ebcdic() {
  static const Ebcdic t;
  return t;
}

inline uint32_t be32(const unsigned char *p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
inline uint16_t be16(const unsigned char *p) { return (uint16_t)((p[0] << 8) | p[1]); }

inline void reverse(unsigned char *p, int n) {
  for (int i = 0; i < n / 2; ++i) {
    const unsigned char t = p[i];
    p[i] = p[n - 1 - i];
    p[n - 1 - i] = t;
  }
}

// The reader's header keys from raw (240 bytes) into h, the first 240 bytes
// of a struct segy, each at its keyindex position.
void header_keys(const unsigned char *raw, unsigned char *h, const keyindex *keyidx) {
  for (int k = 0; k < SEGY_TOTAL_KEY_NUM; ++k) {
    const Key &key = kKeys[k];
    if (!key.read) continue;
    const int start = keyidx[k].start - 1, bytes = keyidx[k].bytes;
    if ((bytes != 2 && bytes != 4) || start < 0 || start + bytes > 240) continue;
    const unsigned char *p = raw + start;
    if (key.kind == I32) {
      const int32_t v = bytes == 4 ? (int32_t)be32(p) : (int32_t)(int16_t)be16(p);
      memcpy(h + key.off, &v, 4);
    } else {
      const uint16_t v = bytes == 2 ? be16(p) : (uint16_t)be32(p);
      memcpy(h + key.off, &v, 2);
    }
  }
}

// ns big-endian samples of `format` into out; false for an unsupported
// format.  One message per call for IBM values beyond binary32.
bool samples_in(const unsigned char *raw, int ns, int format, float *out) {
  if (format == 1) {
    bool range = false;
    for (int i = 0; i < ns; ++i) {
      const uint32_t v = qpsdm_ibm::to_ieee_bits(be32(raw + 4 * (size_t)i), &range);
      memcpy(out + i, &v, 4);
    }
    if (range) fprintf(stderr, "IBM floating point exponent out of range \n");
    return true;
  }
  if (format == 5) {
    for (int i = 0; i < ns; ++i) {
      const uint32_t v = be32(raw + 4 * (size_t)i);
      memcpy(out + i, &v, 4);
    }
    return true;
  }
  fprintf(stderr, "Unsupported SEGY format=%d\n", format);
  return false;
}

inline int32_t field_i32(const unsigned char *h, size_t off) {
  int32_t v;
  memcpy(&v, h + off, 4);
  return v;
}
inline uint16_t field_u16(const unsigned char *h, size_t off) {
  uint16_t v;
  memcpy(&v, h + off, 2);
  return v;
}

bool check_inited(const char *who) {
  if (keyindex_inited) return true;
  fprintf(stderr, "Error: initkeyindex must be called before calling %s\n", who);
  return false;
}

void standard_keys(keyindex *k) {
  for (int i = 0; i < SEGY_TOTAL_KEY_NUM; ++i) {
    k[i].start = kKeys[i].off + 1;
    k[i].bytes = kKeys[i + 1].off - kKeys[i].off;
  }
}

// Index of the first key whose name starts with `key`, or -1.
int find_key(const char *key) {
  const size_t n = strlen(key);
  for (int i = 0; i < kNamed; ++i)
    if (strncmp(key, kKeys[i].name, n) == 0) return i;
  return -1;
}

}  // namespace

void fsgethd(FILE *fp, hd3600 *hd) {
  const size_t r = fread(hd, 1, 3600, fp);
  if (r != 3600)
    fprintf(stderr, "read segy 3600 header fail, we try to get 3600 byte, but only got %d", (int)r);
  const Ebcdic &t = ebcdic();
  for (int i = 0; i < 3200; ++i) hd->text[i] = t.to_ascii[hd->text[i]];
  // binary header: jobid, lino, reno (4 bytes), ntrpr .. vpol (2 bytes)
  unsigned char *b = reinterpret_cast<unsigned char *>(hd);
  for (int off = 3200; off < 3212; off += 4) reverse(b + off, 4);
  for (int off = 3212; off < 3260; off += 2) reverse(b + off, 2);
}

void initkeyindex(keyindex *keyidx) {
  standard_keys(keyidx);
  keyindex_inited = 1;
}

void setkeyindex(keyindex *keyidx, const char *key, int start, int bytes) {
  const int i = find_key(key);
  if (i >= 0 && i < SEGY_TOTAL_KEY_NUM) {
    keyidx[i].start = start;
    keyidx[i].bytes = bytes;
  }
  keyindex_inited = 1;
}

int fsgettr(FILE *fp, segy *tr, short int format, const keyindex *keyidx) {
  if (!check_inited("fsgettr")) return 0;
  unsigned char raw[240];
  if (fread(raw, 240, 1, fp) != 1) return 0;
  header_keys(raw, reinterpret_cast<unsigned char *>(tr), keyidx);
  const int ns = tr->ns;
  if ((int)fread(tr->data, sizeof(float), ns, fp) != ns) return 0;
  unsigned char *d = reinterpret_cast<unsigned char *>(tr->data);
  return samples_in(d, ns, format, tr->data) ? ns : 0;
}

int ssgettr(unsigned char *buf, segy *tr, short int format, const keyindex *keyidx) {
  if (!check_inited("fsgettr")) return 0;
  header_keys(buf, reinterpret_cast<unsigned char *>(tr), keyidx);
  const int ns = tr->ns;
  return samples_in(buf + 240, ns, format, tr->data) ? ns : 0;
}

extern "C" int qpsdm_segy_fgethd(FILE *fp, hd3600 *hd) {
  QPSDM_REQUIRE(fp && hd, "segy_fgethd: null args");
  fsgethd(fp, hd);
  return QPSDM_OK;
}

extern "C" int qpsdm_segy_initkeyindex(keyindex *keyidx) {
  QPSDM_REQUIRE(keyidx, "segy_initkeyindex: null keyindex");
  initkeyindex(keyidx);
  return QPSDM_OK;
}

extern "C" int qpsdm_segy_setkeyindex(keyindex *keyidx, const char *key, int start, int bytes) {
  QPSDM_REQUIRE(keyidx && key && key[0], "segy_setkeyindex: null args");
  const int i = find_key(key);
  QPSDM_REQUIRE(i >= 0 && i < SEGY_TOTAL_KEY_NUM, "segy_setkeyindex: unknown key '%s'", key);
  QPSDM_REQUIRE((bytes == 2 || bytes == 4) && start >= 1 && start - 1 + bytes <= 240,
                "segy_setkeyindex: key '%s' at byte %d width %d does not fit the 240-byte header", key, start,
                bytes);
  setkeyindex(keyidx, key, start, bytes);
  return QPSDM_OK;
}

extern "C" int qpsdm_segy_fgettr(FILE *fp, segy *tr, int format, const keyindex *keyidx) {
  if (!fp || !tr || !keyidx) return 0;
  return fsgettr(fp, tr, (short)format, keyidx);
}

extern "C" int qpsdm_segy_sgettr(const unsigned char *buf, segy *tr, int format, const keyindex *keyidx) {
  if (!buf || !tr || !keyidx) return 0;
  return ssgettr(const_cast<unsigned char *>(buf), tr, (short)format, keyidx);
}

extern "C" long qpsdm_segy_decode(const unsigned char *bytes, size_t nbytes, int ns, int format,
                                  const keyindex *keyidx, float *samples, qpsdm_trhdr *hdrs) {
  QPSDM_REQUIRE(bytes && ns > 0 && ns <= SU_NFLTS, "segy_decode: bad args");
  QPSDM_REQUIRE(format == 1 || format == 5, "segy_decode: format %d (1 = IBM float, 5 = IEEE float)", format);
  keyindex std_keys[SEGY_TOTAL_KEY_NUM];
  if (!keyidx) {
    standard_keys(std_keys);
    keyidx = std_keys;
  }
  const size_t rec = 240 + 4 * (size_t)ns;
  const long n = (long)(nbytes / rec);
  bool range = false;
  for (long i = 0; i < n; ++i) {
    const unsigned char *r = bytes + (size_t)i * rec;
    unsigned char h[240] = {};
    header_keys(r, h, keyidx);
    const int rns = field_u16(h, offsetof(segy, ns));
    if (rns != ns) return qpsdm::set_error(QPSDM_EINVAL, "record %ld has ns=%d, expected %d", i, rns, ns);
    if (samples) {
      float *o = samples + (size_t)i * ns;
      if (format == 1) {
        for (int j = 0; j < ns; ++j) {
          const uint32_t v = qpsdm_ibm::to_ieee_bits(be32(r + 240 + 4 * (size_t)j), &range);
          memcpy(o + j, &v, 4);
        }
      } else {
        for (int j = 0; j < ns; ++j) {
          const uint32_t v = be32(r + 240 + 4 * (size_t)j);
          memcpy(o + j, &v, 4);
        }
      }
    }
    if (hdrs) {
      qpsdm_trhdr &o = hdrs[i];
      o.sx = field_i32(h, offsetof(segy, sx));
      o.sy = field_i32(h, offsetof(segy, sy));
      o.gx = field_i32(h, offsetof(segy, gx));
      o.gy = field_i32(h, offsetof(segy, gy));
      o.fldr = field_i32(h, offsetof(segy, fldr));
      o.cdp = field_i32(h, offsetof(segy, cdp));
      o.offset = field_i32(h, offsetof(segy, offset));
      o.tracl = field_i32(h, offsetof(segy, tracl));
      o.scalco = (int16_t)field_u16(h, offsetof(segy, scalco));
      o.pad = 0;
      o.ns = (uint16_t)rns;
      o.dt = field_u16(h, offsetof(segy, dt));
    }
  }
  if (range) fprintf(stderr, "IBM floating point exponent out of range \n");
  return n;
}
