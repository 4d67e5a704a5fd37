// C-linkage test shim over the reference's own host sources (segy.cc,
// psdmpkg.cc, fft.cc), compiled from /root/reference/src by oracle/Makefile
// into oracle/_ref/libpsdmref.so.  TEST INFRASTRUCTURE ONLY: it lets the CPU
// tests call the reference's SU codec, parameter utilities, taper and PFA FFT
// through ctypes to pin this build's restatements.  Nothing here is shipped.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "psdmpkg.h"
#include "segy.h"
#include "sufft.h"

// Defined in the reference's src/fft.cc (declared there, not in a header).
void pfarc(int isign, int n, float *rz, float *cz);
void pfacr(int isign, int n, float *cz, float *rz);

extern "C" {

int ref_npfar(int n) { return get_opt_n_fftw(n); }

void ref_pfarc(int isign, int n, float *rz, float *cz) { pfarc(isign, n, rz, cz); }

void ref_pfacr(int isign, int n, float *cz, float *rz) { pfacr(isign, n, cz, rz); }

int ref_fetch_and_taper(const float *d, int ns, int it1, int it2, float *o) {
  return fetch_and_taper(d, ns, it1, it2, o);
}

// sgettr on a big-endian SU record; returns ns, copies the host-order header
// (240 B) and samples.
int ref_sgettr(unsigned char *buf, unsigned char *hdr, float *data) {
  segy *tr = (segy *)calloc(1, sizeof(segy));
  int ns = sgettr(buf, tr);
  memcpy(hdr, tr, 240);
  memcpy(data, tr->data, sizeof(float) * ns);
  free(tr);
  return ns;
}

// fgettr over a whole file: returns the number of traces read (<= maxtr),
// headers into hdrs[240*i], samples into data[i*maxns].
int ref_fgettr_all(const char *path, int maxtr, int maxns, unsigned char *hdrs, float *data) {
  FILE *fp = fopen(path, "rb");
  if (!fp) return -1;
  segy *tr = (segy *)calloc(1, sizeof(segy));
  int n = 0;
  while (n < maxtr) {
    int ns = fgettr(fp, tr);
    if (ns == 0) break;
    memcpy(hdrs + 240 * (size_t)n, tr, 240);
    memcpy(data + (size_t)n * maxns, tr->data, sizeof(float) * (ns < maxns ? ns : maxns));
    ++n;
  }
  free(tr);
  fclose(fp);
  return n;
}

// fputtr of one trace given a host-order header and samples.
int ref_fputtr(const char *path, int append, const unsigned char *hdr, const float *data) {
  FILE *fp = fopen(path, append ? "ab" : "wb");
  if (!fp) return -1;
  segy *tr = (segy *)calloc(1, sizeof(segy));
  memcpy(tr, hdr, 240);
  memcpy(tr->data, data, sizeof(float) * tr->ns);
  fputtr(fp, tr);
  free(tr);
  fclose(fp);
  return 0;
}

int *ref_parseint(const char *s, int *n) { return parseint(s, n); }

float *ref_parsetuple(const char *s, int *nt, int *ne) { return parsetuple(s, nt, ne); }

void ref_free(void *p) { free(p); }

int ref_readparfile_i(const char *f, const char *fmt, int *v) { return readparfile(f, fmt, v); }
int ref_readparfile_f(const char *f, const char *fmt, float *v) { return readparfile(f, fmt, v); }
int ref_readparfile_s(const char *f, const char *fmt, char *v) { return readparfile(f, fmt, v); }
int ref_readparfile_ii(const char *f, const char *fmt, int *a, int *b) {
  return readparfile(f, fmt, a, b);
}

float ref_getscale(short s) { return getscale(s); }

float ref_time_offset_mute(float *tm, int n, float off) { return time_offset_mute(tm, n, off); }

void ref_time_offset_mute_santity(float *tm, int n) { time_offset_mute_santity(tm, n); }

void ref_parsepath(char *path, char *proj, char *name, int *iproc, char *task) {
  parsepath(path, proj, name, iproc, task);
}

char *ref_strrtrim(char *s, int c) { return strrtrim(s, c); }

void ref_initargs(int argc, char **argv) { initargs_inner(argc, argv, NULL); }

int ref_getparint(const char *k, int *v) { return getparint(k, v); }
int ref_getparfloat(const char *k, float *v) { return getparfloat(k, v); }
int ref_getpardouble(const char *k, double *v) { return getpardouble(k, v); }
int ref_getparstr(const char *k, char **v) { return getparstr(k, v); }

void ref_get_vel_coor(const char *path, int line, coor_t *c, short *sc, int fci, int nxi,
                      int fcv, int nxv, int flv, int nyv) {
  get_vel_coor(path, line, c, sc, fci, nxi, fcv, nxv, flv, nyv);
}
}

// ---- SEG-Y rev0 (src/segy.cc:83-2752, 2753-4477) ----
// keyidx is int[2*75] (start, bytes) pairs, the layout of keyindex[75].
extern "C" {

void ref_initkeyindex(int *keyidx) { initkeyindex((keyindex *)keyidx); }

void ref_setkeyindex(int *keyidx, const char *key, int start, int bytes) {
  setkeyindex((keyindex *)keyidx, key, start, bytes);
}

// fsgethd of a file's first 3600 bytes into out (3600 B, host order).
int ref_fsgethd(const char *path, unsigned char *out) {
  FILE *fp = fopen(path, "rb");
  if (!fp) return -1;
  hd3600 *hd = (hd3600 *)calloc(1, sizeof(hd3600));
  fsgethd(fp, hd);
  memcpy(out, hd, 3600);
  free(hd);
  fclose(fp);
  return 0;
}

// fsgettr over a file after skip bytes, reusing one trace record as the
// reference's callers do: returns the number of traces read (<= maxtr),
// header bytes into hdrs[240*i], samples into data[i*maxns].
int ref_fsgettr_all(const char *path, long skip, int format, const int *keyidx, int maxtr, int maxns,
                    unsigned char *hdrs, float *data) {
  FILE *fp = fopen(path, "rb");
  if (!fp) return -1;
  fseek(fp, skip, SEEK_SET);
  segy *tr = (segy *)calloc(1, sizeof(segy));
  int n = 0;
  while (n < maxtr) {
    int ns = fsgettr(fp, tr, (short)format, (const keyindex *)keyidx);
    if (ns == 0) break;
    memcpy(hdrs + 240 * (size_t)n, tr, 240);
    memcpy(data + (size_t)n * maxns, tr->data, sizeof(float) * (ns < maxns ? ns : maxns));
    ++n;
  }
  free(tr);
  fclose(fp);
  return n;
}

// Writes a SEG-Y file with the reference's writer: inithd3600 + fsputhd,
// then one fsputtr per trace from host-order headers hdrs[240*i] and
// samples data[i*ns].
int ref_fsput_file(const char *path, int ns, int dt_um, int format, const int *keyidx, int ntr,
                   const unsigned char *hdrs, const float *data) {
  FILE *fp = fopen(path, "wb");
  if (!fp) return -1;
  hd3600 *hd = (hd3600 *)calloc(1, sizeof(hd3600));
  inithd3600(hd, ns, dt_um, (short)format);
  fsputhd(fp, hd);
  free(hd);
  segy *tr = (segy *)calloc(1, sizeof(segy));
  for (int i = 0; i < ntr; ++i) {
    memcpy(tr, hdrs + 240 * (size_t)i, 240);
    memcpy(tr->data, data + (size_t)i * ns, sizeof(float) * ns);
    fsputtr(fp, tr, (short)format, (const keyindex *)keyidx);
  }
  free(tr);
  fclose(fp);
  return 0;
}

}  // extern "C"
