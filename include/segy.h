/*
 * segy.h -- SU trace record and big-endian SU codec of the imaging path.
 *
 * Keeps the reference's drop-in surface (src/segy.h:69-112, 123-610):
 * struct segy is the standard 240-byte SU header (same field names, types
 * and byte offsets, src/segy.cc:49-81) followed by float data[65535];
 * fgettr/fputtr/sgettr/sputtr read and write big-endian SU and byte-swap
 * exactly the header fields the reference enables (src/segy.h:22-65):
 * tracl tracr fldr tracf ep cdp cdpt trid nvs nhs offset gelev selev sdepth
 * gdel sdel scalel scalco sx sy gx gy counit wevel tstat delrt muts mute ns
 * dt minute trwf cdpx cdpy iline xline.  Other header bytes are copied raw
 * on read and written as zero on write, as the reference does.
 * fgettr returns ns, or 0 at end of file / short read.
 *
 * The C++-linkage names match the reference's; the qpsdm_su_* functions are
 * the same operations with C linkage for FFI callers, plus a batch decoder.
 */
#ifndef QPSDM_SEGY_H_
#define QPSDM_SEGY_H_

#include <stdint.h>
#include <stdio.h>

#define SU_NFLTS 65535

typedef struct segy {
  int tracl, tracr, fldr, tracf, ep, cdp, cdpt;
  short trid, nvs, nhs, duse;
  int offset, gelev, selev, sdepth, gdel, sdel, swdep, gwdep;
  short scalel, scalco;
  int sx, sy, gx, gy;
  short counit, wevel, swevel, sut, gut, sstat, gstat, tstat, laga, lagb, delrt, muts, mute;
  unsigned short ns, dt;
  short gain, igc, igi, corr, sfs, sfe, slen, styp, stas, stae, tatyp, afilf, afils, nofilf,
      nofils, lcf, hcf, lcs, hcs, year, day, hour, minute, sec, timbas, trwf, grnors, grnofr,
      grnlof, gaps, otrav;
  int cdpx, cdpy, iline, xline, sp, unscale, ntr;
  short mark, shortpad;
  short unass[14];
  float data[SU_NFLTS];
} segy;

/* ---- SEG-Y rev0 reader (src/segy.h:79-103, 1116-1203; src/segy.cc:83-2752) ----
 * hd3600: the 3600-byte reel header, 3200 text bytes (EBCDIC on disk,
 * converted to the reference's ASCII by fsgethd) then the 400-byte binary
 * header, whose 27 fields fsgethd byte-swaps; same field names, types and
 * offsets as the reference.
 * keyindex: per trace-header key (the 75 keys tracl .. xline of struct segy,
 * in struct order) the 1-based start byte and the width (2 or 4) of the key
 * in the file's 240-byte trace header; initkeyindex sets the standard SU/SEG-Y
 * positions, setkeyindex moves one key (setkeyindex(k, "sx", 181, 4)).
 * fsgettr/ssgettr decode the same header keys as the SU codec, each from its
 * keyindex position (a 2-byte key read into a 4-byte field is sign-extended,
 * unsigned for ns and dt; a 4-byte key read into a 2-byte field keeps its
 * low 16 bits; other widths leave the field as it was), then ns = tr->ns
 * samples of the given format: 1 = 4-byte IBM float, 5 = 4-byte IEEE float,
 * both big-endian.  They return ns, or 0 at end of file, on a short read or
 * an unsupported format.
 * Where the reference exits the process (a reader called before any
 * initkeyindex/setkeyindex) these print the same message to stderr and
 * return 0 instead.  setkeyindex matches the first key whose name starts
 * with `key`; for a name that matches none of the 75 keys it changes
 * nothing, as the reference's does.  A key whose start and width do not fit
 * the 240-byte header is not read (the field keeps its value).  An IBM
 * sample beyond the binary32 range is converted as the reference converts
 * it (exponent field 0), with one message per call rather than per sample. */
#define SEGY_TOTAL_KEY_NUM 75

typedef struct hd3600 {
  unsigned char text[3200];
  int jobid, lino, reno;
  short ntrpr, nart, hdt, dto, hns, nso, format, fold, tsort, vscode, hsfs, hsfe, hslen, hstyp, schn, hstas,
      hstae, htatyp, hcorr, bgrcv, rcvm, mfeet, polyt, vpol;
  unsigned char hunass[340];
} hd3600;

typedef struct keyindex {
  int start; /* 1-based byte of the key in the 240-byte trace header */
  int bytes; /* 2 or 4 */
} keyindex;

#ifdef __cplusplus
extern int keyindex_inited;
void fsgethd(FILE *fp, hd3600 *hd);
void initkeyindex(keyindex *keyidx);
void setkeyindex(keyindex *keyidx, const char *key, int start, int bytes);
int fsgettr(FILE *fp, segy *tr, short int format, const keyindex *keyidx);
int ssgettr(unsigned char *buf, segy *tr, short int format, const keyindex *keyidx);
#endif

#ifdef __cplusplus
int fgettr(FILE *fp, segy *tr);
void fputtr(FILE *fp, const segy *tr);
int sgettr(unsigned char *buf, segy *tr);
void sputtr(unsigned char *buf, const segy *tr);
#define swaptr(buf, tr) sgettr(buf, tr)
extern "C" {
#endif

int qpsdm_su_fgettr(FILE *fp, segy *tr);
int qpsdm_su_fputtr(FILE *fp, const segy *tr);
int qpsdm_su_sgettr(const unsigned char *buf, segy *tr);
int qpsdm_su_sputtr(unsigned char *buf, const segy *tr);

/* Header fields of one decoded trace that the imaging path reads. */
typedef struct {
  int32_t sx, sy, gx, gy;
  int32_t fldr, cdp, offset, tracl;
  int16_t scalco, pad;
  uint16_t ns, dt;
} qpsdm_trhdr;

/* Decode a run of equal-length big-endian SU records (the producer's
 * per-file loop, src/q-psdm.cu:1911-1916): returns the number of records
 * decoded (nbytes / (240 + 4 ns)), samples into samples[i*ns], selected
 * header fields into hdrs[i]; < 0 if a record's ns differs. */
long qpsdm_su_decode(const unsigned char *bytes, size_t nbytes, int ns, float *samples,
                     qpsdm_trhdr *hdrs);

/* SEG-Y rev0 with C linkage: as fsgethd, initkeyindex, setkeyindex,
 * fsgettr and ssgettr above, returning QPSDM_OK / an error code where those
 * return nothing, and an error (nothing changed) for an unknown key or a key
 * that does not fit the 240-byte header. */
int qpsdm_segy_fgethd(FILE *fp, hd3600 *hd);
int qpsdm_segy_initkeyindex(keyindex *keyidx);
int qpsdm_segy_setkeyindex(keyindex *keyidx, const char *key, int start, int bytes);
int qpsdm_segy_fgettr(FILE *fp, segy *tr, int format, const keyindex *keyidx);
int qpsdm_segy_sgettr(const unsigned char *buf, segy *tr, int format, const keyindex *keyidx);

/* Decode a run of equal-length SEG-Y trace records (240-byte header +
 * 4 ns bytes, format 1 or 5) on the host, keys at keyidx (NULL: standard
 * positions): returns the number of records (nbytes / (240 + 4 ns)),
 * samples into samples[i*ns], the header subset into hdrs[i]; < 0 if a
 * record's ns differs.  Same bits as ssgettr per record. */
long qpsdm_segy_decode(const unsigned char *bytes, size_t nbytes, int ns, int format, const keyindex *keyidx,
                       float *samples, qpsdm_trhdr *hdrs);

/* The same decode on the device: ntr records of 240 + 4 ns bytes at
 * raw_dev (4-byte aligned) -> samples_dev [ntr][ns] FP32 and, if hdrs_dev is
 * not NULL, hdrs_dev [ntr]; keys at keyidx (host memory; NULL: standard
 * positions; a key that does not fit the header reads as 0).  Asynchronous
 * on `stream`; bit-identical to qpsdm_segy_decode.  ns is taken as given
 * (the records' own ns key is returned in hdrs_dev for the caller to
 * check). */
int qpsdm_segy_decode_dev(const void *raw_dev, int ntr, int ns, int format, const keyindex *keyidx,
                          float *samples_dev, qpsdm_trhdr *hdrs_dev, void *stream);

#ifdef __cplusplus
}
#endif
#endif
