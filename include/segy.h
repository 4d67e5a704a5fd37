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

#ifdef __cplusplus
}
#endif
#endif
