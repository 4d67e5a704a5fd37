// IBM System/360 single precision -> IEEE 754 binary32, as the reference's
// SEG-Y reader converts format-1 samples (src/segy.cc:1499-1530): an IBM
// word is sign | 7-bit excess-64 base-16 exponent | 24-bit fraction, value
// (-1)^s * 0.F * 16^(E-64).  The fraction is shifted left until its top bit
// (2^23) is set, which makes it the IEEE significand with its implicit one;
// the binary exponent is 4(E-64) minus the shift, biased by 126 (127 for the
// implicit one, -1 for the fraction's leading 0.).  Every IBM value in the
// IEEE normal range converts exactly (24 significant bits either way).  Two
// cases follow the reference rather than IEEE: a zero fraction gives +0
// whatever the sign and exponent, and a biased exponent outside [0, 255]
// (IBM values beyond the binary32 normal range) is replaced by 0, leaving
// sign | fraction bits; a biased exponent of exactly 255 is kept (an
// infinity or NaN bit pattern).  Used by the host codec and the device
// decoder, so both give the same bits.
#ifndef QPSDM_IBM_FLOAT_H_
#define QPSDM_IBM_FLOAT_H_

#include <stdint.h>

#if defined(__HIPCC__)
#define QPSDM_IBM_HD __host__ __device__
#else
#define QPSDM_IBM_HD
#endif

namespace qpsdm_ibm {

// `w` in host byte order.  *out_of_range is set when the exponent had to be
// replaced (the reference prints a message for each such sample).
QPSDM_IBM_HD inline uint32_t to_ieee_bits(uint32_t w, bool *out_of_range) {
  uint32_t frac = w & 0x00ffffffu;
  if (frac == 0) return 0u;
  // leading zero bits of the 24-bit fraction (>= 0, <= 23 here)
  const int shift = __builtin_clz(frac) - 8;
  int e = (int)(((w >> 24) & 0x7fu)) * 4 - 256 - shift + 126;
  frac = (frac << shift) & 0x007fffffu;
  if (e < 0 || e > 255) {
    *out_of_range = true;
    e = 0;
  }
  return (w & 0x80000000u) | ((uint32_t)e << 23) | frac;
}

}  // namespace qpsdm_ibm

#endif  // QPSDM_IBM_FLOAT_H_
