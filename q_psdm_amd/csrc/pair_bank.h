// Element types of the pair banks (QPSDM_F32P, QPSDM_F16P): sample i of a
// level row stored with its forward difference, (b[i], b[i + 1] - b[i]),
// the row's entries 0 and nt - 1 zero.  Written by the Q-bank kernel, read
// by the imaging kernels.
#ifndef QPSDM_PAIR_BANK_H_
#define QPSDM_PAIR_BANK_H_

namespace qpsdm_dev {

struct PairF32 {
  float x, d;
};
// half2 bits: low half b[i], high half the difference
struct PairF16 {
  unsigned v;
};

}  // namespace qpsdm_dev

#endif
