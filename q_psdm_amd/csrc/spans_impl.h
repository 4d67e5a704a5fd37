// The span form's batched launches, shared by the translation units that
// instantiate them for each storage pair (spans.hip: FP32 pair bank with
// FP32 tables; spans_f16.hip: FP16 pair bank with U16Q tables).
#ifndef QPSDM_SPANS_IMPL_H_
#define QPSDM_SPANS_IMPL_H_

#include <hip/hip_fp16.h>
#include <algorithm>
#include <hip/hip_runtime.h>

#include "batch_runs.h"
#include "common.h"

namespace qpsdm_batch {


namespace spans_detail {

template <int NI, int S, int PAT, typename BankT, typename Tab, bool SRC>
__global__ __launch_bounds__(64) void image_batch_spans_kernel(qk::ImgDev P, Tab tab, float *__restrict__ img,
                                                               const float4 *__restrict__ geom,
                                                               const BankT *__restrict__ bank, ShotRuns runs,
                                                               int shot) {
  const OneShot one{runs(shot)};
  qk::image_tile_spans<NI, S, 3, PAT, SRC>(P, tab, img, geom, bank, blockIdx.x, blockIdx.y, one, 1);
}

// One band's tile groups [g0, g1) (ImgDev sp_b, sp_e set), span S.
template <int NI, int PAT, bool SRC, typename BankT, typename Tab>
int launch_band(qk::ImgDev B, int S, int g0, int g1, int ncol, const Tab &tab, const BankT *bank, const Go &go,
                int r) {
  if (g1 <= g0) return QPSDM_OK;
  B.sp_g0 = g0;
  const dim3 grid(ncol, g1 - g0);
  // (spans of 8 intervals run in tiles of 8)
  switch (NI == 8 ? S : S == 8 ? 0 : S) {
#define QPSDM_SPAN(s_)                                                                                          \
  case s_:                                                                                                      \
    hipLaunchKernelGGL((image_batch_spans_kernel<NI, s_, PAT, BankT, Tab, SRC>), grid, dim3(64), 0, go.s, B, tab, \
                       go.img, go.geom, bank, go.runs, r);                                                      \
    return QPSDM_OK;
    QPSDM_SPAN(1)
    QPSDM_SPAN(2)
    QPSDM_SPAN(4)
#undef QPSDM_SPAN
    case 8:
      hipLaunchKernelGGL((image_batch_spans_kernel<8, 8, PAT, BankT, Tab, SRC>), grid, dim3(64), 0, go.s, B, tab,
                         go.img, go.geom, bank, go.runs, r);
      return QPSDM_OK;
    default: return qpsdm::set_error(QPSDM_EINVAL, "image_batch: span of %d intervals in tiles of %d", S, NI);
  }
}

// One band in tiles of NI intervals: tile groups in the 3, 2 slot pattern
// where the range has it (the range's last group with a third slot in its
// last interval when that interval holds one point more, or without the
// pattern when that interval is not the group's last).
template <int NI, bool SRC, typename BankT, typename Tab>
int launch_band_ni(const qk::ImgDev &P, const qk::ImgDev &B, int S, bool last, int ncol, const Tab &tab,
                   const BankT *bank, const Go &go, int r) {
  const int ng = (B.sp_e - B.sp_b + NI - 1) / NI;
  if (!P.sp_pat) return launch_band<NI, 0, SRC>(B, S, 0, ng, ncol, tab, bank, go, r);
  if (!last || P.sp_pat_last == 1) return launch_band<NI, 1, SRC>(B, S, 0, ng, ncol, tab, bank, go, r);
  int rc = launch_band<NI, 1, SRC>(B, S, 0, ng - 1, ncol, tab, bank, go, r);
  if (rc) return rc;
  return P.sp_pat_last == 2 ? launch_band<NI, 2, SRC>(B, S, ng - 1, ng, ncol, tab, bank, go, r)
                            : launch_band<NI, 0, SRC>(B, S, ng - 1, ng, ncol, tab, bank, go, r);
}

// The bands of one shot.  A band's tiles hold tile_z intervals, or one span
// where its spans are longer (8 intervals: two weight nodes per tile).
template <bool SRC, typename BankT, typename Tab>
int launch_spans_t(const qk::ImgDev &P, const Tab &tab, const BankT *bank, const Go &go) {
  if (go.nwg_out) *go.nwg_out = -1;
  if (!go.run) return QPSDM_OK;
  Fanout *F = fanout();
  if (!F) return QPSDM_EHIP;
  std::lock_guard<std::mutex> lock(F->m);
  QPSDM_HIP(hipEventRecord(F->fork, go.s));
  for (int i = 0; i < Fanout::kStreams; ++i) QPSDM_HIP(hipStreamWaitEvent(F->s[i], F->fork, 0));
  const int ncol = ((P.nx + P.tile_x - 1) / P.tile_x) * ((P.ny + P.tile_y - 1) / P.tile_y);
  qk::ImgDev E = P;
  E.tile_z = 8;
  const int n1 = (P.nf_ps + 7) / 8, n2 = (P.nz - P.nf_pe + 7) / 8;
  int rc = QPSDM_OK;
  for (int r = 0; r < go.nshot && !rc; ++r) {
    // the exact body above the node range and past the tables' depth range
    if (n1 > 0) {
      E.zb = 0;
      E.ze = P.nf_ps;
      hipLaunchKernelGGL((image_batch_tiles_kernel<8, BankT, Tab, SRC, true>), dim3(ncol * n1), dim3(64), 0, F->s[0],
                         E, tab, go.img, go.geom, bank, go.runs, r, ncol);
    }
    if (n2 > 0) {
      E.zb = P.nf_pe;
      E.ze = P.nz;
      hipLaunchKernelGGL((image_batch_tiles_kernel<8, BankT, Tab, SRC, true>), dim3(ncol * n2), dim3(64), 0, F->s[0],
                         E, tab, go.img, go.geom, bank, go.runs, r, ncol);
    }
    // each band on its own stream
    for (int k = 0; k < P.sp_n && !rc; ++k) {
      qk::ImgDev B = P;
      B.sp_b = P.sp_w[k];
      B.sp_e = P.sp_w[k + 1];
      const int S = P.sp_s[k];
      const bool last = k == P.sp_n - 1;
      Go gk = go;
      gk.s = F->s[1 + k];
      rc = std::max(P.tile_z, S) == 8 ? launch_band_ni<8, SRC>(P, B, S, last, ncol, tab, bank, gk, r)
                                      : launch_band_ni<4, SRC>(P, B, S, last, ncol, tab, bank, gk, r);
    }
    if (!rc && hipGetLastError() != hipSuccess) rc = qpsdm::set_error(QPSDM_EHIP, "image_batch: span launch failed");
  }
  // the join, also after a failed launch (a capture needs every fork joined)
  for (int i = 0; i < Fanout::kStreams; ++i) {
    const hipError_t e1 = hipEventRecord(F->join[i], F->s[i]);
    const hipError_t e2 = hipStreamWaitEvent(go.s, F->join[i], 0);
    if (!rc && (e1 != hipSuccess || e2 != hipSuccess))
      rc = qpsdm::set_error(QPSDM_EHIP, "image_batch: joining the span streams failed");
  }
  return rc;
}

// Tiles of 4 or 8 intervals.
template <typename BankT, typename Tab>
int launch_spans(const qk::ImgDev &P, const Tab &tab, const BankT *bank, const Go &go) {
  if (P.tile_z != 4 && P.tile_z != 8)
    return qpsdm::set_error(QPSDM_EINVAL, "image_batch: span form with %d intervals per tile", P.tile_z);
  return go.runs.src_hdr ? launch_spans_t<true>(P, tab, bank, go) : launch_spans_t<false>(P, tab, bank, go);
}

}  // namespace spans_detail

// Per storage pair (one translation unit each).
int launch_spans_f32p(const qpsdm_image_params *p, const qk::ImgDev &P, const void *tab, const void *bank,
                      const Go &go);
int launch_spans_f16p(const qpsdm_image_params *p, const qk::ImgDev &P, const void *tab, const void *bank,
                      const Go &go);

}  // namespace qpsdm_batch

#endif
