// The span form's batched launches, shared by the translation units that
// instantiate them for each storage pair (spans.hip: FP32 pair bank with
// FP32 tables; spans_f16.hip: FP16 pair bank with U16Q tables).
#ifndef QPSDM_SPANS_IMPL_H_
#define QPSDM_SPANS_IMPL_H_

#include <hip/hip_fp16.h>
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

template <int NI, int PAT, bool SRC, typename BankT, typename Tab>
int launch_spans_t(const qk::ImgDev &P, const Tab &tab, const BankT *bank, const Go &go) {
  if (go.nwg_out) *go.nwg_out = -1;
  if (!go.run) return QPSDM_OK;
  const int ncol = ((P.nx + P.tile_x - 1) / P.tile_x) * ((P.ny + P.tile_y - 1) / P.tile_y);
  qk::ImgDev E = P;
  E.tile_z = 8;
  const int n1 = (P.nf_ps + 7) / 8, n2 = (P.nz - P.nf_pe + 7) / 8;
  for (int r = 0; r < go.nshot; ++r) {
    if (n1 > 0) {
      E.zb = 0;
      E.ze = P.nf_ps;
      hipLaunchKernelGGL((image_batch_tiles_kernel<8, BankT, Tab, SRC, true>), dim3(ncol * n1), dim3(64), 0, go.s,
                         E, tab, go.img, go.geom, bank, go.runs, r, ncol);
    }
    if (n2 > 0) {
      E.zb = P.nf_pe;
      E.ze = P.nz;
      hipLaunchKernelGGL((image_batch_tiles_kernel<8, BankT, Tab, SRC, true>), dim3(ncol * n2), dim3(64), 0, go.s,
                         E, tab, go.img, go.geom, bank, go.runs, r, ncol);
    }
    for (int k = 0; k < P.sp_n; ++k) {
      qk::ImgDev B = P;
      B.sp_b = P.sp_w[k];
      B.sp_e = P.sp_w[k + 1];
      const dim3 grid(ncol, (B.sp_e - B.sp_b + NI - 1) / NI);
      switch (P.sp_s[k]) {
#define QPSDM_SPAN(s_)                                                                                          \
  case s_:                                                                                                      \
    hipLaunchKernelGGL((image_batch_spans_kernel<NI, s_, PAT, BankT, Tab, SRC>), grid, dim3(64), 0, go.s, B, tab, \
                       go.img, go.geom, bank, go.runs, r);                                                      \
    break;
        QPSDM_SPAN(1)
        QPSDM_SPAN(2)
        QPSDM_SPAN(4)
        QPSDM_SPAN(8)
#undef QPSDM_SPAN
        default: return qpsdm::set_error(QPSDM_EINVAL, "image_batch: span of %d intervals", P.sp_s[k]);
      }
    }
    QPSDM_HIP(hipGetLastError());
  }
  return QPSDM_OK;
}

// Tiles of 4 or 8 intervals; the 3, 2 slot pattern where the node range has it.
template <typename BankT, typename Tab>
int launch_spans(const qk::ImgDev &P, const Tab &tab, const BankT *bank, const Go &go) {
  const bool src = go.runs.src_hdr != nullptr;
  switch (P.tile_z * 2 + (P.sp_pat ? 1 : 0)) {
#define QPSDM_NI(ni, pat)                                                                      \
  case ni * 2 + pat:                                                                           \
    return src ? launch_spans_t<ni, pat, true>(P, tab, bank, go) : launch_spans_t<ni, pat, false>(P, tab, bank, go);
    QPSDM_NI(4, 0)
    QPSDM_NI(4, 1)
    QPSDM_NI(8, 0)
    QPSDM_NI(8, 1)
#undef QPSDM_NI
    default: return qpsdm::set_error(QPSDM_EINVAL, "image_batch: span form with %d intervals per tile", P.tile_z);
  }
}

}  // namespace spans_detail

// Per storage pair (one translation unit each).
int launch_spans_f32p(const qpsdm_image_params *p, const qk::ImgDev &P, const void *tab, const void *bank,
                      const Go &go);
int launch_spans_f16p(const qpsdm_image_params *p, const qk::ImgDev &P, const void *tab, const void *bank,
                      const Go &go);

}  // namespace qpsdm_batch

#endif
