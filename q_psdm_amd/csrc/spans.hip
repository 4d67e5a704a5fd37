// Span form of the batched imaging (weight_nodes = 2): per shot, in shot
// order, the exact body above the node range and beyond the tables' depth
// range, then one launch per band of the node range (one work-group per
// tile of tile_z intervals, the band's span length a compile-time
// parameter).  Each point is imaged by one launch per shot and every launch
// of a shot precedes the next shot's, so per point the sum still runs over
// the shots in order and each shot's traces in order.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "batch_runs.h"
#include "common.h"

namespace qpsdm_batch {

namespace {

template <int NI, int S, typename BankT, typename Tab, bool SRC>
__global__ __launch_bounds__(64) void image_batch_spans_kernel(qk::ImgDev P, Tab tab, float *__restrict__ img,
                                                               const float4 *__restrict__ geom,
                                                               const BankT *__restrict__ bank, ShotRuns runs,
                                                               int shot) {
  const OneShot one{runs(shot)};
  qk::image_tile_spans<NI, S, 3, SRC>(P, tab, img, geom, bank, blockIdx.x, blockIdx.y, one, 1);
}

template <int NI, bool SRC, typename BankT, typename Tab>
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
#define QPSDM_SPAN(s_)                                                                                         \
  case s_:                                                                                                     \
    hipLaunchKernelGGL((image_batch_spans_kernel<NI, s_, BankT, Tab, SRC>), grid, dim3(64), 0, go.s, B, tab, \
                       go.img, go.geom, bank, go.runs, r);                                                     \
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

template <typename BankT, typename Tab>
int launch_spans(const qk::ImgDev &P, const Tab &tab, const BankT *bank, const Go &go) {
  const bool src = go.runs.src_hdr != nullptr;
  switch (P.tile_z) {
#define QPSDM_NI(ni) \
  case ni: return src ? launch_spans_t<ni, true>(P, tab, bank, go) : launch_spans_t<ni, false>(P, tab, bank, go);
    QPSDM_NI(2)
    QPSDM_NI(4)
    QPSDM_NI(8)
#undef QPSDM_NI
    default: return qpsdm::set_error(QPSDM_EINVAL, "image_batch: span form with %d intervals per tile", P.tile_z);
  }
}

}  // namespace

// The span form's storage pairs: the FP32 pair bank with FP32 tables (the
// parity form) and the FP16 pair bank with U16Q tables (the mixed form).
int launch_batch_spans(const qpsdm_image_params *p, const qk::ImgDev &P, const void *tab, int table_dtype,
                       const void *bank, int bank_dtype, const Go &go) {
  if (bank_dtype == QPSDM_F32P && table_dtype == QPSDM_F32)
    return launch_spans(P, qk::make_tab_f32(p->tg, tab), (const qk::PairF32 *)bank, go);
  if (bank_dtype == QPSDM_F16P && table_dtype == QPSDM_U16Q) {
    QPSDM_REQUIRE(qk::u16q_fits(p->tg), "image_batch: U16Q node of %zu bytes (below 2 GiB)",
                  qk::u16q_node_bytes(p->tg));
    return launch_spans(P, qk::make_tab_u16q(p->tg, tab), (const qk::PairF16 *)bank, go);
  }
  return qpsdm::set_error(QPSDM_EINVAL,
                          "image_batch: the span form takes a QPSDM_F32P bank with QPSDM_F32 tables or a QPSDM_F16P "
                          "bank with QPSDM_U16Q tables (got bank %d, tables %d)",
                          bank_dtype, table_dtype);
}

}  // namespace qpsdm_batch
