// Span form of the batched imaging (weight_nodes = 2): per shot, in shot
// order, the exact body above the node range and beyond the tables' depth
// range, then one launch per band of the node range (one work-group per
// tile of tile_z intervals, the band's span length a compile-time
// parameter).  The exact body and each band run on streams of their own
// (they image disjoint points); each point is imaged by one launch per shot
// on one stream, in shot order, so per point the sum still runs over the
// shots in order and each shot's traces in order.  This unit holds the
// dispatch and the FP32 pair bank with FP32 tables (the parity form);
// spans_f16.hip the FP16 pair bank with U16Q tables.
#include "spans_impl.h"

namespace qpsdm_batch {

Fanout *fanout() {
  static std::mutex m;
  static Fanout *sets[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    qpsdm::set_error(QPSDM_EHIP, "image_batch: no current device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(m);
  if (sets[dev]) return sets[dev];
  Fanout *f = new Fanout();
  bool ok = hipEventCreateWithFlags(&f->fork, hipEventDisableTiming) == hipSuccess;
  for (int i = 0; i < Fanout::kStreams && ok; ++i)
    ok = hipStreamCreateWithFlags(&f->s[i], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&f->join[i], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    qpsdm::set_error(QPSDM_EHIP, "image_batch: creating the span-form streams failed");
    return nullptr;  // (the partial set is kept out of use; its handles live until the process ends)
  }
  sets[dev] = f;
  return f;
}

int launch_spans_f32p(const qpsdm_image_params *p, const qk::ImgDev &P, const void *tab, const void *bank,
                      const Go &go) {
  return spans_detail::launch_spans(P, qk::make_tab_f32(p->tg, tab), (const qk::PairF32 *)bank, go);
}

int launch_batch_spans(const qpsdm_image_params *p, const qk::ImgDev &P, const void *tab, int table_dtype,
                       const void *bank, int bank_dtype, const Go &go) {
  if (bank_dtype == QPSDM_F32P && table_dtype == QPSDM_F32) return launch_spans_f32p(p, P, tab, bank, go);
  if (bank_dtype == QPSDM_F16P && table_dtype == QPSDM_U16Q) {
    QPSDM_REQUIRE(qk::u16q_fits(p->tg), "image_batch: U16Q node of %zu bytes (below 2 GiB)",
                  qk::u16q_node_bytes(p->tg));
    return launch_spans_f16p(p, P, tab, bank, go);
  }
  return qpsdm::set_error(QPSDM_EINVAL,
                          "image_batch: the span form takes a QPSDM_F32P bank with QPSDM_F32 tables or a QPSDM_F16P "
                          "bank with QPSDM_U16Q tables (got bank %d, tables %d)",
                          bank_dtype, table_dtype);
}

}  // namespace qpsdm_batch
