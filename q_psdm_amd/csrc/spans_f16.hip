// Span form of the batched imaging with the FP16 pair bank and U16Q tables
// (the mixed form); the dispatch and the FP32 instantiations are in
// spans.hip.
#include "spans_impl.h"

namespace qpsdm_batch {

int launch_spans_f16p(const qpsdm_image_params *p, const qk::ImgDev &P, const void *tab, const void *bank,
                      const Go &go) {
  return spans_detail::launch_spans(P, qk::make_tab_u16q(p->tg, tab), (const qk::PairF16 *)bank, go);
}

}  // namespace qpsdm_batch
