// Shared pieces of the batched imaging launches (batch.hip, spans.hip):
// the per-shot run list, the launch arguments of one batch, and the
// one-work-group-per-tile kernel that also images the node and span forms'
// exact parts.
#ifndef QPSDM_BATCH_RUNS_H_
#define QPSDM_BATCH_RUNS_H_

#include <mutex>

#include "image_spans.h"

namespace qpsdm_batch {

namespace qk = qpsdm_dev;

// Per-shot runs; with src_hdr set, each shot's source leg comes from its own
// source table (header {sx, sy, su0, sv0} per shot, then the tables).
struct ShotRuns {
  const qpsdm_shot *shots;
  float tz;
  const float4 *src_hdr;
  const float2 *src_tab;
  long long src_stride;
  int nus, nvs;
  // The descriptor and header words are read through the constant address
  // space (they do not change during a launch), so they are scalar loads
  // and the run's fields stay in scalar registers.
  __device__ qk::Run operator()(int r) const {
    typedef const __attribute__((address_space(4))) int *cip;
    typedef const __attribute__((address_space(4))) float *cfp;
    const cip si = (cip)(shots + r);
    const cfp sf = (cfp)(shots + r);
    qk::Run run;
    run.tr0 = si[0];
    run.ntr = si[1];
    run.brow0 = run.tr0;
    run.dip = qk::uniform_dip(qk::make_dip(sf[2], sf[3], sf[4], sf[5], tz));
    run.src = nullptr;
    run.ssx = run.ssy = run.su0 = run.sv0 = 0.f;
    run.snu = nus;
    run.snv = nvs;
    if (src_hdr) {
      const cfp h = (cfp)(src_hdr + r);
      run.src = src_tab + (size_t)r * src_stride;
      run.ssx = h[0];
      run.ssy = h[1];
      run.su0 = h[2];
      run.sv0 = h[3];
    }
    return run;
  }
};

// One shot as a run list of length one.
struct OneShot {
  qk::Run run;
  __device__ qk::Run operator()(int) const { return run; }
};

inline ShotRuns shot_runs(const qpsdm_image_params *p, const qpsdm_shot *shots, int nshot, const void *src) {
  ShotRuns R{shots, p->taperzone, nullptr, nullptr, 0, 0, 0};
  if (src) {
    qk::src_subgrid(p, &R.nus, &R.nvs);
    R.src_stride = (long long)R.nvs * R.nus * p->tg.nw;
    R.src_hdr = (const float4 *)src;
    R.src_tab = (const float2 *)((const char *)src + (size_t)nshot * 16);
  }
  return R;
}

// One work-group per tile (nwg = -1): one launch per shot of the batch, in
// shot order, each the tile body for that shot alone (no loop over shots in
// the kernel: its registers are the per-shot kernel's, with no spills), so
// per image point the sum still runs over the shots in order and each
// shot's traces in order.  Scheduled by the hardware dispatcher.
// SRC: the shots carry source tables; RANGE: depth groups of [zb, ze).
template <int ZP, typename BankT, typename Tab, bool SRC, bool RANGE>
__global__ __launch_bounds__(64) void image_batch_tiles_kernel(qk::ImgDev P, Tab tab, float *__restrict__ img,
                                                               const float4 *__restrict__ geom,
                                                               const BankT *__restrict__ bank, ShotRuns runs,
                                                               int shot, int ncol) {
  const int tile = (int)blockIdx.x;
  const OneShot one{runs(shot)};
  qk::image_tile<ZP, SRC, RANGE, true>(P, tab, img, geom, bank, tile % ncol, tile / ncol, one, 1);
}

// Launch arguments of one batch; with run == false only the resident grid
// size is resolved (a query made before graph capture).
struct Go {
  float *img;
  const float4 *geom;
  ShotRuns runs;
  int nshot;
  unsigned *work;
  int nwg;
  hipStream_t s;
  bool run;
  int *nwg_out;
};

// The streams a span-form call spreads its launches over: the exact body
// and each band image disjoint points and are independent, so they run
// concurrently (each one's shots stay in order on its own stream).  One set
// per device, created on first use and kept; a call forks the set from its
// stream with an event and joins it back before returning, under the set's
// mutex, so the call stays ordered on its stream and a captured graph holds
// one fork and one join.
struct Fanout {
  static constexpr int kStreams = 6;  // the exact body + up to five bands
  hipStream_t s[kStreams];
  hipEvent_t fork, join[kStreams];
  std::mutex m;
};
// The current device's set (nullptr with the error set when it cannot be
// created); call it once before a graph capture that holds span launches.
Fanout *fanout();

// Span form of every shot (defined in spans.hip).
int launch_batch_spans(const qpsdm_image_params *p, const qk::ImgDev &P, const void *tab, int table_dtype,
                       const void *bank, int bank_dtype, const Go &go);

}  // namespace qpsdm_batch

#endif
