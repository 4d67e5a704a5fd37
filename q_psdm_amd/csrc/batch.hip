// Batched imaging: shot gathers of uneven trace counts and per-shot
// apertures through one persistent kernel, and captured hipGraphs of
// [Q bank of the batch -> source tables] and [counter reset -> persistent
// imaging].
//
// The persistent grid is sized to the resident capacity of the device; its
// work-groups take image tiles from one atomic counter in depth-slab order
// (neighbouring columns of a slab together, so tiles in flight share table
// and bank lines in L2) until the tiles run out, one dispatch per shot.  A
// tile is owned by one work-group per shot and the shots' dispatches run in
// order on one stream, so the image needs no atomics and the summation
// order (shots, then traces, in order) is fixed.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <mutex>

#include <type_traits>

#include "common.h"
#include "batch_runs.h"

namespace qk = qpsdm_dev;

// Prepares the Q-bank coefficients of a parameter set (defined in qbank.hip).
int qpsdm_qbank_prepare(const qpsdm_bank_params *p);

namespace {

using namespace qpsdm_batch;

__global__ void reset_kernel(unsigned *c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0u;
}

// Persistent grid: one dispatch per shot of the batch, in shot order (the
// tile body for that shot alone, so its registers are the per-shot
// kernel's); the work-groups take the shot's tiles from the counter, which
// is cleared by a kernel before each dispatch.  Per image point the sum runs
// over the shots in order and each shot's traces in order, exactly as
// consecutive per-shot launches.  Held to four waves per SIMD (128 VGPRs):
// uncapped, the tile loop around the body takes 134 and three waves fit.
template <int ZP, typename BankT, typename Tab, bool SRC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
void image_batch_kernel(qk::ImgDev P, Tab tab, float *__restrict__ img, const float4 *__restrict__ geom,
                        const BankT *__restrict__ bank, ShotRuns runs, int shot, unsigned *counter, int ncol,
                        int total) {
  const OneShot one{runs(shot)};
  for (;;) {
    unsigned k = 0;
    if (threadIdx.x == 0) k = atomicAdd(counter, 1u);
    // broadcast as a scalar: the tile and everything derived from it stay
    // wave-uniform (scalar depth walk, scalar buffer offsets)
    const int tile = (int)__builtin_amdgcn_readfirstlane(k);
    if (tile >= total) break;
    // depth group slowest: consecutive tiles are neighbouring columns of one
    // depth slab, which share table and bank lines in L2
    qk::image_tile<ZP, SRC, false, true>(P, tab, img, geom, bank, tile % ncol, tile / ncol, one, 1);
  }
}

// Node form of one shot (weight_nodes = 1): one work-group per node tile
// (blockIdx.x = column tile, blockIdx.y = node group).
template <int NI, int PM, typename BankT, typename Tab, bool SRC>
__global__ __launch_bounds__(64) void image_batch_nodes_kernel(qk::ImgDev P, Tab tab, float *__restrict__ img,
                                                               const float4 *__restrict__ geom,
                                                               const BankT *__restrict__ bank, ShotRuns runs,
                                                               int shot) {
  const OneShot one{runs(shot)};
  qk::image_tile_nodes<NI, PM, SRC>(P, tab, img, geom, bank, blockIdx.x, blockIdx.y, one, 1);
}

template <int ZP>
__global__ __launch_bounds__(64) void count_batch_kernel(qk::ImgDev P, const float4 *__restrict__ geom,
                                                         ShotRuns runs, int nshot,
                                                         unsigned long long *count) {
  const unsigned n = qk::count_tile<ZP>(P, geom, blockIdx.x, blockIdx.y, runs, nshot);
  unsigned long long v = n;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (threadIdx.x == 0 && v) atomicAdd(count, v);
}

template <int ZP, bool SRC, typename BankT, typename Tab>
int launch_batch_t(const qk::ImgDev &P, const Tab &tab, const BankT *bank, const Go &go) {
  auto kern = image_batch_kernel<ZP, BankT, Tab, SRC>;
  const int ncol = ((P.nx + P.tile_x - 1) / P.tile_x) * ((P.ny + P.tile_y - 1) / P.tile_y);
  const int nzg = (P.ze - P.zb + ZP - 1) / ZP;
  if (go.nwg == -1) {
    if (go.nwg_out) *go.nwg_out = -1;
    if (!go.run) return QPSDM_OK;
    for (int r = 0; r < go.nshot; ++r) {
      hipLaunchKernelGGL((image_batch_tiles_kernel<ZP, BankT, Tab, SRC, false>), dim3(ncol * nzg), dim3(64), 0, go.s,
                         P, tab, go.img, go.geom, bank, go.runs, r, ncol);
      QPSDM_HIP(hipGetLastError());
    }
    return QPSDM_OK;
  }
  int nwg = go.nwg;
  if (nwg <= 0) {
    int per_cu = 0, dev = 0, ncu = 0;
    QPSDM_HIP(hipGetDevice(&dev));
    QPSDM_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    QPSDM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, 0));
    nwg = ncu * (per_cu > 0 ? per_cu : 1);
  }
  if (go.nwg_out) *go.nwg_out = nwg;
  if (!go.run) return QPSDM_OK;
  const hipStream_t s = go.s;
  for (int r = 0; r < go.nshot; ++r) {
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, s, go.work, 1);
    hipLaunchKernelGGL(kern, dim3(nwg), dim3(64), 0, s, P, tab, go.img, go.geom, bank, go.runs, r, go.work, ncol,
                       ncol * nzg);
    QPSDM_HIP(hipGetLastError());
  }
  return QPSDM_OK;
}

// Node form of every shot, in shot order: per shot the exact body (one
// work-group per tile of 8 points) above the node range and beyond the
// tables' depth range, then the node body (one work-group per node tile).
template <int NI, int PM, bool SRC, typename BankT, typename Tab>
int launch_batch_nodes_t(const qk::ImgDev &P, const Tab &tab, const BankT *bank, const Go &go) {
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
    hipLaunchKernelGGL((image_batch_nodes_kernel<NI, PM, BankT, Tab, SRC>), dim3(ncol, P.nf_ngroups), dim3(64), 0,
                       go.s, P, tab, go.img, go.geom, bank, go.runs, r);
    QPSDM_HIP(hipGetLastError());
  }
  return QPSDM_OK;
}

template <typename BankT, typename Tab>
int launch_batch_nodes(const qk::ImgDev &P, const Tab &tab, const BankT *bank, const Go &go) {
  constexpr bool ok = (std::is_same<BankT, float>::value || std::is_same<BankT, __half>::value) &&
                      (std::is_same<Tab, qk::TabF32>::value || std::is_same<Tab, qk::TabU16Q>::value);
  if constexpr (!ok) {
    return qpsdm::set_error(QPSDM_EINVAL,
                            "image_batch: the node form takes FP32 or FP16 banks and FP32 or U16Q tables");
  } else {
    const bool src = go.runs.src_hdr != nullptr;
    switch (P.tile_z * 4 + P.nf_pm) {
#define QPSDM_NODES(ni, pm)                                                                  \
  case ni * 4 + pm:                                                                          \
    return src ? launch_batch_nodes_t<ni, pm, true>(P, tab, bank, go)                        \
               : launch_batch_nodes_t<ni, pm, false>(P, tab, bank, go);
      QPSDM_NODES(2, 2)
      QPSDM_NODES(2, 3)
      QPSDM_NODES(3, 2)
      QPSDM_NODES(3, 3)
      QPSDM_NODES(4, 2)
      QPSDM_NODES(4, 3)
#undef QPSDM_NODES
      default:
        return qpsdm::set_error(QPSDM_EINVAL, "image_batch: node form with %d intervals x %d points per tile",
                                P.tile_z, P.nf_pm);
    }
  }
}

template <typename BankT, typename Tab>
int launch_batch_tz(const qk::ImgDev &P, const Tab &tab, const BankT *bank, const Go &go) {
  if (P.nf_ngroups > 0) return launch_batch_nodes(P, tab, bank, go);
  const bool src = go.runs.src_hdr != nullptr;
  switch (P.tile_z) {
#define QPSDM_TZ(zp) \
  case zp: return src ? launch_batch_t<zp, true>(P, tab, bank, go) : launch_batch_t<zp, false>(P, tab, bank, go);
    QPSDM_TZ(2)
    QPSDM_TZ(3)
    QPSDM_TZ(4)
    QPSDM_TZ(6)
    QPSDM_TZ(8)
#undef QPSDM_TZ
    default:
      return qpsdm::set_error(QPSDM_EINVAL, "image_batch: unsupported tile_z=%d (2, 3, 4, 6, 8)", P.tile_z);
  }
}

template <typename Tab>
int launch_batch_bank(const qk::ImgDev &P, const Tab &tab, const void *bank, int bank_dtype, const Go &go) {
  if (bank_dtype == QPSDM_F32) return launch_batch_tz(P, tab, (const float *)bank, go);
  if (bank_dtype == QPSDM_BF16) return launch_batch_tz(P, tab, (const __hip_bfloat16 *)bank, go);
  if (bank_dtype == QPSDM_F16) return launch_batch_tz(P, tab, (const __half *)bank, go);
  return qpsdm::set_error(QPSDM_EINVAL, "image_batch: bank dtype %d", bank_dtype);
}

int batch_dispatch(const qpsdm_image_params *p, const qpsdm_launch *shape, const void *bank, int bank_dtype,
                   const void *tab, int table_dtype, const Go &go) {
  qk::ImgDev P;
  int rc = qk::make_dev(p, shape, 0, 0, &P);
  if (rc) return rc;
  if (p->weight_nodes == 2) return launch_batch_spans(p, P, tab, table_dtype, bank, bank_dtype, go);
  switch (table_dtype) {
    case QPSDM_F32: return launch_batch_bank(P, qk::make_tab_f32(p->tg, tab), bank, bank_dtype, go);
    case QPSDM_F16B: return launch_batch_bank(P, qk::make_tab_b16<false>(p->tg, tab), bank, bank_dtype, go);
    case QPSDM_U16B: return launch_batch_bank(P, qk::make_tab_b16<true>(p->tg, tab), bank, bank_dtype, go);
    case QPSDM_U16X: return launch_batch_bank(P, qk::make_tab_u16x(p->tg, tab), bank, bank_dtype, go);
    case QPSDM_U16Q:
      QPSDM_REQUIRE(qk::u16q_fits(p->tg), "image_batch: U16Q node of %zu bytes (below 2 GiB)",
                    qk::u16q_node_bytes(p->tg));
      return launch_batch_bank(P, qk::make_tab_u16q(p->tg, tab), bank, bank_dtype, go);
    default: return qpsdm::set_error(QPSDM_EINVAL, "image_batch: table dtype %d", table_dtype);
  }
}

}  // namespace

extern "C" int qpsdm_image_batch_src(const qpsdm_image_params *p, const qpsdm_launch *shape, float *img,
                                     const float *geom, const void *bank, int bank_dtype, const void *tab,
                                     int table_dtype, const qpsdm_shot *shots, int nshot, const void *src,
                                     unsigned *work, int nwg, void *stream) {
  QPSDM_REQUIRE(p, "image_batch: null params");
  QPSDM_REQUIRE(img && geom && bank && tab && shots && work && nshot >= 0, "image_batch: null buffer");
  QPSDM_REQUIRE(nwg >= -1, "image_batch: nwg must be > 0, 0 (resident) or -1 (one group per tile)");
  if (nshot == 0) return QPSDM_OK;
  const Go go{img, (const float4 *)geom, shot_runs(p, shots, nshot, src), nshot, work, nwg, (hipStream_t)stream,
              true, nullptr};
  return batch_dispatch(p, shape, bank, bank_dtype, tab, table_dtype, go);
}

extern "C" int qpsdm_image_batch_grid(const qpsdm_image_params *p, const qpsdm_launch *shape, int bank_dtype,
                                      int table_dtype, int *nwg) {
  QPSDM_REQUIRE(p && nwg, "image_batch_grid: null args");
  *nwg = 0;
  const Go q{nullptr, nullptr, ShotRuns{nullptr, p->taperzone, nullptr, nullptr, 0, 0, 0}, 0, nullptr, 0, nullptr,
             false, nwg};
  return batch_dispatch(p, shape, nullptr, bank_dtype, nullptr, table_dtype, q);
}

extern "C" int qpsdm_image_batch(const qpsdm_image_params *p, const qpsdm_launch *shape, float *img,
                                 const float *geom, const void *bank, int bank_dtype, const void *tab,
                                 int table_dtype, const qpsdm_shot *shots, int nshot, unsigned *work,
                                 int nwg, void *stream) {
  return qpsdm_image_batch_src(p, shape, img, geom, bank, bank_dtype, tab, table_dtype, shots, nshot, nullptr,
                               work, nwg, stream);
}

extern "C" int qpsdm_count_batch(const qpsdm_image_params *p, const float *geom, const qpsdm_shot *shots,
                                 int nshot, unsigned long long *count, void *stream) {
  qpsdm_launch sh = {8, 8, 8};  // count kernel shape
  qk::ImgDev P;
  QPSDM_REQUIRE(p, "count_batch: null params");
  const qpsdm_image_params pe = qk::exact_form(p);
  int rc = qk::make_dev(&pe, &sh, 0, 0, &P);
  if (rc) return rc;
  QPSDM_REQUIRE(geom && shots && count && nshot >= 0, "count_batch: null buffer");
  if (nshot == 0) return QPSDM_OK;
  const int nxt = (P.nx + 7) / 8, nyt = (P.ny + 7) / 8;
  hipLaunchKernelGGL(count_batch_kernel<8>, dim3(nxt * nyt, (P.nz + 7) / 8), dim3(64), 0, (hipStream_t)stream,
                     P, (const float4 *)geom, shot_runs(p, shots, nshot, nullptr), nshot, count);
  QPSDM_HIP(hipGetLastError());
  return QPSDM_OK;
}

// ---- captured graph of one batch ---------------------------------------
struct qpsdm_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" int qpsdm_batch_graph_create(const qpsdm_bank_params *bp, const float *traces, int ntr, void *bank,
                                        int bank_dtype, const qpsdm_image_params *ip, const qpsdm_launch *shape,
                                        float *img, const float *geom, const void *tab, int table_dtype,
                                        const qpsdm_shot *shots, int nshot, const qpsdm_surrogate *sur,
                                        const float *tab_f32, const unsigned char *mask, void *src,
                                        unsigned *work, int nwg, int stages, qpsdm_graph **out) {
  QPSDM_REQUIRE(stages > 0 && (stages & ~7) == 0, "batch_graph_create: stages must be a non-empty mask of 1|2|4");
  QPSDM_REQUIRE(!sur || src, "batch_graph_create: a surrogate needs the source-table buffer");
  QPSDM_REQUIRE(out && bp && ip && traces && bank && img && geom && tab && shots && work,
                "batch_graph_create: null args");
  QPSDM_REQUIRE(ntr > 0 && nshot > 0, "batch_graph_create: empty batch");
  // everything that allocates or copies synchronously happens before capture
  int rc = qpsdm_qbank_prepare(bp);
  if (rc) return rc;
  QPSDM_REQUIRE(nwg >= -1, "batch_graph_create: nwg must be > 0, 0 (resident) or -1 (one group per tile)");
  if (nwg == 0) {  // resolve the resident grid size outside the capture
    const Go q{img, (const float4 *)geom, shot_runs(ip, shots, nshot, nullptr), nshot, work, 0, nullptr, false, &nwg};
    rc = batch_dispatch(ip, shape, bank, bank_dtype, tab, table_dtype, q);
    if (rc) return rc;
  }
  if (ip->weight_nodes == 2 && (stages & 4) && !fanout()) return QPSDM_EHIP;  // its streams before the capture
  hipStream_t s;
  QPSDM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  qpsdm_graph *g = new qpsdm_graph();
  auto fail = [&](int code) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    (void)hipStreamDestroy(s);
    return code;
  };
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(qpsdm::set_error(QPSDM_EHIP, "batch_graph_create: begin capture failed"));
  if (stages & 1) rc = qpsdm_qbank(bp, traces, ntr, bank, bank_dtype, s);
  if (!rc && sur && (stages & 2)) rc = qpsdm_surrogate_src_tables(sur, ip, geom, shots, nshot, tab_f32, mask, src, s);
  if (!rc && (stages & 4))
    rc = qpsdm_image_batch_src(ip, shape, img, geom, bank, bank_dtype, tab, table_dtype, shots, nshot,
                               sur ? src : nullptr, work, nwg, s);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &graph);
  g->graph = graph;
  if (rc) return fail(rc);
  if (e != hipSuccess || !graph)
    return fail(qpsdm::set_error(QPSDM_EHIP, "batch_graph_create: end capture failed: %s", hipGetErrorString(e)));
  if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
    return fail(qpsdm::set_error(QPSDM_EHIP, "batch_graph_create: instantiate failed"));
  (void)hipStreamDestroy(s);
  *out = g;
  return QPSDM_OK;
}

extern "C" int qpsdm_graph_launch(qpsdm_graph *g, void *stream) {
  QPSDM_REQUIRE(g && g->exec, "graph_launch: null graph");
  QPSDM_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return QPSDM_OK;
}

extern "C" int qpsdm_graph_destroy(qpsdm_graph *g) {
  if (!g) return QPSDM_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
  return QPSDM_OK;
}
