// Standalone closest-hit / any-hit query kernels on gfx950: persistent waves
// running trace_loop (trace_device.h), a table of kernel variants for the
// autotuner, and the launcher.
//
// Kernel shape: persistent waves take slices of the ray stream from a global
// counter; when enough lanes have finished (wave ballot), the wave refills the
// idle lanes from its slice (ray compaction), so long rays do not hold a wave
// of finished lanes. The per-lane stack lives in LDS ([depth][lane],
// conflict-free) and spills to a global [depth][thread] area past the LDS
// depth of the variant.
#include <algorithm>

#include "trace_device.h"

namespace rq {
namespace {

template <bool ANY, bool COUNT, int BLOCK, int DEPTH, bool TSTACK>
__global__ __launch_bounds__(BLOCK) void k_trace(TraceArgs a) {
  __shared__ int32_t s_node[DEPTH * BLOCK];
  __shared__ float s_tn[TSTACK ? DEPTH * BLOCK : 1];
  trace_loop<ANY, COUNT, BLOCK, DEPTH, TSTACK>(a, &a.counters[0], s_node, s_tn,
                                               blockIdx.x * BLOCK + threadIdx.x,
                                               gridDim.x * BLOCK);
  // last wave out resets the counters for the next launch on this slot
  if ((threadIdx.x & 63) == 0) {
    const uint32_t waves = gridDim.x * (BLOCK / 64);
    const uint32_t old = atomicAdd(&a.counters[1], 1u);
    if (old == waves - 1) {
      atomicExch(&a.counters[0], 0u);
      atomicExch(&a.counters[1], 0u);
    }
  }
}

typedef void (*TraceFn)(TraceArgs);
struct Variant {
  int block, depth;
  bool tstack;
  TraceFn closest, any;
};

#define RQ_VARIANT(B, D, T) \
  { B, D, T, k_trace<false, false, B, D, T>, k_trace<true, false, B, D, false> }
const Variant kVariants[] = {
    RQ_VARIANT(128, 24, false), RQ_VARIANT(128, 16, false), RQ_VARIANT(64, 24, false),
    RQ_VARIANT(256, 16, false), RQ_VARIANT(128, 24, true),  RQ_VARIANT(128, 16, true),
    RQ_VARIANT(64, 16, true),   RQ_VARIANT(64, 32, false),  RQ_VARIANT(64, 12, true),
};
#undef RQ_VARIANT
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

}  // namespace

int trace_num_variants() { return kNumVariants; }

void trace_variant_info(int v, int* block, int* depth, int* tstack) {
  const Variant& x = kVariants[v];
  *block = x.block, *depth = x.depth, *tstack = x.tstack ? 1 : 0;
}

uint32_t trace_grid_blocks(int device, int v) {
  int cus = 0, per = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kVariants[v].closest,
                                                     kVariants[v].block, 0);
  if (cus <= 0) cus = 256;
  if (per <= 0) per = 2;
  return (uint32_t)(cus * per);
}

hipError_t launch_trace(const LaunchCtx& lc, const TraceArgs& a, bool any, bool counted,
                        int variant) {
  if (a.nrays == 0) return hipSuccess;
  // the instrumented kernels always run the canonical configuration (variant 0)
  const Variant& V = kVariants[counted ? 0 : variant];
  uint32_t need = (uint32_t)((a.nrays + V.block - 1) / V.block);
  uint32_t grid = std::max(1u, std::min(a.grid, need));
  dim3 g(grid), blk(V.block);
  hipStream_t s = lc.stream;
  if (any) {
    ProfScope ps(lc, counted ? "trace_any_counted" : "trace_any");
    if (counted)
      hipLaunchKernelGGL((k_trace<true, true, 128, 24, false>), g, blk, 0, s, a);
    else
      hipLaunchKernelGGL(V.any, g, blk, 0, s, a);
  } else {
    ProfScope ps(lc, counted ? "trace_closest_counted" : "trace_closest");
    if (counted)
      hipLaunchKernelGGL((k_trace<false, true, 128, 24, false>), g, blk, 0, s, a);
    else
      hipLaunchKernelGGL(V.closest, g, blk, 0, s, a);
  }
  return hipGetLastError();
}

}  // namespace rq
