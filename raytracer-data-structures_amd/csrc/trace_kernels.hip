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

// WAVES > 0: registers capped for that many waves per SIMD.
template <bool ANY, bool COUNT, int BLOCK, int DEPTH, bool TSTACK, bool WIDE = false, int WAVES = 0>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1)))
void k_trace(TraceArgs a) {
  __shared__ int32_t s_node[DEPTH * BLOCK];
  __shared__ float s_tn[TSTACK ? DEPTH * BLOCK : 1];
  RayStream rs;
  rs.ctr64 = a.counter;
  rs.base = a.ctr_base;
  rs.wave = (blockIdx.x * BLOCK + threadIdx.x) >> 6;
  rs.waves = gridDim.x * (BLOCK / 64);
  trace_loop_split<ANY, COUNT, BLOCK, DEPTH, TSTACK, true, WIDE>(
      a, rs, s_node, s_tn, blockIdx.x * BLOCK + threadIdx.x, gridDim.x * BLOCK);
}

// Device scratch of the allocation-free closest-hit entry
// (rq_closest_hit_noalloc): the launch's ray-stream counter in the first 256
// bytes, then the traversal stack overflow area ([depth][thread] pairs). One
// copy per device, held from module load until the process ends; its address
// comes from hipGetSymbolAddress (noalloc_scratch_address), never from the
// array's name in host code. Calls that overlap on one device share it and
// race.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

__global__ __launch_bounds__(64) void k_counter_reset(unsigned long long* ctr) {
  if (threadIdx.x == 0) *ctr = 0ull;
}

typedef void (*TraceFn)(TraceArgs);
struct Variant {
  int block, depth;
  bool tstack, wide;
  int closest_waves, any_waves;  // waves per SIMD the kernels are capped for (0: none)
  TraceFn closest, any;
};

// Variant 0 is the canonical binary configuration (the counted kernels run
// it); the wide variants need the BVH's 4-wide nodes. A wide any-hit kernel
// whose LDS stack leaves room for 6 waves per SIMD (6 x 4 waves x 64 lanes x
// DEPTH x 4 B <= 160 KB: DEPTH <= 26) is compiled for 6 (80 VGPRs, no spills;
// left to itself the compiler takes 81 and gets 5): it waits on dependent
// node fetches more than it issues, and the sixth wave hides more of that
// wait (shadow rays of a 1.31 M-triangle frame 11 %, uniform rays over 5.24 M
// triangles 6 % faster). The closest-hit walk issues at ~0.8 of the peak and
// spills at that cap, so it keeps its 5 waves. The launch grid follows each
// kernel's own occupancy.
constexpr int any_waves(bool wide, int depth) { return wide && depth <= 26 ? 6 : 0; }
#define RQ_VARIANT(B, D, T, W)                                              \
  {                                                                         \
    B, D, T, W, 0, any_waves(W, D), k_trace<false, false, B, D, T, W>,      \
        k_trace<true, false, B, D, false, W, any_waves(W, D)>               \
  }
// Variants whose closest-hit kernel is capped at six waves per SIMD as well
// (80 VGPRs, 2 spilled). In a 1920x1080x4 frame of a 1.31 M-triangle mesh
// its launches take the same time alone, but beside the concurrent any-hit
// launch (also at six) the two share the CUs better: 3.26 -> 3.21 ms per
// frame.
#define RQ_VARIANT6(B, D)                                                   \
  {                                                                         \
    B, D, false, true, 6, any_waves(true, D),                               \
        k_trace<false, false, B, D, false, true, 6>,                        \
        k_trace<true, false, B, D, false, true, any_waves(true, D)>         \
  }
const Variant kVariants[] = {
    RQ_VARIANT(128, 24, false, false), RQ_VARIANT(128, 16, false, false),
    RQ_VARIANT(64, 24, false, false),  RQ_VARIANT(256, 16, false, false),
    RQ_VARIANT(128, 24, true, false),  RQ_VARIANT(128, 16, true, false),
    RQ_VARIANT(64, 16, true, false),   RQ_VARIANT(64, 32, false, false),
    RQ_VARIANT(64, 12, true, false),   RQ_VARIANT(128, 24, false, true),
    RQ_VARIANT(128, 32, false, true), RQ_VARIANT(64, 32, false, true),
    RQ_VARIANT(128, 24, true, true),  RQ_VARIANT(64, 24, true, true),
    RQ_VARIANT(256, 16, false, true),  RQ_VARIANT(256, 24, false, true),
    RQ_VARIANT(256, 28, false, true),  RQ_VARIANT6(256, 16),
    RQ_VARIANT6(128, 24),
};
#undef RQ_VARIANT
#undef RQ_VARIANT6
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
static_assert(TRACE_DEFAULT_CLOSEST < kNumVariants && TRACE_DEFAULT_ANY < kNumVariants,
              "default variants in the table");

}  // namespace

int trace_num_variants() { return kNumVariants; }

static_assert(sizeof(kVariants) / sizeof(kVariants[0]) <= TRACE_MAX_VARIANTS, "variant table");

void trace_variant_info(int v, int* block, int* depth, int* tstack, int* wide,
                        int* closest_waves, int* any_waves) {
  const Variant& x = kVariants[v];
  *block = x.block, *depth = x.depth, *tstack = x.tstack ? 1 : 0;
  if (wide) *wide = x.wide ? 1 : 0;
  if (closest_waves) *closest_waves = x.closest_waves;
  if (any_waves) *any_waves = x.any_waves;
}

int trace_max_stack(int v) { return kVariants[v].wide ? TRACE_MAX_STACK_WIDE : TRACE_MAX_STACK; }

uint32_t trace_grid_blocks(int device, int v, bool any) {
  int cus = 0, per = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per, any ? kVariants[v].any : kVariants[v].closest, kVariants[v].block, 0);
  if (cus <= 0) cus = 256;
  if (per <= 0) per = 2;
  return (uint32_t)(cus * per);
}

hipError_t noalloc_scratch_address(void** ptr) {
  *ptr = nullptr;
  return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

hipError_t launch_counter_reset(const LaunchCtx& lc, unsigned long long* counter) {
  hipLaunchKernelGGL(k_counter_reset, dim3(1), dim3(64), 0, lc.stream, counter);
  return hipGetLastError();
}

hipError_t launch_trace(const LaunchCtx& lc, const TraceArgs& a, bool any, bool counted,
                        int variant) {
  if (a.nrays == 0) return hipSuccess;
  // the instrumented kernels always run the canonical configuration (variant 0);
  // a wide variant needs the wide nodes
  if (!counted && kVariants[variant].wide && !a.wide) return hipErrorInvalidValue;
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
