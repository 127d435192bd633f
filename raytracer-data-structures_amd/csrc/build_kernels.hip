// Standalone LBVH build kernels: one block per unit of the phases in
// build_device.h, launched in order on the caller's stream.
#include <algorithm>

#include "build_device.h"

namespace rq {
namespace {

__global__ __launch_bounds__(256) void k_build_init(BuildView v) {
  build_init_part(v, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// The two triangle passes gather each triangle's indices and then its
// vertices, a dependent pair of loads per triangle: 1024 threads (4 triangles
// each) per 4096-triangle tile keep more of them in flight than 256 threads
// with 16 each, for the same number of tiles and device atomics.
constexpr int TRI_BLOCK = 1024;
__global__ __launch_bounds__(TRI_BLOCK) void k_centroid_bounds(BuildView v) {
  __shared__ uint32_t lds[6 * (TRI_BLOCK / 64)];
  bounds_tile<TRI_BLOCK>(v, blockIdx.x, lds);
}

__global__ __launch_bounds__(TRI_BLOCK) void k_morton_hist(BuildView v) {
  __shared__ uint32_t lds[SORT_PASSES * 256];
  morton_tile<TRI_BLOCK>(v, blockIdx.x, lds);
}

// The tile id comes from a per-pass ticket, so a tile's look-back only ever
// waits on tiles whose blocks are already running.
template <int PASS>
__global__ __launch_bounds__(BUILD_BLOCK) void k_onesweep(BuildView v) {
  __shared__ uint32_t lds[BUILD_LDS_WORDS];
  __shared__ uint32_t s_tile;
  if (threadIdx.x == 0) s_tile = atomicAdd(&v.B.misc[MISC_TILECTR + PASS], 1u);
  __syncthreads();
  onesweep_tile<PASS>(v, s_tile, lds);
}
// Small builds sort 8 keys per thread (tiles of SORT_TILE_LEAN keys): with
// few tiles a pass costs one tile's latency, and a tile of half the keys
// ranks in half the steps.
template <int PASS>
__global__ __launch_bounds__(BUILD_BLOCK) void k_onesweep_lean(BuildView v) {
  __shared__ uint32_t lds[BUILD_LDS_WORDS];
  __shared__ uint32_t s_tile;
  if (threadIdx.x == 0) s_tile = atomicAdd(&v.B.misc[MISC_TILECTR + PASS], 1u);
  __syncthreads();
  onesweep_tile<PASS, SORT_ITEMS_LEAN>(v, s_tile, lds);
}
typedef void (*OnesweepFn)(BuildView);
const OnesweepFn kOnesweep[SORT_PASSES] = {k_onesweep<0>, k_onesweep<1>, k_onesweep<2>,
                                           k_onesweep<3>};
const OnesweepFn kOnesweepLean[SORT_PASSES] = {k_onesweep_lean<0>, k_onesweep_lean<1>,
                                               k_onesweep_lean<2>, k_onesweep_lean<3>};
constexpr uint32_t SORT_LEAN_MAX_KEYS = 1u << 20;

__global__ __launch_bounds__(BUILD_BLOCK) void k_karras(BuildView v) {
  __shared__ uint32_t lds[karras_lds_words<BUILD_BLOCK>()];
  karras_tile<BUILD_BLOCK>(v, blockIdx.x, lds);
}

template <int CHUNK>
__global__ __launch_bounds__(CHUNK) void k_refit_chunk(BuildView v) {
  __shared__ uint32_t lds[refit_chunk_lds_words<CHUNK>()];
  refit_chunk_tile<CHUNK>(v, blockIdx.x, lds);
}

template <int CHUNK>
__global__ __launch_bounds__(256) void k_refit_span(BuildBuffers B, uint32_t n) {
  refit_span_wave<CHUNK>(B, n, blockIdx.x * blockDim.x + threadIdx.x);
}

// Wide nodes from the finished binary records: every internal node (after a
// blob load), or the nodes spanning refit chunks (the chunk refit writes the
// others).
__global__ __launch_bounds__(256) void k_widen_all(BuildBuffers B, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n - 1) widen_global(B, i);
}
template <int CHUNK>
__global__ __launch_bounds__(256) void k_widen_span(BuildBuffers B, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  if (spans_chunks<CHUNK>(B, i)) widen_global(B, i);
}

// Leaves per chunk (and threads per block) of the standalone chunk-union
// refit. 256-leaf chunks (eight blocks per CU instead of two) measured the
// same for pass 1 on a 1.31 M-triangle mesh (0.139 vs 0.138 ms) and made the
// spanning-node pass slower (0.035 vs 0.020 ms): the pass moves ~370 bytes
// per leaf and is bound by them, not by its occupancy.
constexpr int BUILD_REFIT_CHUNK = REFIT_CHUNK;

// Tree height: each block's deepest leaf to B.flags[block] (free once the
// refit is done), k_height reduces them (a device atomic per wave on one word
// serialises across the XCDs).
__global__ __launch_bounds__(256) void k_height_walk(BuildBuffers B, uint32_t n) {
  __shared__ uint32_t s_h[4];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t h = i < n - 1 ? leaf_depth(B, n, i) : 0u;
  for (int d = 32; d > 0; d >>= 1) h = max(h, (uint32_t)__shfl_xor((int)h, d, 64));
  if ((threadIdx.x & 63) == 0) s_h[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x == 0) B.flags[blockIdx.x] = max(max(s_h[0], s_h[1]), max(s_h[2], s_h[3]));
}

__global__ __launch_bounds__(1024) void k_height(BuildBuffers B, uint32_t nblocks) {
  __shared__ uint32_t s_h[16];
  uint32_t h = 0;
  for (uint32_t b = threadIdx.x; b < nblocks; b += 1024) h = max(h, B.flags[b]);
  for (int d = 32; d > 0; d >>= 1) h = max(h, (uint32_t)__shfl_xor((int)h, d, 64));
  if ((threadIdx.x & 63) == 0) s_h[threadIdx.x >> 6] = h;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 16; ++w) h = max(h, s_h[w]);
    B.misc[MISC_HEIGHT] = h;
  }
}

}  // namespace

hipError_t launch_build(const LaunchCtx& lc, const float* vx, const float* vy, const float* vz,
                        const uint32_t* idx, uint32_t n, int mode, BuildBuffers& B, const float* vr) {
  if (n == 0) return hipSuccess;
  hipStream_t s = lc.stream;
  const BuildView v{vx, vy, vz, idx, n, mode, B, vr, vr != nullptr};
  const bool lean = n < SORT_LEAN_MAX_KEYS;
  const uint32_t ntiles = tiles_of(n, lean ? SORT_TILE_LEAN : SORT_TILE);
  const uint32_t nstatus = SORT_PASSES * ntiles * 256;
  {
    ProfScope ps(lc, "build_init");
    hipLaunchKernelGGL(k_build_init, dim3(std::min<uint32_t>(1024u, tiles_of(nstatus, 256) + 1)),
                       dim3(256), 0, s, v);
  }
  const uint32_t ttiles = tiles_of(n, TRI_TILE);
  if (mode == RQ_MORTON_SCENE_AABB) {
    ProfScope ps(lc, "centroid_bounds");
    hipLaunchKernelGGL(k_centroid_bounds, dim3(ttiles), dim3(TRI_BLOCK), 0, s, v);
  }
  {
    ProfScope ps(lc, "morton_hist");
    hipLaunchKernelGGL(k_morton_hist, dim3(ttiles), dim3(TRI_BLOCK), 0, s, v);
  }
  for (int pass = 0; pass < SORT_PASSES; ++pass) {
    ProfScope ps(lc, "radix_onesweep");
    hipLaunchKernelGGL((lean ? kOnesweepLean : kOnesweep)[pass], dim3(ntiles), dim3(BUILD_BLOCK), 0,
                       s, v);
  }
  if (n >= 2) {
    ProfScope ps(lc, "karras");
    hipLaunchKernelGGL(k_karras, dim3(tiles_of(n - 1, BUILD_BLOCK)), dim3(BUILD_BLOCK), 0, s, v);
  }
  // chunk-union refit (pass 1 with the chunk-internal wide records, pass 2
  // for the nodes spanning chunks, then their wide records)
  constexpr int C = BUILD_REFIT_CHUNK;
  const uint32_t cchunks = tiles_of(n, C);
  {
    ProfScope ps(lc, "refit");
    hipLaunchKernelGGL(k_refit_chunk<C>, dim3(cchunks), dim3(C), 0, s, v);
  }
  if (cchunks > 1) {
    ProfScope ps(lc, "refit_top");
    hipLaunchKernelGGL(k_refit_span<C>, dim3(tiles_of(n - 1, 256)), dim3(256), 0, s, B, n);
  }
  if (cchunks > 1 && B.wide) {
    ProfScope ps(lc, "widen");
    hipLaunchKernelGGL(k_widen_span<C>, dim3(tiles_of(n - 1, 256)), dim3(256), 0, s, B, n);
  }
  return hipGetLastError();
}

hipError_t launch_widen(const LaunchCtx& lc, BuildBuffers& B, uint32_t n) {
  if (n < 2 || !B.wide) return hipGetLastError();
  ProfScope ps(lc, "widen");
  hipLaunchKernelGGL(k_widen_all, dim3(tiles_of(n - 1, 256)), dim3(256), 0, lc.stream, B, n);
  return hipGetLastError();
}

hipError_t launch_height(const LaunchCtx& lc, BuildBuffers& B, uint32_t n) {
  if (n < 2) return hipGetLastError();
  const uint32_t nblocks = tiles_of(n - 1, 256);  // <= n - 1 entries of B.flags
  hipLaunchKernelGGL(k_height_walk, dim3(nblocks), dim3(256), 0, lc.stream, B, n);
  hipLaunchKernelGGL(k_height, dim3(1), dim3(1024), 0, lc.stream, B, nblocks);
  return hipGetLastError();
}

}  // namespace rq
