// Treelet restructuring of a built LBVH (gfx950): one round of the method of
// Karras & Aila, "Fast Parallel Construction of High-Quality Bounding Volume
// Hierarchies" (HPG 2013, section 4), with one primitive per leaf. Optional:
// the reference never optimises its trees; rq_bvh_treelets applies it to a
// built BVH, trading build time for fewer node visits per ray.
//
// Cost of a tree = sum of its internal nodes' half areas (dx*dy + dy*dz) +
// dz*dx. Every internal node with at least 7 leaves below it roots a treelet,
// children before parents: starting from the node's two children, the treelet
// leaf with the largest half area (the first in slot order on ties) that is an
// internal node is replaced by its two children (the second one appended as
// the last slot) until there are 7. The optimal tree over the 7 leaves comes
// from dynamic programming over their 127 subsets; it replaces the treelet
// when its cost is strictly lower and its height is not greater (the tree
// height never grows, so the traversal stack bound holds). The root keeps its
// index; the other internal nodes are handed out in formation order to the
// new internal nodes in pre-order, the child holding the lowest slot first.
//
// Schedule: a node's outcome depends only on its own subtree, and the
// treelets rooted at one depth have disjoint subtrees, so one launch per
// depth, deepest first, gives the same tree as any children-first order (the
// oracle's post-order). Per depth, a scan gives every node at that depth its
// height and leaf count from its children (written by the level below) and
// lists the treelet roots; then persistent waves take the roots in turn, all
// 64 lanes of a wave on one treelet. Node depths come from a walk up the
// parent links.
//
// Per-node scratch (free once the LBVH is built): B.flags = depth, B.slot_h =
// height (first n - 1 words) and leaf count (next n - 1 words), B.frontier =
// the treelet roots of the current depth, B.status[d] = their number.
#include <algorithm>

#include "rq_internal.h"

namespace rq {
namespace {

constexpr int TL_LEAVES = 7;
constexpr int TL_SUBSETS = 1 << TL_LEAVES;  // 128 (index 0 unused)
constexpr uint32_t ERR_TREELET_DEPTH = 32u;
constexpr int TL_MAX_DEPTHS = 64;      // internal-node depths (the tree is at most 64 deep)
constexpr uint32_t TL_PROC_WAVES = 8192;

// (subset, sub) pairs of the subset DP by subset size (2..7): the split of
// s is p = lowest(s) | sub, sub a proper subset of s - lowest(s). 966 pairs.
struct DpTables {
  uint16_t pair[966];     // s | sub << 7
  uint16_t pair_off[9];   // pairs of size k: [pair_off[k], pair_off[k + 1])
  uint8_t subset[120];    // subsets of size 2..7
  uint16_t subset_off[9];
};
constexpr int popc7(int s) {
  int c = 0;
  for (int k = 0; k < 7; ++k) c += (s >> k) & 1;
  return c;
}
constexpr DpTables make_tables() {
  DpTables t{};
  int np = 0, ns = 0;
  for (int k = 0; k <= 8; ++k) {
    if (k < 9) t.pair_off[k] = (uint16_t)np, t.subset_off[k] = (uint16_t)ns;
    if (k < 2 || k > 7) continue;
    for (int s = 1; s < TL_SUBSETS; ++s) {
      if (popc7(s) != k) continue;
      t.subset[ns++] = (uint8_t)s;
      const int low = s & -s, rest = s ^ low;
      for (int sub = (rest - 1) & rest;; sub = (sub - 1) & rest) {
        t.pair[np++] = (uint16_t)(s | (sub << 7));
        if (sub == 0) break;
      }
    }
  }
  return t;
}
__constant__ DpTables kTables = make_tables();
static_assert(make_tables().pair_off[8] == 966 && make_tables().subset_off[8] == 120,
              "subset DP tables");
static_assert(sizeof(DpTables) % 4 == 0, "DP tables copied in words");

__device__ inline float half_area(const float* b) {
  const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return (dx * dy + dy * dz) + dz * dx;
}
__device__ inline int32_t link_lo(const NodeRec& r) { return __float_as_int(r.lo0.w); }
__device__ inline int32_t link_hi(const NodeRec& r) { return __float_as_int(r.hi0.w); }
__device__ inline void child_box(const NodeRec& r, int side, float* b) {
  const float4 lo = side ? r.lo1 : r.lo0, hi = side ? r.hi1 : r.hi0;
  b[0] = lo.x, b[1] = lo.y, b[2] = lo.z, b[3] = hi.x, b[4] = hi.y, b[5] = hi.z;
}
__device__ inline uint32_t link_slot(int32_t c, uint32_t n) {
  return c >= 0 ? (uint32_t)c : (n - 1) + (uint32_t)~c;
}

// Per-wave LDS of one treelet.
struct alignas(8) TreeletLds {
  unsigned long long best[TL_SUBSETS];  // (cost bits << 32) | (127 - sub), min wins
  float A[TL_SUBSETS], C[TL_SUBSETS];
  uint8_t P[TL_SUBSETS], H[TL_SUBSETS];
  float box[TL_LEAVES][6];
  int32_t L[TL_LEAVES];
  uint32_t cnt[TL_LEAVES];
  int32_t I[TL_LEAVES - 1], IL[TL_LEAVES - 1], IR[TL_LEAVES - 1];
  int32_t at_node[TL_LEAVES - 1];  // new internal node j: index and subset
  uint8_t at_set[TL_LEAVES - 1];
};

// The lanes of one wave exchange data through its LDS area: order the LDS
// accesses before and after.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup", "local");
}

// Depth of every internal node (walk up the parent links, at most 64 steps);
// clears the per-depth root counters.
__global__ __launch_bounds__(256) void k_tl_depth(BuildBuffers B, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (uint32_t)TL_MAX_DEPTHS) B.status[i] = 0u;
  if (i >= n - 1) return;
  uint32_t d = 0, cur = i;
  while (cur != 0u) {
    const int32_t pe = B.parent[cur];
    cur = (uint32_t)pe >> 1;
    if (pe < 0 || cur >= n - 1 || ++d > 64) {
      atomicOr(&B.misc[MISC_ERR], ERR_TREELET_DEPTH);
      d = 0xFFFFFFFFu;  // never matches a level: the node is left as it is
      break;
    }
  }
  B.flags[i] = d;
}

// One treelet rooted at `root` (height hroot, >= 7 leaves), all 64 lanes.
__device__ void treelet(BuildBuffers& B, uint32_t n, uint32_t root, uint32_t hroot,
                        const DpTables& T, TreeletLds& S, uint32_t lane) {
  uint32_t* height = B.slot_h;
  uint32_t* count = B.slot_h + (n - 1);
  // formation: lane k < 7 holds treelet slot k (its link and box)
  {
    const NodeRec r0 = B.nodes[root];
    int32_t my_link = 0;
    float my_box[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (lane < 2) {
      my_link = lane ? link_hi(r0) : link_lo(r0);
      child_box(r0, (int)lane, my_box);
    }
    if (lane == 0) S.I[0] = (int32_t)root, S.IL[0] = link_lo(r0), S.IR[0] = link_hi(r0);
    for (uint32_t m = 2; m < (uint32_t)TL_LEAVES; ++m) {
      // the largest half area among the slots holding an internal node, the
      // first slot on ties (areas are >= +0, so their bits order them)
      const bool open = lane < m && my_link >= 0;
      uint32_t ab = open ? __float_as_uint(half_area(my_box)) : 0u, mx = ab;
      for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
      const unsigned long long cand = __ballot(open && ab == mx);
      const int best = cand ? __ffsll((long long)cand) - 1 : 0;  // >= 7 leaves: never empty
      const int32_t j = __shfl(my_link, best, 64);
      const NodeRec r = B.nodes[j < 0 ? 0 : j];
      if ((int)lane == best) my_link = link_lo(r), child_box(r, 0, my_box);
      if (lane == m) my_link = link_hi(r), child_box(r, 1, my_box);
      if (lane == 0) S.I[m - 1] = j, S.IL[m - 1] = link_lo(r), S.IR[m - 1] = link_hi(r);
    }
    if (lane < (uint32_t)TL_LEAVES) {
      S.L[lane] = my_link;
#pragma unroll
      for (int q = 0; q < 6; ++q) S.box[lane][q] = my_box[q];
    }
  }
  wave_sync();
  // subset areas, singletons, DP init
  for (uint32_t s = lane + 1; s < TL_SUBSETS; s += 64) {
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
#pragma unroll
    for (int k = 0; k < TL_LEAVES; ++k)
      if (s >> k & 1u) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          b[q] = fminf(b[q], S.box[k][q]);
          b[3 + q] = fmaxf(b[3 + q], S.box[k][3 + q]);
        }
      }
    S.A[s] = half_area(b);
    S.best[s] = ~0ull;
  }
  if (lane < TL_LEAVES) {
    const int32_t c = S.L[lane];
    const uint32_t s = 1u << lane;
    S.C[s] = 0.0f, S.P[s] = 0;
    S.H[s] = (uint8_t)(c >= 0 ? height[c] : 0u);
    S.cnt[lane] = c >= 0 ? count[c] : 1u;
  }
  wave_sync();
  for (int k = 2; k <= TL_LEAVES; ++k) {
    for (uint32_t t = T.pair_off[k] + lane; t < T.pair_off[k + 1]; t += 64) {
      const uint32_t e = T.pair[t], s = e & 127u, sub = e >> 7;
      const uint32_t p = (s & (0u - s)) | sub;
      const float c = S.C[p] + S.C[s ^ p];
      const unsigned long long key =
          ((unsigned long long)__float_as_uint(c) << 32) | (unsigned long long)(127u - sub);
      atomicMin(&S.best[s], key);
    }
    wave_sync();
    for (uint32_t t = T.subset_off[k] + lane; t < T.subset_off[k + 1]; t += 64) {
      const uint32_t s = T.subset[t];
      const unsigned long long key = S.best[s];
      const uint32_t p = (s & (0u - s)) | (127u - (uint32_t)(key & 0xFFFFFFFFull));
      S.C[s] = S.A[s] + __uint_as_float((uint32_t)(key >> 32));
      S.P[s] = (uint8_t)p;
      S.H[s] = (uint8_t)(1u + max((uint32_t)S.H[p], (uint32_t)S.H[s ^ p]));
    }
    wave_sync();
  }
  // cost of the treelet as it stands (reverse formation order: a node's
  // internal children were opened after it)
  int mask[TL_LEAVES - 1];
  float cost[TL_LEAVES - 1];
#pragma unroll
  for (int j = TL_LEAVES - 2; j >= 0; --j) {
    int mk[2] = {0, 0};
    float cc[2] = {0.0f, 0.0f};
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int32_t c = side ? S.IR[j] : S.IL[j];
      bool found = false;
#pragma unroll
      for (int k = 0; k < TL_LEAVES; ++k)
        if (!found && S.L[k] == c) mk[side] = 1 << k, cc[side] = 0.0f, found = true;
#pragma unroll
      for (int q = 1; q < TL_LEAVES - 1; ++q)
        if (!found && q > j && S.I[q] == c) mk[side] = mask[q], cc[side] = cost[q], found = true;
    }
    mask[j] = mk[0] | mk[1];
    cost[j] = S.A[mask[j]] + (cc[0] + cc[1]);
  }
  const bool rebuild = S.C[127] < cost[0] && (uint32_t)S.H[127] <= hroot;
  if (!rebuild) return;  // height and leaf count of the root are already written
  // new internal nodes in pre-order (lane 0)
  if (lane == 0) {
    uint8_t stack[TL_LEAVES];
    int sp = 0, next = 0;
    stack[sp++] = 127;
    while (sp > 0) {
      const uint32_t s = stack[--sp];
      S.at_node[next] = S.I[next], S.at_set[next] = (uint8_t)s;
      ++next;
      const uint32_t p = S.P[s], q = s ^ p;
      if (q & (q - 1u)) stack[sp++] = (uint8_t)q;
      if (p & (p - 1u)) stack[sp++] = (uint8_t)p;
    }
  }
  wave_sync();
  // lane j writes new internal node j: child links and boxes, its children's
  // parent links, its height and leaf count
  if (lane < TL_LEAVES - 1) {
    const uint32_t s = S.at_set[lane];
    const int32_t idx = S.at_node[lane];
    const uint32_t part[2] = {S.P[s], s ^ S.P[s]};
    int32_t lk[2];
    float bb[2][6];
    uint32_t cn = 0;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const uint32_t q = part[side];
      if ((q & (q - 1u)) == 0u) {
        lk[side] = S.L[__builtin_ctz(q)];
      } else {
        int32_t f = 0;
#pragma unroll
        for (int j = 0; j < TL_LEAVES - 1; ++j)
          if (S.at_set[j] == q) f = S.at_node[j];
        lk[side] = f;
      }
      float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
#pragma unroll
      for (int k = 0; k < TL_LEAVES; ++k)
        if (q >> k & 1u) {
          cn += S.cnt[k];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            b[a] = fminf(b[a], S.box[k][a]);
            b[3 + a] = fmaxf(b[3 + a], S.box[k][3 + a]);
          }
        }
#pragma unroll
      for (int q2 = 0; q2 < 6; ++q2) bb[side][q2] = b[q2];
    }
    NodeRec r;
    r.lo0 = make_float4(bb[0][0], bb[0][1], bb[0][2], __int_as_float(lk[0]));
    r.hi0 = make_float4(bb[0][3], bb[0][4], bb[0][5], __int_as_float(lk[1]));
    r.lo1 = make_float4(bb[1][0], bb[1][1], bb[1][2], 0.0f);
    r.hi1 = make_float4(bb[1][3], bb[1][4], bb[1][5], 0.0f);
    B.nodes[idx] = r;
    B.parent[link_slot(lk[0], n)] = (int32_t)((uint32_t)idx << 1);
    B.parent[link_slot(lk[1], n)] = (int32_t)(((uint32_t)idx << 1) | 1u);
    height[idx] = S.H[s];
    count[idx] = cn;
  }
}

// All internal nodes at depth d: height and leaf count from the children
// (written by the level below), and the list of treelet roots (>= 7 leaves)
// for k_tl_proc. A block takes TL_SCAN_NODES nodes and gathers its roots in
// LDS, so the per-depth counter (cleared by k_tl_depth) gets one device
// atomic per block: atomics on one address serialise across the XCDs.
constexpr int TL_SCAN_BLOCK = 256;
constexpr uint32_t TL_SCAN_NODES = 4096;
__global__ __launch_bounds__(TL_SCAN_BLOCK) void k_tl_scan(BuildBuffers B, uint32_t n, uint32_t d) {
  __shared__ uint32_t s_roots[TL_SCAN_NODES];
  __shared__ uint32_t s_count, s_base;
  uint32_t* height = B.slot_h;
  uint32_t* count = B.slot_h + (n - 1);
  if (threadIdx.x == 0) s_count = 0u;
  __syncthreads();
  const uint32_t first = blockIdx.x * TL_SCAN_NODES;
  for (uint32_t k = threadIdx.x; k < TL_SCAN_NODES; k += TL_SCAN_BLOCK) {
    const uint32_t i = first + k;
    if (i >= n - 1 || B.flags[i] != d) continue;
    const NodeRec r = B.nodes[i];
    const int32_t a = link_lo(r), b = link_hi(r);
    const uint32_t c = (a >= 0 ? count[a] : 1u) + (b >= 0 ? count[b] : 1u);
    height[i] = 1u + max(a >= 0 ? height[a] : 0u, b >= 0 ? height[b] : 0u);
    count[i] = c;
    if (c >= (uint32_t)TL_LEAVES) s_roots[atomicAdd(&s_count, 1u)] = i;
  }
  __syncthreads();
  const uint32_t total = s_count;
  if (total == 0u) return;
  if (threadIdx.x == 0) s_base = atomicAdd(&B.status[d], total);
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < total; k += TL_SCAN_BLOCK) B.frontier[s_base + k] = s_roots[k];
}

// The treelets of depth d's list, one wave per treelet, the waves striding
// over the list. The DP tables are copied to LDS once per block (read from
// global memory they would put a load latency on every DP step).
constexpr int TL_PROC_BLOCK = 256;
__global__ __launch_bounds__(TL_PROC_BLOCK) void k_tl_proc(BuildBuffers B, uint32_t n, uint32_t d) {
  __shared__ TreeletLds s_tl[TL_PROC_BLOCK / 64];
  __shared__ DpTables s_tab;
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&kTables);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&s_tab);
    for (uint32_t k = threadIdx.x; k < sizeof(DpTables) / 4; k += TL_PROC_BLOCK) dst[k] = src[k];
  }
  __syncthreads();
  const uint32_t total = B.status[d], lane = threadIdx.x & 63u;
  const uint32_t waves = gridDim.x * (TL_PROC_BLOCK / 64);
  TreeletLds& S = s_tl[threadIdx.x >> 6];
  for (uint32_t e = blockIdx.x * (TL_PROC_BLOCK / 64) + (threadIdx.x >> 6); e < total; e += waves) {
    const uint32_t root = B.frontier[e];
    treelet(B, n, root, B.slot_h[root], s_tab, S, lane);
    wave_sync();
  }
}

}  // namespace

hipError_t launch_treelets(const LaunchCtx& lc, BuildBuffers& B, uint32_t n, uint32_t height,
                           int rounds) {
  if (n < (uint32_t)TL_LEAVES || height == 0) return hipSuccess;
  if (height > (uint32_t)TL_MAX_DEPTHS) return hipErrorInvalidValue;
  const dim3 grid((n - 1 + 255) / 256), block(256);
  const dim3 sgrid((n - 1 + TL_SCAN_NODES - 1) / TL_SCAN_NODES), sblock(TL_SCAN_BLOCK);
  // persistent treelet waves: enough to fill the device, at most one per 7 nodes
  const uint32_t pwaves = std::min<uint32_t>(TL_PROC_WAVES, (n - 1) / TL_LEAVES + 1);
  const dim3 pgrid((pwaves + TL_PROC_BLOCK / 64 - 1) / (TL_PROC_BLOCK / 64)), pblock(TL_PROC_BLOCK);
  for (int r = 0; r < rounds; ++r) {
    {
      ProfScope ps(lc, "treelet_depth");
      hipLaunchKernelGGL(k_tl_depth, grid, block, 0, lc.stream, B, n);
    }
    // internal nodes lie at depths 0 .. height - 1; the height never grows
    for (uint32_t d = height; d-- > 0;) {
      {
        ProfScope ps(lc, "treelet_scan");
        hipLaunchKernelGGL(k_tl_scan, sgrid, sblock, 0, lc.stream, B, n, d);
      }
      ProfScope ps(lc, "treelet_proc");
      hipLaunchKernelGGL(k_tl_proc, pgrid, pblock, 0, lc.stream, B, n, d);
    }
  }
  return hipGetLastError();
}

}  // namespace rq
