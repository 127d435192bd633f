// LBVH build on gfx950: Morton codes -> onesweep 8-bit radix sort -> Karras
// internal-node emission -> bottom-up atomic refit.
//
// Replaces constructLBVHTree (accelerators.h:570-586): its Morton step
// (morton3D :385-394 on (centre + 30)/1000, :577), its sort (the decimal LSD
// radixSort :348-369, which sorts a discarded copy) and its tree step (the
// median split :452-564) become four kernels whose results are bit-exact
// with the CPU restatement in oracle/.
#include "rq_internal.h"

namespace rq {
namespace {

// ---- small device helpers --------------------------------------------------
__device__ inline uint32_t f2ord(float f) {
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ inline float ord2f(uint32_t u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// Spread the 10 low bits of v to bit positions 3k (shift-or-mask form).
__device__ inline uint32_t spread3(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

// morton3D semantics (accelerators.h:387-393): clamp(a*1024, 0, 1023) with
// the std::max/std::min comparisons, truncate, x is the high bit of a triple.
__device__ inline uint32_t quant10(float a) {
  a = a * 1024.0f;
  a = (a < 0.0f) ? 0.0f : a;
  a = (1023.0f < a) ? 1023.0f : a;
  return (a != a) ? 0u : (uint32_t)a;
}
__device__ inline uint32_t morton30(float x, float y, float z) {
  return (spread3(quant10(x)) << 2) | (spread3(quant10(y)) << 1) | spread3(quant10(z));
}

__device__ inline float3 centroid(const float* __restrict__ vx, const float* __restrict__ vy,
                                  const float* __restrict__ vz, const uint32_t* __restrict__ idx,
                                  uint32_t t) {
  uint32_t a = idx[3 * t], b = idx[3 * t + 1], c = idx[3 * t + 2];
  float3 r;
  r.x = ((vx[a] + vx[b]) + vx[c]) / 3.0f + 0.0f;
  r.y = ((vy[a] + vy[b]) + vy[c]) / 3.0f + 0.0f;
  r.z = ((vz[a] + vz[b]) + vz[c]) / 3.0f + 0.0f;
  return r;
}

__device__ inline uint32_t ld_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st_agent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Inclusive scan across the 64 lanes of a wave.
__device__ inline uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// ---- kernels -----------------------------------------------------------------
__global__ void k_build_init(uint32_t* __restrict__ misc, uint32_t* __restrict__ status,
                             uint32_t nstatus) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = i; k < nstatus; k += stride) status[k] = 0u;
  if (i < MISC_BOUNDS) misc[i] = 0u;  // histograms + tile counters
  if (i == 0) misc[MISC_ERR] = 0u;
  if (i < 3) {
    misc[MISC_BOUNDS + i] = 0xFFFFFFFFu;  // lo (ordered min)
    misc[MISC_BOUNDS + 3 + i] = 0u;        // hi (ordered max)
  }
}

__global__ __launch_bounds__(256) void k_centroid_bounds(const float* __restrict__ vx,
                                                          const float* __restrict__ vy,
                                                          const float* __restrict__ vz,
                                                          const uint32_t* __restrict__ idx,
                                                          uint32_t n, uint32_t* __restrict__ misc) {
  uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    float3 c = centroid(vx, vy, vz, idx, t);
    uint32_t o[3] = {f2ord(c.x), f2ord(c.y), f2ord(c.z)};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = min(lo[k], o[k]);
      hi[k] = max(hi[k], o[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      lo[k] = min(lo[k], (uint32_t)__shfl_xor((int)lo[k], off, 64));
      hi[k] = max(hi[k], (uint32_t)__shfl_xor((int)hi[k], off, 64));
    }
  }
  __shared__ uint32_t s_red[6][4];
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) s_red[k][wave] = lo[k], s_red[3 + k][wave] = hi[k];
  }
  __syncthreads();
  if (threadIdx.x < 6) {  // one device atomic per bound per block
    const int k = threadIdx.x;
    uint32_t v = s_red[k][0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
      v = k < 3 ? min(v, s_red[k][w]) : max(v, s_red[k][w]);
    if (k < 3)
      atomicMin(&misc[MISC_BOUNDS + k], v);
    else
      atomicMax(&misc[MISC_BOUNDS + k], v);
  }
}

// Morton code per triangle + the four 8-bit digit histograms of all codes.
__global__ __launch_bounds__(256) void k_morton_hist(const float* __restrict__ vx,
                                                      const float* __restrict__ vy,
                                                      const float* __restrict__ vz,
                                                      const uint32_t* __restrict__ idx, uint32_t n,
                                                      int mode, uint32_t* __restrict__ misc,
                                                      uint32_t* __restrict__ codes) {
  __shared__ uint32_t s_hist[SORT_PASSES * 256];
  for (int i = threadIdx.x; i < SORT_PASSES * 256; i += blockDim.x) s_hist[i] = 0u;
  float lo[3] = {0, 0, 0}, ext[3] = {0, 0, 0};
  if (mode == RQ_MORTON_SCENE_AABB) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = ord2f(misc[MISC_BOUNDS + k]);
      ext[k] = ord2f(misc[MISC_BOUNDS + 3 + k]) - lo[k];
    }
  }
  __syncthreads();
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    float3 c = centroid(vx, vy, vz, idx, t);
    float q[3] = {c.x, c.y, c.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (mode == RQ_MORTON_SCENE_AABB)
        q[k] = ext[k] > 0.0f ? (q[k] - lo[k]) / ext[k] : 0.0f;
      else
        q[k] = (q[k] + 30.0f) / 1000.0f;
    }
    uint32_t code = morton30(q[0], q[1], q[2]);
    codes[t] = code;
#pragma unroll
    for (int p = 0; p < SORT_PASSES; ++p) atomicAdd(&s_hist[p * 256 + ((code >> (8 * p)) & 255u)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < SORT_PASSES * 256; i += blockDim.x) {
    uint32_t v = s_hist[i];
    if (v) atomicAdd(&misc[MISC_HIST + i], v);
  }
}

// One stable 8-bit LSD pass (onesweep: ranking in LDS, decoupled look-back
// over per-tile digit counts, LDS-staged scatter). Tile = SORT_TILE keys;
// wave w of a tile owns the contiguous range [w*64*ITEMS, (w+1)*64*ITEMS) and
// lane l its keys i*64 + l, so ranking wave by wave, item by item, keeps the
// input order among equal digits.
__global__ __launch_bounds__(SORT_BLOCK) void k_onesweep(const uint32_t* __restrict__ kin,
                                                          const uint32_t* __restrict__ vin,
                                                          uint32_t* __restrict__ kout,
                                                          uint32_t* __restrict__ vout, uint32_t n,
                                                          int pass, uint32_t* __restrict__ misc,
                                                          uint32_t* __restrict__ status) {
  static_assert(SORT_BLOCK == 256, "one thread per digit in the look-back");
  __shared__ uint32_t s_cnt[SORT_WAVES][256];
  __shared__ uint32_t s_keys[SORT_TILE];
  __shared__ uint32_t s_vals[SORT_TILE];
  __shared__ uint32_t s_excl[256];
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_wsum[2][SORT_WAVES];
  __shared__ uint32_t s_tile;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int shift = 8 * pass;
  if (tid == 0) s_tile = atomicAdd(&misc[MISC_TILECTR + pass], 1u);
  for (int i = tid; i < SORT_WAVES * 256; i += SORT_BLOCK) (&s_cnt[0][0])[i] = 0u;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t base = tile * SORT_TILE;
  const uint32_t wbase = base + wave * (64 * SORT_ITEMS);
  const uint64_t lt = (1ull << lane) - 1ull;

  uint32_t key[SORT_ITEMS], val[SORT_ITEMS], rank[SORT_ITEMS];
#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i) {
    uint32_t g = wbase + i * 64 + lane;
    bool ok = g < n;
    key[i] = ok ? kin[g] : 0u;
    val[i] = ok ? (vin ? vin[g] : g) : 0u;
  }
#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i) {
    bool ok = wbase + i * 64 + lane < n;
    uint32_t d = (key[i] >> shift) & 255u;
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint64_t bal = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    uint32_t before = (uint32_t)__popcll(peers & lt);
    uint32_t c = ok ? s_cnt[wave][d] : 0u;
    rank[i] = c + before;
    if (ok && before == 0) s_cnt[wave][d] = c + (uint32_t)__popcll(peers);
  }
  __syncthreads();

  // per digit (thread = digit): wave-exclusive prefix and tile total
  const uint32_t d = (uint32_t)tid;
  uint32_t total = 0;
#pragma unroll
  for (int w = 0; w < SORT_WAVES; ++w) {
    uint32_t c = s_cnt[w][d];
    s_cnt[w][d] = total;
    total += c;
  }
  uint32_t* st = status + (size_t)tile * 256 + d;
  st_agent(st, (tile == 0 ? LB_INC : LB_AGG) | total);

  // block-exclusive scan of the tile totals and of the global histogram
  uint32_t hcount = misc[MISC_HIST + pass * 256 + d];
  uint32_t inc_t = wave_incl_scan(total, lane);
  uint32_t inc_h = wave_incl_scan(hcount, lane);
  if (lane == 63) {
    s_wsum[0][wave] = inc_t;
    s_wsum[1][wave] = inc_h;
  }
  __syncthreads();
  uint32_t pre_t = 0, pre_h = 0;
  for (int w = 0; w < wave; ++w) pre_t += s_wsum[0][w], pre_h += s_wsum[1][w];
  const uint32_t blk_excl = pre_t + inc_t - total;
  const uint32_t glb_excl = pre_h + inc_h - hcount;

  // decoupled look-back over earlier tiles for this digit
  uint32_t prefix = 0;
  if (tile > 0) {
    int64_t j = (int64_t)tile - 1;
    uint32_t spins = 0;
    for (;;) {
      uint32_t s = ld_agent(status + (size_t)j * 256 + d);
      uint32_t f = s & ~LB_MASK;
      if (f == 0u) {
        if (++spins > LOOKBACK_SPIN_LIMIT) {  // bounded wait: report, never hang
          atomicOr(&misc[MISC_ERR], ERR_LOOKBACK_SPIN);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      prefix += s & LB_MASK;
      if (f == LB_INC) break;
      --j;
    }
    st_agent(st, LB_INC | (prefix + total));
  }
  s_excl[d] = blk_excl;
  s_base[d] = glb_excl + prefix - blk_excl;
  __syncthreads();

#pragma unroll
  for (int i = 0; i < SORT_ITEMS; ++i) {
    if (wbase + i * 64 + lane < n) {
      uint32_t dd = (key[i] >> shift) & 255u;
      uint32_t pos = s_excl[dd] + s_cnt[wave][dd] + rank[i];
      if (pos < (uint32_t)SORT_TILE) {
        s_keys[pos] = key[i];
        s_vals[pos] = val[i];
      } else {
        atomicOr(&misc[MISC_ERR], ERR_SCATTER_RANGE);
      }
    }
  }
  __syncthreads();
  const uint32_t cnt = min((uint32_t)SORT_TILE, n - base);
  for (uint32_t p = tid; p < cnt; p += SORT_BLOCK) {
    uint32_t k = s_keys[p];
    uint32_t g = s_base[(k >> shift) & 255u] + p;
    if (g < n) {
      kout[g] = k;
      vout[g] = s_vals[p];
    } else {
      atomicOr(&misc[MISC_ERR], ERR_SCATTER_RANGE);
    }
  }
}

// Karras 2012 emission: internal node i in [0, n-2] from the sorted codes;
// equal codes fall back to the index (augmented key code<<32 | i).
__device__ inline int kdelta(const uint32_t* __restrict__ k, int64_t n, int64_t i, uint32_t ki,
                             int64_t j) {
  if (j < 0 || j >= n) return -1;
  uint32_t kj = k[j];
  return ki != kj ? __clz(ki ^ kj) : 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

__global__ __launch_bounds__(256) void k_karras(const uint32_t* __restrict__ keys, uint32_t n,
                                                 NodeRec* __restrict__ nodes,
                                                 int32_t* __restrict__ parent,
                                                 uint32_t* __restrict__ flags,
                                                 uint2* __restrict__ range,
                                                 uint32_t* __restrict__ misc) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t N = n;
  if (i >= N - 1) return;
  const uint32_t ki = keys[i];
  const int d = kdelta(keys, N, i, ki, i + 1) > kdelta(keys, N, i, ki, i - 1) ? 1 : -1;
  const int dmin = kdelta(keys, N, i, ki, i - d);
  int64_t lmax = 2;
  while (kdelta(keys, N, i, ki, i + lmax * d) > dmin) lmax <<= 1;
  int64_t l = 0;
  for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
    if (kdelta(keys, N, i, ki, i + (l + t) * d) > dmin) l += t;
  const int64_t j = i + l * d;
  const int dnode = kdelta(keys, N, i, ki, j);
  int64_t s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (kdelta(keys, N, i, ki, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int64_t g = i + s * d + min(d, 0);
  const int32_t left = (min(i, j) == g) ? ~(int32_t)g : (int32_t)g;
  const int32_t right = (max(i, j) == g + 1) ? ~(int32_t)(g + 1) : (int32_t)(g + 1);
  reinterpret_cast<int32_t*>(&nodes[i].lo0)[3] = left;
  reinterpret_cast<int32_t*>(&nodes[i].hi0)[3] = right;
  const int64_t ul = left >= 0 ? left : (N - 1) + ~left;
  const int64_t ur = right >= 0 ? right : (N - 1) + ~right;
  flags[i] = 0u;
  range[i] = make_uint2((uint32_t)min(i, j), (uint32_t)max(i, j));
  if (g < 0 || g + 1 > N - 1 || ul >= 2 * N - 1 || ur >= 2 * N - 1) {
    atomicOr(&misc[MISC_ERR], ERR_TOPOLOGY);
    return;
  }
  parent[ul] = (int32_t)(i << 1);
  parent[ur] = (int32_t)((i << 1) | 1);
}

// Bottom-up refit, pass 1 (one block per REFIT_CHUNK leaves): each leaf
// writes its triangle record and its box into the parent's child slot; the
// second arriver at a node (atomic counter) unions the two child boxes and
// climbs. A node whose leaf range lies inside the block's chunk is completed
// by two threads of this block, so its counter needs only work-group scope.
// A node spanning chunks is left to pass 2: the arriving thread deposits its
// box, counts itself with a relaxed device atomic, and the second of two
// arrivals lists the node in the block's frontier. min/max is
// order-independent and zeros are canonicalised, so the boxes are bit-exact
// whatever the order.
__device__ inline void put_slot(NodeRec* nodes, uint32_t node, uint32_t side, const float* lo,
                                const float* hi) {
  float* slo = side ? &nodes[node].lo1.x : &nodes[node].lo0.x;
  float* shi = side ? &nodes[node].hi1.x : &nodes[node].hi0.x;
  slo[0] = lo[0], slo[1] = lo[1], slo[2] = lo[2];
  shi[0] = hi[0], shi[1] = hi[1], shi[2] = hi[2];
}
__device__ inline void union_slot(const NodeRec* nodes, uint32_t node, uint32_t side, float* lo,
                                  float* hi) {
  const float* olo = side ? &nodes[node].lo1.x : &nodes[node].lo0.x;
  const float* ohi = side ? &nodes[node].hi1.x : &nodes[node].hi0.x;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    lo[q] = fminf(lo[q], olo[q]);
    hi[q] = fmaxf(hi[q], ohi[q]);
  }
}
__device__ inline void put_root(uint32_t* misc, const float* lo, const float* hi, uint32_t h) {
  float* root = reinterpret_cast<float*>(misc + MISC_ROOT);
  for (int q = 0; q < 3; ++q) root[q] = lo[q], root[3 + q] = hi[q];
  misc[MISC_HEIGHT] = h;
}

__global__ __launch_bounds__(REFIT_BLOCK) void k_refit_local(
    const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz,
    const uint32_t* __restrict__ idx, const uint32_t* __restrict__ sorted_idx, uint32_t n,
    NodeRec* nodes, TriRec* __restrict__ tris, const int32_t* __restrict__ parent,
    uint32_t* flags, uint32_t* slot_h, const uint2* __restrict__ range,
    uint32_t* __restrict__ frontier, uint32_t* __restrict__ fcount, uint32_t* __restrict__ misc) {
  __shared__ uint32_t s_nf;
  if (threadIdx.x == 0) s_nf = 0u;
  __syncthreads();
  const uint32_t chunk_lo = blockIdx.x * REFIT_CHUNK;
  const uint32_t chunk_hi = min(chunk_lo + REFIT_CHUNK, n) - 1;
  for (int j = 0; j < REFIT_CHUNK / REFIT_BLOCK; ++j) {
    const uint32_t k = chunk_lo + j * REFIT_BLOCK + threadIdx.x;
    if (k >= n) break;
    const uint32_t p = sorted_idx[k];
    const uint32_t a = idx[3 * p], b = idx[3 * p + 1], c = idx[3 * p + 2];
    const float3 v0 = make_float3(vx[a], vy[a], vz[a]);
    const float3 v1 = make_float3(vx[b], vy[b], vz[b]);
    const float3 v2 = make_float3(vx[c], vy[c], vz[c]);
    TriRec tr;
    tr.v0 = make_float4(v0.x, v0.y, v0.z, __int_as_float((int32_t)p));
    tr.e1 = make_float4(v1.x - v0.x, v1.y - v0.y, v1.z - v0.z, 0.0f);
    tr.e2 = make_float4(v2.x - v0.x, v2.y - v0.y, v2.z - v0.z, 0.0f);
    tris[k] = tr;
    float lo[3] = {fminf(fminf(v0.x, v1.x), v2.x) + 0.0f, fminf(fminf(v0.y, v1.y), v2.y) + 0.0f,
                   fminf(fminf(v0.z, v1.z), v2.z) + 0.0f};
    float hi[3] = {fmaxf(fmaxf(v0.x, v1.x), v2.x) + 0.0f, fmaxf(fmaxf(v0.y, v1.y), v2.y) + 0.0f,
                   fmaxf(fmaxf(v0.z, v1.z), v2.z) + 0.0f};
    if (n == 1) {
      put_root(misc, lo, hi, 0u);
      break;
    }
    uint32_t h = 0, cur = (n - 1) + k;
    for (int level = 0;; ++level) {
      const int32_t pe = parent[cur];
      const uint32_t node = (uint32_t)pe >> 1, side = (uint32_t)pe & 1u;
      if (pe < 0 || node >= n - 1 || level > 64) {  // height <= 63 for 30-bit codes
        atomicOr(&misc[MISC_ERR], ERR_REFIT_DEPTH);
        break;
      }
      put_slot(nodes, node, side, lo, hi);
      slot_h[2 * node + side] = h;
      const uint2 rg = range[node];
      if (rg.x < chunk_lo || rg.y > chunk_hi) {
        // spans chunks: finished by the top pass once both children are in
        const uint32_t old =
            __hip_atomic_fetch_add(&flags[node], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old == 1u) frontier[chunk_lo + atomicAdd(&s_nf, 1u)] = node;
        break;
      }
      const uint32_t old =
          __hip_atomic_fetch_add(&flags[node], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (old == 0u) break;
      union_slot(nodes, node, side ^ 1u, lo, hi);
      h = 1u + max(h, slot_h[2 * node + (side ^ 1u)]);
      if (node == 0u) {
        put_root(misc, lo, hi, h);
        break;
      }
      cur = node;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) fcount[blockIdx.x] = s_nf;
}

// Refit pass 2, one work-group: the nodes spanning chunks, bottom-up from the
// frontier lists of pass 1 (their child slots were written before the kernel
// boundary); counters shared inside the work-group need only its scope.
__global__ __launch_bounds__(REFIT_TOP_BLOCK) void k_refit_top(
    NodeRec* nodes, const int32_t* __restrict__ parent, uint32_t* flags, uint32_t* slot_h,
    const uint32_t* __restrict__ frontier, const uint32_t* __restrict__ fcount, uint32_t nchunks,
    uint32_t n, uint32_t* __restrict__ misc) {
  for (uint32_t bch = threadIdx.x; bch < nchunks; bch += REFIT_TOP_BLOCK) {
    const uint32_t cnt = fcount[bch];
    for (uint32_t e = 0; e < cnt; ++e) {
      uint32_t node = frontier[bch * REFIT_CHUNK + e];
      float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      union_slot(nodes, node, 0u, lo, hi);
      union_slot(nodes, node, 1u, lo, hi);
      uint32_t h = 1u + max(slot_h[2 * node], slot_h[2 * node + 1]);
      for (int level = 0;; ++level) {
        if (node == 0u) {
          put_root(misc, lo, hi, h);
          break;
        }
        const int32_t pe = parent[node];
        const uint32_t q = (uint32_t)pe >> 1, side = (uint32_t)pe & 1u;
        if (pe < 0 || q >= n - 1 || level > 64) {
          atomicOr(&misc[MISC_ERR], ERR_REFIT_DEPTH);
          break;
        }
        put_slot(nodes, q, side, lo, hi);
        slot_h[2 * q + side] = h;
        const uint32_t old =
            __hip_atomic_fetch_add(&flags[q], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (old == 0u) break;
        union_slot(nodes, q, side ^ 1u, lo, hi);
        h = 1u + max(h, slot_h[2 * q + (side ^ 1u)]);
        node = q;
      }
    }
  }
}

inline uint32_t cdiv(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

hipError_t launch_build(const LaunchCtx& lc, const float* vx, const float* vy, const float* vz,
                        const uint32_t* idx, uint32_t n, int mode, BuildBuffers& B) {
  if (n == 0) return hipSuccess;
  hipStream_t s = lc.stream;
  const uint32_t ntiles = cdiv(n, SORT_TILE);
  const uint32_t nstatus = SORT_PASSES * ntiles * 256;
  {
    ProfScope ps(lc, "build_init");
    uint32_t blocks = std::min<uint32_t>(1024u, cdiv(nstatus, 256) + 1);
    hipLaunchKernelGGL(k_build_init, dim3(blocks), dim3(256), 0, s, B.misc, B.status, nstatus);
  }
  const uint32_t gblocks = std::min<uint32_t>(2048u, cdiv(n, 256));
  if (mode == RQ_MORTON_SCENE_AABB) {
    ProfScope ps(lc, "centroid_bounds");
    hipLaunchKernelGGL(k_centroid_bounds, dim3(std::min<uint32_t>(gblocks, 512u)), dim3(256), 0, s,
                       vx, vy, vz, idx, n, B.misc);
  }
  {
    ProfScope ps(lc, "morton_hist");
    hipLaunchKernelGGL(k_morton_hist, dim3(gblocks), dim3(256), 0, s, vx, vy, vz, idx, n, mode,
                       B.misc, B.codes);
  }
  const uint32_t* kin = B.codes;
  const uint32_t* vin = nullptr;
  for (int pass = 0; pass < SORT_PASSES; ++pass) {
    uint32_t* kout = (pass & 1) ? B.keys_a : B.keys_b;
    uint32_t* vout = (pass & 1) ? B.vals_a : B.vals_b;
    ProfScope ps(lc, "radix_onesweep");
    hipLaunchKernelGGL(k_onesweep, dim3(ntiles), dim3(SORT_BLOCK), 0, s, kin, vin, kout, vout, n,
                       pass, B.misc, B.status + (size_t)pass * ntiles * 256);
    kin = kout;
    vin = vout;
  }
  if (n >= 2) {
    ProfScope ps(lc, "karras");
    hipLaunchKernelGGL(k_karras, dim3(cdiv(n - 1, 256)), dim3(256), 0, s, B.keys_a, n, B.nodes,
                       B.parent, B.flags, B.range, B.misc);
  }
  const uint32_t nchunks = cdiv(n, REFIT_CHUNK);
  {
    ProfScope ps(lc, "refit");
    hipLaunchKernelGGL(k_refit_local, dim3(nchunks), dim3(REFIT_BLOCK), 0, s, vx, vy, vz, idx,
                       B.vals_a, n, B.nodes, B.tris, B.parent, B.flags, B.slot_h, B.range,
                       B.frontier, B.fcount, B.misc);
  }
  if (nchunks > 1) {
    ProfScope ps(lc, "refit_top");
    hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(REFIT_TOP_BLOCK), 0, s, B.nodes, B.parent,
                       B.flags, B.slot_h, B.frontier, B.fcount, nchunks, n, B.misc);
  }
  return hipGetLastError();
}

}  // namespace rq
