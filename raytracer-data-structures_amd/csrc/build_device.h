// Block-level device functions of the LBVH build (Morton -> onesweep radix
// sort -> Karras emission -> two-pass refit). Each function is one unit of
// work for one block of BUILD_BLOCK (256) threads (the triangle passes take
// the block size as a parameter): the standalone kernels in
// build_kernels.hip launch one block per unit, the persistent batch kernel in
// batch_kernels.hip pulls the same units from its work queue. LDS scratch is
// passed in so that a kernel calling several phases holds one buffer.
//
// Replaces constructLBVHTree (accelerators.h:570-586): its Morton step
// (morton3D :385-394 on (centre + 30)/1000, :577), its sort (the decimal LSD
// radixSort :348-369, which sorts a discarded copy) and its tree step (the
// median split :452-564); results are bit-exact with the CPU restatement in
// oracle/.
#pragma once
#include "rq_internal.h"

namespace rq {

constexpr int BUILD_BLOCK = 256;
constexpr int TRI_TILE = 4096;   // triangles per bounds / Morton unit
constexpr int NODE_TILE = 4096;  // internal nodes per Karras unit
// LDS words needed by the largest phase (the onesweep tile)
__host__ __device__ constexpr int sort_lds_words(int items) {
  return SORT_WAVES * 256 + SORT_BLOCK * items + 256 + 256 + 2 * SORT_WAVES;
}
constexpr int BUILD_LDS_WORDS = sort_lds_words(SORT_ITEMS);

// Everything one mesh build reads and writes.
// Triangles: vx/vy/vz vertices, idx 3 per primitive, vr null. Spheres
// (the reference's primitive, accelerators.h:54-92): vx/vy/vz centres, vr
// radii, idx unused.
struct BuildView {
  const float *vx, *vy, *vz;
  const uint32_t* idx;
  uint32_t n;
  int mode;
  BuildBuffers B;
  const float* vr = nullptr;
  int32_t sphere = 0;  // vr is set (an integer flag: device code tests it, not the pointer)
};

// ---- helpers ------------------------------------------------------------------
// Element i of a per-primitive array (i < MAX_PRIMS < 2^30): the byte offset
// fits 32 bits, so the access takes a scalar base and a 32-bit lane offset
// (one VGPR per address instead of a 64-bit pair).
template <typename T>
__device__ inline T& el32(T* base, uint32_t i) {
  return *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (uint32_t)(i * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ inline const T& el32(const T* base, uint32_t i) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) +
                                     (uint32_t)(i * (uint32_t)sizeof(T)));
}
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

// Triangle centroid ((v0 + v1) + v2) / 3 per axis; a sphere's centre
// (main.cpp:680-683 places each object at its centre). Zeros canonicalised.
__device__ inline float3 centroid(const BuildView& v, uint32_t t) {
  if (v.sphere) return make_float3(v.vx[t] + 0.0f, v.vy[t] + 0.0f, v.vz[t] + 0.0f);
  const uint32_t a = v.idx[3 * t], b = v.idx[3 * t + 1], c = v.idx[3 * t + 2];
  float3 r;
  r.x = ((v.vx[a] + v.vx[b]) + v.vx[c]) / 3.0f + 0.0f;
  r.y = ((v.vy[a] + v.vy[b]) + v.vy[c]) / 3.0f + 0.0f;
  r.z = ((v.vz[a] + v.vz[b]) + v.vz[c]) / 3.0f + 0.0f;
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

__host__ __device__ inline uint32_t tiles_of(uint32_t n, uint32_t tile) {
  return (n + tile - 1) / tile;
}

// ---- phase 0: reset the build's counters, histograms and look-back words --------
// Threads [first, first + stride) of the caller's grid share the work.
__device__ inline void build_init_part(const BuildView& v, uint32_t first, uint32_t stride) {
  uint32_t* misc = v.B.misc;
  const uint32_t nstatus = SORT_PASSES * tiles_of(v.n, SORT_TILE_LEAN) * 256;  // both tilings
  for (uint32_t k = first; k < nstatus; k += stride) v.B.status[k] = 0u;
  for (uint32_t k = first; k < (uint32_t)MISC_BOUNDS; k += stride) misc[k] = 0u;  // hist, tickets
  if (first == 0) misc[MISC_ERR] = 0u, misc[MISC_HEIGHT] = HEIGHT_UNKNOWN;
  if (first < 3) {
    misc[MISC_BOUNDS + first] = 0xFFFFFFFFu;  // lo (ordered min)
    misc[MISC_BOUNDS + 3 + first] = 0u;        // hi (ordered max)
  }
}

// ---- phase 1: centroid bounds of one tile (SCENE_AABB map) ------------------------
template <int BLOCK = BUILD_BLOCK>
__device__ inline void bounds_tile(const BuildView& v, uint32_t tile, uint32_t* lds) {
  constexpr int W = BLOCK / 64;  // waves; lds holds 6 * W words
  uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
  const uint32_t end = min((tile + 1) * (uint32_t)TRI_TILE, v.n);
  for (uint32_t t = tile * TRI_TILE + threadIdx.x; t < end; t += BLOCK) {
    const float3 c = centroid(v, t);
    const uint32_t o[3] = {f2ord(c.x), f2ord(c.y), f2ord(c.z)};
#pragma unroll
    for (int k = 0; k < 3; ++k) lo[k] = min(lo[k], o[k]), hi[k] = max(hi[k], o[k]);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      lo[k] = min(lo[k], (uint32_t)__shfl_xor((int)lo[k], off, 64));
      hi[k] = max(hi[k], (uint32_t)__shfl_xor((int)hi[k], off, 64));
    }
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) lds[k * W + wave] = lo[k], lds[(3 + k) * W + wave] = hi[k];
  }
  __syncthreads();
  if (threadIdx.x < 6) {  // one device atomic per bound per tile
    const int k = threadIdx.x;
    uint32_t m = lds[k * W];
    for (int w = 1; w < W; ++w) m = k < 3 ? min(m, lds[k * W + w]) : max(m, lds[k * W + w]);
    if (k < 3)
      atomicMin(&v.B.misc[MISC_BOUNDS + k], m);
    else
      atomicMax(&v.B.misc[MISC_BOUNDS + k], m);
  }
  __syncthreads();
}

// ---- phase 2: Morton codes of one tile + the four 8-bit digit histograms ------
template <int BLOCK = BUILD_BLOCK>
__device__ inline void morton_tile(const BuildView& v, uint32_t tile, uint32_t* lds) {
  uint32_t* s_hist = lds;  // SORT_PASSES * 256
  for (int i = threadIdx.x; i < SORT_PASSES * 256; i += BLOCK) s_hist[i] = 0u;
  float lo[3] = {0, 0, 0}, ext[3] = {0, 0, 0};
  if (v.mode == RQ_MORTON_SCENE_AABB) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo[k] = ord2f(v.B.misc[MISC_BOUNDS + k]);
      ext[k] = ord2f(v.B.misc[MISC_BOUNDS + 3 + k]) - lo[k];
    }
  }
  __syncthreads();
  const uint32_t end = min((tile + 1) * (uint32_t)TRI_TILE, v.n);
  for (uint32_t t = tile * TRI_TILE + threadIdx.x; t < end; t += BLOCK) {
    const float3 c = centroid(v, t);
    float q[3] = {c.x, c.y, c.z};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (v.mode == RQ_MORTON_SCENE_AABB)
        q[k] = ext[k] > 0.0f ? (q[k] - lo[k]) / ext[k] : 0.0f;
      else
        q[k] = (q[k] + 30.0f) / 1000.0f;
    }
    const uint32_t code = morton30(q[0], q[1], q[2]);
    v.B.codes[t] = code;
#pragma unroll
    for (int p = 0; p < SORT_PASSES; ++p)
      atomicAdd(&s_hist[p * 256 + ((code >> (8 * p)) & 255u)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < SORT_PASSES * 256; i += BLOCK) {
    const uint32_t c = s_hist[i];
    if (c) atomicAdd(&v.B.misc[MISC_HIST + i], c);
  }
  __syncthreads();
}

// ---- phases 3-6: one stable 8-bit LSD pass over one tile ------------------------
// onesweep: wave w of the tile owns keys [w*64*ITEMS, (w+1)*64*ITEMS), lane l
// its keys i*64 + l; ranking wave by wave, item by item (8 ballots give the
// peers with the same digit) keeps input order among equal digits. Tile digit
// counts go to a decoupled look-back (flag in the top 2 bits of one word,
// agent-scope relaxed atomics, bounded wait), keys and then values are staged
// in one LDS tile buffer so the global writes of each digit are contiguous.
// Every earlier tile of the pass must already be running (ticket order).
template <int PASS, int ITEMS = SORT_ITEMS>
__device__ inline void onesweep_tile(const BuildView& v, uint32_t tile, uint32_t* lds) {
  constexpr int pass = PASS;  // static: buffer selection never indexes the view
  constexpr int TILE = SORT_BLOCK * ITEMS;
  static_assert(BUILD_BLOCK == SORT_BLOCK && SORT_BLOCK == 256, "one thread per digit");
  uint32_t* s_cnt = lds;                        // [SORT_WAVES][256]
  uint32_t* s_stage = s_cnt + SORT_WAVES * 256;  // [TILE] keys, then values
  uint32_t* s_excl = s_stage + TILE;         // [256]
  uint32_t* s_base = s_excl + 256;              // [256]
  uint32_t* s_wsum = s_base + 256;              // [2][SORT_WAVES]

  const uint32_t n = v.n;
  const uint32_t* kin = pass == 0 ? v.B.codes : ((pass & 1) ? v.B.keys_b : v.B.keys_a);
  const uint32_t* vin = pass == 0 ? nullptr : ((pass & 1) ? v.B.vals_b : v.B.vals_a);
  uint32_t* kout = (pass & 1) ? v.B.keys_a : v.B.keys_b;
  uint32_t* vout = (pass & 1) ? v.B.vals_a : v.B.vals_b;
  static_assert(PASS >= 0 && PASS < SORT_PASSES, "30-bit codes: four 8-bit passes");
  uint32_t* status = v.B.status + (size_t)pass * tiles_of(n, TILE) * 256;
  uint32_t* misc = v.B.misc;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int shift = 8 * pass;
  for (int i = tid; i < SORT_WAVES * 256; i += SORT_BLOCK) s_cnt[i] = 0u;
  __syncthreads();
  const uint32_t base = tile * TILE;
  const uint32_t wbase = base + wave * (64 * ITEMS);
  const uint64_t lt = (1ull << lane) - 1ull;

  // Full tiles load the values with the keys (their latency hides behind
  // the ranking); lean tiles, whose register budget is tight, read them only
  // when they are staged, after the keys went out.
  constexpr bool PREFETCH = ITEMS >= 16;
  uint32_t key[ITEMS], rank[ITEMS], val[PREFETCH ? ITEMS : 1];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t g = wbase + i * 64 + lane;
    key[i] = g < n ? el32(kin, g) : 0u;
    if (PREFETCH) val[i] = g < n ? (vin ? el32(vin, g) : g) : 0u;
  }
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const bool ok = wbase + i * 64 + lane < n;
    const uint32_t d = (key[i] >> shift) & 255u;
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bal = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bal : ~bal;
    }
    const uint32_t before = (uint32_t)__popcll(peers & lt);
    const uint32_t c = ok ? s_cnt[wave * 256 + d] : 0u;
    rank[i] = c + before;
    if (ok && before == 0) s_cnt[wave * 256 + d] = c + (uint32_t)__popcll(peers);
  }
  __syncthreads();
  // Only key[] and rank[] stay live across the scans below: hiding the keys
  // and the lane id from the optimiser makes it recompute digits, indices and
  // LDS addresses after the scans instead of holding them in registers.
  uint32_t lane_r = (uint32_t)lane;
  asm volatile("" : "+v"(lane_r));
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) asm volatile("" : "+v"(key[i]));

  // per digit (thread = digit): wave-exclusive prefix and tile total
  const uint32_t d = (uint32_t)tid;
  uint32_t total = 0;
#pragma unroll
  for (int w = 0; w < SORT_WAVES; ++w) {
    const uint32_t c = s_cnt[w * 256 + d];
    s_cnt[w * 256 + d] = total;
    total += c;
  }
  uint32_t* st = status + (size_t)tile * 256 + d;
  st_agent(st, (tile == 0 ? LB_INC : LB_AGG) | total);

  // block-exclusive scans of the tile totals and of the global histogram
  const uint32_t hcount = misc[MISC_HIST + pass * 256 + d];
  const uint32_t inc_t = wave_incl_scan(total, lane);
  const uint32_t inc_h = wave_incl_scan(hcount, lane);
  if (lane == 63) s_wsum[wave] = inc_t, s_wsum[SORT_WAVES + wave] = inc_h;
  __syncthreads();
  uint32_t pre_t = 0, pre_h = 0;
  for (int w = 0; w < wave; ++w) pre_t += s_wsum[w], pre_h += s_wsum[SORT_WAVES + w];
  const uint32_t blk_excl = pre_t + inc_t - total;
  const uint32_t glb_excl = pre_h + inc_h - hcount;

  // decoupled look-back over earlier tiles for this digit, LB_WINDOW status
  // words per round: the loads of a window are in flight together, and they
  // are consumed in order up to the first inclusive prefix (done) or the first
  // word not yet published (polled again from there). Tile 0 publishes an
  // inclusive prefix, so the walk ends at tile 0 at the latest.
  constexpr int LB_WINDOW = 8;
  uint32_t prefix = 0;
  if (tile > 0) {
    int64_t j = (int64_t)tile - 1;
    uint32_t spins = 0;
    for (;;) {
      uint32_t w[LB_WINDOW];
#pragma unroll
      for (int q = 0; q < LB_WINDOW; ++q)
        w[q] = j - q >= 0 ? ld_agent(status + (size_t)(j - q) * 256 + d) : 0u;
      int used = 0;
      bool done = false;
#pragma unroll
      for (int q = 0; q < LB_WINDOW; ++q) {
        if (done || used < q || j - q < 0) break;
        const uint32_t f = w[q] & ~LB_MASK;
        if (f == 0u) break;
        prefix += w[q] & LB_MASK;
        ++used;
        done = f == LB_INC;
      }
      if (done) break;
      j -= used;
      if (used == 0) {
        if (++spins > LOOKBACK_SPIN_LIMIT) {  // bounded wait: report, never hang
          atomicOr(&misc[MISC_ERR], ERR_LOOKBACK_SPIN);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    st_agent(st, LB_INC | (prefix + total));
  }
  s_excl[d] = blk_excl;
  s_base[d] = glb_excl + prefix - blk_excl;
  __syncthreads();

  // keys: rank[i] becomes the tile position, dst[j] the global position of
  // staged slot tid + j * SORT_BLOCK (reused for the values)
  static_assert(TILE % SORT_BLOCK == 0, "whole rounds");
  constexpr int ROUNDS = TILE / SORT_BLOCK;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t dd = (key[i] >> shift) & 255u;
    rank[i] += s_excl[dd] + s_cnt[wave * 256 + dd];
    if (wbase + i * 64 + lane_r < n) {
      if (rank[i] < (uint32_t)TILE)
        s_stage[rank[i]] = key[i];
      else
        atomicOr(&misc[MISC_ERR], ERR_SCATTER_RANGE);
    }
  }
  __syncthreads();
  const uint32_t cnt = min((uint32_t)TILE, n - base);
  uint32_t dst[ROUNDS];
#pragma unroll
  for (int j = 0; j < ROUNDS; ++j) {
    const uint32_t p = tid + j * SORT_BLOCK;
    dst[j] = n;
    if (p < cnt) {
      const uint32_t k = s_stage[p];
      const uint32_t g = s_base[(k >> shift) & 255u] + p;
      if (g < n) {
        el32(kout, g) = k;
        dst[j] = g;
      } else {
        atomicOr(&misc[MISC_ERR], ERR_SCATTER_RANGE);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t g = wbase + i * 64 + lane_r;
    if (g < n && rank[i] < (uint32_t)TILE)
      s_stage[rank[i]] = PREFETCH ? val[i] : (vin ? el32(vin, g) : g);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < ROUNDS; ++j)
    if (dst[j] < n) el32(vout, dst[j]) = s_stage[tid + j * SORT_BLOCK];
  __syncthreads();
}

// Child links of the internal nodes as Karras emits them, (left, right) per
// node: kept in the slot_h words (2(n-1), free until a treelet pass uses them
// as scratch) until the refit writes each node record whole.
__device__ inline int2* links_of(const BuildBuffers& B) { return reinterpret_cast<int2*>(B.slot_h); }

// ---- phase 7: Karras 2012 emission for a tile of internal nodes ----------------
// Internal node i in [0, n-2] from the sorted codes; equal codes fall back to
// the index (augmented key code<<32 | i).
//
// A node's searches read the keys around it, a chain of dependent loads (the
// direction, the doubling bound, the binary searches for the far end and the
// split); most nodes span a few leaves. A unit first stages the keys of its
// nodes and KARRAS_HALO on each side in LDS, and a search reads a key from
// there when it lies in that window, from global memory otherwise.
constexpr int KARRAS_HALO = 256;
typedef __attribute__((address_space(3))) const uint32_t* LdsWords;  // an LDS array
struct KeyWindow {
  const uint32_t* __restrict__ keys;  // all n sorted keys (global)
  LdsWords win;                       // keys[lo, hi), read with LDS loads
  int64_t lo, hi;
  // The LDS word is read at a clamped index whatever j, and replaced by the
  // global one only outside the window: with one read per branch the
  // compiler merges them into a generic (flat) load, which waits on both the
  // LDS and the vector memory counters.
  __device__ uint32_t operator()(int64_t j) const {
    uint32_t k = win[min(max(j - lo, (int64_t)0), hi - lo - 1)];
    if (j < lo || j >= hi) k = keys[j];
    return k;
  }
};
// Stages keys [lo, hi) of the n keys into win (the block's threads); the
// caller synchronises before the searches read it.
__device__ inline KeyWindow stage_keys(const uint32_t* __restrict__ keys, int64_t n, int64_t lo,
                                       int64_t hi, uint32_t* win) {
  lo = max(lo, (int64_t)0), hi = min(hi, n);
  for (int64_t j = lo + threadIdx.x; j < hi; j += blockDim.x) win[j - lo] = keys[j];
  return KeyWindow{keys, (LdsWords)win, lo, hi};
}

template <class K>
__device__ inline int kdelta(const K& k, int64_t n, int64_t i, uint32_t ki, int64_t j) {
  if (j < 0 || j >= n) return -1;
  const uint32_t kj = k(j);
  return ki != kj ? __clz(ki ^ kj) : 32 + __clz((uint32_t)i ^ (uint32_t)j);
}

template <class K>
__device__ inline void karras_node(const BuildView& v, const K& keys, int64_t i) {
  const int64_t N = v.n;
  const uint32_t ki = keys(i);
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
  // the child links go to a compact array (consecutive nodes, whole lines);
  // the refit writes them into the node records with the boxes
  links_of(v.B)[i] = make_int2(left, right);
  const int64_t ul = left >= 0 ? left : (N - 1) + ~left;
  const int64_t ur = right >= 0 ? right : (N - 1) + ~right;
  v.B.flags[i] = 0u;
  v.B.range[i] = make_uint2((uint32_t)min(i, j), (uint32_t)max(i, j));
  if (g < 0 || g + 1 > N - 1 || ul >= 2 * N - 1 || ur >= 2 * N - 1) {
    atomicOr(&v.B.misc[MISC_ERR], ERR_TOPOLOGY);
    return;
  }
  v.B.parent[ul] = (int32_t)(i << 1);
  v.B.parent[ur] = (int32_t)((i << 1) | 1);
}

// LDS words of a Karras unit of TILE nodes: its keys and the halo.
template <int TILE>
constexpr int karras_lds_words() {
  return TILE + 2 * KARRAS_HALO;
}

// TILE internal nodes per unit: the batch kernel queues NODE_TILE, the
// standalone kernel one node per thread (each node is an independent chain of
// dependent key loads, so more units hide more latency). lds holds
// karras_lds_words<TILE>() words.
template <int TILE = NODE_TILE>
__device__ inline void karras_tile(const BuildView& v, uint32_t tile, uint32_t* lds) {
  if (v.n < 2) return;
  const int64_t begin = (int64_t)tile * TILE;
  const int64_t end = min(begin + TILE, (int64_t)v.n - 1);
  const KeyWindow w = stage_keys(v.B.keys_a, v.n, begin - KARRAS_HALO, end + 1 + KARRAS_HALO, lds);
  __syncthreads();
  for (int64_t i = begin + threadIdx.x; i < end; i += blockDim.x) karras_node(v, w, i);
}

// ---- phases 8-9: bottom-up refit ---------------------------------------------------
// Helpers of the chunk-union refit below: a node's box goes into its parent's
// child slot (put_slot); the root box and height into misc (put_root).
__device__ inline void put_slot(NodeRec* nodes, uint32_t node, uint32_t side, const float* lo,
                                const float* hi) {
  float* slo = side ? &nodes[node].lo1.x : &nodes[node].lo0.x;
  float* shi = side ? &nodes[node].hi1.x : &nodes[node].hi0.x;
  slo[0] = lo[0], slo[1] = lo[1], slo[2] = lo[2];
  shi[0] = hi[0], shi[1] = hi[1], shi[2] = hi[2];
}
__device__ inline void put_root(uint32_t* misc, const float* lo, const float* hi, uint32_t h) {
  float* root = reinterpret_cast<float*>(misc + MISC_ROOT);
  for (int q = 0; q < 3; ++q) root[q] = lo[q], root[3 + q] = hi[q];
  misc[MISC_HEIGHT] = h;
}

// Leaf record of sorted position k and its box (the leaf's slot value).
__device__ inline void leaf_record(const BuildView& v, uint32_t k, TriRec& tr, float* lo,
                                   float* hi) {
  const uint32_t p = v.B.vals_a[k];
  if (v.sphere) {
    // sphere leaf: centre and radius, box centre -/+ radius (main.cpp:687-689);
    // the record keeps r and r*r (the reference's radius2)
    const float cx = v.vx[p], cy = v.vy[p], cz = v.vz[p], r = v.vr[p];
    tr.v0 = make_float4(cx, cy, cz, __int_as_float((int32_t)p));
    tr.e1 = make_float4(r, r * r, 0.0f, 0.0f);
    tr.e2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    lo[0] = (cx - r) + 0.0f, lo[1] = (cy - r) + 0.0f, lo[2] = (cz - r) + 0.0f;
    hi[0] = (cx + r) + 0.0f, hi[1] = (cy + r) + 0.0f, hi[2] = (cz + r) + 0.0f;
  } else {
    const uint32_t a = v.idx[3 * p], b = v.idx[3 * p + 1], c = v.idx[3 * p + 2];
    const float3 v0 = make_float3(v.vx[a], v.vy[a], v.vz[a]);
    const float3 v1 = make_float3(v.vx[b], v.vy[b], v.vz[b]);
    const float3 v2 = make_float3(v.vx[c], v.vy[c], v.vz[c]);
    tr.v0 = make_float4(v0.x, v0.y, v0.z, __int_as_float((int32_t)p));
    tr.e1 = make_float4(v1.x - v0.x, v1.y - v0.y, v1.z - v0.z, 0.0f);
    tr.e2 = make_float4(v2.x - v0.x, v2.y - v0.y, v2.z - v0.z, 0.0f);
    lo[0] = fminf(fminf(v0.x, v1.x), v2.x) + 0.0f, lo[1] = fminf(fminf(v0.y, v1.y), v2.y) + 0.0f;
    lo[2] = fminf(fminf(v0.z, v1.z), v2.z) + 0.0f;
    hi[0] = fmaxf(fmaxf(v0.x, v1.x), v2.x) + 0.0f, hi[1] = fmaxf(fmaxf(v0.y, v1.y), v2.y) + 0.0f;
    hi[2] = fmaxf(fmaxf(v0.z, v1.z), v2.z) + 0.0f;
  }
}

// ---- 4-wide collapse ------------------------------------------------------------
// The wide node of binary internal node i holds, in binary order, for each
// child c of i: c's box if c is a leaf (one slot, the next one unused), else
// the boxes of c's two children (two slots). Boxes are copied exactly, so a
// slab test against a wide slot is the test the binary traversal makes
// against the same subtree. Every internal node gets its wide node (only the
// root and even-depth nodes are reached from the root).
struct WideSlots {
  float lo[4][3], hi[4][3];
  int32_t link[4];
};
__device__ inline void wide_empty(WideSlots& w, int k) {
  w.lo[k][0] = w.lo[k][1] = w.lo[k][2] = INFINITY;
  w.hi[k][0] = w.hi[k][1] = w.hi[k][2] = -INFINITY;
  w.link[k] = 0;
}
__device__ inline void wide_store(WideNode* W, const WideSlots& w) {
  W->lox = make_float4(w.lo[0][0], w.lo[1][0], w.lo[2][0], w.lo[3][0]);
  W->hix = make_float4(w.hi[0][0], w.hi[1][0], w.hi[2][0], w.hi[3][0]);
  W->loy = make_float4(w.lo[0][1], w.lo[1][1], w.lo[2][1], w.lo[3][1]);
  W->hiy = make_float4(w.hi[0][1], w.hi[1][1], w.hi[2][1], w.hi[3][1]);
  W->loz = make_float4(w.lo[0][2], w.lo[1][2], w.lo[2][2], w.lo[3][2]);
  W->hiz = make_float4(w.hi[0][2], w.hi[1][2], w.hi[2][2], w.hi[3][2]);
  W->link = make_int4(w.link[0], w.link[1], w.link[2], w.link[3]);
  W->pad = make_int4(0, 0, 0, 0);
}
// From the finished binary node records in global memory.
__device__ inline void widen_global(const BuildBuffers& B, uint32_t i) {
  const NodeRec N = B.nodes[i];
  WideSlots w;
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const int32_t c = __float_as_int(side ? N.hi0.w : N.lo0.w);
    const float4 clo = side ? N.lo1 : N.lo0, chi = side ? N.hi1 : N.hi0;
    if (c < 0) {
      w.lo[2 * side][0] = clo.x, w.lo[2 * side][1] = clo.y, w.lo[2 * side][2] = clo.z;
      w.hi[2 * side][0] = chi.x, w.hi[2 * side][1] = chi.y, w.hi[2 * side][2] = chi.z;
      w.link[2 * side] = c;
      wide_empty(w, 2 * side + 1);
    } else {
      const NodeRec C = B.nodes[c];
      w.lo[2 * side][0] = C.lo0.x, w.lo[2 * side][1] = C.lo0.y, w.lo[2 * side][2] = C.lo0.z;
      w.hi[2 * side][0] = C.hi0.x, w.hi[2 * side][1] = C.hi0.y, w.hi[2 * side][2] = C.hi0.z;
      w.lo[2 * side + 1][0] = C.lo1.x, w.lo[2 * side + 1][1] = C.lo1.y;
      w.lo[2 * side + 1][2] = C.lo1.z;
      w.hi[2 * side + 1][0] = C.hi1.x, w.hi[2 * side + 1][1] = C.hi1.y;
      w.hi[2 * side + 1][2] = C.hi1.z;
      w.link[2 * side] = __float_as_int(C.lo0.w), w.link[2 * side + 1] = __float_as_int(C.hi0.w);
    }
  }
  wide_store(B.wide + i, w);
}

// Tree height on demand (the chunk-union refit leaves it open): the depth of
// node i's leaf children, d + 1 for a node at depth d with a leaf child (walk
// up the parent links, at most 64 steps), else 0; the height is the maximum.
__device__ inline uint32_t leaf_depth(const BuildBuffers& B, uint32_t n, uint32_t i) {
  const NodeRec N = B.nodes[i];
  if (__float_as_int(N.lo0.w) >= 0 && __float_as_int(N.hi0.w) >= 0) return 0u;
  uint32_t d = 0, cur = i;
  while (cur != 0u) {
    const int32_t pe = B.parent[cur];
    cur = (uint32_t)pe >> 1;
    if (pe < 0 || cur >= n - 1 || ++d > 64) {
      atomicOr(&B.misc[MISC_ERR], ERR_WIDEN_DEPTH);
      return 0u;
    }
  }
  return d + 1u;
}

// ---- chunk-union refit ------------------------------------------------------------
// Pass 1, one block of CHUNK threads per chunk of CHUNK leaves (the standalone
// build: REFIT_CHUNK = 1024; the batch kernel: its 256-thread blocks):
// each thread writes its leaf record, the block scans the leaf boxes into
// per-leaf prefix and suffix unions of the chunk (and the chunk's union), and
// each leaf climbs while the node's leaf range lies inside the chunk: the
// second of a node's two arrivals (an arrival counter) unions the sibling's
// box and climbs on. Such a node's index lies in the chunk too (a Karras
// node i is an end of its own range), so its counter and child boxes live in
// LDS. A node spanning chunks only receives its child's box. Pass 2 gives
// every spanning node [a, b] its box directly: suffix(a) | chunks strictly
// between | prefix(b) (min / max are order-independent, so this is the box a
// climb would build). The height is left to the on-demand depth walk.
__device__ inline void box_union(float* b, const float* o) {
  b[0] = fminf(b[0], o[0]), b[1] = fminf(b[1], o[1]), b[2] = fminf(b[2], o[2]);
  b[3] = fmaxf(b[3], o[3]), b[4] = fmaxf(b[4], o[4]), b[5] = fmaxf(b[5], o[5]);
}
__device__ inline void box_store(float* dst, const float* b) {
  float2* d = reinterpret_cast<float2*>(dst);  // 24-byte records, 8-byte aligned
  d[0] = make_float2(b[0], b[1]), d[1] = make_float2(b[2], b[3]), d[2] = make_float2(b[4], b[5]);
}
__device__ inline void box_load(const float* src, float* b) {
  const float2* q = reinterpret_cast<const float2*>(src);
  const float2 a = q[0], c = q[1], e = q[2];
  b[0] = a.x, b[1] = a.y, b[2] = c.x, b[3] = c.y, b[4] = e.x, b[5] = e.y;
}

// LDS: CHUNK arrival counters, 2 * CHUNK / 64 wave unions (6 floats each), the
// child boxes of the chunk's nodes ([node - chunk start][side][6]), and the
// parent links of the chunk's leaves and internal nodes with the internal
// nodes' leaf ranges (loaded once, so the climb reads no global memory).
template <int CHUNK>
constexpr int refit_chunk_lds_words() {
  return CHUNK + 2 * (CHUNK / 64) * 6 + CHUNK * 12 + CHUNK * 6;
}
constexpr int REFIT_CHUNK_LDS_WORDS = refit_chunk_lds_words<REFIT_CHUNK>();

// Whether internal node i's leaf range crosses a boundary of CHUNK-leaf chunks.
template <int CHUNK>
__device__ inline bool spans_chunks(const BuildBuffers& B, uint32_t i) {
  const uint2 rg = B.range[i];
  return rg.x / CHUNK != rg.y / CHUNK;
}

template <int CHUNK = REFIT_CHUNK>
__device__ inline void refit_chunk_tile(const BuildView& v, uint32_t chunk, uint32_t* lds) {
  static_assert(CHUNK % 64 == 0 && CHUNK <= 1024, "one leaf per thread, whole waves");
  static_assert(CHUNK >= REFIT_CHUNK_MIN, "the chunk unions are allocated for REFIT_CHUNK_MIN");
  uint32_t* s_flag = lds;
  float* s_wpre = reinterpret_cast<float*>(lds + CHUNK);  // [wave][6] exclusive prefix
  float* s_wsuf = s_wpre + (CHUNK / 64) * 6;               // [wave][6] exclusive suffix
  float* s_box = s_wsuf + (CHUNK / 64) * 6;                // [node][side][6] child boxes
  int32_t* s_pleaf = reinterpret_cast<int32_t*>(s_box + CHUNK * 12);  // parent of leaf
  int32_t* s_pint = s_pleaf + CHUNK;                                   // parent of node
  uint2* s_rng = reinterpret_cast<uint2*>(s_pint + CHUNK);             // range of node
  int2* s_kid = reinterpret_cast<int2*>(s_rng + CHUNK);                // children of node
  const BuildBuffers& B = v.B;
  const uint32_t n = v.n;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint32_t chunk_lo = chunk * CHUNK;
  const uint32_t chunk_hi = min(chunk_lo + CHUNK, n) - 1;
  const uint32_t k = chunk_lo + tid;
  const bool valid = k <= chunk_hi;
  s_flag[tid] = 0u;
  // the chunk's parent links and ranges (internal node ids chunk_lo + tid)
  if (valid && n > 1) s_pleaf[tid] = B.parent[(n - 1) + k];
  if (k < n - 1) {
    s_pint[tid] = k == 0u ? -1 : B.parent[k];  // the root has no parent
    s_rng[tid] = B.range[k];
    s_kid[tid] = links_of(B)[k];
  }
  float box[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  if (valid) {
    TriRec tr;
    leaf_record(v, k, tr, box, box + 3);
    B.tris[k] = tr;
  }
  // wave-level inclusive prefix (pre) and suffix (suf) unions
  float pre[6], suf[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) pre[q] = suf[q] = box[q];
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    float up[6], dn[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) up[q] = __shfl_up(pre[q], d, 64), dn[q] = __shfl_down(suf[q], d, 64);
    if (lane >= (uint32_t)d) box_union(pre, up);
    if (lane + d < 64u) box_union(suf, dn);
  }
  // wave unions -> exclusive prefix / suffix over the waves of the block
  if (lane == 63) {
#pragma unroll
    for (int q = 0; q < 6; ++q) s_wpre[wave * 6 + q] = pre[q];  // the wave's union
  }
  __syncthreads();
  float ep[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  float es[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  if (tid < CHUNK / 64) {
    for (uint32_t w = 0; w < CHUNK / 64; ++w) {
      if (w < tid) box_union(ep, s_wpre + w * 6);
      if (w > tid) box_union(es, s_wpre + w * 6);
    }
  }
  __syncthreads();  // every wave union read before it is overwritten
  if (tid < CHUNK / 64) {
#pragma unroll
    for (int q = 0; q < 6; ++q) s_wpre[tid * 6 + q] = ep[q], s_wsuf[tid * 6 + q] = es[q];
  }
  __syncthreads();
  box_union(pre, s_wpre + wave * 6);
  box_union(suf, s_wsuf + wave * 6);
  if (valid) {
    box_store(B.pre + 6ull * k, pre);
    box_store(B.suf + 6ull * k, suf);
  }
  if (tid == 0) box_store(B.cbox + 6ull * chunk, suf);  // leaf 0's suffix: the chunk's union
  if (valid && n == 1) put_root(B.misc, box, box + 3, 0u);
  float* lo = box;
  float* hi = box + 3;
  // cur: this thread's leaf (level 0), then the internal node it completed,
  // which lies in the chunk (its index is in its own range)
  uint32_t cur = k;
  for (int level = 0; valid && n > 1; ++level) {
    const int32_t pe = level == 0 ? s_pleaf[tid] : s_pint[cur - chunk_lo];
    const uint32_t node = (uint32_t)pe >> 1, side = (uint32_t)pe & 1u;
    if (pe < 0 || node >= n - 1 || level > 64) {
      atomicOr(&B.misc[MISC_ERR], ERR_REFIT_DEPTH);
      break;
    }
    // a node outside the chunk's index span cannot have its range inside it;
    // a spanning parent takes the box in its record's slot (pass 2 finishes
    // it), a parent inside the chunk in LDS (its record is written whole below)
    if (node < chunk_lo || node > chunk_hi || s_rng[node - chunk_lo].x < chunk_lo ||
        s_rng[node - chunk_lo].y > chunk_hi) {
      put_slot(B.nodes, node, side, lo, hi);
      break;
    }
    float* mine = s_box + (node - chunk_lo) * 12;
#pragma unroll
    for (int q = 0; q < 6; ++q) mine[6 * side + q] = box[q];
    // Everything another thread of the climb reads is in LDS (the sibling's
    // box), so the arrival counter orders LDS only: the node-record stores
    // to global memory are read by later kernels and need not complete
    // before the count (an all-memory release would wait for every
    // outstanding store at each level).
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    const uint32_t old = __hip_atomic_fetch_add(&s_flag[node - chunk_lo], 1u, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    if (old == 0u) break;
    box_union(box, mine + 6 * (side ^ 1u));  // the sibling's box, stored before its arrival
    if (node == 0u) {  // the whole tree lies in this chunk
      float* root = reinterpret_cast<float*>(B.misc + MISC_ROOT);
      for (int q = 0; q < 3; ++q) root[q] = lo[q], root[3 + q] = hi[q];
      break;
    }
    cur = node;
  }
  // Node records and wide records of the chunk's completed nodes (both
  // children arrived, so every box below them is final in LDS): thread t
  // writes node chunk start + t, so the 64- and 128-byte records go out in
  // consecutive lines. An internal child of such a node lies in the chunk too.
  __syncthreads();
  if (k >= n - 1 || s_flag[tid] != 2u) return;
  const float* mine = s_box + tid * 12;
  const int2 kid = s_kid[tid];
  NodeRec nr;
  nr.lo0 = make_float4(mine[0], mine[1], mine[2], __int_as_float(kid.x));
  nr.hi0 = make_float4(mine[3], mine[4], mine[5], __int_as_float(kid.y));
  nr.lo1 = make_float4(mine[6], mine[7], mine[8], 0.0f);
  nr.hi1 = make_float4(mine[9], mine[10], mine[11], 0.0f);
  B.nodes[k] = nr;
  if (!B.wide) return;
  WideSlots ws;
#pragma unroll
  for (int sd = 0; sd < 2; ++sd) {
    const int32_t c = sd ? kid.y : kid.x;
    if (c < 0) {
      const float* cb = mine + 6 * sd;
      ws.lo[2 * sd][0] = cb[0], ws.lo[2 * sd][1] = cb[1], ws.lo[2 * sd][2] = cb[2];
      ws.hi[2 * sd][0] = cb[3], ws.hi[2 * sd][1] = cb[4], ws.hi[2 * sd][2] = cb[5];
      ws.link[2 * sd] = c;
      wide_empty(ws, 2 * sd + 1);
    } else {
      const float* gb = s_box + ((uint32_t)c - chunk_lo) * 12;
      const int2 gk = s_kid[(uint32_t)c - chunk_lo];
#pragma unroll
      for (int g = 0; g < 2; ++g) {
        ws.lo[2 * sd + g][0] = gb[6 * g], ws.lo[2 * sd + g][1] = gb[6 * g + 1];
        ws.lo[2 * sd + g][2] = gb[6 * g + 2], ws.hi[2 * sd + g][0] = gb[6 * g + 3];
        ws.hi[2 * sd + g][1] = gb[6 * g + 4], ws.hi[2 * sd + g][2] = gb[6 * g + 5];
      }
      ws.link[2 * sd] = gk.x, ws.link[2 * sd + 1] = gk.y;
    }
  }
  wide_store(B.wide + k, ws);
}

// Pass 2: the lanes of a wave each take internal node i; the wave then works
// through its spanning nodes one at a time, all 64 lanes striding over the
// chunks strictly inside the span (a node near the root spans them all).
// Every lane must reach this call (the wave shuffles).
template <int CHUNK = REFIT_CHUNK>
__device__ inline void refit_span_wave(const BuildBuffers& B, uint32_t n, uint32_t i) {
  const uint32_t lane = threadIdx.x & 63u;
  uint2 rg = make_uint2(0u, 0u);
  bool span = false;
  if (i < n - 1) {
    rg = B.range[i];
    span = rg.x / CHUNK != rg.y / CHUNK;
  }
  uint64_t m = __ballot(span);
  while (m) {
    const int src = __ffsll((unsigned long long)m) - 1;
    m &= m - 1ull;
    const uint32_t ca = (uint32_t)__shfl((int)(rg.x / CHUNK), src, 64);
    const uint32_t cb = (uint32_t)__shfl((int)(rg.y / CHUNK), src, 64);
    float box[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY}, o[6];
    for (uint32_t c = ca + 1 + lane; c < cb; c += 64) {
      box_load(B.cbox + 6ull * c, o);
      box_union(box, o);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
#pragma unroll
      for (int q = 0; q < 6; ++q) o[q] = __shfl_xor(box[q], d, 64);
      box_union(box, o);
    }
    if ((int)lane != src) continue;
    // the spanning node's own links (its slots were filled by its children)
    const int2 kid = links_of(B)[i];
    reinterpret_cast<int32_t*>(&B.nodes[i].lo0)[3] = kid.x;
    reinterpret_cast<int32_t*>(&B.nodes[i].hi0)[3] = kid.y;
    box_load(B.suf + 6ull * rg.x, o);
    box_union(box, o);
    box_load(B.pre + 6ull * rg.y, o);
    box_union(box, o);
    if (i == 0u) {
      float* root = reinterpret_cast<float*>(B.misc + MISC_ROOT);
      for (int q = 0; q < 6; ++q) root[q] = box[q];
      continue;
    }
    const int32_t pe = B.parent[i];
    const uint32_t node = (uint32_t)pe >> 1, side = (uint32_t)pe & 1u;
    if (pe < 0 || node >= n - 1) {
      atomicOr(&B.misc[MISC_ERR], ERR_REFIT_DEPTH);
      continue;
    }
    put_slot(B.nodes, node, side, box, box + 3);
  }
}


}  // namespace rq
