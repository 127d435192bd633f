// Device side of the LBVH traversal: ray state, slab and triangle tests, and
// the persistent per-wave loop shared by the standalone query kernels
// (trace_kernels.hip) and the batch kernel (batch_kernels.hip).
//
// Replaces the per-ray recursion of boxIntersect (accelerators.h:668-690)
// and the candidate loop of castRay (main.cpp:343-358): instead of collecting
// every candidate and then testing them, one lane owns one ray and walks the
// tree with a stack, nearer child first, culling boxes against the best hit
// so far. The result is the (t, primitive id) lexicographic minimum, i.e. the
// NONE loop of main.cpp:373-387 (lowest index wins ties).
#pragma once
#include "rq_internal.h"

namespace rq {

// Where a wave takes its slices of the ray stream from. The standalone kernels
// number the slices (a.chunk rays each for the first a.big_slices, then
// a.chunk_small): wave w starts on slice w (no atomic), and takes slice
// waves + k from the k-th add to a 64-bit counter that is never reset
// (k = counter add - base). When the waves' first slices cover the launch no
// wave touches the counter; otherwise each wave makes exactly one failing add
// and every later slice is one add, so the launch moves the counter by
// exactly its number of slices, which the host adds to the base of the
// slot's next launch. The batch kernel counts rays from 0 in LDS.
struct RayStream {
  uint32_t* ctr32 = nullptr;            // counter of rays from 0 (batch, LDS), or
  unsigned long long* ctr64 = nullptr;  // the slot's monotonic slice counter
  uint64_t base = 0;                    // ctr64 at launch
  uint32_t wave = 0, waves = 0;         // this wave's index, waves in the grid
};
// The launch's ray count and slice sizes, in scalar registers: from the
// kernel arguments, or computed here from a count read from device memory
// (TraceArgs::nrays_dev).
struct Schedule {
  uint32_t n, chunk, small, big;
};
__device__ inline Schedule schedule_of(const TraceArgs& a, uint32_t waves) {
  if (!a.nrays_dev) return Schedule{a.nrays, a.chunk, a.chunk_small, a.big_slices};
  const uint32_t n = min(a.nrays, __builtin_amdgcn_readfirstlane(*a.nrays_dev));
  const SliceSchedule q = slice_schedule(n, trace_chunk(n, waves, a.chunk));
  return Schedule{n, q.chunk, q.small, q.big};
}
__device__ inline uint32_t slice_begin(const Schedule& a, uint64_t s) {
  return s <= a.big ? (uint32_t)min(s * a.chunk, (uint64_t)a.n)
                    : (uint32_t)min((uint64_t)a.big * a.chunk + (s - a.big) * a.small,
                                    (uint64_t)a.n);
}

// A wave takes its slices from the RayStream unless the caller gives a slice
// source: Src::next(b, e) returns the wave's next slice [b, e) of a.rays, or
// false when the wave is to finish its rays and return. It is called with
// every lane of the wave active and its result is wave-uniform.
struct NoSlices {
  static constexpr bool kActive = false;
  __device__ bool next(uint32_t&, uint32_t&) { return false; }
};

struct RayState {
  float ox, oy, oz, dx, dy, dz;
  float ix, iy, iz;  // 1/d
  bool sx, sy, sz;   // sign bit of 1/d: the near plane of that axis is hi
  float tmin, tmax;
};

// Relative slack on the far bound: a triangle's rounded Möller–Trumbore t can
// fall a few ulps outside the exact box of its vertices (hits on edges and
// vertices); culling only beyond far * (1 + 2^-15) keeps every box that holds
// the brute-force closest hit.
constexpr float SLAB_SLACK = 1.000030517578125f;
// Sphere scenes: the reference's sphere test computes d2 = l.l - tca^2, whose
// cancellation at a distance L moves a grazing hit's t by up to about
// L * 2^-11 (and can accept a ray just outside the box); culling at the best
// hit with a 2^-9 relative slack keeps every box whose sphere's computed t
// can still beat the best hit.
constexpr float SPHERE_SLACK = 1.001953125f;
// Wide walk: the near / far rows are loaded at byte offsets node * 128 + row
// held in 32 bits (a scalar base plus one vector offset; the row is scalar
// when the wave's octant is uniform, per lane otherwise), so the node count
// must stay below 2^25; larger trees select the rows per value.
constexpr uint32_t WIDE_UNIFORM_MAX_NODES = 1u << 25;

// Slab test (same operations and order as the oracle): near/far planes by
// the sign of 1/d, t = (plane - o) * (1/d). A zero direction component gives
// +-inf, or NaN with the origin on the plane, which the NaN-ignoring max
// (near) / min (far) read as "inside". Hit when near <= far * SLAB_SLACK.
__device__ inline bool slab(const float4& lo, const float4& hi, const RayState& r, float thi,
                            float slack, float& tn) {
  const float nx = ((r.sx ? hi.x : lo.x) - r.ox) * r.ix, fx = ((r.sx ? lo.x : hi.x) - r.ox) * r.ix;
  const float ny = ((r.sy ? hi.y : lo.y) - r.oy) * r.iy, fy = ((r.sy ? lo.y : hi.y) - r.oy) * r.iy;
  const float nz = ((r.sz ? hi.z : lo.z) - r.oz) * r.iz, fz = ((r.sz ? lo.z : hi.z) - r.oz) * r.iz;
  tn = fmaxf(fmaxf(nx, ny), fmaxf(nz, r.tmin));
  const float tf = fminf(fminf(fx, fy), fminf(fz, thi));
  return tn <= tf * slack;
}

// Slot test of a wide node on pre-selected near / far planes: the same
// operations as slab(), so a slot culls exactly where the binary walk culls
// the same box.
__device__ inline bool slab_w(float nx, float ny, float nz, float fx, float fy, float fz,
                              const RayState& r, float thi, float slack, float& tn) {
  const float an = (nx - r.ox) * r.ix, bn = (ny - r.oy) * r.iy, cn = (nz - r.oz) * r.iz;
  const float af = (fx - r.ox) * r.ix, bf = (fy - r.oy) * r.iy, cf = (fz - r.oz) * r.iz;
  tn = fmaxf(fmaxf(an, bn), fmaxf(cn, r.tmin));
  const float tf = fminf(fminf(af, bf), fminf(cf, thi));
  return tn <= tf * slack;
}

__device__ inline void cas_pair(uint32_t& ka, int32_t& ca, uint32_t& kb, int32_t& cb) {
  const bool sw = kb < ka;
  const uint32_t k0 = sw ? kb : ka, k1 = sw ? ka : kb;
  const int32_t c0 = sw ? cb : ca, c1 = sw ? ca : cb;
  ka = k0, kb = k1, ca = c0, cb = c1;
}
// The same exchange for hit / miss keys (any hit): a miss before a hit swaps.
// The flags stay lane masks, so only the two link selects are vector work.
__device__ inline void cas_hit(bool& ha, int32_t& ca, bool& hb, int32_t& cb) {
  const bool sw = !ha && hb;
  const int32_t c0 = sw ? cb : ca, c1 = sw ? ca : cb;
  const bool h0 = ha || hb, h1 = ha && hb;
  ha = h0, hb = h1, ca = c0, cb = c1;
}

// Möller–Trumbore, main.cpp:137-162 (v_0 read as v0), t restricted to
// (tmin, tmax). Dot products (x*x' + y*y') + z*z', no contraction.
__device__ inline bool mt(const TriRec& T, const RayState& r, float& t, float& u, float& v) {
  const float px = r.dy * T.e2.z - r.dz * T.e2.y;
  const float py = r.dz * T.e2.x - r.dx * T.e2.z;
  const float pz = r.dx * T.e2.y - r.dy * T.e2.x;
  const float det = (T.e1.x * px + T.e1.y * py) + T.e1.z * pz;
  if (fabsf(det) < 1e-6f) return false;
  const float inv = 1.0f / det;
  const float tx = r.ox - T.v0.x, ty = r.oy - T.v0.y, tz = r.oz - T.v0.z;
  const float uu = ((tx * px + ty * py) + tz * pz) * inv;
  if (uu < 0.0f || uu > 1.0f) return false;
  const float qx = ty * T.e1.z - tz * T.e1.y;
  const float qy = tz * T.e1.x - tx * T.e1.z;
  const float qz = tx * T.e1.y - ty * T.e1.x;
  const float vv = ((r.dx * qx + r.dy * qy) + r.dz * qz) * inv;
  if (vv < 0.0f || uu + vv > 1.0f) return false;
  const float tt = ((T.e2.x * qx + T.e2.y * qy) + T.e2.z * qz) * inv;
  if (!(tt > r.tmin && tt < r.tmax)) return false;
  t = tt, u = uu, v = vv;
  return true;
}

// Geometric ray-sphere test of the reference (accelerators.h:79-92): a centre
// behind the origin (tca < 0) or a miss distance beyond the radius rejects;
// otherwise t0 = tca - thc, t1 = tca + thc, and castRay takes t1 when t0 < 0
// (main.cpp:349-353). The reference accepts that t whatever its value (it is
// >= 0 once tca >= 0); here it must lie in [tmin, tmax).
__device__ inline bool sphere_hit(const TriRec& S, const RayState& r, float& t) {
  const float lx = S.v0.x - r.ox, ly = S.v0.y - r.oy, lz = S.v0.z - r.oz;
  const float tca = (lx * r.dx + ly * r.dy) + lz * r.dz;
  if (tca < 0.0f) return false;
  const float d2 = ((lx * lx + ly * ly) + lz * lz) - tca * tca;
  const float r2 = S.e1.y;
  if (d2 > r2) return false;
  const float thc = sqrtf(r2 - d2);
  const float t0 = tca - thc, t1 = tca + thc;
  const float tt = t0 < 0.0f ? t1 : t0;
  if (!(tt >= r.tmin && tt < r.tmax)) return false;
  t = tt;
  return true;
}

// Persistent traversal loop of one wave. BLOCK threads per block; DEPTH stack entries
// per lane in LDS ([depth][lane], conflict-free), deeper entries in the global
// overflow area ([depth][thread]); TSTACK: entries carry their entry distance
// and a closest-hit pop skips entries now beyond the best hit (the box test
// redone against the current best). Leaves reached from a node are tested in
// the same loop iteration; the order of node visits and triangle tests is the
// canonical one unless TSTACK skips an entry. Rays are taken a.chunk at a time
// from *ctr (global or LDS), or a slice at a time from *src; the loop returns
// when the stream has no rays left for the wave and its lanes are idle.
// SPHERES: the leaves may be spheres (a.sphere); false compiles triangles only.
// WIDE: walk the 4-wide nodes (a.wide) instead of the binary ones: the slots
// a ray's box test accepts are visited nearest first (closest hit; any hit in
// slot order), the others pushed farthest first. The same boxes are tested
// with the same slack, so culling is as conservative as the binary walk and
// the result (the lexicographic minimum) is the same; visit order differs.
// The wide records as a raw buffer resource (scalar registers): row loads at
// a 32-bit offset from it, with the row's fixed part as a scalar or immediate
// offset, need no 64-bit address arithmetic. Dword 3 selects the 32-bit data
// format that gfx9 buffer instructions require.
__device__ inline __amdgpu_buffer_rsrc_t wide_rsrc(const TraceArgs& a) {
  const uint64_t bytes = (uint64_t)a.nprims * sizeof(WideNode);
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<WideNode*>(a.wide), (short)0,
                                           (int)(uint32_t)min(bytes, (uint64_t)0xFFFFFFFFu),
                                           0x00020000);
}
__device__ inline float4 row_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0);
  return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                     __uint_as_float(v[3]));
}
__device__ inline int4 link_load(__amdgpu_buffer_rsrc_t r, uint32_t voff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, 96, 0);
  return make_int4((int)v[0], (int)v[1], (int)v[2], (int)v[3]);
}

// ROWS, how the wide walk addresses a node's rows: a tree of
// WIDE_UNIFORM_MAX_NODES nodes or more has rows beyond a 32-bit byte offset,
// addressed in 64 bits with the near / far rows selected per value (ROWS_BIG);
// smaller trees address them by 32-bit offsets from the array's base
// (ROWS_SMALL); ROWS_ANY decides per launch inside one code path.
constexpr int ROWS_SMALL = 0, ROWS_BIG = 1, ROWS_ANY = 2;
template <bool ANY, bool COUNT, int BLOCK, int DEPTH, bool TSTACK, bool SPHERES, bool WIDE,
          int ROWS, class Src>
__device__ inline void trace_loop_impl(const TraceArgs& a, const RayStream& rs, int32_t* s_node,
                                       float* s_tn, uint32_t gtid, uint32_t gstride, Src* src) {
  static_assert(!(WIDE && COUNT), "counted traversal walks the binary nodes");
  const bool big = ROWS == ROWS_ANY ? a.nprims > WIDE_UNIFORM_MAX_NODES : ROWS == ROWS_BIG;
  const int tid = threadIdx.x, lane = tid & 63;
  const uint64_t lt = (1ull << lane) - 1ull;
  const int max_sp = DEPTH + (int)a.gdepth;
  const Schedule sc = schedule_of(a, (uint32_t)rs.waves);
  const float slack = (SPHERES && a.sphere) ? SPHERE_SLACK : SLAB_SLACK;

  RayState r{};  // lanes that never take a ray still read their (unused) octant
  bool active = false;
  // wave-private slice of the ray stream [cnext, cend), refilled CHUNK rays at
  // a time from the global counter (one atomic per slice, not per refill)
  uint32_t cnext = 0, cend = 0;
  bool wave_out = false;  // the global stream is exhausted for this wave
  if (rs.ctr64) {
    cnext = slice_begin(sc, rs.wave), cend = slice_begin(sc, rs.wave + 1ull);
    wave_out = cnext >= sc.n;
  }
  uint32_t ray = 0;
  int32_t node = 0, best = -1;
  int sp = 0;
  float bt = 0.0f, bu = 0.0f, bv = 0.0f;
  uint32_t vis = 0, tst = 0;

  auto finish = [&]() {
    const uint32_t oi = a.out_index ? a.out_index[ray] : ray;
    if (ANY) {
      a.occ[oi] = best >= 0 ? 1 : 0;
    } else {
      a.prim[oi] = best;
      if (a.t) a.t[oi] = best >= 0 ? bt : INFINITY;
      if (a.u) a.u[oi] = bu;
      if (a.v) a.v[oi] = bv;
    }
    if (COUNT) {
      a.visits[oi] = vis;
      a.tests[oi] = tst;
    }
    active = false;
  };
  auto push = [&](int32_t c, float tn) {
    if (sp < DEPTH) {
      s_node[sp * BLOCK + tid] = c;
      if (TSTACK) s_tn[sp * BLOCK + tid] = tn;
    } else if (sp < max_sp) {
      const size_t g = (size_t)(sp - DEPTH) * gstride + gtid;
      a.gstack[2 * g] = c;
      if (TSTACK) a.gstack[2 * g + 1] = __float_as_int(tn);
    } else {
      atomicOr(a.err, 1u);  // deeper than any LBVH of 30-bit codes: reported
    }
    sp = min(sp + 1, max_sp);
  };
  // Pops the next entry into `node`; false when the stack is empty. The LDS
  // entry is read unconditionally (at a clamped depth) and replaced by the
  // overflow entry only past the LDS depth: with one read per branch the
  // compiler merges them into a generic (flat) load, which takes the vector
  // memory path and waits on both counters for every pop.
  auto pop = [&]() -> bool {
    while (sp > 0) {
      --sp;
      const int ls = min(sp, DEPTH - 1);
      node = s_node[ls * BLOCK + tid];
      float tn = TSTACK ? s_tn[ls * BLOCK + tid] : 0.0f;
      if (sp >= DEPTH) {
        const size_t g = (size_t)(sp - DEPTH) * gstride + gtid;
        node = a.gstack[2 * g];
        if (TSTACK) tn = __int_as_float(a.gstack[2 * g + 1]);
      }
      if (!TSTACK || ANY || tn <= bt * slack) return true;
    }
    return false;
  };

  for (;;) {
    // ---- refill idle lanes from the wave's slice of the ray stream -------
    uint64_t m = __ballot(!active);
    for (int round = 0; round < 2 && m && !wave_out; ++round) {
      if (Src::kActive && cnext >= cend) {
        uint32_t b = 0, e = 0;
        if (!src->next(b, e)) {
          wave_out = true;
          break;
        }
        cnext = b, cend = e;
      } else if (cnext >= cend) {
        if (rs.ctr64 && slice_begin(sc, rs.waves) >= sc.n) {  // first slices cover all
          wave_out = true;
          break;
        }
        uint64_t s = 0;
        if (lane == 0)
          s = rs.ctr64 ? atomicAdd(rs.ctr64, 1ull) - rs.base + rs.waves
                       : atomicAdd(rs.ctr32, sc.chunk);
        s = __shfl(s, 0, 64);
        const uint32_t b = rs.ctr64 ? slice_begin(sc, s) : (uint32_t)min(s, (uint64_t)sc.n);
        if (b >= sc.n) {
          wave_out = true;
          break;
        }
        cnext = b;
        cend = rs.ctr64 ? slice_begin(sc, s + 1) : min(b + sc.chunk, sc.n);
      }
      const uint32_t take = min((uint32_t)__popcll(m), cend - cnext);
      const uint32_t rank = (uint32_t)__popcll(m & lt);
      if (!active && rank < take) {
        const uint32_t rr = cnext + rank;
        ray = rr;
        r.ox = a.rays.ox[rr], r.oy = a.rays.oy[rr], r.oz = a.rays.oz[rr];
        r.dx = a.rays.dx[rr], r.dy = a.rays.dy[rr], r.dz = a.rays.dz[rr];
        r.tmin = a.rays.tmin[rr], r.tmax = a.rays.tmax[rr];
        r.ix = 1.0f / r.dx, r.iy = 1.0f / r.dy, r.iz = 1.0f / r.dz;
        r.sx = signbit(r.ix), r.sy = signbit(r.iy), r.sz = signbit(r.iz);
        best = -1, bt = r.tmax, bu = 0.0f, bv = 0.0f, vis = 0, tst = 0, sp = 0;
        node = a.nprims == 1 ? ~0 : 0;
        active = true;
        if (!(r.tmin < r.tmax) || a.nprims == 0) finish();
      }
      cnext += take;
      m = __ballot(!active);
    }
    if (__ballot(active) == 0ull && wave_out) break;

    // Wide walk: when every active lane's direction has the same signs (an
    // octant), the near / far rows are picked by scalar byte offsets instead
    // of a select per value. Lanes that finish keep it so; a refill
    // recomputes it.
    uint32_t uni_rows = 0u, s_oct = 0u;
    uint64_t mx = 0, my = 0, mz = 0;  // lanes whose 1/d has that sign bit set
    const __amdgpu_buffer_rsrc_t wrs = wide_rsrc(a);
    if (WIDE) {
      const uint32_t oct = (uint32_t)r.sx | ((uint32_t)r.sy << 1) | ((uint32_t)r.sz << 2);
      const uint64_t am = __ballot(active);
      const int first = am ? __ffsll((long long)am) - 1 : 0;
      const uint32_t oref = (uint32_t)__shfl((int)oct, first, 64);
      const bool uni = __ballot(active && oct != oref) == 0ull && !big;
      uni_rows = __builtin_amdgcn_readfirstlane((uint32_t)uni);
      s_oct = __builtin_amdgcn_readfirstlane(oref);
      mx = __ballot(r.sx), my = __ballot(r.sy), mz = __ballot(r.sz);
    }

    // ---- traverse until enough lanes are idle to be worth a refill -------
    for (;;) {
      if (WIDE && active && node >= 0) {
        float4 nx, fx, ny, fy, nz, fz;
        const WideNode* w = a.wide + node;
        // the node's row offset, shared by both offset paths (computed here, a
        // path that only issues loads needs no wait behind the other's loads)
        const uint32_t base = (uint32_t)node << 7;
        int4 ln;
        if (!big && !uni_rows) {
          // mixed octants: the same rows at per-lane byte offsets (an OR into
          // the node's 128-byte row, cheaper than a select per value). The
          // offsets are formed from the wave's sign masks at each visit
          // (volatile: not hoisted out of the loop, where they would hold
          // three registers across it; the closest-hit kernel would reach 91
          // and leave less room beside a concurrent any-hit launch).
          uint32_t qx, qy, qz;
          asm volatile("v_cndmask_b32_e64 %0, 0, 16, %1" : "=v"(qx) : "s"(mx));
          asm volatile("v_cndmask_b32_e64 %0, 0, 16, %1" : "=v"(qy) : "s"(my));
          asm volatile("v_cndmask_b32_e64 %0, 0, 16, %1" : "=v"(qz) : "s"(mz));
          nx = row_load(wrs, base | qx, 0u), fx = row_load(wrs, base | (qx ^ 16u), 0u);
          ny = row_load(wrs, base | qy, 32u), fy = row_load(wrs, base | (qy ^ 16u), 32u);
          nz = row_load(wrs, base | qz, 64u), fz = row_load(wrs, base | (qz ^ 16u), 64u);
          if (ROWS == ROWS_ANY) ln = link_load(wrs, base);
        } else if (ROWS == ROWS_ANY ? uni_rows != 0u : !big) {
          // rows at byte offsets lox 0, hix 16, loy 32, hiy 48, loz 64, hiz 80;
          // the near row of an axis is hi when its sign bit is set
          const uint32_t ox = (s_oct & 1u) << 4, oy = 32u + ((s_oct & 2u) << 3),
                         oz = 64u + ((s_oct & 4u) << 2);
          nx = row_load(wrs, base, ox), fx = row_load(wrs, base, ox ^ 16u);
          ny = row_load(wrs, base, oy), fy = row_load(wrs, base, oy ^ 16u);
          nz = row_load(wrs, base, oz), fz = row_load(wrs, base, oz ^ 16u);
          if (ROWS == ROWS_ANY) ln = link_load(wrs, base);
        } else {
          const float4 lx = w->lox, hx = w->hix, ly = w->loy, hy = w->hiy, lz = w->loz,
                       hz = w->hiz;
          // near / far rows by the direction's sign (hi is near when 1/d < 0)
          nx = r.sx ? hx : lx, fx = r.sx ? lx : hx;
          ny = r.sy ? hy : ly, fy = r.sy ? ly : hy;
          nz = r.sz ? hz : lz, fz = r.sz ? lz : hz;
          if (ROWS != ROWS_SMALL) ln = w->link;
        }
        // the link row after the rows: the uniform path, which statically
        // follows the per-lane path's loads, then starts with nothing in
        // flight to wait on. (A walk that decides the row form per launch
        // loads it in each branch instead: one load after them would be a
        // select between a buffer and a 64-bit load.)
        if (ROWS == ROWS_SMALL) ln = link_load(wrs, base);
        float t0, t1, t2, t3;
        const bool h0 = slab_w(nx.x, ny.x, nz.x, fx.x, fy.x, fz.x, r, bt, slack, t0) & (ln.x != 0);
        const bool h1 = slab_w(nx.y, ny.y, nz.y, fx.y, fy.y, fz.y, r, bt, slack, t1) & (ln.y != 0);
        const bool h2 = slab_w(nx.z, ny.z, nz.z, fx.z, fy.z, fz.z, r, bt, slack, t2) & (ln.z != 0);
        const bool h3 = slab_w(nx.w, ny.w, nz.w, fx.w, fy.w, fz.w, r, bt, slack, t3) & (ln.w != 0);
        const int k = (int)h0 + (int)h1 + (int)h2 + (int)h3;
        // sort the slots by entry distance, misses last: a hit's key is its
        // distance's bits (ordered for tn >= +0, i.e. tmin >= 0; otherwise
        // only the visit order suffers), a miss's 0xFFFFFFFF. Any hit: hits
        // first, distances unused.
        auto key = [](bool h, float t) -> uint32_t {
          return h ? (ANY ? 0u : __float_as_uint(t)) : 0xFFFFFFFFu;
        };
        uint32_t k0, k1, k2, k3;
        int32_t c0 = ln.x, c1 = ln.y, c2 = ln.z, c3 = ln.w;
        if (ANY) {
          // the same network on the hit flags (every pushed entry is a hit,
          // key 0)
          bool g0 = h0, g1 = h1, g2 = h2, g3 = h3;
          cas_hit(g0, c0, g1, c1), cas_hit(g2, c2, g3, c3);
          cas_hit(g0, c0, g2, c2), cas_hit(g1, c1, g3, c3);
          cas_hit(g1, c1, g2, c2);
          k0 = k1 = k2 = k3 = 0u;
        } else {
          k0 = key(h0, t0), k1 = key(h1, t1), k2 = key(h2, t2), k3 = key(h3, t3);
          cas_pair(k0, c0, k1, c1), cas_pair(k2, c2, k3, c3);
          cas_pair(k0, c0, k2, c2), cas_pair(k1, c1, k3, c3);
          cas_pair(k1, c1, k2, c2);
        }
        if (k == 0) {
          if (!pop()) finish();
        } else {
          // push the k-1 farther hits, farthest deepest: entry sp+j holds
          // sorted slot k-1-j; with three LDS entries free all three are
          // written (entries above the new top are dead)
          if (k > 1 && sp + 3 <= DEPTH) {
            const int32_t e0 = k == 4 ? c3 : (k == 3 ? c2 : c1), e1 = k == 4 ? c2 : c1;
            s_node[sp * BLOCK + tid] = e0;
            s_node[(sp + 1) * BLOCK + tid] = e1;
            s_node[(sp + 2) * BLOCK + tid] = c1;
            if (TSTACK) {
              const uint32_t d0 = k == 4 ? k3 : (k == 3 ? k2 : k1), d1 = k == 4 ? k2 : k1;
              s_tn[sp * BLOCK + tid] = __uint_as_float(d0);
              s_tn[(sp + 1) * BLOCK + tid] = __uint_as_float(d1);
              s_tn[(sp + 2) * BLOCK + tid] = __uint_as_float(k1);
            }
            sp += k - 1;
          } else {
            if (k > 3) push(c3, __uint_as_float(k3));
            if (k > 2) push(c2, __uint_as_float(k2));
            if (k > 1) push(c1, __uint_as_float(k1));
          }
          node = c0;
        }
      } else if (!WIDE && active && node >= 0) {
        const float4* q = reinterpret_cast<const float4*>(a.nodes + node);
        const float4 lo0 = q[0], hi0 = q[1], lo1 = q[2], hi1 = q[3];
        if (COUNT) vis += 2;
        float tn0, tn1;
        const bool h0 = slab(lo0, hi0, r, bt, slack, tn0);
        const bool h1 = slab(lo1, hi1, r, bt, slack, tn1);
        const int32_t c0 = __float_as_int(lo0.w), c1 = __float_as_int(hi0.w);
        if (h0 && h1) {
          const bool swap = tn1 < tn0;
          push(swap ? c0 : c1, swap ? tn0 : tn1);
          node = swap ? c1 : c0;
        } else if (h0) {
          node = c0;
        } else if (h1) {
          node = c1;
        } else if (!pop()) {
          finish();
        }
      }
      // leaves reached from here are tested before the next node fetch
      while (active && node < 0) {
        const TriRec T = a.tris[~node];
        if (COUNT) ++tst;
        float t, u = 0.0f, v = 0.0f;
        bool done = false;
        if ((SPHERES && a.sphere) ? sphere_hit(T, r, t) : mt(T, r, t, u, v)) {
          const int32_t prim = __float_as_int(T.v0.w);
          if (ANY) {
            best = prim, bt = t;
            done = true;
          } else if (best < 0 || t < bt || (t == bt && prim < best)) {
            best = prim, bt = t, bu = u, bv = v;
          }
        }
        if (done || !pop()) finish();
      }
      const uint64_t idle = __ballot(!active);
      if (__ballot(active) == 0ull || (!wave_out && __popcll(idle) >= a.refill_idle)) break;
    }
  }

}

// One code path for trees of any size (the batch and serving kernels, whose
// launches walk many trees and hold their build phases in the same kernel).
template <bool ANY, bool COUNT, int BLOCK, int DEPTH, bool TSTACK, bool SPHERES = true,
          bool WIDE = false, class Src = NoSlices>
__device__ inline void trace_loop(const TraceArgs& a, const RayStream& rs, int32_t* s_node,
                                  float* s_tn, uint32_t gtid, uint32_t gstride,
                                  Src* src = nullptr) {
  trace_loop_impl<ANY, COUNT, BLOCK, DEPTH, TSTACK, SPHERES, WIDE, ROWS_ANY, Src>(
      a, rs, s_node, s_tn, gtid, gstride, src);
}
// A path per row form, chosen once per launch (the query kernels: one tree).
// Below 2^25 nodes every row, the link row included, is a load at a 32-bit
// offset from the scalar base, with no 64-bit address arithmetic; the closest-
// hit kernels also drop their register spills.
template <bool ANY, bool COUNT, int BLOCK, int DEPTH, bool TSTACK, bool SPHERES = true,
          bool WIDE = false>
__device__ inline void trace_loop_split(const TraceArgs& a, const RayStream& rs, int32_t* s_node,
                                        float* s_tn, uint32_t gtid, uint32_t gstride) {
  if (WIDE && a.nprims > WIDE_UNIFORM_MAX_NODES)
    trace_loop_impl<ANY, COUNT, BLOCK, DEPTH, TSTACK, SPHERES, WIDE, ROWS_BIG, NoSlices>(
        a, rs, s_node, s_tn, gtid, gstride, nullptr);
  else
    trace_loop_impl<ANY, COUNT, BLOCK, DEPTH, TSTACK, SPHERES, WIDE, ROWS_SMALL, NoSlices>(
        a, rs, s_node, s_tn, gtid, gstride, nullptr);
}

}  // namespace rq
