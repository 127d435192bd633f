// CPU BASELINE PORT — test/benchmark infrastructure only (never linked by the
// product library). A restatement, on triangles, of the reference CPU path
// that GPU_LBVH replaces:
//   build : constructBVHNew, accelerators.h:246-337 (median split on the
//           longest axis, std::partition + std::nth_element, 1 primitive/leaf)
//   query : boxIntersect, accelerators.h:668-690 (recursive, collects every
//           object whose leaf box is hit, each child box tested twice) with the
//           slab test of accelerators.h:588-626 (divisions, no t clip), then
//           the candidate loop of main.cpp:343-358 (strict <, so the first
//           candidate in DFS order wins ties).
//   build : constructLBVHTree, accelerators.h:570-586 (port_lbvh_build): the
//           reference's "LBVH", timed as the reference LBVH build
// Deviations, each needed for triangles: the centroid bounds start empty
// (the reference seeds them with the node bounds, :275, which for unequal
// primitive boxes can leave every centroid on one side and recurse without
// end); a coincident-centroid leaf keeps the objects [start, end) (:290-291
// indexes from 0); nodes live in an index array, not shared_ptr copies.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <list>
#include <thread>
#include <vector>

namespace {
struct V3 {
  float x, y, z;
  float operator[](int i) const { return i == 0 ? x : i == 1 ? y : z; }
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

struct Box {
  V3 lo{3.4e38f, 3.4e38f, 3.4e38f}, hi{-3.4e38f, -3.4e38f, -3.4e38f};
};
inline Box join(const Box& a, const Box& b) {  // JoinBounds :201-214
  return {{std::min(a.lo.x, b.lo.x), std::min(a.lo.y, b.lo.y), std::min(a.lo.z, b.lo.z)},
          {std::max(a.hi.x, b.hi.x), std::max(a.hi.y, b.hi.y), std::max(a.hi.z, b.hi.z)}};
}
inline Box joinp(const Box& a, V3 p) {  // JoinBoxPopintBounds :216-229
  return {{std::min(a.lo.x, p.x), std::min(a.lo.y, p.y), std::min(a.lo.z, p.z)},
          {std::max(a.hi.x, p.x), std::max(a.hi.y, p.y), std::max(a.hi.z, p.z)}};
}
inline int max_axis(const Box& b) {  // GeMaximumAxis :190-199
  V3 e = sub(b.hi, b.lo);
  if (e.x > e.y && e.x > e.z) return 0;
  if (e.y > e.z) return 1;
  return 2;
}

struct Obj {  // SceneObject :173-188 reduced to what the path reads
  int id;
  V3 center;
  Box box;
};
struct Node {
  Box box;
  int left = -1, right = -1;
  std::vector<int> objs;  // leaf objects
};

struct PortBVH {
  std::vector<Node> nodes;
  std::vector<float> tri;  // per primitive id: v0, e1, e2
  int root = -1;
};

int build(std::vector<Obj>& o, int s, int e, PortBVH& t) {
  int me = (int)t.nodes.size();
  t.nodes.emplace_back();
  Box bounds;
  for (int i = s; i < e; ++i) bounds = join(bounds, o[i].box);
  t.nodes[me].box = bounds;
  int n = e - s;
  if (n <= 1) {
    t.nodes[me].objs.push_back(o[s].id);
    return me;
  }
  Box cb;
  for (int i = s; i < e; ++i) cb = joinp(cb, o[i].center);
  int dim = max_axis(cb);
  if (cb.hi[dim] == cb.lo[dim]) {
    for (int i = s; i < e; ++i) t.nodes[me].objs.push_back(o[i].id);
    return me;
  }
  float pmid = (cb.lo[dim] + cb.hi[dim]) / 2;
  Obj* mid = std::partition(&o[s], &o[e - 1] + 1, [dim, pmid](const Obj& p) {
    return p.center[dim] < pmid;
  });
  int m = (int)(mid - &o[0]);
  if (m != s && m != e) {
    m = (s + e) / 2;
    std::nth_element(&o[s], &o[m], &o[e - 1] + 1,
                     [dim](const Obj& a, const Obj& b) { return a.center[dim] < b.center[dim]; });
  }
  if (m == s || m == e) m = (s + e) / 2;  // unreachable with true centroid bounds
  int l = build(o, s, m, t);
  int r = build(o, m, e, t);
  t.nodes[me].left = l;
  t.nodes[me].right = r;
  return me;
}

// ---- constructLBVHTree (accelerators.h:570-586) ------------------------------
// expandBits / morton3D, accelerators.h:374-394 (multiply-mask spread, clamp to
// [0, 1023] after scaling by 1024, x in the high bit of each triple).
inline uint32_t expand_bits(uint32_t v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}
inline uint32_t morton3d(float x, float y, float z) {
  x = std::min(std::max(x * 1024.0f, 0.0f), 1023.0f);
  y = std::min(std::max(y * 1024.0f, 0.0f), 1023.0f);
  z = std::min(std::max(z * 1024.0f, 0.0f), 1023.0f);
  return expand_bits((uint32_t)x) * 4 + expand_bits((uint32_t)y) * 2 + expand_bits((uint32_t)z);
}
// radixSort, accelerators.h:348-369: a decimal LSD sort of `max` passes into ten
// std::list pockets (push_back per element, then erase from the front), on a
// vector taken by value, so the sorted copy is discarded. The reference's
// int m = pow(10, i + 1) overflows at i = 9; the pass divisors here are 64-bit
// (same work, no undefined behaviour). Returns a checksum so the work is kept.
uint64_t radix_sort_copy(std::vector<uint32_t> codes, int max) {
  std::list<uint32_t> pocket[10];
  uint64_t p = 1;
  for (int i = 0; i < max; ++i, p *= 10) {
    for (uint32_t c : codes) pocket[(c % (p * 10)) / p].push_back(c);
    size_t k = 0;
    for (int j = 0; j < 10; ++j)
      while (!pocket[j].empty()) {
        codes[k++] = *pocket[j].begin();
        pocket[j].erase(pocket[j].begin());
      }
  }
  return codes.empty() ? 0 : codes.front() + codes.back();
}
// findSplit, accelerators.h:397-449 (vector taken by value; leading zeros of
// the low 8 bits only); constructLBVHNew discards its result (:518).
int find_split(std::vector<uint32_t> codes, int first, int last) {
  auto prefix8 = [](uint32_t c) {
    int len = -1;
    while ((++len < 8) && ((c & 0x80) == 0)) c <<= 1;
    return len;
  };
  const uint32_t fc = codes[first], lc = codes[last];
  if (fc == lc) return (first + last) >> 1;
  const int common = prefix8(fc ^ lc);
  int split = first, step = last - first;
  do {
    step = (step + 1) >> 1;
    const int ns = split + step;
    if (ns < last && prefix8(fc ^ codes[ns]) > common) split = ns;
  } while (step > 1);
  return split;
}
// Appends tree b (built on its own) to t, its node indices shifted.
int append_tree(PortBVH& t, PortBVH&& b) {
  const int off = (int)t.nodes.size();
  for (Node& n : b.nodes) {
    if (n.left >= 0) n.left += off, n.right += off;
    t.nodes.push_back(std::move(n));
  }
  return b.root + off;
}

// accelerators.h:588-626
inline bool slab(V3 p, V3 d, const Box& b) {
  float tmin = (b.lo.x - p.x) / d.x, tmax = (b.hi.x - p.x) / d.x;
  if (tmin > tmax) std::swap(tmin, tmax);
  float tymin = (b.lo.y - p.y) / d.y, tymax = (b.hi.y - p.y) / d.y;
  if (tymin > tymax) std::swap(tymin, tymax);
  if ((tmin > tymax) || (tymin > tmax)) return false;
  if (tymin > tmin) tmin = tymin;
  if (tymax < tmax) tmax = tymax;
  float tzmin = (b.lo.z - p.z) / d.z, tzmax = (b.hi.z - p.z) / d.z;
  if (tzmin > tzmax) std::swap(tzmin, tzmax);
  if ((tmin > tzmax) || (tzmin > tmax)) return false;
  return true;
}

// accelerators.h:668-690
void box_intersect(const PortBVH& t, V3 p, V3 d, int node, std::vector<int>& out) {
  const Node& n = t.nodes[node];
  if (!slab(p, d, n.box)) return;
  if (n.left < 0) {
    for (int id : n.objs) out.push_back(id);
    return;
  }
  if (slab(p, d, t.nodes[n.left].box)) box_intersect(t, p, d, n.left, out);
  if (slab(p, d, t.nodes[n.right].box)) box_intersect(t, p, d, n.right, out);
}

inline bool mt(const float* T, V3 o, V3 d, float tmin, float tmax, float* t) {
  V3 v0{T[0], T[1], T[2]}, e1{T[3], T[4], T[5]}, e2{T[6], T[7], T[8]};
  V3 p = cross(d, e2);
  float det = dot(e1, p);
  if (std::fabs(det) < 1e-6f) return false;
  float inv = 1 / det;
  V3 tv = sub(o, v0);
  float u = dot(tv, p) * inv;
  if (u < 0 || u > 1) return false;
  V3 q = cross(tv, e1);
  float v = dot(d, q) * inv;
  if (v < 0 || u + v > 1) return false;
  float tt = dot(e2, q) * inv;
  if (!(tt > tmin && tt < tmax)) return false;
  *t = tt;
  return true;
}
}  // namespace

extern "C" {

void* port_bvh_build(const float* vx, const float* vy, const float* vz, const uint32_t* idx,
                     uint32_t ntri, uint32_t* n_nodes) {
  auto* t = new PortBVH();
  std::vector<Obj> objs(ntri);
  t->tri.resize((size_t)ntri * 9);
  for (uint32_t p = 0; p < ntri; ++p) {
    V3 a{vx[idx[3 * p]], vy[idx[3 * p]], vz[idx[3 * p]]};
    V3 b{vx[idx[3 * p + 1]], vy[idx[3 * p + 1]], vz[idx[3 * p + 1]]};
    V3 c{vx[idx[3 * p + 2]], vy[idx[3 * p + 2]], vz[idx[3 * p + 2]]};
    Obj& o = objs[p];
    o.id = (int)p;
    o.center = {((a.x + b.x) + c.x) / 3.0f, ((a.y + b.y) + c.y) / 3.0f, ((a.z + b.z) + c.z) / 3.0f};
    o.box = joinp(joinp(joinp(Box(), a), b), c);
    V3 e1 = sub(b, a), e2 = sub(c, a);
    float* T = &t->tri[9 * (size_t)p];
    T[0] = a.x, T[1] = a.y, T[2] = a.z, T[3] = e1.x, T[4] = e1.y, T[5] = e1.z, T[6] = e2.x,
    T[7] = e2.y, T[8] = e2.z;
  }
  if (ntri) {
    t->nodes.reserve(2 * (size_t)ntri);
    t->root = build(objs, 0, (int)ntri, *t);
  }
  if (n_nodes) *n_nodes = (uint32_t)t->nodes.size();
  return t;
}

// constructLBVHTree (accelerators.h:570-586) on the same objects: Morton codes
// of the centres mapped by (c + 30) / 1000 (:576-580), radixSort of a copy
// (:581, discarded), then constructLBVHNew (:452-567) at the root: bounds,
// centroid bounds, findSplit on a copy (discarded), the median split, and the
// two halves built by constructBVHNew, on two std::threads when the root holds
// more than 500 objects (:543-548). Every object is kept (the reference passes
// size - 1 as the end, :583, and drops the last one). Queries run on the
// result like on port_bvh_build's.
void* port_lbvh_build(const float* vx, const float* vy, const float* vz, const uint32_t* idx,
                      uint32_t ntri, uint32_t* n_nodes) {
  auto* t = new PortBVH();
  std::vector<Obj> objs(ntri);
  t->tri.resize((size_t)ntri * 9);
  for (uint32_t p = 0; p < ntri; ++p) {
    V3 a{vx[idx[3 * p]], vy[idx[3 * p]], vz[idx[3 * p]]};
    V3 b{vx[idx[3 * p + 1]], vy[idx[3 * p + 1]], vz[idx[3 * p + 1]]};
    V3 c{vx[idx[3 * p + 2]], vy[idx[3 * p + 2]], vz[idx[3 * p + 2]]};
    Obj& o = objs[p];
    o.id = (int)p;
    o.center = {((a.x + b.x) + c.x) / 3.0f, ((a.y + b.y) + c.y) / 3.0f, ((a.z + b.z) + c.z) / 3.0f};
    o.box = joinp(joinp(joinp(Box(), a), b), c);
    V3 e1 = sub(b, a), e2 = sub(c, a);
    float* T = &t->tri[9 * (size_t)p];
    T[0] = a.x, T[1] = a.y, T[2] = a.z, T[3] = e1.x, T[4] = e1.y, T[5] = e1.z, T[6] = e2.x,
    T[7] = e2.y, T[8] = e2.z;
  }
  const int n = (int)ntri;
  std::vector<uint32_t> codes(ntri);
  for (int i = 0; i < n; ++i) {
    const V3 m = objs[i].center;
    codes[i] = morton3d((m.x + 30) / 1000, (m.y + 30) / 1000, (m.z + 30) / 1000);
  }
  volatile uint64_t keep = radix_sort_copy(codes, 10);
  (void)keep;
  if (n == 0) {
    if (n_nodes) *n_nodes = 0;
    return t;
  }
  t->nodes.reserve(2 * (size_t)ntri);
  t->nodes.emplace_back();
  t->root = 0;
  Box bounds;
  for (int i = 0; i < n; ++i) bounds = join(bounds, objs[i].box);
  t->nodes[0].box = bounds;
  Box cb;
  for (int i = 0; i < n; ++i) cb = joinp(cb, objs[i].center);
  const int dim = max_axis(cb);
  if (n <= 1 || cb.hi[dim] == cb.lo[dim]) {
    for (int i = 0; i < n; ++i) t->nodes[0].objs.push_back(objs[i].id);
    if (n_nodes) *n_nodes = 1;
    return t;
  }
  volatile int split = find_split(codes, 0, n - 1);
  (void)split;
  const float pmid = (cb.lo[dim] + cb.hi[dim]) / 2;
  Obj* mid = std::partition(&objs[0], &objs[n - 1] + 1,
                            [dim, pmid](const Obj& p) { return p.center[dim] < pmid; });
  int m = (int)(mid - &objs[0]);
  if (m != 0 && m != n) {
    m = n / 2;
    std::nth_element(&objs[0], &objs[m], &objs[n - 1] + 1,
                     [dim](const Obj& a, const Obj& b) { return a.center[dim] < b.center[dim]; });
  }
  if (m == 0 || m == n) m = n / 2;
  PortBVH lt, rt;
  lt.nodes.reserve(2 * (size_t)m);
  rt.nodes.reserve(2 * (size_t)(n - m));
  if (n > 500) {
    std::thread tl([&] { lt.root = build(objs, 0, m, lt); });
    std::thread tr([&] { rt.root = build(objs, m, n, rt); });
    tl.join();
    tr.join();
  } else {
    lt.root = build(objs, 0, m, lt);
    rt.root = build(objs, m, n, rt);
  }
  const int l = append_tree(*t, std::move(lt));
  const int r = append_tree(*t, std::move(rt));
  t->nodes[0].left = l;
  t->nodes[0].right = r;
  if (n_nodes) *n_nodes = (uint32_t)t->nodes.size();
  return t;
}

void port_bvh_free(void* h) { delete static_cast<PortBVH*>(h); }

// Closest hit for rays [0, n); prim = -1 on miss. nthreads >= 1 splits the
// rays into contiguous chunks, one std::thread each.
void port_closest(const void* h, const float* ox, const float* oy, const float* oz,
                  const float* dx, const float* dy, const float* dz, const float* tmin,
                  const float* tmax, uint32_t n, int32_t* prim, float* tout, int nthreads) {
  const PortBVH& t = *static_cast<const PortBVH*>(h);
  auto body = [&](uint32_t b, uint32_t e) {
    std::vector<int> cand;
    for (uint32_t i = b; i < e; ++i) {
      V3 o{ox[i], oy[i], oz[i]}, d{dx[i], dy[i], dz[i]};
      cand.clear();
      if (t.root >= 0) box_intersect(t, o, d, t.root, cand);
      int best = -1;
      float tn = tmax[i];
      for (int c : cand) {
        float tt;
        if (mt(&t.tri[9 * (size_t)c], o, d, tmin[i], tmax[i], &tt) && tt < tn) {
          tn = tt;
          best = c;
        }
      }
      prim[i] = best;
      if (tout) tout[i] = best >= 0 ? tn : INFINITY;
    }
  };
  if (nthreads <= 1) {
    body(0, n);
    return;
  }
  std::vector<std::thread> th;
  uint32_t chunk = (n + nthreads - 1) / nthreads;
  for (int k = 0; k < nthreads; ++k) {
    uint32_t b = k * chunk, e = std::min<uint32_t>(n, b + chunk);
    if (b >= e) break;
    th.emplace_back(body, b, e);
  }
  for (auto& x : th) x.join();
}

}  // extern "C"
