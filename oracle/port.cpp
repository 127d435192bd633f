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
// Deviations, each needed for triangles: the centroid bounds start empty
// (the reference seeds them with the node bounds, :275, which for unequal
// primitive boxes can leave every centroid on one side and recurse without
// end); a coincident-centroid leaf keeps the objects [start, end) (:290-291
// indexes from 0); nodes live in an index array, not shared_ptr copies.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
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
