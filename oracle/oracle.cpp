// CPU ORACLE — test infrastructure only (see oracle.h). Never linked by the
// product library. Compiled with -ffp-contract=off so every float operation is
// a single IEEE-754 rounding, in the order written.
//
// Reference files cited below are relative to the reference checkout
// (project/raytracer/...). Operation order of dot and cross products follows
// geometry.h:94-95 ((x*vx + y*vy) + z*vz) and geometry.h:109-110.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <random>
#include <thread>
#include <vector>

namespace {

struct V3 {
  float x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
// geometry.h:94-95
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// geometry.h:109-110
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// geometry.h:125-134 (normalize: factor = 1/sqrt(norm), then multiply). The
// template's unqualified sqrt is the C double sqrt (geometry.h includes
// <cmath> without std:: overloads in scope), so the factor is computed in
// double and rounded once to float.
inline V3 normalize(V3 a) {
  float n = a.x * a.x + a.y * a.y + a.z * a.z;
  if (n > 0) {
    float f = (float)(1.0 / ::sqrt((double)n));
    a.x *= f, a.y *= f, a.z *= f;
  }
  return a;
}
// NaN-ignoring min/max (C99 fmin/fmax semantics, as v_min_f32/v_max_f32).
inline float fmn(float a, float b) { return (a < b || b != b) ? a : b; }
inline float fmx(float a, float b) { return (a > b || b != b) ? a : b; }

inline V3 vtx(const float* vx, const float* vy, const float* vz, uint32_t i) {
  return {vx[i], vy[i], vz[i]};
}

template <class F>
void parallel_for(uint32_t n, int nthreads, F f) {
  if (nthreads <= 1 || n < 1024) {
    for (uint32_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  uint32_t chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    uint32_t b = t * chunk, e = std::min<uint32_t>(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=] {
      for (uint32_t i = b; i < e; ++i) f(i);
    });
  }
  for (auto& x : th) x.join();
}

inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }
inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

}  // namespace

// ---------------------------------------------------------------------------
// Morton codes: accelerators.h:374-381 (expandBits), :385-394 (morton3D)
// ---------------------------------------------------------------------------
extern "C" uint32_t or_expand_bits(uint32_t v) {
  // Restated as a plain bit loop (bit k of v -> bit 3k), which is the
  // function the multiply/mask sequence at accelerators.h:376-379 computes
  // for v < 1024.
  uint32_t r = 0;
  for (int k = 0; k < 10; ++k) r |= ((v >> k) & 1u) << (3 * k);
  return r;
}

extern "C" uint32_t or_morton3d(float x, float y, float z) {
  // accelerators.h:387-393: clamp(x*1024, 0, 1023) with std::min/std::max,
  // truncate to unsigned, interleave x as the most significant bit of each
  // triple (xx*4 + yy*2 + zz). NaN is undefined in the reference; it maps to 0.
  auto q = [](float a) -> uint32_t {
    a = a * 1024.0f;
    a = (a < 0.0f) ? 0.0f : a;          // std::max(a, 0.0f)
    a = (1023.0f < a) ? 1023.0f : a;    // std::min(a, 1023.0f)
    if (a != a) return 0;
    return (uint32_t)a;
  };
  return or_expand_bits(q(x)) * 4 + or_expand_bits(q(y)) * 2 + or_expand_bits(q(z));
}

extern "C" void or_centroid(const float* vx, const float* vy, const float* vz,
                            const uint32_t* idx, uint32_t t, float c[3]) {
  // Triangle centroid (the reference has none; its spheres use the centre,
  // main.cpp:680-683): ((v0 + v1) + v2) / 3 per axis.
  V3 a = vtx(vx, vy, vz, idx[3 * t]), b = vtx(vx, vy, vz, idx[3 * t + 1]),
     d = vtx(vx, vy, vz, idx[3 * t + 2]);
  c[0] = ((a.x + b.x) + d.x) / 3.0f;
  c[1] = ((a.y + b.y) + d.y) / 3.0f;
  c[2] = ((a.z + b.z) + d.z) / 3.0f;
}

namespace {
// Morton codes of n centroids (3 floats each, zeros canonicalised in place).
void codes_from_centroids(std::vector<float>& cs, uint32_t ntri, int mode, uint32_t* codes) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t t = 0; t < ntri; ++t) {
    float* c = &cs[3 * (size_t)t];
    for (int k = 0; k < 3; ++k) {
      c[k] = c[k] + 0.0f;  // canonical zero
      lo[k] = std::min(lo[k], c[k]);
      hi[k] = std::max(hi[k], c[k]);
    }
  }
  for (uint32_t t = 0; t < ntri; ++t) {
    const float* c = &cs[3 * (size_t)t];
    float n[3];
    for (int k = 0; k < 3; ++k) {
      if (mode == 0) {
        n[k] = (c[k] + 30.0f) / 1000.0f;  // accelerators.h:577
      } else {
        float e = hi[k] - lo[k];
        n[k] = e > 0.0f ? (c[k] - lo[k]) / e : 0.0f;
      }
    }
    codes[t] = or_morton3d(n[0], n[1], n[2]);
  }
}
}  // namespace

extern "C" int or_morton_codes(const float* vx, const float* vy, const float* vz, uint32_t nv,
                               const uint32_t* idx, uint32_t ntri, int mode, uint32_t* codes) {
  (void)nv;
  std::vector<float> cs((size_t)ntri * 3);
  for (uint32_t t = 0; t < ntri; ++t) or_centroid(vx, vy, vz, idx, t, &cs[3 * (size_t)t]);
  codes_from_centroids(cs, ntri, mode, codes);
  return 0;
}

// ---------------------------------------------------------------------------
// Sort: the reference's radixSort (accelerators.h:348-369) sorts a copy and
// discards it; what it would produce is an ascending sort of the codes. The
// values carried are the primitive indices, stable in index order.
// ---------------------------------------------------------------------------
extern "C" void or_sort_pairs(const uint32_t* codes, uint32_t n, uint32_t* sorted_codes,
                              uint32_t* sorted_idx) {
  std::vector<uint32_t> p(n);
  std::iota(p.begin(), p.end(), 0u);
  std::stable_sort(p.begin(), p.end(), [&](uint32_t a, uint32_t b) { return codes[a] < codes[b]; });
  for (uint32_t i = 0; i < n; ++i) {
    sorted_idx[i] = p[i];
    sorted_codes[i] = codes[p[i]];
  }
}

// ---------------------------------------------------------------------------
// Karras 2012 radix-tree emission (cited by accelerators.h:371,568-569; the
// reference's findSplit :397-449 is its top-down sibling, result discarded).
// ---------------------------------------------------------------------------
namespace {
struct Delta {
  const uint32_t* k;
  int64_t n;
  int operator()(int64_t i, int64_t j) const {
    if (j < 0 || j >= n) return -1;
    uint32_t a = k[i], b = k[j];
    if (a != b) return clz32(a ^ b);
    return 32 + clz32((uint32_t)i ^ (uint32_t)j);
  }
};
}  // namespace

extern "C" void or_karras(const uint32_t* sc, uint32_t n, int32_t* left, int32_t* right) {
  if (n < 2) return;
  Delta delta{sc, (int64_t)n};
  for (int64_t i = 0; i < (int64_t)n - 1; ++i) {
    int d = delta(i, i + 1) > delta(i, i - 1) ? 1 : -1;
    int dmin = delta(i, i - d);
    int64_t lmax = 2;
    while (delta(i, i + lmax * d) > dmin) lmax <<= 1;
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
      if (delta(i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = delta(i, j);
    int64_t s = 0, t = l;
    do {
      t = (t + 1) >> 1;
      if (delta(i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int64_t g = i + s * d + std::min(d, 0);
    left[i] = (std::min(i, j) == g) ? ~(int32_t)g : (int32_t)g;
    right[i] = (std::max(i, j) == g + 1) ? ~(int32_t)(g + 1) : (int32_t)(g + 1);
  }
}

// Independent top-down builder on 64-bit augmented keys (code << 32 | index):
// split = last index sharing more leading bits with the first key than the
// last key does (the binary search of accelerators.h:425-446, on full keys).
extern "C" void or_topdown(const uint32_t* sc, uint32_t n, int32_t* left, int32_t* right) {
  if (n < 2) return;
  auto key = [&](int64_t i) { return ((uint64_t)sc[i] << 32) | (uint64_t)i; };
  struct R {
    int64_t first, last, node;
  };
  std::vector<R> st;
  st.push_back({0, (int64_t)n - 1, 0});
  while (!st.empty()) {
    R r = st.back();
    st.pop_back();
    uint64_t kf = key(r.first);
    int common = clz64(kf ^ key(r.last));
    int64_t split = r.first, step = r.last - r.first;
    do {
      step = (step + 1) >> 1;
      int64_t ns = split + step;
      if (ns < r.last && clz64(kf ^ key(ns)) > common) split = ns;
    } while (step > 1);
    if (split == r.first) {
      left[r.node] = ~(int32_t)split;
    } else {
      left[r.node] = (int32_t)split;
      st.push_back({r.first, split, split});
    }
    if (split + 1 == r.last) {
      right[r.node] = ~(int32_t)(split + 1);
    } else {
      right[r.node] = (int32_t)(split + 1);
      st.push_back({split + 1, r.last, split + 1});
    }
  }
}

void refit_internal(uint32_t n, const int32_t* left, const int32_t* right, float* boxes);

// Bottom-up AABB refit: leaf box = min/max of the triangle's vertices
// (JoinBoxPopintBounds, accelerators.h:216-229), internal box = union of the
// children (JoinBounds, :201-214). Zeros are canonicalised to +0.
extern "C" void or_refit(const float* vx, const float* vy, const float* vz, const uint32_t* idx,
                         const uint32_t* sorted_idx, uint32_t n, const int32_t* left,
                         const int32_t* right, float* boxes) {
  if (n == 0) return;
  auto leafbox = [&](uint32_t k) { return boxes + 6 * ((size_t)(n - 1) + k); };
  for (uint32_t k = 0; k < n; ++k) {
    uint32_t p = sorted_idx[k];
    float* b = leafbox(k);
    for (int a = 0; a < 3; ++a) {
      const float* c = a == 0 ? vx : a == 1 ? vy : vz;
      float x0 = c[idx[3 * p]], x1 = c[idx[3 * p + 1]], x2 = c[idx[3 * p + 2]];
      b[a] = std::min(std::min(x0, x1), x2) + 0.0f;
      b[3 + a] = std::max(std::max(x0, x1), x2) + 0.0f;
    }
  }
  refit_internal(n, left, right, boxes);
}

// Unions child boxes bottom-up (leaf boxes already in place).
void refit_internal(uint32_t n, const int32_t* left, const int32_t* right, float* boxes) {
  if (n < 2) return;
  auto leafbox = [&](uint32_t k) { return boxes + 6 * ((size_t)(n - 1) + k); };
  auto box = [&](int32_t c) -> float* { return c >= 0 ? boxes + 6 * (size_t)c : leafbox(~c); };
  // post-order over internal nodes
  std::vector<std::pair<int32_t, int>> st;
  st.push_back({0, 0});
  while (!st.empty()) {
    auto& top = st.back();
    int32_t i = top.first;
    if (top.second == 0) {
      top.second = 1;
      if (right[i] >= 0) st.push_back({right[i], 0});
      if (left[i] >= 0) st.push_back({left[i], 0});
      continue;
    }
    st.pop_back();
    float *b = boxes + 6 * (size_t)i, *l = box(left[i]), *r = box(right[i]);
    for (int a = 0; a < 3; ++a) {
      b[a] = std::min(l[a], r[a]);
      b[3 + a] = std::max(l[3 + a], r[3 + a]);
    }
  }
}

extern "C" uint32_t or_tree_height(const int32_t* left, const int32_t* right, uint32_t n) {
  if (n < 2) return 0;
  uint32_t h = 0;
  std::vector<std::pair<int32_t, uint32_t>> st{{0, 0}};
  while (!st.empty()) {
    auto [i, d] = st.back();
    st.pop_back();
    for (int32_t c : {left[i], right[i]}) {
      if (c >= 0)
        st.push_back({c, d + 1});
      else
        h = std::max(h, d + 1);
    }
  }
  return h;
}

// ---------------------------------------------------------------------------
// Möller–Trumbore, main.cpp:137-162 (MOLLER_TRUMBORE branch, non-culling),
// with the v_0 typo at :152 read as v0 and the hit restricted to t in
// (tmin, tmax). EPS = 1e-6 (main.cpp:59).
// ---------------------------------------------------------------------------
namespace {
inline bool mt_e(V3 v0, V3 e1, V3 e2, V3 o, V3 d, float tmin, float tmax, float* t, float* u,
                 float* v) {
  V3 p = cross(d, e2);
  float det = dot(e1, p);
  if (std::fabs(det) < 1e-6f) return false;
  float inv = 1 / det;
  V3 tv = sub(o, v0);
  float uu = dot(tv, p) * inv;
  if (uu < 0 || uu > 1) return false;
  V3 q = cross(tv, e1);
  float vv = dot(d, q) * inv;
  if (vv < 0 || uu + vv > 1) return false;
  float tt = dot(e2, q) * inv;
  if (!(tt > tmin && tt < tmax)) return false;
  *t = tt, *u = uu, *v = vv;
  return true;
}
}  // namespace

extern "C" int or_mt(const float v0[3], const float v1[3], const float v2[3], const float o[3],
                     const float d[3], float tmin, float tmax, float* t, float* u, float* v) {
  V3 a{v0[0], v0[1], v0[2]}, b{v1[0], v1[1], v1[2]}, c{v2[0], v2[1], v2[2]};
  return mt_e(a, sub(b, a), sub(c, a), {o[0], o[1], o[2]}, {d[0], d[1], d[2]}, tmin, tmax, t, u,
              v)
             ? 1
             : 0;
}

// accelerators.h:79-92 (geometric sphere test; rejects tca < 0)
extern "C" int or_ray_sphere(const float c[3], float r, const float o[3], const float d[3],
                             float* t0, float* t1) {
  V3 l = sub({c[0], c[1], c[2]}, {o[0], o[1], o[2]});
  V3 dd{d[0], d[1], d[2]};
  float tca = dot(l, dd);
  if (tca < 0) return 0;
  float d2 = dot(l, l) - tca * tca;
  float r2 = r * r;
  if (d2 > r2) return 0;
  float thc = std::sqrt(r2 - d2);
  *t0 = tca - thc;
  *t1 = tca + thc;
  return 1;
}

namespace {
// A sphere leaf as castRay uses it (main.cpp:349-353): the reference's test,
// then t = t0, or t1 when t0 < 0; accepted here when t is in [tmin, tmax).
inline bool sphere_e(V3 c, float r2, V3 o, V3 d, float tmin, float tmax, float* t) {
  V3 l = sub(c, o);
  float tca = dot(l, d);
  if (tca < 0) return false;
  float d2 = dot(l, l) - tca * tca;
  if (d2 > r2) return false;
  float thc = std::sqrt(r2 - d2);
  float t0 = tca - thc, t1 = tca + thc;
  float tt = t0 < 0 ? t1 : t0;
  if (!(tt >= tmin && tt < tmax)) return false;
  *t = tt;
  return true;
}
}  // namespace

// ---------------------------------------------------------------------------
// Brute-force queries: the NONE branch of castRay (main.cpp:373-387) walks
// every primitive in index order and keeps strictly smaller t, so the lowest
// index wins ties.
// ---------------------------------------------------------------------------
namespace {
inline void load_ray(const or_rays* r, uint32_t i, V3* o, V3* d, float* tmin, float* tmax) {
  *o = {r->ox[i], r->oy[i], r->oz[i]};
  *d = {r->dx[i], r->dy[i], r->dz[i]};
  *tmin = r->tmin[i];
  *tmax = r->tmax[i];
}
}  // namespace

extern "C" void or_closest_brute(const float* vx, const float* vy, const float* vz,
                                 const uint32_t* idx, uint32_t ntri, const or_rays* rays,
                                 uint32_t nrays, or_hits* hits, int nthreads) {
  std::vector<float> tri((size_t)ntri * 9);
  for (uint32_t p = 0; p < ntri; ++p) {
    V3 a = vtx(vx, vy, vz, idx[3 * p]), b = vtx(vx, vy, vz, idx[3 * p + 1]),
       c = vtx(vx, vy, vz, idx[3 * p + 2]);
    V3 e1 = sub(b, a), e2 = sub(c, a);
    float* T = &tri[9 * (size_t)p];
    T[0] = a.x, T[1] = a.y, T[2] = a.z, T[3] = e1.x, T[4] = e1.y, T[5] = e1.z, T[6] = e2.x,
    T[7] = e2.y, T[8] = e2.z;
  }
  parallel_for(nrays, nthreads, [&](uint32_t i) {
    V3 o, d;
    float tmin, tmax;
    load_ray(rays, i, &o, &d, &tmin, &tmax);
    int32_t best = -1;
    float bt = tmax, bu = 0, bv = 0;
    for (uint32_t p = 0; p < ntri; ++p) {
      const float* T = &tri[9 * (size_t)p];
      float t, u, v;
      if (mt_e({T[0], T[1], T[2]}, {T[3], T[4], T[5]}, {T[6], T[7], T[8]}, o, d, tmin, tmax, &t,
               &u, &v) &&
          (best < 0 || t < bt)) {
        best = (int32_t)p, bt = t, bu = u, bv = v;
      }
    }
    hits->prim[i] = best;
    if (hits->t) hits->t[i] = best >= 0 ? bt : INFINITY;
    if (hits->u) hits->u[i] = bu;
    if (hits->v) hits->v[i] = bv;
  });
}

extern "C" void or_any_brute(const float* vx, const float* vy, const float* vz,
                             const uint32_t* idx, uint32_t ntri, const or_rays* rays,
                             uint32_t nrays, uint8_t* occ, int nthreads) {
  parallel_for(nrays, nthreads, [&](uint32_t i) {
    V3 o, d;
    float tmin, tmax;
    load_ray(rays, i, &o, &d, &tmin, &tmax);
    uint8_t hit = 0;
    for (uint32_t p = 0; p < ntri && !hit; ++p) {
      V3 a = vtx(vx, vy, vz, idx[3 * p]), b = vtx(vx, vy, vz, idx[3 * p + 1]),
         c = vtx(vx, vy, vz, idx[3 * p + 2]);
      float t, u, v;
      if (mt_e(a, sub(b, a), sub(c, a), o, d, tmin, tmax, &t, &u, &v)) hit = 1;
    }
    occ[i] = hit;
  });
}

extern "C" void or_any_brute_spheres(const float* cx, const float* cy, const float* cz,
                                     const float* r, uint32_t n, const or_rays* rays,
                                     uint32_t nrays, uint8_t* occ, int nthreads) {
  parallel_for(nrays, nthreads, [&](uint32_t i) {
    V3 o, d;
    float tmin, tmax;
    load_ray(rays, i, &o, &d, &tmin, &tmax);
    uint8_t hit = 0;
    for (uint32_t p = 0; p < n && !hit && tmin < tmax; ++p) {
      float t;
      if (sphere_e({cx[p], cy[p], cz[p]}, r[p] * r[p], o, d, tmin, tmax, &t)) hit = 1;
    }
    occ[i] = hit;
  });
}

// ---------------------------------------------------------------------------
// Oracle LBVH and its canonical traversal.
// ---------------------------------------------------------------------------
struct or_bvh {
  uint32_t n = 0;
  std::vector<uint32_t> codes, sorted_codes, sorted_idx;
  std::vector<int32_t> left, right;
  std::vector<float> boxes;  // (2n-1)*6
  std::vector<float> tri;    // n*9 in leaf order: v0, e1, e2 (spheres: centre, r, r*r)
  uint32_t height = 0;
  int kind = 0;              // 0 triangles, 1 spheres
};

extern "C" or_bvh* or_bvh_build(const float* vx, const float* vy, const float* vz, uint32_t nv,
                                const uint32_t* idx, uint32_t ntri, int morton_mode) {
  auto* b = new or_bvh();
  b->n = ntri;
  b->codes.resize(ntri);
  b->sorted_codes.resize(ntri);
  b->sorted_idx.resize(ntri);
  b->left.assign(ntri > 1 ? ntri - 1 : 0, 0);
  b->right.assign(ntri > 1 ? ntri - 1 : 0, 0);
  b->boxes.assign(ntri ? (2 * (size_t)ntri - 1) * 6 : 0, 0.0f);
  b->tri.resize((size_t)ntri * 9);
  if (ntri == 0) return b;
  or_morton_codes(vx, vy, vz, nv, idx, ntri, morton_mode, b->codes.data());
  or_sort_pairs(b->codes.data(), ntri, b->sorted_codes.data(), b->sorted_idx.data());
  or_karras(b->sorted_codes.data(), ntri, b->left.data(), b->right.data());
  or_refit(vx, vy, vz, idx, b->sorted_idx.data(), ntri, b->left.data(), b->right.data(),
           b->boxes.data());
  for (uint32_t k = 0; k < ntri; ++k) {
    uint32_t p = b->sorted_idx[k];
    V3 a = vtx(vx, vy, vz, idx[3 * p]), c1 = vtx(vx, vy, vz, idx[3 * p + 1]),
       c2 = vtx(vx, vy, vz, idx[3 * p + 2]);
    V3 e1 = sub(c1, a), e2 = sub(c2, a);
    float* T = &b->tri[9 * (size_t)k];
    T[0] = a.x, T[1] = a.y, T[2] = a.z, T[3] = e1.x, T[4] = e1.y, T[5] = e1.z, T[6] = e2.x,
    T[7] = e2.y, T[8] = e2.z;
  }
  b->height = or_tree_height(b->left.data(), b->right.data(), ntri);
  return b;
}

extern "C" or_bvh* or_bvh_build_spheres(const float* cx, const float* cy, const float* cz,
                                        const float* r, uint32_t n, int morton_mode) {
  auto* b = new or_bvh();
  b->kind = 1;
  b->n = n;
  b->codes.resize(n);
  b->sorted_codes.resize(n);
  b->sorted_idx.resize(n);
  b->left.assign(n > 1 ? n - 1 : 0, 0);
  b->right.assign(n > 1 ? n - 1 : 0, 0);
  b->boxes.assign(n ? (2 * (size_t)n - 1) * 6 : 0, 0.0f);
  b->tri.assign((size_t)n * 9, 0.0f);
  if (n == 0) return b;
  // centroid = centre (main.cpp:680-683 / accelerators.h:576)
  std::vector<float> cs((size_t)n * 3);
  for (uint32_t p = 0; p < n; ++p) cs[3 * p] = cx[p], cs[3 * p + 1] = cy[p], cs[3 * p + 2] = cz[p];
  codes_from_centroids(cs, n, morton_mode, b->codes.data());
  or_sort_pairs(b->codes.data(), n, b->sorted_codes.data(), b->sorted_idx.data());
  or_karras(b->sorted_codes.data(), n, b->left.data(), b->right.data());
  for (uint32_t k = 0; k < n; ++k) {
    uint32_t p = b->sorted_idx[k];
    float* bx = &b->boxes[6 * ((size_t)(n - 1) + k)];
    const float c[3] = {cx[p], cy[p], cz[p]};
    for (int a = 0; a < 3; ++a) bx[a] = (c[a] - r[p]) + 0.0f, bx[3 + a] = (c[a] + r[p]) + 0.0f;
    float* T = &b->tri[9 * (size_t)k];
    T[0] = c[0], T[1] = c[1], T[2] = c[2], T[3] = r[p], T[4] = r[p] * r[p];
  }
  refit_internal(n, b->left.data(), b->right.data(), b->boxes.data());
  b->height = or_tree_height(b->left.data(), b->right.data(), n);
  return b;
}

extern "C" void or_bvh_free(or_bvh* b) { delete b; }

extern "C" void or_bvh_arrays(const or_bvh* b, uint32_t* codes, uint32_t* sorted_idx,
                              int32_t* left, int32_t* right, float* boxes) {
  size_t n = b->n;
  if (codes) std::copy(b->codes.begin(), b->codes.end(), codes);
  if (sorted_idx) std::copy(b->sorted_idx.begin(), b->sorted_idx.end(), sorted_idx);
  if (left) std::copy(b->left.begin(), b->left.end(), left);
  if (right) std::copy(b->right.begin(), b->right.end(), right);
  if (boxes && n) std::copy(b->boxes.begin(), b->boxes.end(), boxes);
}

extern "C" uint32_t or_bvh_height(const or_bvh* b) { return b->height; }

// ---------------------------------------------------------------------------
// Treelet restructuring of a built LBVH: one round of the method of Karras &
// Aila, "Fast Parallel Construction of High-Quality Bounding Volume
// Hierarchies" (HPG 2013, section 4), with one primitive per leaf. Not part
// of the reference (its trees are never optimised); the GPU's optional
// rq_bvh_treelets restates this function and must match it node for node.
//
// Cost of a tree = sum of its internal nodes' half areas (dx*dy + dy*dz) +
// dz*dx (with one primitive per leaf the leaves' cost is the same for every
// arrangement). Every internal node with at least 7 leaves below it, children
// before parents, roots a treelet: starting from its two children, the
// treelet leaf with the largest half area (the first in slot order on ties)
// that is an internal node is replaced by its two children (the second one
// appended as the last slot) until there are 7. The optimal tree over the 7
// leaves comes from dynamic programming over subsets (treelet_dp); it
// replaces the treelet when its cost is strictly lower than the current
// one's and its height is not greater (so the tree height never grows and
// the traversal stack bound holds). The internal nodes are reused: the root
// keeps its index, the others are handed out in formation order to the new
// internal nodes in pre-order (a node before its first, then its second
// child), the child holding the lowest leaf slot first. Each node's outcome
// depends only on its own subtree, so any children-first schedule gives this
// tree.
namespace {
inline float half_area(const float* b) {
  const float dx = b[3] - b[0], dy = b[4] - b[1], dz = b[5] - b[2];
  return (dx * dy + dy * dz) + dz * dx;
}
struct TreeletDP {
  float A[128], C[128];
  uint8_t P[128], H[128];  // split (the part with the lowest slot), height
};
// Subsets in numeric order (every proper subset of s is smaller than s). For
// s with lowest slot `low` the splits are p = low | sub over the proper
// subsets sub of s - low, in decreasing sub; the first minimum of
// C[p] + C[s - p] wins (costs are >= +0, so this is the minimum of the pair
// (cost bits, 127 - sub)).
void treelet_dp(int m, const float (*box)[6], const uint8_t* hleaf, TreeletDP& d) {
  const int S = 1 << m;
  for (int s = 1; s < S; ++s) {
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int k = 0; k < m; ++k)
      if (s >> k & 1)
        for (int a = 0; a < 3; ++a) {
          b[a] = std::min(b[a], box[k][a]);
          b[3 + a] = std::max(b[3 + a], box[k][3 + a]);
        }
    d.A[s] = half_area(b);
    if ((s & (s - 1)) == 0) {
      const int k = __builtin_ctz(s);
      d.C[s] = 0.0f, d.P[s] = 0, d.H[s] = hleaf[k];
      continue;
    }
    const int low = s & -s, rest = s ^ low;
    float best = 0.0f;
    int bp = 0;
    for (int sub = (rest - 1) & rest;; sub = (sub - 1) & rest) {
      const int p = low | sub;
      const float c = d.C[p] + d.C[s ^ p];
      if (bp == 0 || c < best) best = c, bp = p;
      if (sub == 0) break;
    }
    d.C[s] = d.A[s] + best;
    d.P[s] = (uint8_t)bp;
    d.H[s] = (uint8_t)(1 + std::max(d.H[bp], d.H[s ^ bp]));
  }
}
}  // namespace

extern "C" float or_treelet_dp(const float* boxes, int m, uint8_t* split, uint8_t* height) {
  TreeletDP d;
  uint8_t h0[7] = {};
  treelet_dp(m, reinterpret_cast<const float(*)[6]>(boxes), h0, d);
  for (int s = 1; s < (1 << m); ++s) {
    if (split) split[s] = d.P[s];
    if (height) height[s] = d.H[s];
  }
  return d.C[(1 << m) - 1];
}

extern "C" uint32_t or_bvh_treelets(or_bvh* b) {
  const uint32_t n = b->n;
  if (n < 7) return 0;
  int32_t* left = b->left.data();
  int32_t* right = b->right.data();
  auto box = [&](int32_t c) -> float* {
    return &b->boxes[6 * (size_t)(c >= 0 ? c : (int32_t)(n - 1) + ~c)];
  };
  std::vector<uint32_t> cnt(n - 1), ht(n - 1);
  auto count = [&](int32_t c) { return c >= 0 ? cnt[c] : 1u; };
  auto height = [&](int32_t c) { return c >= 0 ? ht[c] : 0u; };
  // children-first order of the tree as built
  std::vector<int32_t> order;
  order.reserve(n - 1);
  std::vector<std::pair<int32_t, int>> st{{0, 0}};
  while (!st.empty()) {
    auto& t = st.back();
    const int32_t i = t.first;
    if (t.second == 0) {
      t.second = 1;
      if (left[i] >= 0) st.push_back({left[i], 0});
    } else if (t.second == 1) {
      t.second = 2;
      if (right[i] >= 0) st.push_back({right[i], 0});
    } else {
      order.push_back(i);
      st.pop_back();
    }
  }
  uint32_t rebuilt = 0;
  for (const int32_t i : order) {
    cnt[i] = count(left[i]) + count(right[i]);
    ht[i] = 1 + std::max(height(left[i]), height(right[i]));
    if (cnt[i] < 7) continue;
    int32_t L[7] = {left[i], right[i]}, I[6] = {i};
    int m = 2, ni = 1;
    while (m < 7) {
      int best = -1;
      float ba = 0.0f;
      for (int k = 0; k < m; ++k)
        if (L[k] >= 0) {
          const float a = half_area(box(L[k]));
          if (best < 0 || a > ba) best = k, ba = a;
        }
      const int32_t j = L[best];  // a node with >= 7 leaves always has one to open
      I[ni++] = j;
      L[best] = left[j];
      L[m++] = right[j];
    }
    float lb[7][6];
    uint8_t hl[7];
    for (int k = 0; k < 7; ++k) {
      std::copy(box(L[k]), box(L[k]) + 6, lb[k]);
      hl[k] = (uint8_t)height(L[k]);
    }
    TreeletDP d;
    treelet_dp(7, lb, hl, d);
    // cost of the treelet as it stands (the internal nodes in reverse
    // formation order: a node's internal children were opened after it)
    int mask[6];
    float cost[6];
    auto part = [&](int32_t c, int* mk, float* cc) {
      for (int k = 0; k < 7; ++k)
        if (L[k] == c) {
          *mk = 1 << k, *cc = 0.0f;
          return;
        }
      for (int j = 1; j < 6; ++j)
        if (I[j] == c) {
          *mk = mask[j], *cc = cost[j];
          return;
        }
    };
    for (int j = 5; j >= 0; --j) {
      int ml = 0, mr = 0;
      float cl = 0.0f, cr = 0.0f;
      part(left[I[j]], &ml, &cl);
      part(right[I[j]], &mr, &cr);
      mask[j] = ml | mr;
      cost[j] = d.A[mask[j]] + (cl + cr);
    }
    if (!(d.C[127] < cost[0]) || d.H[127] > ht[i]) continue;
    ++rebuilt;
    // rebuild in pre-order
    int next = 0;
    std::vector<int> stack{127};
    std::vector<int32_t> at(128, -1);
    while (!stack.empty()) {
      const int s = stack.back();
      stack.pop_back();
      at[s] = I[next++];
      const int p = d.P[s], q = s ^ p;
      if (q & (q - 1)) stack.push_back(q);
      if (p & (p - 1)) stack.push_back(p);
    }
    // links and boxes, children before parents (higher pre-order first)
    std::vector<int> subsets;
    for (int s = 1; s < 128; ++s)
      if (at[s] >= 0) subsets.push_back(s);
    std::sort(subsets.begin(), subsets.end(),
              [](int x, int y) { return __builtin_popcount(x) < __builtin_popcount(y); });
    auto link = [&](int s) { return (s & (s - 1)) ? at[s] : L[__builtin_ctz(s)]; };
    for (const int s : subsets) {
      const int32_t idx = at[s], l = link(d.P[s]), r = link(s ^ d.P[s]);
      left[idx] = l, right[idx] = r;
      float* bo = box(idx);
      const float *x = box(l), *y = box(r);
      for (int a = 0; a < 3; ++a) {
        bo[a] = std::min(x[a], y[a]);
        bo[3 + a] = std::max(x[3 + a], y[3 + a]);
      }
      cnt[idx] = count(l) + count(r);
      ht[idx] = d.H[s];
    }
  }
  b->height = or_tree_height(left, right, n);
  return rebuilt;
}

namespace {
// Slab test shared with the GPU kernels (same operations, same order); hit
// when near <= far * SLAB_SLACK (inclusive, so equal-t ties are never culled).
struct RayPre {
  V3 o, d, inv;
  bool nx, ny, nz;  // sign bit of 1/d per axis: near plane is hi when set
  float tmin, tmax;
};
inline RayPre prep(V3 o, V3 d, float tmin, float tmax) {
  RayPre r;
  r.o = o, r.d = d, r.tmin = tmin, r.tmax = tmax;
  r.inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
  r.nx = std::signbit(r.inv.x), r.ny = std::signbit(r.inv.y), r.nz = std::signbit(r.inv.z);
  return r;
}
// Relative slack on the far bound: a triangle's rounded Möller–Trumbore t can
// fall a few ulps outside the exact box of its vertices (hits on edges and
// vertices); culling only beyond far * (1 + 2^-15) keeps every box that holds
// the brute-force closest hit.
constexpr float SLAB_SLACK = 1.000030517578125f;
// Sphere scenes: the reference's sphere test computes d2 = l.l - tca^2, whose
// cancellation at a distance L moves a grazing hit's t by up to about
// L * 2^-11 (and can accept a ray just outside the box); a 2^-9 relative
// slack keeps every box whose sphere's computed t can still beat the best hit.
constexpr float SPHERE_SLACK = 1.001953125f;
// Near/far planes chosen by the sign of 1/d, t = (plane - o) * (1/d). A zero
// direction component gives +-inf, or NaN when the origin lies on the plane;
// the NaN-ignoring max (near) and min (far) then treat that axis as "inside",
// which is exact for a ray parallel to the slab.
inline bool slab(const RayPre& r, const float* b, float thi, float* tnear,
                 float slack = SLAB_SLACK) {
  float nx = ((r.nx ? b[3] : b[0]) - r.o.x) * r.inv.x, fx = ((r.nx ? b[0] : b[3]) - r.o.x) * r.inv.x;
  float ny = ((r.ny ? b[4] : b[1]) - r.o.y) * r.inv.y, fy = ((r.ny ? b[1] : b[4]) - r.o.y) * r.inv.y;
  float nz = ((r.nz ? b[5] : b[2]) - r.o.z) * r.inv.z, fz = ((r.nz ? b[2] : b[5]) - r.o.z) * r.inv.z;
  float tn = fmx(fmx(nx, ny), fmx(nz, r.tmin));
  float tf = fmn(fmn(fx, fy), fmn(fz, thi));
  *tnear = tn;
  return tn <= tf * slack;
}

}  // namespace

// NONE loop over spheres (main.cpp:376-386): index order, strictly smaller t.
// box_filter != 0 keeps only spheres whose box (centre -/+ r) the ray hits by
// the slab test (sphere slack) with the far bound at tmax: the candidates of the reference's
// BVH path (boxIntersect, main.cpp:343-358). The two differ only where the
// sphere test's d2 = l.l - tca^2 loses its precision to cancellation and
// accepts a ray that misses the sphere's box.
extern "C" void or_closest_brute_spheres(const float* cx, const float* cy, const float* cz,
                                         const float* r, uint32_t n, const or_rays* rays,
                                         uint32_t nrays, or_hits* hits, int box_filter,
                                         int nthreads) {
  parallel_for(nrays, nthreads, [&](uint32_t i) {
    V3 o, d;
    float tmin, tmax;
    load_ray(rays, i, &o, &d, &tmin, &tmax);
    const RayPre rp = prep(o, d, tmin, tmax);
    int32_t best = -1;
    float bt = tmax;
    for (uint32_t p = 0; p < n && tmin < tmax; ++p) {
      float t;
      if (!sphere_e({cx[p], cy[p], cz[p]}, r[p] * r[p], o, d, tmin, tmax, &t)) continue;
      if (box_filter) {
        const float b[6] = {(cx[p] - r[p]) + 0.0f, (cy[p] - r[p]) + 0.0f, (cz[p] - r[p]) + 0.0f,
                            (cx[p] + r[p]) + 0.0f, (cy[p] + r[p]) + 0.0f, (cz[p] + r[p]) + 0.0f};
        float tn;
        if (!slab(rp, b, tmax, &tn, SPHERE_SLACK)) continue;
      }
      if (best < 0 || t < bt) best = (int32_t)p, bt = t;
    }
    hits->prim[i] = best;
    if (hits->t) hits->t[i] = best >= 0 ? bt : INFINITY;
    if (hits->u) hits->u[i] = 0.0f;
    if (hits->v) hits->v[i] = 0.0f;
  });
}

namespace {
// Canonical traversal: start at internal node 0; for an internal node load
// both child records (2 node visits), cull each against [tmin, best t], go to
// the nearer hit child and push the other (ties go left); a leaf is one
// triangle test. Closest hit keeps the (t, primitive id) lexicographic minimum;
// any-hit stops at the first valid t.
template <bool ANY>
int32_t trace(const or_bvh* b, const RayPre& r, float* bt, float* bu, float* bv, uint32_t* vis,
              uint32_t* tst, uint32_t* max_stack = nullptr) {
  uint32_t n = b->n;
  int32_t best = -1;
  float t_best = r.tmax, u_best = 0, v_best = 0;
  uint32_t visits = 0, tests = 0;
  if (n == 0) {
    *bt = INFINITY;
    return -1;
  }
  int32_t stack[128];
  int sp = 0, sp_max = 0;
  int32_t node = n == 1 ? ~0 : 0;
  for (;;) {
    if (node >= 0) {
      visits += 2;
      int32_t c0 = b->left[node], c1 = b->right[node];
      const float* b0 = &b->boxes[6 * (size_t)(c0 >= 0 ? c0 : (int32_t)(n - 1) + ~c0)];
      const float* b1 = &b->boxes[6 * (size_t)(c1 >= 0 ? c1 : (int32_t)(n - 1) + ~c1)];
      float tn0, tn1;
      const float slack = b->kind == 1 ? SPHERE_SLACK : SLAB_SLACK;
      bool h0 = slab(r, b0, t_best, &tn0, slack), h1 = slab(r, b1, t_best, &tn1, slack);
      if (h0 && h1) {
        if (tn1 < tn0) {
          stack[sp++] = c0;
          node = c1;
        } else {
          stack[sp++] = c1;
          node = c0;
        }
        sp_max = std::max(sp_max, sp);
        continue;
      } else if (h0) {
        node = c0;
        continue;
      } else if (h1) {
        node = c1;
        continue;
      }
    } else {
      uint32_t k = (uint32_t)~node;
      ++tests;
      const float* T = &b->tri[9 * (size_t)k];
      float t, u = 0, v = 0;
      const bool hit = b->kind == 1
                           ? sphere_e({T[0], T[1], T[2]}, T[4], r.o, r.d, r.tmin, r.tmax, &t)
                           : mt_e({T[0], T[1], T[2]}, {T[3], T[4], T[5]}, {T[6], T[7], T[8]}, r.o,
                                  r.d, r.tmin, r.tmax, &t, &u, &v);
      if (hit) {
        int32_t prim = (int32_t)b->sorted_idx[k];
        if (ANY) {
          best = prim;
          t_best = t;
          break;
        }
        if (best < 0 || t < t_best || (t == t_best && prim < best)) {
          best = prim, t_best = t, u_best = u, v_best = v;
        }
      }
    }
    if (sp == 0) break;
    node = stack[--sp];
  }
  *bt = best >= 0 ? t_best : INFINITY;
  *bu = u_best, *bv = v_best;
  if (vis) *vis = visits;
  if (tst) *tst = tests;
  if (max_stack) *max_stack = (uint32_t)sp_max;
  return best;
}
}  // namespace

// The deepest stack of each ray's canonical closest-hit walk (entries held at
// once: the far child is pushed when both children's boxes are hit).
extern "C" void or_closest_bvh_stack(const or_bvh* b, const or_rays* rays, uint32_t nrays,
                                     uint32_t* max_stack, int nthreads) {
  parallel_for(nrays, nthreads, [&](uint32_t i) {
    V3 o, d;
    float tmin, tmax, t, u, v;
    load_ray(rays, i, &o, &d, &tmin, &tmax);
    max_stack[i] = 0;
    if (tmin < tmax)
      trace<false>(b, prep(o, d, tmin, tmax), &t, &u, &v, nullptr, nullptr, &max_stack[i]);
  });
}

extern "C" void or_closest_bvh(const or_bvh* b, const or_rays* rays, uint32_t nrays, or_hits* hits,
                               uint32_t* visits, uint32_t* tests, int nthreads) {
  parallel_for(nrays, nthreads, [&](uint32_t i) {
    V3 o, d;
    float tmin, tmax;
    load_ray(rays, i, &o, &d, &tmin, &tmax);
    float t, u, v;
    uint32_t vi = 0, te = 0;
    int32_t p = -1;
    if (tmin < tmax)
      p = trace<false>(b, prep(o, d, tmin, tmax), &t, &u, &v, &vi, &te);
    else
      t = INFINITY, u = v = 0;
    hits->prim[i] = p;
    if (hits->t) hits->t[i] = t;
    if (hits->u) hits->u[i] = u;
    if (hits->v) hits->v[i] = v;
    if (visits) visits[i] = vi;
    if (tests) tests[i] = te;
  });
}

extern "C" void or_any_bvh(const or_bvh* b, const or_rays* rays, uint32_t nrays, uint8_t* occ,
                           uint32_t* visits, uint32_t* tests, int nthreads) {
  parallel_for(nrays, nthreads, [&](uint32_t i) {
    V3 o, d;
    float tmin, tmax;
    load_ray(rays, i, &o, &d, &tmin, &tmax);
    float t, u, v;
    uint32_t vi = 0, te = 0;
    int32_t p = -1;
    if (tmin < tmax) p = trace<true>(b, prep(o, d, tmin, tmax), &t, &u, &v, &vi, &te);
    occ[i] = p >= 0 ? 1 : 0;
    if (visits) visits[i] = vi;
    if (tests) tests[i] = te;
  });
}

// ---------------------------------------------------------------------------
// Random numbers and ray generation.
// ---------------------------------------------------------------------------
extern "C" void or_mt19937_raw(uint32_t n, uint32_t* out) {
  std::mt19937 g;  // default seed 5489, as main.cpp:506
  for (uint32_t i = 0; i < n; ++i) out[i] = g();
}

extern "C" void or_mt19937_jitter(uint32_t n, double* out) {
  // main.cpp:503-508: function-local statics, default-seeded.
  std::uniform_real_distribution<double> dist(0.0, 1.0);
  std::mt19937 g;
  for (uint32_t i = 0; i < n; ++i) out[i] = dist(g);
}

// Philox4x32-10 (Salmon et al., SC'11; Random123 reference constants).
extern "C" void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) k0 += 0x9E3779B9u, k1 += 0xBB67AE85u;
    uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0, c1 = n1, c2 = n2, c3 = n3;
  }
  out[0] = c0, out[1] = c1, out[2] = c2, out[3] = c3;
}

namespace {
inline void philox(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t o[4]) {
  uint32_t ctr[4] = {a, b, c, d}, key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  or_philox4x32_10(ctr, key, o);
}
inline double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
inline float u24(uint32_t a) { return (float)(a >> 8) * (1.0f / 16777216.0f); }

// main.cpp:544-557 (fov hard-coded 30, origin 0, direction (xx, yy, -1)
// normalised). The double/float mix follows the C++ promotions of the
// reference expression: (x + r) is double, invWidth/angle/aspect are floats
// promoted to double, the result is narrowed to float.
struct Cam {
  float invW, invH, aspect, angle;
};
inline Cam cam_setup(uint32_t w, uint32_t h) {
  Cam c;
  c.invW = 1 / float(w);
  c.invH = 1 / float(h);
  float fov = 30;
  c.aspect = w / float(h);
  c.angle = (float)std::tan(3.141592653589793 * 0.5 * fov / 180.);
  return c;
}
inline V3 cam_dir(const Cam& c, uint32_t x, uint32_t y, double r1, double r2) {
  float xx = (float)((2 * ((x + r1) * (double)c.invW) - 1) * (double)c.angle * (double)c.aspect);
  float yy = (float)((1 - 2 * ((y + r2) * (double)c.invH)) * (double)c.angle);
  return normalize({xx, yy, -1.0f});
}
}  // namespace

extern "C" void or_camera_rays_ref(uint32_t w, uint32_t h, uint32_t spp, const double* jit,
                                   float* ox, float* oy, float* oz, float* dx, float* dy,
                                   float* dz, float* tmin, float* tmax) {
  Cam c = cam_setup(w, h);
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t x = 0; x < w; ++x)
      for (uint32_t s = 0; s < spp; ++s) {
        size_t i = ((size_t)y * w + x) * spp + s;
        V3 d = cam_dir(c, x, y, jit[2 * i], jit[2 * i + 1]);
        ox[i] = oy[i] = oz[i] = 0.0f;
        dx[i] = d.x, dy[i] = d.y, dz[i] = d.z;
        tmin[i] = 0.0f, tmax[i] = INFINITY;
      }
}

extern "C" void or_camera_rays_philox(uint32_t w, uint32_t h, uint32_t spp, uint64_t seed,
                                      float* ox, float* oy, float* oz, float* dx, float* dy,
                                      float* dz, float* tmin, float* tmax) {
  Cam c = cam_setup(w, h);
  for (uint32_t y = 0; y < h; ++y)
    for (uint32_t x = 0; x < w; ++x)
      for (uint32_t s = 0; s < spp; ++s) {
        size_t i = ((size_t)y * w + x) * spp + s;
        uint32_t o[4];
        philox(seed, (uint32_t)i, (uint32_t)(i >> 32), 0u, 0x43414D31u, o);
        V3 d = cam_dir(c, x, y, u53(o[0], o[1]), u53(o[2], o[3]));
        ox[i] = oy[i] = oz[i] = 0.0f;
        dx[i] = d.x, dy[i] = d.y, dz[i] = d.z;
        tmin[i] = 0.0f, tmax[i] = INFINITY;
      }
}

namespace {
// Ray i of the uniform generator: origin uniform in [lo, hi) by Philox words
// (seed, counter (i, 0, 0, 'ORIG')); direction uniform on S^2 by rejection
// from the cube [-1,1)^3, at most 64 attempts (counter word 1 = attempt,
// word 3 'DIR!'), then +z. r2 summed as ((x*x + y*y) + z*z).
inline void uniform_ray(uint32_t i, uint64_t seed, const float lo[3], const float ext[3], V3* o3,
                        V3* d3) {
  uint32_t o[4];
  philox(seed, i, 0u, 0u, 0x4F524947u, o);
  *o3 = {lo[0] + ext[0] * u24(o[0]), lo[1] + ext[1] * u24(o[1]), lo[2] + ext[2] * u24(o[2])};
  V3 d{0.0f, 0.0f, 1.0f};
  for (uint32_t a = 0; a < 64; ++a) {
    philox(seed, i, a, 1u, 0x44495221u, o);
    float x = u24(o[0]) * 2.0f - 1.0f, y = u24(o[1]) * 2.0f - 1.0f, z = u24(o[2]) * 2.0f - 1.0f;
    float r2 = x * x + y * y + z * z;
    if (r2 > 1e-6f && r2 <= 1.0f) {
      float f = 1 / std::sqrt(r2);
      d = {x * f, y * f, z * f};
      break;
    }
  }
  *d3 = d;
}
}  // namespace

extern "C" void or_uniform_rays(uint32_t n, uint64_t seed, const float lo[3], const float hi[3],
                                float* ox, float* oy, float* oz, float* dx, float* dy, float* dz,
                                float* tmin, float* tmax) {
  or_uniform_rays_at(nullptr, n, seed, lo, hi, ox, oy, oz, dx, dy, dz, tmin, tmax);
}

extern "C" void or_uniform_rays_at(const uint32_t* ids, uint32_t n, uint64_t seed,
                                   const float lo[3], const float hi[3], float* ox, float* oy,
                                   float* oz, float* dx, float* dy, float* dz, float* tmin,
                                   float* tmax) {
  const float ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  for (uint32_t k = 0; k < n; ++k) {
    V3 o, d;
    uniform_ray(ids ? ids[k] : k, seed, lo, ext, &o, &d);
    ox[k] = o.x, oy[k] = o.y, oz[k] = o.z;
    dx[k] = d.x, dy[k] = d.y, dz[k] = d.z;
    tmin[k] = 0.0f, tmax[k] = INFINITY;
  }
}

// The serving form's pinhole camera (no reference counterpart: the reference
// renders one fixed camera per process, main.cpp:541-558). Basis on the host:
// f = normalize(look_at - eye), r = normalize(f x up), u = r x f, each
// normalised by dividing by the float sqrt of ((x*x + y*y) + z*z). Image
// w = max(1, width) by h = ceil(n / w); ray i goes through pixel
// (i % w, i / w) jittered by the first two words of Philox(seed,
// (i, 0, 0, 'PINH')) as 24-bit fractions:
//   sx = ((2 * ((x + j0) * (1/w)) - 1) * tan_half) * (w/h),
//   sy = (1 - 2 * ((y + j1) * (1/h))) * tan_half,
//   d  = normalize((f + sx*r) + sy*u)  (per component, double-sqrt normalize),
// tan_half = tanf(0.5f * fov_deg * 3.14159265358979f / 180.0f) (float, left to right).
extern "C" void or_pinhole_rays_at(const uint32_t* ids, uint32_t count, uint32_t n,
                                   uint64_t seed, const float eye[3], const float look_at[3],
                                   const float up_in[3], float fov_deg, uint32_t width,
                                   float* ox, float* oy, float* oz, float* dx, float* dy,
                                   float* dz, float* tmin, float* tmax) {
  auto nrm = [](float* v) {
    const float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 0.0f) v[0] /= l, v[1] /= l, v[2] /= l;
  };
  float f[3] = {look_at[0] - eye[0], look_at[1] - eye[1], look_at[2] - eye[2]};
  nrm(f);
  float r[3] = {f[1] * up_in[2] - f[2] * up_in[1], f[2] * up_in[0] - f[0] * up_in[2],
                f[0] * up_in[1] - f[1] * up_in[0]};
  nrm(r);
  const float u[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2],
                      r[0] * f[1] - r[1] * f[0]};
  const uint32_t w = std::max<uint32_t>(1, width);
  const uint32_t h = std::max<uint32_t>(1, (uint32_t)(((uint64_t)n + w - 1) / w));
  const float tan_half = std::tan(0.5f * fov_deg * 3.14159265358979f / 180.0f);
  const float aspect = (float)w / (float)h, inv_w = 1.0f / (float)w, inv_h = 1.0f / (float)h;
  for (uint32_t k = 0; k < count; ++k) {
    const uint32_t i = ids ? ids[k] : k;
    const uint32_t x = i % w, y = i / w;
    uint32_t o[4];
    philox(seed, i, 0u, 0u, 0x50494E48u, o);
    const float sx = (2.0f * (((float)x + u24(o[0])) * inv_w) - 1.0f) * tan_half * aspect;
    const float sy = (1.0f - 2.0f * (((float)y + u24(o[1])) * inv_h)) * tan_half;
    const V3 d = normalize({(f[0] + sx * r[0]) + sy * u[0], (f[1] + sx * r[1]) + sy * u[1],
                            (f[2] + sx * r[2]) + sy * u[2]});
    ox[k] = eye[0], oy[k] = eye[1], oz[k] = eye[2];
    dx[k] = d.x, dy[k] = d.y, dz[k] = d.z;
    tmin[k] = 0.0f, tmax[k] = INFINITY;
  }
}

// ---------------------------------------------------------------------------
// Shading: castRay's DIFFUSE_AND_GLOSSY branch (main.cpp:447-492) with a real
// shadow any-hit in place of the trace_more stub (:235-245), and an optional
// mirror bounce weighted by fresnel (:85-103, ior 2 as :444), reflect :218-221.
// ---------------------------------------------------------------------------
namespace {
inline float clampf(float lo, float hi, float v) { return std::max(lo, std::min(hi, v)); }
inline float fresnel(V3 I, V3 N, float ior) {
  float cosi = clampf(-1, 1, dot(I, N));
  float etai = 1, etat = ior;
  if (cosi > 0) std::swap(etai, etat);
  float sint = etai / etat * std::sqrt(std::max(0.f, 1 - cosi * cosi));
  if (sint >= 1) return 1;
  float cost = std::sqrt(std::max(0.f, 1 - sint * sint));
  cosi = std::fabs(cosi);
  float Rs = ((etat * cosi) - (etai * cost)) / ((etat * cosi) + (etai * cost));
  float Rp = ((etai * cosi) - (etat * cost)) / ((etai * cosi) + (etat * cost));
  return (Rs * Rs + Rp * Rp) / 2;
}
inline V3 reflect(V3 I, V3 N) {
  float s = 2 * dot(I, N);
  return sub(I, mul(N, s));
}
// refract (main.cpp:223-233): eta * I + (eta * cosi - sqrt(k)) * n, or 0 on
// total internal reflection.
inline V3 refract(V3 I, V3 N, float ior) {
  float cosi = clampf(-1, 1, dot(I, N));
  float etai = 1, etat = ior;
  V3 n = N;
  if (cosi < 0) {
    cosi = -cosi;
  } else {
    std::swap(etai, etat);
    n = {-N.x, -N.y, -N.z};
  }
  float eta = etai / etat;
  float k = 1 - eta * eta * (1 - cosi * cosi);
  if (k < 0) return {0, 0, 0};
  return add(mul(I, eta), mul(n, eta * cosi - std::sqrt(k)));
}

struct HitGeo {
  V3 p, n;  // hit point, normal facing against the ray
};
inline HitGeo hit_geo(const or_bvh* b, uint32_t leaf, V3 o, V3 d, float t) {
  const float* T = &b->tri[9 * (size_t)leaf];
  HitGeo g;
  g.p = add(o, mul(d, t));
  if (b->kind == 1)
    g.n = normalize(sub(g.p, {T[0], T[1], T[2]}));  // main.cpp:397-398
  else
    g.n = normalize(cross({T[3], T[4], T[5]}, {T[6], T[7], T[8]}));
  if (dot(d, g.n) > 0) g.n = {-g.n.x, -g.n.y, -g.n.z};
  return g;
}
inline V3 shadow_origin(V3 d, const HitGeo& g, float bias) {
  return dot(d, g.n) < 0 ? add(g.p, mul(g.n, bias)) : sub(g.p, mul(g.n, bias));
}
// Direct light at a hit: one point light, Lambert + Phong^25, diffuse
// constant getDiffuseColor(0.2) = (0.815, 0.235, 0.031) (accelerators.h:94-99).
inline V3 direct(V3 d, const HitGeo& g, const or_shade_params* sp, int32_t prim,
                 bool in_shadow) {
  V3 L{sp->light[0], sp->light[1], sp->light[2]}, E{sp->emission[0], sp->emission[1],
                                                    sp->emission[2]};
  V3 ld = sub(L, g.p);
  ld = normalize(ld);
  float LdotN = std::max(0.f, dot(ld, g.n));
  float vis = (float)(1 - (int)in_shadow);
  V3 amt{(E.x * vis) * LdotN, (E.y * vis) * LdotN, (E.z * vis) * LdotN};
  V3 R = reflect({-ld.x, -ld.y, -ld.z}, g.n);
  float sp25 = std::pow(std::max(0.f, -dot(R, d)), 25.0f);
  const float diff[3] = {0.815f * 0.8f, 0.235f * 0.8f, 0.031f * 0.8f};
  V3 c;
  // sphere->surfaceColor (main.cpp:490): per primitive when given
  const float* S = sp->prim_rgb ? sp->prim_rgb + 3 * (size_t)prim : sp->surface;
  c.x = ((amt.x * diff[0]) / 2.0f + (sp25 * E.x) * 0.5f) + S[0];
  c.y = ((amt.y * diff[1]) / 2.0f + (sp25 * E.y) * 0.5f) + S[1];
  c.z = ((amt.z * diff[2]) / 2.0f + (sp25 * E.z) * 0.5f) + S[2];
  return c;
}
inline RayPre shadow_ray(V3 d, const HitGeo& g, const or_shade_params* sp) {
  V3 L{sp->light[0], sp->light[1], sp->light[2]};
  V3 so = shadow_origin(d, g, sp->bias);
  V3 ld = sub(L, g.p);
  float dist2 = dot(ld, ld);
  ld = normalize(ld);
  return prep(so, ld, 0.0f, std::sqrt(dist2));
}
}  // namespace

extern "C" void or_render_samples(const or_bvh* b, const float* vx, const float* vy,
                                  const float* vz, const uint32_t* idx, const or_rays* prim,
                                  uint32_t nrays, const or_shade_params* sp, float* rgb,
                                  int32_t* hit0, int32_t* hit1, uint8_t* sh0, uint8_t* sh1,
                                  uint64_t* counts, int nthreads) {
  (void)vx, (void)vy, (void)vz, (void)idx;
  std::vector<uint32_t> leaf(b->n);
  for (uint32_t k = 0; k < b->n; ++k) leaf[b->sorted_idx[k]] = k;
  // counts: [0] rays traced, [1] closest-hit node visits, [2] closest-hit tests,
  //         [3] any-hit node visits, [4] any-hit tests, [5] closest rays, [6] any rays
  std::vector<uint64_t> cnt((size_t)nthreads * 8, 0);
  int nt = std::max(1, nthreads);
  std::vector<std::thread> th;
  uint32_t chunk = (nrays + nt - 1) / nt;
  auto body = [&](int tid, uint32_t beg, uint32_t end) {
    uint64_t* C = &cnt[(size_t)tid * 8];
    for (uint32_t i = beg; i < end; ++i) {
      V3 o, d;
      float tmin, tmax;
      load_ray(prim, i, &o, &d, &tmin, &tmax);
      float t, u, v;
      uint32_t vi, te;
      int32_t p0 = -1;
      if (sp->mode != 1) {
        p0 = trace<false>(b, prep(o, d, tmin, tmax), &t, &u, &v, &vi, &te);
        C[0]++, C[1] += vi, C[2] += te, C[5]++;
      }
      int32_t p1 = -1;
      uint8_t s0 = 0, s1 = 0;
      V3 col{sp->sky[0], sp->sky[1], sp->sky[2]};
      const V3 sky = col;
      auto mat = [&](int32_t p) -> int { return sp->prim_mat ? sp->prim_mat[p] : 0; };
      if (sp->mode == 1) {
        // castRay's KDTREE case: black on any hit of the primary ray, else sky
        float st, su, sv;
        s0 = trace<true>(b, prep(o, d, tmin, tmax), &st, &su, &sv, &vi, &te) >= 0;
        C[0]++, C[3] += vi, C[4] += te, C[6]++;
        if (s0) col = {0, 0, 0};
      } else if (p0 >= 0 && mat(p0) != 0) {
        // REFLECTION_AND_REFRACTION / REFLECTION (main.cpp:418-446): the
        // reflected colour is traced by the reference but never used, so it is
        // not traced here; fresnel takes ior 2, refract ior 3, as written there
        HitGeo g = hit_geo(b, leaf[p0], o, d, t);
        const float kr = fresnel(d, g.n, 2.0f);
        if (mat(p0) == 2) {
          col = {1 - kr, 1 - kr, 1 - kr};
        } else {
          V3 rd = normalize(refract(d, g.n, 3.0f));
          V3 ro = dot(rd, g.n) < 0 ? sub(g.p, mul(g.n, sp->bias)) : add(g.p, mul(g.n, sp->bias));
          float t1, u1, v1;
          p1 = trace<false>(b, prep(ro, rd, 0.0f, INFINITY), &t1, &u1, &v1, &vi, &te);
          C[0]++, C[1] += vi, C[2] += te, C[5]++;
          V3 rc = sky;  // depth 2 (main.cpp:311: beyond depth 2 the sky)
          if (p1 >= 0) {
            HitGeo g1 = hit_geo(b, leaf[p1], ro, rd, t1);
            if (mat(p1) == 0) {
              if (sp->shadows) {
                float st, su, sv;
                RayPre sr1 = shadow_ray(rd, g1, sp);
                s1 = trace<true>(b, sr1, &st, &su, &sv, &vi, &te) >= 0;
                C[0]++, C[3] += vi, C[4] += te, C[6]++;
              }
              rc = direct(rd, g1, sp, p1, s1);
            } else {
              const float kr1 = fresnel(rd, g1.n, 2.0f);
              rc = mat(p1) == 2 ? V3{1 - kr1, 1 - kr1, 1 - kr1}
                                : V3{sky.x * (1 - kr1), sky.y * (1 - kr1), sky.z * (1 - kr1)};
            }
          }
          const float w = 1 - kr;
          col = {rc.x * w, rc.y * w, rc.z * w};
        }
      } else if (p0 >= 0) {
        HitGeo g = hit_geo(b, leaf[p0], o, d, t);
        float st, su, sv;
        if (sp->shadows) {
          RayPre sr = shadow_ray(d, g, sp);
          s0 = trace<true>(b, sr, &st, &su, &sv, &vi, &te) >= 0;
          C[0]++, C[3] += vi, C[4] += te, C[6]++;
        }
        col = direct(d, g, sp, p0, s0);
        if (sp->bounce) {
          V3 rd = normalize(reflect(d, g.n));
          V3 ro = dot(rd, g.n) < 0 ? sub(g.p, mul(g.n, sp->bias)) : add(g.p, mul(g.n, sp->bias));
          float kr = fresnel(d, g.n, 2.0f);
          float t1, u1, v1;
          p1 = trace<false>(b, prep(ro, rd, 0.0f, INFINITY), &t1, &u1, &v1, &vi, &te);
          C[0]++, C[1] += vi, C[2] += te, C[5]++;
          V3 bc{sp->sky[0], sp->sky[1], sp->sky[2]};
          if (p1 >= 0) {
            HitGeo g1 = hit_geo(b, leaf[p1], ro, rd, t1);
            if (sp->shadows) {
              RayPre sr1 = shadow_ray(rd, g1, sp);
              s1 = trace<true>(b, sr1, &st, &su, &sv, &vi, &te) >= 0;
              C[0]++, C[3] += vi, C[4] += te, C[6]++;
            }
            bc = direct(rd, g1, sp, p1, s1);
          }
          col = {col.x + kr * bc.x, col.y + kr * bc.y, col.z + kr * bc.z};
        }
      }
      if (rgb) rgb[3 * (size_t)i] = col.x, rgb[3 * (size_t)i + 1] = col.y,
                   rgb[3 * (size_t)i + 2] = col.z;
      if (hit0) hit0[i] = p0;
      if (hit1) hit1[i] = p1;
      if (sh0) sh0[i] = s0;
      if (sh1) sh1[i] = s1;
    }
  };
  for (int t = 0; t < nt; ++t) {
    uint32_t b0 = t * chunk, e0 = std::min<uint32_t>(nrays, b0 + chunk);
    if (b0 >= e0) break;
    th.emplace_back(body, t, b0, e0);
  }
  for (auto& x : th) x.join();
  if (counts) {
    for (int k = 0; k < 8; ++k) counts[k] = 0;
    for (int t = 0; t < nt; ++t)
      for (int k = 0; k < 8; ++k) counts[k] += cnt[(size_t)t * 8 + k];
  }
}

extern "C" void or_resolve_ppm(const float* rgb, uint32_t npix, uint32_t spp, uint8_t* bytes) {
  // Sum the spp samples of each pixel in sample order (main.cpp:551-560),
  // then (unsigned char)(min(1, c / aa) * 255) per channel (:521-523).
  for (uint32_t p = 0; p < npix; ++p) {
    float s[3] = {0, 0, 0};
    for (uint32_t k = 0; k < spp; ++k)
      for (int c = 0; c < 3; ++c) s[c] += rgb[3 * ((size_t)p * spp + k) + c];
    for (int c = 0; c < 3; ++c)
      bytes[3 * (size_t)p + c] = (uint8_t)(std::min(float(1), s[c] / spp) * 255);
  }
}
