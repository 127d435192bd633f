// Ray generation and shading kernels of the wavefront renderer (gfx950).
//
// Replace the per-pixel loop of render() (main.cpp:549-564) and the shading
// part of castRay (main.cpp:394-497). Arithmetic follows the reference's
// operation order (Vec3 ops of geometry.h:84-134, no contraction); only
// powf (specular, main.cpp:475) is not correctly rounded on either side.
#include <algorithm>

#include "rq/rq_render.h"
#include "rq_render_internal.h"

namespace rq {
namespace {

struct V3 {
  float x, y, z;
};
__device__ inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Vec3::normalize (geometry.h:125-134): its unqualified sqrt inside the
// template resolves to the C double sqrt, so factor = (float)(1.0 / sqrt(n))
// is computed in double (one rounding to float), then multiplied in float.
__device__ inline V3 normalize(V3 a) {
  float n = (a.x * a.x + a.y * a.y) + a.z * a.z;
  if (n > 0.0f) {
    const float f = (float)(1.0 / sqrt((double)n));
    a.x *= f, a.y *= f, a.z *= f;
  }
  return a;
}
// std::max(0.f, x) / std::min semantics (comparison-based, not fmaxf).
__device__ inline float max0(float x) { return (0.0f < x) ? x : 0.0f; }

// ---- Philox4x32-10 (Salmon et al. 2011) -------------------------------------
__device__ inline void philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint64_t seed,
                              uint32_t out[4]) {
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0, c1 = lo1, c2 = n2, c3 = lo0;
    k0 += 0x9E3779B9u, k1 += 0xBB67AE85u;
  }
  out[0] = c0, out[1] = c1, out[2] = c2, out[3] = c3;
}
__device__ inline double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
__device__ inline float u24(uint32_t a) { return (float)(a >> 8) * (1.0f / 16777216.0f); }

// ---- ray generation ---------------------------------------------------------
// main.cpp:554-557 with the C++ promotions of the reference expression:
// (x + r) is double; invWidth, angle, aspectratio are floats widened to
// double; the product is narrowed to float; then Vec3::normalize in float.
__global__ __launch_bounds__(256) void k_raygen_camera(CamParams cp, const double* __restrict__ jit,
                                                        float* ox, float* oy, float* oz, float* dx,
                                                        float* dy, float* dz, float* tmin,
                                                        float* tmax) {
  const uint64_t n = (uint64_t)cp.w * cp.h * cp.spp;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t pix = i / cp.spp;
  const uint32_t x = (uint32_t)(pix % cp.w), y = (uint32_t)(pix / cp.w);
  double r1, r2;
  if (jit) {
    r1 = jit[2 * i], r2 = jit[2 * i + 1];
  } else {
    uint32_t o[4];
    philox((uint32_t)i, (uint32_t)(i >> 32), 0u, 0x43414D31u, cp.seed, o);
    r1 = u53(o[0], o[1]), r2 = u53(o[2], o[3]);
  }
  const float xx =
      (float)((2.0 * (((double)x + r1) * (double)cp.invW) - 1.0) * (double)cp.angle * (double)cp.aspect);
  const float yy = (float)((1.0 - 2.0 * (((double)y + r2) * (double)cp.invH)) * (double)cp.angle);
  const V3 d = normalize({xx, yy, -1.0f});
  // null origin / interval arrays: the caller keeps shared constant ones
  // (every camera ray starts at the origin with [0, inf))
  if (ox) ox[i] = 0.0f, oy[i] = 0.0f, oz[i] = 0.0f;
  dx[i] = d.x, dy[i] = d.y, dz[i] = d.z;
  if (tmin) tmin[i] = 0.0f, tmax[i] = INFINITY;
}

__global__ __launch_bounds__(256) void k_raygen_uniform(uint32_t n, uint64_t seed, float lox,
                                                         float loy, float loz, float ex, float ey,
                                                         float ez, float* ox, float* oy, float* oz,
                                                         float* dx, float* dy, float* dz,
                                                         float* tmin, float* tmax) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t o[4];
  philox(i, 0u, 0u, 0x4F524947u, seed, o);
  ox[i] = lox + ex * u24(o[0]);
  oy[i] = loy + ey * u24(o[1]);
  oz[i] = loz + ez * u24(o[2]);
  V3 d{0.0f, 0.0f, 1.0f};
  for (uint32_t a = 0; a < 64; ++a) {
    philox(i, a, 1u, 0x44495221u, seed, o);
    const float x = u24(o[0]) * 2.0f - 1.0f, y = u24(o[1]) * 2.0f - 1.0f,
                z = u24(o[2]) * 2.0f - 1.0f;
    const float r2 = (x * x + y * y) + z * z;
    if (r2 > 1e-6f && r2 <= 1.0f) {
      const float f = 1.0f / sqrtf(r2);
      d = {x * f, y * f, z * f};
      break;
    }
  }
  dx[i] = d.x, dy[i] = d.y, dz[i] = d.z;
  tmin[i] = 0.0f, tmax[i] = INFINITY;
}

// Pinhole camera of the serving requests: ray i through pixel (i % w, i / w)
// of a w x h image, jittered by Philox(seed, i); d = normalize(f + sx*r + sy*u)
// with the camera basis computed on the host.
__global__ __launch_bounds__(256) void k_raygen_pinhole(PinholeParams p, float* ox, float* oy,
                                                         float* oz, float* dx, float* dy,
                                                         float* dz, float* tmin, float* tmax) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  const uint32_t x = i % p.w, y = i / p.w;
  uint32_t o[4];
  philox(i, 0u, 0u, 0x50494E48u, p.seed, o);
  const float sx = (2.0f * (((float)x + u24(o[0])) * p.inv_w) - 1.0f) * p.tan_half * p.aspect;
  const float sy = (1.0f - 2.0f * (((float)y + u24(o[1])) * p.inv_h)) * p.tan_half;
  const V3 d = normalize({(p.f[0] + sx * p.r[0]) + sy * p.u[0], (p.f[1] + sx * p.r[1]) + sy * p.u[1],
                          (p.f[2] + sx * p.r[2]) + sy * p.u[2]});
  ox[i] = p.eye[0], oy[i] = p.eye[1], oz[i] = p.eye[2];
  dx[i] = d.x, dy[i] = d.y, dz[i] = d.z;
  tmin[i] = 0.0f, tmax[i] = INFINITY;
}

// ---- shading ---------------------------------------------------------------
struct Geo {
  V3 p, n;
};
__device__ inline V3 vtx(const SceneView& s, uint32_t i) { return {s.vx[i], s.vy[i], s.vz[i]}; }
// Hit point o + d*t and the normal turned to face against the ray
// (main.cpp:396-406): a sphere's normalize(p - centre), a triangle's
// geometric normal normalize(e1 x e2).
__device__ inline Geo hit_geo(const SceneView& s, int32_t prim, V3 o, V3 d, float t) {
  Geo g;
  g.p = add(o, mul(d, t));
  if (s.r) {
    g.n = normalize(sub(g.p, vtx(s, (uint32_t)prim)));
  } else {
    const uint32_t a = s.idx[3 * prim], b = s.idx[3 * prim + 1], c = s.idx[3 * prim + 2];
    const V3 v0 = vtx(s, a), v1 = vtx(s, b), v2 = vtx(s, c);
    g.n = normalize(cross(sub(v1, v0), sub(v2, v0)));
  }
  if (dot(d, g.n) > 0.0f) g.n = {-g.n.x, -g.n.y, -g.n.z};
  return g;
}
// sphere->surfaceColor (main.cpp:490): per primitive when the scene has colours.
__device__ inline V3 surface_of(const SceneView& s, const ShadeParams& sp, int32_t prim) {
  if (s.rgb) return {s.rgb[3 * prim], s.rgb[3 * prim + 1], s.rgb[3 * prim + 2]};
  return {sp.surface[0], sp.surface[1], sp.surface[2]};
}
__device__ inline V3 reflect(V3 I, V3 N) { return sub(I, mul(N, 2.0f * dot(I, N))); }
__device__ inline int mat_of(const SceneView& s, int32_t prim) { return s.mat ? s.mat[prim] : 0; }
// main.cpp:85-103
__device__ inline float fresnel(V3 I, V3 N, float ior) {
  float c = dot(I, N);
  c = (c < 1.0f) ? c : 1.0f;
  c = (-1.0f < c) ? c : -1.0f;
  float etai = 1.0f, etat = ior;
  if (c > 0.0f) {
    float tmp = etai;
    etai = etat, etat = tmp;
  }
  const float sint = etai / etat * sqrtf(max0(1.0f - c * c));
  if (sint >= 1.0f) return 1.0f;
  const float cost = sqrtf(max0(1.0f - sint * sint));
  c = fabsf(c);
  const float Rs = ((etat * c) - (etai * cost)) / ((etat * c) + (etai * cost));
  const float Rp = ((etai * c) - (etat * cost)) / ((etai * c) + (etat * cost));
  return (Rs * Rs + Rp * Rp) / 2.0f;
}
// main.cpp:463-490 for one light; diffuse = getDiffuseColor(0.2) * 0.8. The
// terms that do not depend on the shadow test (max(0, L.N) and the Phong
// term) are computed where the shadow ray is spawned and kept per sample;
// the colour is finished once the shadow flag is known.
struct LitTerms {
  float LdotN, s25;
};
__device__ inline LitTerms lit_terms(V3 d, const Geo& g, const ShadeParams& sp) {
  const V3 L{sp.light[0], sp.light[1], sp.light[2]};
  const V3 ld = normalize(sub(L, g.p));
  LitTerms t;
  t.LdotN = max0(dot(ld, g.n));
  const V3 R = reflect({-ld.x, -ld.y, -ld.z}, g.n);
  t.s25 = powf(max0(-dot(R, d)), 25.0f);
  return t;
}
__device__ inline V3 lit_color(LitTerms t, const ShadeParams& sp, V3 surf, bool shadow) {
  const float vis = shadow ? 0.0f : 1.0f;
  const V3 amt{(sp.emission[0] * vis) * t.LdotN, (sp.emission[1] * vis) * t.LdotN,
               (sp.emission[2] * vis) * t.LdotN};
  const float k0 = 0.815f * 0.8f, k1 = 0.235f * 0.8f, k2 = 0.031f * 0.8f;
  V3 c;
  c.x = ((amt.x * k0) / 2.0f + (t.s25 * sp.emission[0]) * 0.5f) + surf.x;
  c.y = ((amt.y * k1) / 2.0f + (t.s25 * sp.emission[1]) * 0.5f) + surf.y;
  c.z = ((amt.z * k2) / 2.0f + (t.s25 * sp.emission[2]) * 0.5f) + surf.z;
  return c;
}
__device__ inline V3 direct(V3 d, const Geo& g, const ShadeParams& sp, V3 surf, bool shadow) {
  return lit_color(lit_terms(d, g, sp), sp, surf, shadow);
}
// Every ray the renderer spawns starts its interval at 0; a batch written
// with R.tmin null keeps a shared array of zeros for it (not written here).
__device__ inline void put_ray(const RaySoA& R, uint64_t i, V3 o, V3 d, float tmin, float tmax) {
  R.ox[i] = o.x, R.oy[i] = o.y, R.oz[i] = o.z;
  R.dx[i] = d.x, R.dy[i] = d.y, R.dz[i] = d.z;
  if (R.tmin) R.tmin[i] = tmin;
  R.tmax[i] = tmax;
}
__device__ inline void put_inactive(const RaySoA& R, uint64_t i) {
  if (R.tmin) R.tmin[i] = 0.0f;
  R.tmax[i] = 0.0f;
}
struct ShadowRay {
  V3 o, d;
  float tmax;
};
__device__ inline ShadowRay make_shadow(V3 d, const Geo& g, const ShadeParams& sp) {
  const V3 L{sp.light[0], sp.light[1], sp.light[2]};
  const V3 so = dot(d, g.n) < 0.0f ? add(g.p, mul(g.n, sp.bias)) : sub(g.p, mul(g.n, sp.bias));
  V3 ld = sub(L, g.p);
  const float dist2 = dot(ld, ld);
  ld = normalize(ld);
  return ShadowRay{so, ld, sqrtf(dist2)};
}
__device__ inline void shadow_ray(const RaySoA& R, uint64_t i, V3 d, const Geo& g,
                                  const ShadeParams& sp) {
  const ShadowRay r = make_shadow(d, g, sp);
  put_ray(R, i, r.o, r.d, 0.0f, r.tmax);
}
// refract (main.cpp:223-233): eta * I + (eta * cosi - sqrt(k)) * n, or 0 on
// total internal reflection; clamp as main.cpp:76-79.
__device__ inline V3 refract(V3 I, V3 N, float ior) {
  float c = dot(I, N);
  c = (c < 1.0f) ? c : 1.0f;
  c = (-1.0f < c) ? c : -1.0f;
  float etai = 1.0f, etat = ior;
  V3 n = N;
  if (c < 0.0f) {
    c = -c;
  } else {
    const float tmp = etai;
    etai = etat, etat = tmp;
    n = {-N.x, -N.y, -N.z};
  }
  const float eta = etai / etat;
  const float k = 1.0f - eta * eta * (1.0f - c * c);
  if (k < 0.0f) return {0.0f, 0.0f, 0.0f};
  return add(mul(I, eta), mul(n, eta * c - sqrtf(k)));
}
// A batch read with R.ox null starts every ray at the origin (camera rays).
__device__ inline V3 get_o(const RaySoA& R, uint64_t i) {
  if (!R.ox) return {0.0f, 0.0f, 0.0f};
  return {R.ox[i], R.oy[i], R.oz[i]};
}
__device__ inline V3 get_d(const RaySoA& R, uint64_t i) { return {R.dx[i], R.dy[i], R.dz[i]}; }

// Frame ray counts on demand: the active rays (tmin < tmax) of one batch.
// One atomic per block.
__global__ __launch_bounds__(256) void k_count_active(const float* __restrict__ tmin,
                                                       const float* __restrict__ tmax,
                                                       uint64_t n, uint32_t* counter) {
  __shared__ uint32_t s_c;
  if (threadIdx.x == 0) s_c = 0u;
  __syncthreads();
  uint32_t c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    c += tmin[i] < tmax[i];
  atomicAdd(&s_c, c);
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(counter, s_c);
}

// Stage 1: from primary hits spawn the shadow ray (slot i) and the second
// ray (slot i): the mirror bounce, or a REFLECTION_AND_REFRACTION hit's
// refraction ray (main.cpp:420-429; its reflection ray is traced there but
// its colour never used, so it is not traced here).
__global__ __launch_bounds__(256) void k_shade_primary(SceneView s, ShadeParams sp, uint32_t n,
                                                        RaySoA prim, const int32_t* __restrict__ h0,
                                                        const float* __restrict__ t0, RaySoA shad,
                                                        RaySoA bnc, float4* __restrict__ lit0) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = i < n;
  if (!in) return;
  const bool second = sp.bounce || sp.materials;
  const int32_t p = h0[i];
  if (p < 0) {
    put_inactive(shad, i);
    if (second) put_inactive(bnc, i);
    return;
  }
  const V3 o = get_o(prim, i), d = get_d(prim, i);
  const Geo g = hit_geo(s, p, o, d, t0[i]);
  const int m = mat_of(s, p);
  if (m != 0) {
    put_inactive(shad, i);
    if (m == 1) {
      const V3 rd = normalize(refract(d, g.n, 3.0f));
      const V3 ro =
          dot(rd, g.n) < 0.0f ? sub(g.p, mul(g.n, sp.bias)) : add(g.p, mul(g.n, sp.bias));
      put_ray(bnc, i, ro, rd, 0.0f, INFINITY);
    } else {
      put_inactive(bnc, i);
    }
    return;
  }
  if (sp.shadows) shadow_ray(shad, i, d, g, sp);
  // the final colour's shadow-independent terms, and the bounce weight
  const LitTerms lt = lit_terms(d, g, sp);
  lit0[i] = make_float4(lt.LdotN, lt.s25, sp.bounce ? fresnel(d, g.n, 2.0f) : 0.0f, 0.0f);
  if (sp.materials) put_inactive(bnc, i);
  if (sp.bounce) {
    const V3 rd = normalize(reflect(d, g.n));
    const V3 ro =
        dot(rd, g.n) < 0.0f ? sub(g.p, mul(g.n, sp.bias)) : add(g.p, mul(g.n, sp.bias));
    put_ray(bnc, i, ro, rd, 0.0f, INFINITY);
  }
}

// Stage 2: from bounce hits spawn their shadow ray (slot n + i) and keep
// their shadow-independent colour terms.
// With a compact output (co.rays.ox set), each shadow ray is also appended
// to co.rays at a slot taken from co.count (one atomic per wave) with its
// sample in co.idx, and samples without one get a 0 flag in co.occ: the
// any-hit launch then traces only the live rays (a few thousand of millions
// of samples on a mesh that mostly reflects to the sky).
__device__ inline void shade_bounce_sample(const SceneView& s, const ShadeParams& sp, uint32_t n,
                                           const RaySoA& bnc, const int32_t* __restrict__ h1,
                                           const float* __restrict__ t1, const RaySoA& shad,
                                           float2* __restrict__ lit1, const CompactOut& co,
                                           uint32_t i) {
  const int32_t p = h1[i];
  if (p < 0 || mat_of(s, p) != 0) {  // a second refracting / reflecting hit casts no shadow ray
    put_inactive(shad, (uint64_t)n + i);
    if (co.rays.ox) co.occ[i] = 0;
    return;
  }
  const V3 o = get_o(bnc, i), d = get_d(bnc, i);
  const Geo g = hit_geo(s, p, o, d, t1[i]);
  if (sp.shadows) {
    const ShadowRay r = make_shadow(d, g, sp);
    put_ray(shad, (uint64_t)n + i, r.o, r.d, 0.0f, r.tmax);
    if (co.rays.ox) {
      // every lane still here casts one: wave-aggregated slot
      const uint64_t m = __ballot(1);
      const int leader = __ffsll((unsigned long long)m) - 1;
      const uint32_t lane = threadIdx.x & 63u;
      uint32_t base = 0;
      if ((int)lane == leader) base = atomicAdd(co.count, (uint32_t)__popcll(m));
      base = (uint32_t)__shfl((int)base, leader, 64);
      const uint32_t k = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
      put_ray(co.rays, k, r.o, r.d, 0.0f, r.tmax);
      co.idx[k] = i;
    }
  }
  const LitTerms lt = lit_terms(d, g, sp);
  lit1[i] = make_float2(lt.LdotN, lt.s25);
}

// Grid-stride: the launch keeps a bounded grid (it runs beside the primary
// shadow rays' persistent any-hit launch, which holds most of the device's
// wave slots, and few resident waves each taking many samples get through
// faster there than a block per 256 samples).
__global__ __launch_bounds__(256) void k_shade_bounce(SceneView s, ShadeParams sp, uint32_t n,
                                                       RaySoA bnc, const int32_t* __restrict__ h1,
                                                       const float* __restrict__ t1, RaySoA shad,
                                                       float2* __restrict__ lit1, CompactOut co) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    shade_bounce_sample(s, sp, n, bnc, h1, t1, shad, lit1, co, i);
}

__global__ __launch_bounds__(64) void k_zero_word(uint32_t* w) {
  if (threadIdx.x == 0) *w = 0u;
}

// Colour of sample i (main.cpp:418-490): sky on a miss; the material switch
// (refraction colour * (1 - kr), or the grey 1 - kr); or the diffuse colour
// plus kr times the bounce's, from the terms the primary and bounce shading
// kept (the same operations on the same hit geometry as computing them here).
struct FinalArgs {
  SceneView s;
  ShadeParams sp;
  uint32_t npix, spp;
  RaySoA prim, bnc;
  const int32_t *h0, *h1;
  const float *t0, *t1;
  const uint8_t* occ;
  const float4* lit0;
  const float2* lit1;
  float *srgb, *img;
};
__device__ inline V3 sample_color(const FinalArgs& f, uint32_t i) {
  const SceneView& s = f.s;
  const ShadeParams& sp = f.sp;
  const uint32_t n = f.npix * f.spp;
  const V3 sky{sp.sky[0], sp.sky[1], sp.sky[2]};
  const int32_t p = f.h0[i];
  if (p < 0) return sky;
  const int m = mat_of(s, p);
  if (m != 0) {
    const V3 o = get_o(f.prim, i), d = get_d(f.prim, i);
    const Geo g = hit_geo(s, p, o, d, f.t0[i]);
    const float kr = fresnel(d, g.n, 2.0f);
    if (m == 2) return {1.0f - kr, 1.0f - kr, 1.0f - kr};
    V3 rc = sky;  // depth 2; a refracting surface there shows the sky (depth 3)
    const int32_t q = f.h1[i];
    if (q >= 0) {
      if (mat_of(s, q) == 0) {
        const float2 l1 = f.lit1[i];
        rc = lit_color({l1.x, l1.y}, sp, surface_of(s, sp, q), sp.shadows && f.occ[n + i] != 0);
      } else {
        const V3 o1 = get_o(f.bnc, i), d1 = get_d(f.bnc, i);
        const Geo g1 = hit_geo(s, q, o1, d1, f.t1[i]);
        const float kr1 = fresnel(d1, g1.n, 2.0f);
        rc = mat_of(s, q) == 2
                 ? V3{1.0f - kr1, 1.0f - kr1, 1.0f - kr1}
                 : V3{sky.x * (1.0f - kr1), sky.y * (1.0f - kr1), sky.z * (1.0f - kr1)};
      }
    }
    const float w = 1.0f - kr;
    return {rc.x * w, rc.y * w, rc.z * w};
  }
  const float4 l0 = f.lit0[i];
  V3 col = lit_color({l0.x, l0.y}, sp, surface_of(s, sp, p), sp.shadows && f.occ[i] != 0);
  if (sp.bounce) {
    const float kr = l0.z;
    V3 bc = sky;
    const int32_t q = f.h1[i];
    if (q >= 0) {
      const float2 l1 = f.lit1[i];
      bc = lit_color({l1.x, l1.y}, sp, surface_of(s, sp, q), sp.shadows && f.occ[n + i] != 0);
    }
    col = {col.x + kr * bc.x, col.y + kr * bc.y, col.z + kr * bc.z};
  }
  return col;
}
__device__ inline void put_srgb(const FinalArgs& f, uint32_t i, V3 c) {
  if (f.srgb) f.srgb[3 * (size_t)i] = c.x, f.srgb[3 * (size_t)i + 1] = c.y, f.srgb[3 * (size_t)i + 2] = c.z;
}

// Final, one thread per pixel: colour its samples and sum them in sample
// order (main.cpp:551-560). Any samples per pixel.
__global__ __launch_bounds__(256) void k_shade_final(FinalArgs f) {
  const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  if (px >= f.npix) return;
  V3 acc{0.0f, 0.0f, 0.0f};
  for (uint32_t k = 0; k < f.spp; ++k) {
    const uint32_t i = px * f.spp + k;
    const V3 col = sample_color(f, i);
    put_srgb(f, i, col);
    acc = {acc.x + col.x, acc.y + col.y, acc.z + col.z};
  }
  f.img[3 * (size_t)px] = acc.x, f.img[3 * (size_t)px + 1] = acc.y, f.img[3 * (size_t)px + 2] = acc.z;
}

// Final, one thread per sample (samples per pixel a power of two up to 64):
// coalesced loads and stores; a pixel's spp consecutive lanes are summed in
// sample order by its first lane, as the per-pixel loop sums them.
__global__ __launch_bounds__(256) void k_shade_final_samples(FinalArgs f) {
  const uint32_t n = f.npix * f.spp;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = (int)(threadIdx.x & 63u);
  V3 col{0.0f, 0.0f, 0.0f};
  if (i < n) {
    col = sample_color(f, i);
    put_srgb(f, i, col);
  }
  V3 acc{0.0f, 0.0f, 0.0f};
  for (uint32_t j = 0; j < f.spp; ++j) {
    const int src = lane + (int)j;  // the first lane of a pixel reads its group
    acc.x = acc.x + __shfl(col.x, src, 64);
    acc.y = acc.y + __shfl(col.y, src, 64);
    acc.z = acc.z + __shfl(col.z, src, 64);
  }
  if (i < n && (i & (f.spp - 1u)) == 0u) {
    const size_t px = i / f.spp;
    f.img[3 * px] = acc.x, f.img[3 * px + 1] = acc.y, f.img[3 * px + 2] = acc.z;
  }
}

// castRay's KDTREE case (main.cpp:362-370): black where the primary ray hits
// anything, the sky elsewhere; per pixel the samples summed in order.
__global__ __launch_bounds__(256) void k_shade_occlusion(ShadeParams sp, uint32_t npix,
                                                          uint32_t spp,
                                                          const uint8_t* __restrict__ occ,
                                                          int32_t* __restrict__ h0,
                                                          float* __restrict__ srgb,
                                                          float* __restrict__ img) {
  const uint32_t px = blockIdx.x * blockDim.x + threadIdx.x;
  if (px >= npix) return;
  V3 acc{0.0f, 0.0f, 0.0f};
  for (uint32_t k = 0; k < spp; ++k) {
    const uint32_t i = px * spp + k;
    const V3 col = occ[i] ? V3{0.0f, 0.0f, 0.0f} : V3{sp.sky[0], sp.sky[1], sp.sky[2]};
    h0[i] = -1;
    if (srgb) srgb[3 * (size_t)i] = col.x, srgb[3 * (size_t)i + 1] = col.y,
                  srgb[3 * (size_t)i + 2] = col.z;
    acc = {acc.x + col.x, acc.y + col.y, acc.z + col.z};
  }
  img[3 * (size_t)px] = acc.x, img[3 * (size_t)px + 1] = acc.y, img[3 * (size_t)px + 2] = acc.z;
}

inline uint32_t cdiv(uint64_t a, uint32_t b) { return (uint32_t)((a + b - 1) / b); }

// Plain device-to-device copy, 16 bytes per lane (also the known-size kernel
// that calibrates the HBM byte counters).
__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src,
                                                 uint4* __restrict__ dst, uint64_t n16) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

hipError_t launch_copy16(const LaunchCtx& lc, const void* src, void* dst, uint64_t bytes) {
  if (bytes == 0) return hipSuccess;
  ProfScope ps(lc, "copy16");
  const uint64_t n16 = bytes / 16;
  hipLaunchKernelGGL(k_copy16, dim3(std::min<uint64_t>((n16 + 255) / 256, 8192)), dim3(256), 0,
                     lc.stream, (const uint4*)src, (uint4*)dst, n16);
  return hipGetLastError();
}

hipError_t launch_raygen_camera(const LaunchCtx& lc, const CamParams& cp, const double* jit,
                                const RaySoA& R) {
  const uint64_t n = (uint64_t)cp.w * cp.h * cp.spp;
  if (n == 0) return hipSuccess;
  ProfScope ps(lc, "raygen_camera");
  hipLaunchKernelGGL(k_raygen_camera, dim3(cdiv(n, 256)), dim3(256), 0, lc.stream, cp, jit, R.ox,
                     R.oy, R.oz, R.dx, R.dy, R.dz, R.tmin, R.tmax);
  return hipGetLastError();
}

hipError_t launch_raygen_uniform(const LaunchCtx& lc, uint32_t n, uint64_t seed, const float lo[3],
                                 const float hi[3], const RaySoA& R) {
  if (n == 0) return hipSuccess;
  ProfScope ps(lc, "raygen_uniform");
  const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
  hipLaunchKernelGGL(k_raygen_uniform, dim3(cdiv(n, 256)), dim3(256), 0, lc.stream, n, seed, lo[0],
                     lo[1], lo[2], ex, ey, ez, R.ox, R.oy, R.oz, R.dx, R.dy, R.dz, R.tmin, R.tmax);
  return hipGetLastError();
}

hipError_t launch_raygen_pinhole(const LaunchCtx& lc, const PinholeParams& p, const RaySoA& R) {
  if (p.n == 0) return hipSuccess;
  ProfScope ps(lc, "raygen_pinhole");
  hipLaunchKernelGGL(k_raygen_pinhole, dim3(cdiv(p.n, 256)), dim3(256), 0, lc.stream, p, R.ox, R.oy,
                     R.oz, R.dx, R.dy, R.dz, R.tmin, R.tmax);
  return hipGetLastError();
}

hipError_t launch_shade_primary(const LaunchCtx& lc, const SceneView& s, const ShadeParams& sp,
                                uint32_t n, const RaySoA& prim, const int32_t* h0, const float* t0,
                                const RaySoA& shad, const RaySoA& bnc, float4* lit0) {
  ProfScope ps(lc, "shade_primary");
  hipLaunchKernelGGL(k_shade_primary, dim3(cdiv(n, 256)), dim3(256), 0, lc.stream, s, sp, n, prim,
                     h0, t0, shad, bnc, lit0);
  return hipGetLastError();
}

hipError_t launch_shade_bounce(const LaunchCtx& lc, const SceneView& s, const ShadeParams& sp,
                               uint32_t n, const RaySoA& bnc, const int32_t* h1, const float* t1,
                               const RaySoA& shad, float2* lit1, const CompactOut& co) {
  if (co.rays.ox) hipLaunchKernelGGL(k_zero_word, dim3(1), dim3(64), 0, lc.stream, co.count);
  ProfScope ps(lc, "shade_bounce");
  hipLaunchKernelGGL(k_shade_bounce, dim3(std::min<uint32_t>(cdiv(n, 256), 2048)), dim3(256), 0,
                     lc.stream, s, sp, n, bnc, h1, t1, shad, lit1, co);
  return hipGetLastError();
}

hipError_t launch_count_active(const LaunchCtx& lc, const RaySoA& R, uint64_t n,
                               uint32_t* counter) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_count_active, dim3(std::min<uint32_t>(cdiv(n, 256), 1024u)), dim3(256), 0,
                     lc.stream, R.tmin, R.tmax, n, counter);
  return hipGetLastError();
}

hipError_t launch_shade_occlusion(const LaunchCtx& lc, const ShadeParams& sp, uint32_t npix,
                                  uint32_t spp, const uint8_t* occ, int32_t* h0, float* srgb,
                                  float* img) {
  ProfScope ps(lc, "shade_occlusion");
  hipLaunchKernelGGL(k_shade_occlusion, dim3(cdiv(npix, 256)), dim3(256), 0, lc.stream, sp, npix,
                     spp, occ, h0, srgb, img);
  return hipGetLastError();
}

hipError_t launch_shade_final(const LaunchCtx& lc, const SceneView& s, const ShadeParams& sp,
                              uint32_t npix, uint32_t spp, const RaySoA& prim, const RaySoA& bnc,
                              const int32_t* h0, const float* t0, const int32_t* h1,
                              const float* t1, const uint8_t* occ, const float4* lit0,
                              const float2* lit1, float* srgb, float* img) {
  ProfScope ps(lc, "shade_final");
  const FinalArgs f{s, sp, npix, spp, prim, bnc, h0, h1, t0, t1, occ, lit0, lit1, srgb, img};
  if (spp >= 1 && spp <= 64 && (spp & (spp - 1)) == 0) {
    const uint64_t n = (uint64_t)npix * spp;
    hipLaunchKernelGGL(k_shade_final_samples, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0,
                       lc.stream, f);
  } else {
    hipLaunchKernelGGL(k_shade_final, dim3(cdiv(npix, 256)), dim3(256), 0, lc.stream, f);
  }
  return hipGetLastError();
}

}  // namespace rq
