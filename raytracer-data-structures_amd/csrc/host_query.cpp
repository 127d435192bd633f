// Host-pointer queries (rq_closest_hit_host, rq_any_hit_host) and the per-ray
// forms (rq_cast_ray, rq_occluded_ray): the caller's host arrays are staged
// through one pinned buffer and one device buffer per context, in chunks of
// at most HOST_CHUNK rays, around the device queries of rq_api.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "rq_ctx.h"

using namespace rq;

namespace {

constexpr uint32_t HOST_CHUNK = 1u << 20;  // rays per staged launch
constexpr size_t RAY_BYTES = 8 * sizeof(float);
constexpr size_t HIT_BYTES = sizeof(int32_t) + 3 * sizeof(float);

uint32_t round64(uint32_t m) { return (m + 63u) & ~63u; }

// Grows the context's pinned and device staging buffers to `bytes` each.
int stage_reserve(rq_ctx* c, size_t bytes) {
  if (bytes <= c->stage_bytes) return RQ_OK;
  if (c->stage_host) (void)hipHostFree(c->stage_host);
  c->stage_host = nullptr;
  if (c->stage_dev) ctx_free(c, c->stage_dev);
  c->stage_dev = nullptr;
  c->stage_bytes = 0;
  const hipError_t e = hipHostMalloc(&c->stage_host, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    c->stage_host = nullptr;
    return hip_err(c, e, "pinned staging buffer");
  }
  const int rc = ctx_alloc(c, bytes, &c->stage_dev);
  if (rc) {
    (void)hipHostFree(c->stage_host);
    c->stage_host = nullptr;
    c->stage_dev = nullptr;
    return rc;
  }
  c->stage_bytes = bytes;
  return RQ_OK;
}

bool rays_ok(const rq_rays* r) {
  return r && r->ox && r->oy && r->oz && r->dx && r->dy && r->dz && r->tmin && r->tmax;
}

// One staged query over host arrays: rays in, closest hits or occlusion out.
int host_query(rq_ctx* c, const rq_bvh* b, const rq_rays* rays, uint32_t n, rq_hits* hits,
               uint8_t* occ, bool any, void* stream) {
  if (!c) return RQ_E_INVALID;
  if (!rays_ok(rays)) return set_err(c, RQ_E_INVALID, "null host ray array");
  if (any ? !occ : (!hits || !hits->prim)) return set_err(c, RQ_E_INVALID, "null output array");
  if (n == 0) return RQ_OK;
  (void)hipSetDevice(c->device);
  const uint32_t cap = round64(std::min(n, HOST_CHUNK));
  int rc = stage_reserve(c, (size_t)cap * (RAY_BYTES + HIT_BYTES));
  if (rc) return rc;
  const hipStream_t s = (hipStream_t)stream;
  const float* src[8] = {rays->ox, rays->oy, rays->oz, rays->dx,
                         rays->dy, rays->dz, rays->tmin, rays->tmax};
  for (uint32_t i0 = 0; i0 < n; i0 += HOST_CHUNK) {
    const uint32_t m = std::min(HOST_CHUNK, n - i0), ms = round64(m);
    // packed [8][ms] ray arrays, then the outputs
    char* h = static_cast<char*>(c->stage_host);
    char* d = static_cast<char*>(c->stage_dev);
    for (int k = 0; k < 8; ++k) std::memcpy(h + (size_t)k * ms * 4, src[k] + i0, (size_t)m * 4);
    hipError_t e = hipMemcpyAsync(d, h, (size_t)8 * ms * 4, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_err(c, e, "staging copy in");
    const float* df = reinterpret_cast<const float*>(d);
    const rq_rays dr{df, df + ms, df + 2 * (size_t)ms, df + 3 * (size_t)ms,
                     df + 4 * (size_t)ms, df + 5 * (size_t)ms, df + 6 * (size_t)ms,
                     df + 7 * (size_t)ms};
    char* dout = d + (size_t)8 * ms * 4;
    char* hout = h + (size_t)8 * ms * 4;
    size_t out_bytes = 0;
    if (any) {
      rc = rq_any_hit(c, b, &dr, m, reinterpret_cast<uint8_t*>(dout), stream);
      out_bytes = m;
    } else {
      rq_hits dh{reinterpret_cast<int32_t*>(dout),
                 hits->t ? reinterpret_cast<float*>(dout + (size_t)ms * 4) : nullptr,
                 hits->u ? reinterpret_cast<float*>(dout + (size_t)ms * 8) : nullptr,
                 hits->v ? reinterpret_cast<float*>(dout + (size_t)ms * 12) : nullptr};
      rc = rq_closest_hit(c, b, &dr, m, &dh, stream);
      out_bytes = (size_t)ms * 4 * (hits->v ? 4 : hits->u ? 3 : hits->t ? 2 : 1);
    }
    if (rc) return rc;
    e = hipMemcpyAsync(hout, dout, out_bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_err(c, e, "staging copy out");
    if (any) {
      std::memcpy(occ + i0, hout, m);
    } else {
      std::memcpy(hits->prim + i0, hout, (size_t)m * 4);
      if (hits->t) std::memcpy(hits->t + i0, hout + (size_t)ms * 4, (size_t)m * 4);
      if (hits->u) std::memcpy(hits->u + i0, hout + (size_t)ms * 8, (size_t)m * 4);
      if (hits->v) std::memcpy(hits->v + i0, hout + (size_t)ms * 12, (size_t)m * 4);
    }
  }
  return RQ_OK;
}

}  // namespace

namespace rq {
void stage_release(rq_ctx* c) {
  if (c->stage_host) (void)hipHostFree(c->stage_host);
  c->stage_host = nullptr;
  c->stage_dev = nullptr;  // freed with the context's allocations
  c->stage_bytes = 0;
}
}  // namespace rq

extern "C" {

int rq_closest_hit_host(rq_ctx* c, const rq_bvh* b, const rq_rays* rays, uint32_t n,
                        rq_hits* hits, void* stream) {
  return host_query(c, b, rays, n, hits, nullptr, false, stream);
}

int rq_any_hit_host(rq_ctx* c, const rq_bvh* b, const rq_rays* rays, uint32_t n, uint8_t* occ,
                    void* stream) {
  return host_query(c, b, rays, n, nullptr, occ, true, stream);
}

int rq_cast_ray(rq_ctx* c, const rq_bvh* b, const float orig[3], const float dir[3], float tmin,
                float tmax, int32_t* prim, float* t, float* uv) {
  if (!c) return RQ_E_INVALID;
  if (!orig || !dir || !prim) return set_err(c, RQ_E_INVALID, "null ray or output");
  const rq_rays r{orig, orig + 1, orig + 2, dir, dir + 1, dir + 2, &tmin, &tmax};
  float tt = 0.0f, u = 0.0f, v = 0.0f;
  rq_hits h{prim, &tt, uv ? &u : nullptr, uv ? &v : nullptr};
  const int rc = host_query(c, b, &r, 1, &h, nullptr, false, nullptr);
  if (rc) return rc;
  if (t) *t = tt;
  if (uv) uv[0] = u, uv[1] = v;
  return RQ_OK;
}

int rq_occluded_ray(rq_ctx* c, const rq_bvh* b, const float orig[3], const float dir[3],
                    float tmin, float tmax, uint8_t* occluded) {
  if (!c) return RQ_E_INVALID;
  if (!orig || !dir || !occluded) return set_err(c, RQ_E_INVALID, "null ray or output");
  const rq_rays r{orig, orig + 1, orig + 2, dir, dir + 1, dir + 2, &tmin, &tmax};
  return host_query(c, b, &r, 1, nullptr, occluded, true, nullptr);
}

}  // extern "C"
