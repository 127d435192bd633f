// Wavefront renderer orchestration (include/rq/rq_render.h): the GPU_LBVH
// replacement of render() + castRay (main.cpp:291-566). Host code; one
// launch per stage over every sample of the frame.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "rq/rq_render.h"
#include "rq_ctx.h"
#include "rq_render_internal.h"

using namespace rq;

struct rq_renderer {
  rq_render_params p;
  uint32_t n = 0, npix = 0;
  CamParams cam;
  ShadeParams sh;
  RaySoA prim, bnc, shad;  // shad holds 2n rays when bounce, else n
  int32_t *h0 = nullptr, *h1 = nullptr;
  float *t0 = nullptr, *t1 = nullptr;
  // the second hits' shadow rays compacted (side-stream frames): rays, their
  // samples and their count
  RaySoA shadc{};
  // shared constant arrays, written once at creation: zeros (2n floats when a
  // second ray per sample is traced) serve as every batch's tmin and as the
  // camera rays' origins; infs (n floats) as the camera rays' tmax
  float* zeros = nullptr;
  float* infs = nullptr;
  uint32_t* shidx = nullptr;
  uint32_t* shcount = nullptr;
  uint8_t* occ = nullptr;
  float4* lit0 = nullptr;  // per sample: a diffuse primary hit's L.N, Phong term, bounce weight
  float2* lit1 = nullptr;  // per sample: a diffuse second hit's L.N and Phong term
  float* srgb = nullptr;
  float* img = nullptr;
  double* jit = nullptr;
  uint32_t* counters = nullptr;
  // the primary hits' shadow rays are traced on `side` while the bounce rays
  // are traced on the frame's stream (default priority, ordered by events)
  hipStream_t side = nullptr;
  hipEvent_t ev_primary = nullptr, ev_shadow = nullptr;
  // recorded on the frame's stream after its last closest-hit launch (or the
  // occlusion frame's any-hit launch): what is left of the frame then is
  // shading and the few second-hit shadow rays (rq_render_wait_traced)
  hipEvent_t ev_traced = nullptr;
  bool traced_recorded = false;
  std::vector<void*> allocs;
  // a second ray per sample: the mirror bounce, or a refraction ray
  bool second() const { return p.mode == RQ_RENDER_SHADE && (p.bounce || p.materials); }
};

namespace {

int ralloc(rq_ctx* c, rq_renderer* r, size_t bytes, void** p) {
  int rc = ctx_alloc(c, bytes, p);
  if (rc == RQ_OK && *p) r->allocs.push_back(*p);
  return rc;
}

// A ray batch whose tmin is the renderer's shared array of zeros (every ray
// the renderer spawns starts its interval at 0).
int alloc_rays(rq_ctx* c, rq_renderer* r, size_t n, RaySoA& R) {
  float** f[7] = {&R.ox, &R.oy, &R.oz, &R.dx, &R.dy, &R.dz, &R.tmax};
  for (auto* q : f) {
    int rc = ralloc(c, r, n * 4, (void**)q);
    if (rc) return rc;
  }
  R.tmin = r->zeros;
  return RQ_OK;
}
// The batch as the shading kernels write it: tmin left to the shared zeros.
RaySoA write_view(RaySoA R) {
  R.tmin = nullptr;
  return R;
}

void free_renderer(rq_ctx* c, rq_renderer* r) {
  for (void* p : r->allocs) ctx_free(c, p);
  if (r->ev_primary) (void)hipEventDestroy(r->ev_primary);
  if (r->ev_shadow) (void)hipEventDestroy(r->ev_shadow);
  if (r->ev_traced) (void)hipEventDestroy(r->ev_traced);
  if (r->side) (void)rq_stream_destroy(c, r->side);  // and the stream's stack overflow area
  delete r;
}

rq_rays as_rays(const RaySoA& R, size_t off = 0) {
  return rq_rays{R.ox + off, R.oy + off, R.oz + off, R.dx + off,
                 R.dy + off, R.dz + off, R.tmin + off, R.tmax + off};
}

CamParams cam_params(const rq_render_params* p) {
  CamParams cp;
  cp.w = p->width, cp.h = p->height, cp.spp = p->spp;
  cp.invW = 1 / float(p->width);
  cp.invH = 1 / float(p->height);
  float fov = 30;  // main.cpp:545
  cp.aspect = p->width / float(p->height);
  cp.angle = (float)std::tan(3.141592653589793 * 0.5 * fov / 180.);
  cp.seed = p->seed;
  return cp;
}

}  // namespace

extern "C" {

void rq_render_defaults(rq_render_params* p) {
  std::memset(p, 0, sizeof *p);
  p->width = 640, p->height = 480, p->spp = 1;
  p->camera = RQ_CAMERA_REF_JITTER;
  p->seed = 0;
  p->light[0] = 0, p->light[1] = 3, p->light[2] = 30;
  p->emission[0] = p->emission[1] = p->emission[2] = 1;
  p->sky[0] = 0.6f, p->sky[1] = 0.8f, p->sky[2] = 1.0f;
  p->bias = 1e-4f;
  p->bounce = 0;
  p->shadows = 1;
  p->materials = 0;
  p->mode = RQ_RENDER_SHADE;
}

int rq_renderer_create(rq_ctx* c, const rq_render_params* p, rq_renderer** out) {
  if (!c || !p || !out) return RQ_E_INVALID;
  *out = nullptr;
  uint64_t n = (uint64_t)p->width * p->height * p->spp;
  if (n == 0 || 2 * n > 0xFFFFFFFFull) return set_err(c, RQ_E_INVALID, "bad image size");
  if (p->camera != RQ_CAMERA_REF_JITTER && p->camera != RQ_CAMERA_PHILOX)
    return set_err(c, RQ_E_INVALID, "unknown camera %d", p->camera);
  if (p->mode != RQ_RENDER_SHADE && p->mode != RQ_RENDER_OCCLUSION)
    return set_err(c, RQ_E_INVALID, "unknown render mode %d", p->mode);
  if (p->materials && p->bounce)
    return set_err(c, RQ_E_INVALID, "materials and the mirror bounce are exclusive");
  hipSetDevice(c->device);
  auto* r = new rq_renderer();
  r->p = *p;
  r->n = (uint32_t)n;
  r->npix = p->width * p->height;
  r->cam = cam_params(p);
  for (int k = 0; k < 3; ++k) {
    r->sh.light[k] = p->light[k], r->sh.emission[k] = p->emission[k];
    r->sh.surface[k] = p->surface[k], r->sh.sky[k] = p->sky[k];
  }
  r->sh.bias = p->bias;
  r->sh.bounce = p->bounce;
  r->sh.shadows = p->shadows;
  r->sh.materials = p->materials;
  const bool second = r->second();
  const size_t ns = second ? 2 * n : n;
  // shared constant arrays: zeros (every tmin, the camera rays' origins) and
  // +inf (the camera rays' tmax), written once here
  int rc = ralloc(c, r, ns * 4, (void**)&r->zeros);
  if (!rc) rc = ralloc(c, r, n * 4, (void**)&r->infs);
  if (!rc) {
    hipError_t e = hipMemset(r->zeros, 0, ns * 4);
    if (e == hipSuccess) e = hipMemsetD32((hipDeviceptr_t)r->infs, 0x7F800000, n);
    if (e != hipSuccess) rc = hip_err(c, e, "renderer constants");
  }
  // camera rays: only their directions are written per frame
  if (!rc) rc = ralloc(c, r, n * 4, (void**)&r->prim.dx);
  if (!rc) rc = ralloc(c, r, n * 4, (void**)&r->prim.dy);
  if (!rc) rc = ralloc(c, r, n * 4, (void**)&r->prim.dz);
  r->prim.ox = r->prim.oy = r->prim.oz = r->prim.tmin = r->zeros;
  r->prim.tmax = r->infs;
  if (!rc && second) rc = alloc_rays(c, r, n, r->bnc);
  if (!rc) rc = alloc_rays(c, r, ns, r->shad);
  if (!rc) rc = ralloc(c, r, n * 4, (void**)&r->h0);
  if (!rc) rc = ralloc(c, r, n * 4, (void**)&r->t0);
  if (!rc && second) rc = ralloc(c, r, n * 4, (void**)&r->h1);
  if (!rc && second) rc = ralloc(c, r, n * 4, (void**)&r->t1);
  if (!rc) rc = ralloc(c, r, ns, (void**)&r->occ);
  if (!rc && p->mode == RQ_RENDER_SHADE) rc = ralloc(c, r, n * 16, (void**)&r->lit0);
  if (!rc && second) rc = ralloc(c, r, n * 8, (void**)&r->lit1);
  if (!rc) rc = ralloc(c, r, n * 12, (void**)&r->srgb);
  if (!rc) rc = ralloc(c, r, (size_t)r->npix * 12, (void**)&r->img);
  if (!rc) rc = ralloc(c, r, 16, (void**)&r->counters);
  if (!rc && p->camera == RQ_CAMERA_REF_JITTER) {
    // main.cpp:503-508: one default-seeded stream, two draws per sample
    // (xx first, then yy) in pixel-major order.
    std::vector<double> j(2 * n);
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    std::mt19937 g;
    for (auto& v : j) v = dist(g);
    rc = ralloc(c, r, 16 * n, (void**)&r->jit);
    if (!rc) {
      hipError_t e = hipMemcpy(r->jit, j.data(), 16 * n, hipMemcpyHostToDevice);
      if (e != hipSuccess) rc = hip_err(c, e, "jitter upload");
    }
  }
  if (!rc) {
    hipError_t e = hipMemset(r->counters, 0, 16);
    if (e != hipSuccess) rc = hip_err(c, e, "counter init");
  }
  if (!rc && r->second() && p->shadows) rc = alloc_rays(c, r, n, r->shadc);
  if (!rc && r->second() && p->shadows) rc = ralloc(c, r, n * 4, (void**)&r->shidx);
  if (!rc && r->second() && p->shadows) rc = ralloc(c, r, 4, (void**)&r->shcount);
  if (!rc) {
    const hipError_t e = hipEventCreateWithFlags(&r->ev_traced, hipEventDisableTiming);
    if (e != hipSuccess) rc = hip_err(c, e, "renderer event");
  }
  if (!rc && r->second() && p->shadows) {
    hipError_t e = hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ev_primary, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ev_shadow, hipEventDisableTiming);
    if (e != hipSuccess) rc = hip_err(c, e, "renderer streams");
  }
  if (rc) {
    free_renderer(c, r);
    return rc;
  }
  *out = r;
  return RQ_OK;
}

int rq_renderer_destroy(rq_ctx* c, rq_renderer* r) {
  if (!c || !r) return RQ_E_INVALID;
  (void)hipDeviceSynchronize();
  free_renderer(c, r);
  return RQ_OK;
}

int rq_render_frame(rq_ctx* c, rq_renderer* r, const rq_bvh* b, void* stream) {
  if (!c || !r || !b) return RQ_E_INVALID;
  LaunchCtx lc = launch_ctx(c, stream);
  const rq_scene* s = b->scene;
  SceneView sv{s->vx, s->vy, s->vz, s->idx, s->r, s->rgb, r->p.materials ? s->mat : nullptr};
  // the camera rays' directions only: their origins and intervals are the
  // shared constant arrays
  const RaySoA gen{nullptr, nullptr, nullptr, r->prim.dx, r->prim.dy, r->prim.dz, nullptr, nullptr};
  hipError_t e = launch_raygen_camera(lc, r->cam, r->jit, gen);
  if (e != hipSuccess) return hip_err(c, e, "raygen");
  rq_rays pr = as_rays(r->prim);
  auto traced = [&]() -> int {
    const hipError_t er = hipEventRecord(r->ev_traced, (hipStream_t)stream);
    if (er != hipSuccess) return hip_err(c, er, "traced event");
    r->traced_recorded = true;
    return RQ_OK;
  };
  if (r->p.mode == RQ_RENDER_OCCLUSION) {
    int rc = rq_any_hit(c, b, &pr, r->n, r->occ, stream);
    if (!rc) rc = traced();
    if (rc) return rc;
    e = launch_shade_occlusion(lc, r->sh, r->npix, r->p.spp, r->occ, r->h0, r->srgb, r->img);
    return e == hipSuccess ? RQ_OK : hip_err(c, e, "shade occlusion");
  }
  rq_hits h0{r->h0, r->t0, nullptr, nullptr};
  int rc = rq_closest_hit(c, b, &pr, r->n, &h0, stream);
  if (!rc && !r->second()) rc = traced();
  if (rc) return rc;
  // the shading kernels read the camera rays with a null origin (all at the
  // origin) and write the spawned rays without their tmin (all zero)
  RaySoA cam = r->prim;
  cam.ox = cam.oy = cam.oz = nullptr;
  e = launch_shade_primary(lc, sv, r->sh, r->n, cam, r->h0, r->t0, write_view(r->shad),
                           write_view(r->bnc), r->lit0);
  if (e != hipSuccess) return hip_err(c, e, "shade primary");
  if (r->side) {
    // the primary shadow rays on the side stream, concurrently with the
    // second rays and their shadow rays here: each launch's last waves
    // overlap the other's
    e = hipEventRecord(r->ev_primary, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(r->side, r->ev_primary, 0);
    if (e != hipSuccess) return hip_err(c, e, "frame streams");
    rq_rays sr = as_rays(r->shad);
    rc = rq_any_hit(c, b, &sr, r->n, r->occ, r->side);
    if (rc) return rc;
    e = hipEventRecord(r->ev_shadow, r->side);
    if (e != hipSuccess) return hip_err(c, e, "frame streams");
  }
  if (r->second()) {
    rq_rays br = as_rays(r->bnc);
    rq_hits h1{r->h1, r->t1, nullptr, nullptr};
    rc = rq_closest_hit(c, b, &br, r->n, &h1, stream);
    if (!rc) rc = traced();
    if (rc) return rc;
    const CompactOut co = r->side ? CompactOut{write_view(r->shadc), r->shidx, r->shcount,
                                               r->occ + r->n}
                                  : CompactOut{};
    e = launch_shade_bounce(lc, sv, r->sh, r->n, r->bnc, r->h1, r->t1, write_view(r->shad),
                            r->lit1, co);
    if (e != hipSuccess) return hip_err(c, e, "shade bounce");
  }
  if (r->side) {
    // the second hits' shadow rays, compacted: their count is on the device
    const rq_rays sr = as_rays(r->shadc);
    rc = any_hit_indexed(c, b, &sr, r->n, r->shcount, r->shidx, r->occ + r->n, stream);
    if (rc) return rc;
    e = hipStreamWaitEvent((hipStream_t)stream, r->ev_shadow, 0);
    if (e != hipSuccess) return hip_err(c, e, "frame streams");
  } else if (r->p.shadows) {
    rq_rays sr = as_rays(r->shad);
    rc = rq_any_hit(c, b, &sr, r->second() ? 2 * r->n : r->n, r->occ, stream);
    if (rc) return rc;
  }
  e = launch_shade_final(lc, sv, r->sh, r->npix, r->p.spp, cam, r->bnc, r->h0, r->t0, r->h1,
                         r->t1, r->occ, r->lit0, r->lit1, r->srgb, r->img);
  if (e != hipSuccess) return hip_err(c, e, "shade final");
  return RQ_OK;
}

int rq_render_wait_traced(rq_ctx* c, rq_renderer* r, void* stream) {
  if (!c || !r) return RQ_E_INVALID;
  if (!r->traced_recorded) return RQ_OK;  // no frame yet: nothing to wait for
  const hipError_t e = hipStreamWaitEvent((hipStream_t)stream, r->ev_traced, 0);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "wait traced");
}

int rq_render_counts(rq_ctx* c, rq_renderer* r, rq_frame_counts* out) {
  if (!c || !r || !out) return RQ_E_INVALID;
  uint32_t k[4];
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemset(r->counters, 0, 16);
  const LaunchCtx lc{nullptr, nullptr};
  const bool shade = r->p.mode == RQ_RENDER_SHADE;
  if (e == hipSuccess && r->second()) e = launch_count_active(lc, r->bnc, r->n, r->counters);
  if (e == hipSuccess && shade && r->p.shadows)
    e = launch_count_active(lc, r->shad, r->second() ? 2ull * r->n : r->n, r->counters + 1);
  if (e == hipSuccess) e = hipMemcpy(k, r->counters, 16, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_err(c, e, "counts");
  out->primary = r->n;
  out->bounce = r->second() ? k[0] : 0;
  out->shadow = shade && r->p.shadows ? k[1] : 0;
  return RQ_OK;
}

int rq_render_read_image(rq_ctx* c, rq_renderer* r, float* rgb, uint8_t* ppm) {
  if (!c || !r) return RQ_E_INVALID;
  std::vector<float> img((size_t)r->npix * 3);
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(img.data(), r->img, img.size() * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_err(c, e, "image read");
  if (rgb) std::memcpy(rgb, img.data(), img.size() * 4);
  if (ppm)
    for (size_t k = 0; k < img.size(); ++k)
      ppm[k] = (uint8_t)(std::min(float(1), img[k] / r->p.spp) * 255);
  return RQ_OK;
}

int rq_render_read_samples(rq_ctx* c, rq_renderer* r, int32_t* hit0, int32_t* hit1,
                           uint8_t* sh0, uint8_t* sh1, float* sample_rgb) {
  if (!c || !r) return RQ_E_INVALID;
  const size_t n = r->n;
  hipError_t e = hipDeviceSynchronize();
  if (e == hipSuccess && hit0) e = hipMemcpy(hit0, r->h0, n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && hit1) {
    if (r->second())
      e = hipMemcpy(hit1, r->h1, n * 4, hipMemcpyDeviceToHost);
    else
      for (size_t i = 0; i < n; ++i) hit1[i] = -1;
  }
  if (e == hipSuccess && sh0) {
    if (r->p.shadows || r->p.mode == RQ_RENDER_OCCLUSION)
      e = hipMemcpy(sh0, r->occ, n, hipMemcpyDeviceToHost);
    else
      std::memset(sh0, 0, n);
  }
  if (e == hipSuccess && sh1) {
    if (r->second() && r->p.shadows)
      e = hipMemcpy(sh1, r->occ + n, n, hipMemcpyDeviceToHost);
    else
      std::memset(sh1, 0, n);
  }
  if (e == hipSuccess && sample_rgb)
    e = hipMemcpy(sample_rgb, r->srgb, n * 12, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_err(c, e, "sample read");
  return RQ_OK;
}

int rq_render_primary_rays(rq_ctx* c, rq_renderer* r, rq_rays* out, uint32_t* n) {
  if (!c || !r || !out) return RQ_E_INVALID;
  *out = as_rays(r->prim);
  if (n) *n = r->n;
  return RQ_OK;
}

int rq_render_rays(rq_ctx* c, rq_renderer* r, int which, rq_rays* out, uint32_t* n) {
  if (!c || !r || !out) return RQ_E_INVALID;
  if (which == 0) {
    *out = as_rays(r->prim);
    if (n) *n = r->n;
  } else if (which == 1 && r->second()) {
    *out = as_rays(r->bnc);
    if (n) *n = r->n;
  } else if (which == 2) {
    *out = as_rays(r->shad);
    if (n) *n = r->second() ? 2 * r->n : r->n;
  } else {
    return set_err(c, RQ_E_INVALID, "no ray batch %d", which);
  }
  return RQ_OK;
}

int rq_raygen_camera(rq_ctx* c, const rq_render_params* p, const double* jit, float* ox,
                     float* oy, float* oz, float* dx, float* dy, float* dz, float* tmin,
                     float* tmax, void* stream) {
  if (!c || !p) return RQ_E_INVALID;
  if (p->camera == RQ_CAMERA_REF_JITTER && !jit)
    return set_err(c, RQ_E_INVALID, "REF_JITTER camera needs the jitter array");
  RaySoA R{ox, oy, oz, dx, dy, dz, tmin, tmax};
  hipError_t e = launch_raygen_camera(launch_ctx(c, stream), cam_params(p),
                                      p->camera == RQ_CAMERA_REF_JITTER ? jit : nullptr, R);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "raygen camera");
}

int rq_raygen_uniform(rq_ctx* c, uint32_t n, uint64_t seed, const float lo[3], const float hi[3],
                      float* ox, float* oy, float* oz, float* dx, float* dy, float* dz,
                      float* tmin, float* tmax, void* stream) {
  if (!c || !lo || !hi) return RQ_E_INVALID;
  RaySoA R{ox, oy, oz, dx, dy, dz, tmin, tmax};
  hipError_t e = launch_raygen_uniform(launch_ctx(c, stream), n, seed, lo, hi, R);
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "raygen uniform");
}

int rq_write_ppm(const char* path, uint32_t w, uint32_t h, const uint8_t* bytes) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return RQ_E_IO;
  std::fprintf(f, "P6\n%u %u\n255\n", w, h);
  size_t nb = (size_t)w * h * 3;
  size_t wr = std::fwrite(bytes, 1, nb, f);
  int rc = std::fclose(f);
  return (wr == nb && rc == 0) ? RQ_OK : RQ_E_IO;
}

}  // extern "C"
