// A C++ program that links librq.so and drives GPU_LBVH the way a
// reference-side integration would: the build-side case of main.cpp:791-843
// (OBJ read, scene upload, rq_build_lbvh) and the query-side cases of castRay
// (main.cpp:330-388): a batch of device rays, the same rays from host arrays,
// and the per-ray calls of the reference's per-pixel loop.
//
//   caller <mesh.obj> <rays.bin> <out.bin>
//
// rays.bin: uint32 n, then the 8 float32 arrays ox oy oz dx dy dz tmin tmax.
// out.bin:  uint32 n, uint32 k (rays also cast one by one, the first k), then
//           device batch   prim[n] int32, t[n] f32, occluded[n] u8,
//           host batch     prim[n], t[n], occluded[n],
//           per ray        prim[k], t[k], occluded[k],
//           tree           n_prims, height (uint32 each).
// Exit status 0 on success; a failing call prints rq_last_error and exits 1.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rq/rq.h"
#include "rq/rq_scene.h"
#include "rq/settings.h"

namespace {

rq_ctx* g_ctx = nullptr;

void check(int rc, const char* what) {
  if (rc == RQ_OK) return;
  std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, rq_last_error(g_ctx));
  std::exit(1);
}

struct HostRays {
  uint32_t n = 0;
  std::vector<float> a[8];
  rq_rays view() const {
    return rq_rays{a[0].data(), a[1].data(), a[2].data(), a[3].data(),
                   a[4].data(), a[5].data(), a[6].data(), a[7].data()};
  }
};

bool read_rays(const char* path, HostRays* r) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  bool ok = std::fread(&r->n, 4, 1, f) == 1;
  for (int k = 0; ok && k < 8; ++k) {
    r->a[k].resize(r->n);
    ok = std::fread(r->a[k].data(), 4, r->n, f) == r->n;
  }
  std::fclose(f);
  return ok;
}

template <typename T>
void put(FILE* f, const std::vector<T>& v) {
  std::fwrite(v.data(), sizeof(T), v.size(), f);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s mesh.obj rays.bin out.bin\n", argv[0]);
    return 2;
  }
  HostRays rays;
  if (!read_rays(argv[2], &rays)) {
    std::fprintf(stderr, "cannot read %s\n", argv[2]);
    return 2;
  }
  const uint32_t n = rays.n, k = n < 1000 ? n : 1000;
  const rq_settings settings = rq_settings_default();

  std::vector<int32_t> prim(n), prim_h(n), prim_r(k);
  std::vector<float> t(n), t_h(n), t_r(k);
  std::vector<uint8_t> occ(n), occ_h(n), occ_r(k);
  rq_bvh_info info{};

  switch (settings.data_structure) {
    case RQ_GPU_LBVH: {
      // ---- build side (the case main.cpp:791-843 gains)
      check(rq_create(&g_ctx, 0, 0), "rq_create");
      rq_mesh m;
      check(rq_obj_read(argv[1], &m), "rq_obj_read");
      rq_scene* scene = nullptr;
      check(rq_scene_upload(g_ctx, m.vx, m.vy, m.vz, m.nv, m.idx, m.ntri, RQ_PRIM_TRIANGLE, &scene,
                            nullptr),
            "rq_scene_upload");
      rq_mesh_free(&m);
      rq_bvh* bvh = nullptr;
      check(rq_build_lbvh(g_ctx, scene, RQ_MORTON_SCENE_AABB, &bvh, nullptr), "rq_build_lbvh");
      check(rq_sync(g_ctx, nullptr), "rq_sync");
      check(rq_bvh_info_get(g_ctx, bvh, &info), "rq_bvh_info_get");

      // ---- query side, batched: device SoA rays and hits
      void* d[8];
      for (int j = 0; j < 8; ++j) {
        check(rq_device_alloc(g_ctx, (size_t)n * 4 + 4, &d[j]), "rq_device_alloc");
        check(rq_memcpy_h2d(g_ctx, d[j], rays.a[j].data(), (size_t)n * 4, nullptr), "h2d");
      }
      void *d_prim, *d_t, *d_occ;
      check(rq_device_alloc(g_ctx, (size_t)n * 4 + 4, &d_prim), "rq_device_alloc");
      check(rq_device_alloc(g_ctx, (size_t)n * 4 + 4, &d_t), "rq_device_alloc");
      check(rq_device_alloc(g_ctx, (size_t)n + 4, &d_occ), "rq_device_alloc");
      const rq_rays dr{(float*)d[0], (float*)d[1], (float*)d[2], (float*)d[3],
                       (float*)d[4], (float*)d[5], (float*)d[6], (float*)d[7]};
      rq_hits dh{(int32_t*)d_prim, (float*)d_t, nullptr, nullptr};
      check(rq_closest_hit(g_ctx, bvh, &dr, n, &dh, nullptr), "rq_closest_hit");
      check(rq_any_hit(g_ctx, bvh, &dr, n, (uint8_t*)d_occ, nullptr), "rq_any_hit");
      check(rq_sync(g_ctx, nullptr), "rq_sync");
      check(rq_memcpy_d2h(g_ctx, prim.data(), d_prim, (size_t)n * 4, nullptr), "d2h");
      check(rq_memcpy_d2h(g_ctx, t.data(), d_t, (size_t)n * 4, nullptr), "d2h");
      check(rq_memcpy_d2h(g_ctx, occ.data(), d_occ, n, nullptr), "d2h");
      check(rq_sync(g_ctx, nullptr), "rq_sync");
      for (int j = 0; j < 8; ++j) check(rq_device_free(g_ctx, d[j]), "rq_device_free");
      check(rq_device_free(g_ctx, d_prim), "rq_device_free");
      check(rq_device_free(g_ctx, d_t), "rq_device_free");
      check(rq_device_free(g_ctx, d_occ), "rq_device_free");

      // ---- the same rays from host arrays
      const rq_rays hr = rays.view();
      rq_hits hh{prim_h.data(), t_h.data(), nullptr, nullptr};
      check(rq_closest_hit_host(g_ctx, bvh, &hr, n, &hh, nullptr), "rq_closest_hit_host");
      check(rq_any_hit_host(g_ctx, bvh, &hr, n, occ_h.data(), nullptr), "rq_any_hit_host");

      // ---- per ray, as castRay's loop over pixels calls boxIntersect and
      // kdtreeIntersect with one ray each
      for (uint32_t i = 0; i < k; ++i) {
        const float o[3] = {rays.a[0][i], rays.a[1][i], rays.a[2][i]};
        const float dir[3] = {rays.a[3][i], rays.a[4][i], rays.a[5][i]};
        float uv[2];
        check(rq_cast_ray(g_ctx, bvh, o, dir, rays.a[6][i], rays.a[7][i], &prim_r[i], &t_r[i], uv),
              "rq_cast_ray");
        check(rq_occluded_ray(g_ctx, bvh, o, dir, rays.a[6][i], rays.a[7][i], &occ_r[i]),
              "rq_occluded_ray");
      }
      uint32_t bits = 0;
      check(rq_device_errors(g_ctx, &bits), "rq_device_errors");
      if (bits) {
        std::fprintf(stderr, "device error bits %u\n", bits);
        return 1;
      }
      check(rq_release(g_ctx, bvh), "rq_release");
      check(rq_release(g_ctx, scene), "rq_release");
      check(rq_destroy(g_ctx), "rq_destroy");
      g_ctx = nullptr;
      break;
    }
    default:
      std::fprintf(stderr, "this caller only runs GPU_LBVH\n");
      return 2;
  }

  FILE* f = std::fopen(argv[3], "wb");
  if (!f) return 2;
  std::fwrite(&n, 4, 1, f);
  std::fwrite(&k, 4, 1, f);
  put(f, prim), put(f, t), put(f, occ);
  put(f, prim_h), put(f, t_h), put(f, occ_h);
  put(f, prim_r), put(f, t_r), put(f, occ_r);
  std::fwrite(&info.n_prims, 4, 1, f);
  std::fwrite(&info.height, 4, 1, f);
  std::fclose(f);
  std::printf("caller ok: %u rays, %u cast one by one, %u triangles\n", n, k, info.n_prims);
  return 0;
}
