// Cached BVH blob (include/rq/rq.h, rq_bvh_save / rq_bvh_load): a built LBVH
// as bytes, so that a scene loaded again skips its build. The reference keeps
// nothing between runs (its only artefact is output.ppm); this is the cache in
// front of the build path.
//
// Layout (little-endian):
//   header, 64 bytes: magic "RQBVHBL1", version, primitive kind, Morton mode,
//     primitive count n, scene element count, scene fingerprint (FNV-1a 64 of
//     the scene's device arrays), payload bytes, payload checksum (FNV-1a 64)
//   payload: codes[n], sorted primitive index[n], parent[2n-1] (u32 each),
//     internal nodes[n-1] (64 B), leaf records[n] (48 B), root box (6 floats),
//     height (u32)
// Loading checks magic, version, sizes and checksum (RQ_E_IO on a damaged
// blob) and that the scene matches the fingerprint (RQ_E_INVALID otherwise),
// then copies the arrays into a new BVH of that scene.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "rq_ctx.h"

using namespace rq;

namespace {

constexpr char kMagic[8] = {'R', 'Q', 'B', 'V', 'H', 'B', 'L', '1'};
constexpr uint32_t kVersion = 1;

struct Header {
  char magic[8];
  uint32_t version;
  int32_t kind;
  int32_t mode;
  uint32_t n;
  uint32_t scene_elems;
  uint32_t pad0;
  uint64_t scene_fp;
  uint64_t payload_bytes;
  uint64_t payload_sum;
  uint64_t pad1;
};
static_assert(sizeof(Header) == 64, "64-byte blob header");

uint64_t fnv1a(const void* p, size_t n, uint64_t h = 1469598103934665603ull) {
  const uint8_t* b = static_cast<const uint8_t*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

size_t payload_bytes(uint32_t n) {
  const size_t ni = n > 1 ? n - 1 : 0;
  return 4ull * n + 4ull * n + 4ull * (n ? 2ull * n - 1 : 0) + 64 * ni + 48ull * n + 6 * 4 + 4;
}

// FNV-1a of the scene's device arrays (vertices + indices, or centres + radii).
int scene_fingerprint(rq_ctx* c, const rq_scene* s, uint64_t* fp, uint32_t* elems) {
  uint64_t h = 1469598103934665603ull;
  std::vector<uint8_t> buf;
  auto add = [&](const void* dev, size_t bytes) -> hipError_t {
    if (!bytes) return hipSuccess;
    buf.resize(bytes);
    hipError_t e = hipMemcpy(buf.data(), dev, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess) h = fnv1a(buf.data(), bytes, h);
    return e;
  };
  hipError_t e = add(s->vx, 4ull * s->nv);
  if (e == hipSuccess) e = add(s->vy, 4ull * s->nv);
  if (e == hipSuccess) e = add(s->vz, 4ull * s->nv);
  if (e == hipSuccess && s->kind == RQ_PRIM_SPHERE) e = add(s->r, 4ull * s->ntri);
  if (e == hipSuccess && s->kind == RQ_PRIM_TRIANGLE) e = add(s->idx, 12ull * s->ntri);
  if (e != hipSuccess) return hip_err(c, e, "scene fingerprint");
  *fp = h;
  *elems = s->nv;
  return RQ_OK;
}

}  // namespace

extern "C" {

int rq_bvh_blob_size(rq_ctx* c, const rq_bvh* b, size_t* bytes) {
  if (!c || !b || b->magic != BVH_MAGIC || !bytes) return RQ_E_INVALID;
  *bytes = sizeof(Header) + payload_bytes(b->n);
  return RQ_OK;
}

int rq_bvh_save(rq_ctx* c, const rq_bvh* b, void* blob, size_t cap, size_t* written) {
  if (!c || !b || b->magic != BVH_MAGIC || !blob) return set_err(c, RQ_E_INVALID, "bad argument");
  const uint32_t n = b->n;
  const size_t pb = payload_bytes(n), total = sizeof(Header) + pb;
  if (cap < total) return set_err(c, RQ_E_INVALID, "blob buffer too small (%zu < %zu)", cap, total);
  int hrc = ensure_height(c, b);
  if (hrc) return hrc;
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_err(c, e, "save sync");
  Header h{};
  std::memcpy(h.magic, kMagic, 8);
  h.version = kVersion;
  h.kind = b->scene->kind;
  h.mode = b->mode;
  h.n = n;
  int rc = scene_fingerprint(c, b->scene, &h.scene_fp, &h.scene_elems);
  if (rc) return rc;
  uint8_t* p = static_cast<uint8_t*>(blob) + sizeof(Header);
  uint8_t* q = p;
  const size_t ni = n > 1 ? n - 1 : 0, np = n ? 2ull * n - 1 : 0;
  auto get = [&](const void* dev, size_t bytes) {
    if (e == hipSuccess && bytes) e = hipMemcpy(q, dev, bytes, hipMemcpyDeviceToHost);
    q += bytes;
  };
  get(b->B.codes, 4ull * n);
  get(b->B.vals_a, 4ull * n);
  get(b->B.parent, 4 * np);
  get(b->B.nodes, 64 * ni);
  get(b->B.tris, 48ull * n);
  uint32_t misc[MISC_WORDS] = {};
  if (e == hipSuccess && n) e = hipMemcpy(misc, b->B.misc, sizeof misc, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_err(c, e, "save copy");
  if (n && misc[MISC_ERR]) return set_err(c, RQ_E_INVALID, "the BVH's build reported errors");
  std::memcpy(q, misc + MISC_ROOT, 24), q += 24;
  std::memcpy(q, misc + MISC_HEIGHT, 4), q += 4;
  h.payload_bytes = pb;
  h.payload_sum = fnv1a(p, pb);
  std::memcpy(blob, &h, sizeof h);
  if (written) *written = total;
  return RQ_OK;
}

int rq_bvh_load(rq_ctx* c, const rq_scene* s, const void* blob, size_t bytes, rq_bvh** out,
                void* stream) {
  if (!c || !out) return RQ_E_INVALID;
  *out = nullptr;
  if (!s || s->magic != SCENE_MAGIC) return set_err(c, RQ_E_INVALID, "bad scene handle");
  if (!blob || bytes < sizeof(Header)) return set_err(c, RQ_E_IO, "blob shorter than its header");
  Header h;
  std::memcpy(&h, blob, sizeof h);
  if (std::memcmp(h.magic, kMagic, 8) != 0 || h.version != kVersion)
    return set_err(c, RQ_E_IO, "not a BVH blob of this version");
  if (h.n > MAX_PRIMS || h.payload_bytes != payload_bytes(h.n) ||
      bytes < sizeof(Header) + h.payload_bytes)
    return set_err(c, RQ_E_IO, "blob sizes inconsistent");
  const uint8_t* p = static_cast<const uint8_t*>(blob) + sizeof(Header);
  if (fnv1a(p, h.payload_bytes) != h.payload_sum)
    return set_err(c, RQ_E_IO, "blob checksum mismatch");
  if (h.kind != s->kind || h.n != s->ntri) return set_err(c, RQ_E_INVALID, "blob is for another scene");
  uint64_t fp;
  uint32_t elems;
  int rc = scene_fingerprint(c, s, &fp, &elems);
  if (rc) return rc;
  if (fp != h.scene_fp || elems != h.scene_elems)
    return set_err(c, RQ_E_INVALID, "blob is for another scene (fingerprint)");
  hipSetDevice(c->device);
  auto* b = new rq_bvh();
  b->magic = BVH_MAGIC;
  b->scene = s;
  b->mode = h.mode;
  b->n = h.n;
  rc = bvh_alloc(c, b, b->n, true);
  if (rc) {
    delete b;
    return rc;
  }
  const uint32_t n = h.n;
  const size_t ni = n > 1 ? n - 1 : 0, np = n ? 2ull * n - 1 : 0;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipSuccess;
  auto put = [&](void* dev, size_t nbytes) {
    if (e == hipSuccess && nbytes) e = hipMemcpyAsync(dev, p, nbytes, hipMemcpyHostToDevice, st);
    p += nbytes;
  };
  put(b->B.codes, 4ull * n);
  put(b->B.vals_a, 4ull * n);
  put(b->B.parent, 4 * np);
  put(b->B.nodes, 64 * ni);
  put(b->B.tris, 48ull * n);
  uint32_t misc[MISC_WORDS] = {};
  std::memcpy(misc + MISC_ROOT, p, 24), p += 24;
  std::memcpy(misc + MISC_HEIGHT, p, 4), p += 4;
  if (e == hipSuccess && n) e = hipMemcpyAsync(b->B.misc, misc, sizeof misc, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = launch_widen(launch_ctx(c, stream), b->B, n);  // derived, not stored
  if (e == hipSuccess) e = hipStreamSynchronize(st);  // the blob may go away after return
  if (e != hipSuccess) {
    ctx_free(c, b->base);
    delete b;
    return hip_err(c, e, "blob load");
  }
  *out = b;
  return RQ_OK;
}

int rq_bvh_save_file(rq_ctx* c, const rq_bvh* b, const char* path) {
  if (!path) return RQ_E_INVALID;
  size_t bytes = 0;
  int rc = rq_bvh_blob_size(c, b, &bytes);
  if (rc) return rc;
  std::vector<uint8_t> buf(bytes);
  rc = rq_bvh_save(c, b, buf.data(), bytes, nullptr);
  if (rc) return rc;
  FILE* f = std::fopen(path, "wb");
  if (!f) return set_err(c, RQ_E_IO, "cannot write %s", path);
  const size_t w = std::fwrite(buf.data(), 1, bytes, f);
  const int cl = std::fclose(f);
  return (w == bytes && cl == 0) ? RQ_OK : set_err(c, RQ_E_IO, "short write to %s", path);
}

int rq_bvh_load_file(rq_ctx* c, const rq_scene* s, const char* path, rq_bvh** out) {
  if (!path || !out) return RQ_E_INVALID;
  FILE* f = std::fopen(path, "rb");
  if (!f) return set_err(c, RQ_E_IO, "cannot read %s", path);
  std::fseek(f, 0, SEEK_END);
  const long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  std::vector<uint8_t> buf(len > 0 ? (size_t)len : 0);
  const size_t got = len > 0 ? std::fread(buf.data(), 1, buf.size(), f) : 0;
  std::fclose(f);
  if (got != buf.size()) return set_err(c, RQ_E_IO, "short read of %s", path);
  return rq_bvh_load(c, s, buf.data(), buf.size(), out, nullptr);
}

}  // extern "C"
