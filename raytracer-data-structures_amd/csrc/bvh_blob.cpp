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
// blob) and that the scene matches the fingerprint (RQ_E_INVALID otherwise).
// A checksum only detects accidents (anyone editing the file can recompute
// it), so the payload is then checked against the scene on the host, in
// O(n), before anything reaches the device (RQ_E_IO on any failure): every
// internal node and every leaf is reached exactly once from the root through
// in-range child links, the parent links agree with the child links, the
// stored height is the tree's and fits the traversal stack, the sorted
// primitive order is a permutation, every leaf record is the one the build
// writes for its primitive, and every node's child boxes are exactly the
// boxes below it (so the traversal culls where a fresh build would). Saving
// writes the two words no kernel reads (the root's parent, the unused lanes of
// the second child box) as -1 and 0, so a blob is a function of the tree.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
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

uint64_t fnv1a(const void* p, size_t n, uint64_t h = 14695981039346656037ull) {
  const uint8_t* b = static_cast<const uint8_t*>(p);
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
  return h;
}

size_t payload_bytes(uint32_t n) {
  const size_t ni = n > 1 ? n - 1 : 0;
  return 4ull * n + 4ull * n + 4ull * (n ? 2ull * n - 1 : 0) + 64 * ni + 48ull * n + 6 * 4 + 4;
}

// A scene's primitives on the host: triangles (vertices + 3 indices each) or
// spheres (centres + radii).
struct HostPrims {
  int kind = RQ_PRIM_TRIANGLE;
  const float *vx = nullptr, *vy = nullptr, *vz = nullptr;
  uint32_t nv = 0;
  const uint32_t* idx = nullptr;  // triangles
  const float* r = nullptr;       // spheres
  uint32_t n = 0;
};

// FNV-1a 64 over the arrays in device order: vx, vy, vz, then the radii
// (spheres) or the indices (triangles).
uint64_t prims_fingerprint(const HostPrims& s) {
  uint64_t h = 14695981039346656037ull;
  h = fnv1a(s.vx, 4ull * s.nv, h);
  h = fnv1a(s.vy, 4ull * s.nv, h);
  h = fnv1a(s.vz, 4ull * s.nv, h);
  if (s.kind == RQ_PRIM_SPHERE) h = fnv1a(s.r, 4ull * s.n, h);
  if (s.kind == RQ_PRIM_TRIANGLE) h = fnv1a(s.idx, 12ull * s.n, h);
  return h;
}

// The scene's device arrays copied to the host (for the fingerprint and the
// payload check).
struct HostScene {
  std::vector<float> vx, vy, vz, r;
  std::vector<uint32_t> idx;
  HostPrims view(const rq_scene* s) const {
    HostPrims p;
    p.kind = s->kind;
    p.vx = vx.data(), p.vy = vy.data(), p.vz = vz.data();
    p.nv = s->nv;
    p.idx = idx.empty() ? nullptr : idx.data();
    p.r = r.empty() ? nullptr : r.data();
    p.n = s->ntri;
    return p;
  }
};

int download_scene(rq_ctx* c, const rq_scene* s, HostScene* h) {
  h->vx.resize(s->nv), h->vy.resize(s->nv), h->vz.resize(s->nv);
  hipError_t e = hipDeviceSynchronize();
  auto get = [&](void* dst, const void* dev, size_t bytes) {
    if (e == hipSuccess && bytes) e = hipMemcpy(dst, dev, bytes, hipMemcpyDeviceToHost);
  };
  get(h->vx.data(), s->vx, 4ull * s->nv);
  get(h->vy.data(), s->vy, 4ull * s->nv);
  get(h->vz.data(), s->vz, 4ull * s->nv);
  if (s->kind == RQ_PRIM_SPHERE) {
    h->r.resize(s->ntri);
    get(h->r.data(), s->r, 4ull * s->ntri);
  } else {
    h->idx.resize(3ull * s->ntri);
    get(h->idx.data(), s->idx, 12ull * s->ntri);
  }
  return e == hipSuccess ? RQ_OK : hip_err(c, e, "scene download");
}

// The build's leaf record of primitive p and its box (leaf_record in the
// device build: e = v - v0 per component, box = fminf / fmaxf of the three
// vertices (centre -/+ r for spheres) with zeros canonicalised by + 0.0f).
void leaf_of(const HostPrims& s, uint32_t p, float rec[12], float box[6]) {
  std::memset(rec, 0, 12 * sizeof(float));
  if (s.kind == RQ_PRIM_SPHERE) {
    const float cx = s.vx[p], cy = s.vy[p], cz = s.vz[p], r = s.r[p];
    rec[0] = cx, rec[1] = cy, rec[2] = cz;
    rec[4] = r, rec[5] = r * r;
    box[0] = (cx - r) + 0.0f, box[1] = (cy - r) + 0.0f, box[2] = (cz - r) + 0.0f;
    box[3] = (cx + r) + 0.0f, box[4] = (cy + r) + 0.0f, box[5] = (cz + r) + 0.0f;
  } else {
    const uint32_t a = s.idx[3ull * p], b = s.idx[3ull * p + 1], c = s.idx[3ull * p + 2];
    const float v0[3] = {s.vx[a], s.vy[a], s.vz[a]}, v1[3] = {s.vx[b], s.vy[b], s.vz[b]},
                v2[3] = {s.vx[c], s.vy[c], s.vz[c]};
    for (int q = 0; q < 3; ++q) {
      rec[q] = v0[q];
      rec[4 + q] = v1[q] - v0[q];
      rec[8 + q] = v2[q] - v0[q];
      box[q] = std::fmin(std::fmin(v0[q], v1[q]), v2[q]) + 0.0f;
      box[3 + q] = std::fmax(std::fmax(v0[q], v1[q]), v2[q]) + 0.0f;
    }
  }
}

bool same_bits(const float* a, const float* b, int k) { return std::memcmp(a, b, 4 * k) == 0; }

struct Why {
  std::string* out;
  int fail(int code, const char* fmt, ...) {
    if (out) {
      char buf[256];
      va_list ap;
      va_start(ap, fmt);
      std::vsnprintf(buf, sizeof buf, fmt, ap);
      va_end(ap);
      *out = buf;
    }
    return code;
  }
};

// Header and payload of a blob against the scene's primitives: RQ_OK,
// RQ_E_IO (damaged or not a well-formed LBVH of these primitives) or
// RQ_E_INVALID (a blob of another scene).
int check_blob(const void* blob, size_t bytes, const HostPrims& s, Header* hdr_out,
               std::string* why_out) {
  Why why{why_out};
  if (!blob || bytes < sizeof(Header)) return why.fail(RQ_E_IO, "blob shorter than its header");
  Header h;
  std::memcpy(&h, blob, sizeof h);
  if (std::memcmp(h.magic, kMagic, 8) != 0 || h.version != kVersion)
    return why.fail(RQ_E_IO, "not a BVH blob of this version");
  if (h.n > MAX_PRIMS || h.payload_bytes != payload_bytes(h.n) ||
      bytes < sizeof(Header) + h.payload_bytes)
    return why.fail(RQ_E_IO, "blob sizes inconsistent");
  const uint8_t* p = static_cast<const uint8_t*>(blob) + sizeof(Header);
  if (fnv1a(p, h.payload_bytes) != h.payload_sum)
    return why.fail(RQ_E_IO, "blob checksum mismatch");
  if (h.kind != s.kind || h.n != s.n) return why.fail(RQ_E_INVALID, "blob is for another scene");
  if (prims_fingerprint(s) != h.scene_fp || h.scene_elems != s.nv)
    return why.fail(RQ_E_INVALID, "blob is for another scene (fingerprint)");
  if (hdr_out) *hdr_out = h;
  const uint32_t n = h.n;
  if (n == 0) return RQ_OK;
  const size_t ni = n - 1;
  std::vector<uint32_t> sidx(n);
  std::vector<int32_t> parent(2ull * n - 1);
  std::vector<NodeRec> nodes(ni);
  std::vector<TriRec> tris(n);
  p += 4ull * n;  // codes: data only (exported, never used as an index)
  std::memcpy(sidx.data(), p, 4ull * n), p += 4ull * n;
  std::memcpy(parent.data(), p, 4 * parent.size()), p += 4 * parent.size();
  if (ni) std::memcpy(nodes.data(), p, 64 * ni);
  p += 64 * ni;
  std::memcpy(tris.data(), p, 48ull * n), p += 48ull * n;
  float root[6];
  uint32_t height;
  std::memcpy(root, p, 24), p += 24;
  std::memcpy(&height, p, 4);

  // sorted primitive order: a permutation; leaf records: the build's
  std::vector<float> lbox(6ull * n);
  std::vector<uint8_t> seen(n, 0);
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t q = sidx[k];
    if (q >= n || seen[q]) return why.fail(RQ_E_IO, "sorted order is not a permutation at %u", k);
    seen[q] = 1;
    float rec[12];
    leaf_of(s, q, rec, &lbox[6ull * k]);
    const float* t = &tris[k].v0.x;
    int32_t id;
    std::memcpy(&id, &tris[k].v0.w, 4);
    if (id != (int32_t)q || !same_bits(t, rec, 3) || !same_bits(t + 4, rec + 4, 4) ||
        !same_bits(t + 8, rec + 8, 4))
      return why.fail(RQ_E_IO, "leaf record %u differs from its primitive %u", k, q);
  }
  if (n == 1) {
    if (!same_bits(root, lbox.data(), 6)) return why.fail(RQ_E_IO, "root box");
    return height < TRACE_MAX_STACK ? RQ_OK : why.fail(RQ_E_IO, "height %u", height);
  }
  // topology from the root: in-range links, each node and leaf reached once,
  // parents consistent; the order of first visits, children after parents
  std::vector<uint8_t> vint(ni, 0);
  std::fill(seen.begin(), seen.end(), 0);
  std::vector<uint32_t> order;
  order.reserve(ni);
  std::vector<std::pair<uint32_t, uint32_t>> stack{{0u, 0u}};
  vint[0] = 1;
  uint32_t deepest = 0, leaves = 0;
  while (!stack.empty()) {
    const auto [i, depth] = stack.back();
    stack.pop_back();
    order.push_back(i);
    int32_t link[2];
    std::memcpy(&link[0], &nodes[i].lo0.w, 4);
    std::memcpy(&link[1], &nodes[i].hi0.w, 4);
    for (int side = 0; side < 2; ++side) {
      const int32_t c = link[side];
      const int32_t pe = (int32_t)((i << 1) | (uint32_t)side);
      if (c >= 0) {
        if ((uint32_t)c >= ni || c == 0 || vint[c])
          return why.fail(RQ_E_IO, "node %u: child link %d out of range or repeated", i, c);
        if (parent[c] != pe) return why.fail(RQ_E_IO, "parent link of node %d", c);
        vint[c] = 1;
        stack.push_back({(uint32_t)c, depth + 1});
      } else {
        const uint32_t k = ~(uint32_t)c;
        if (k >= n || seen[k])
          return why.fail(RQ_E_IO, "node %u: leaf link %u out of range or repeated", i, k);
        if (parent[ni + k] != pe) return why.fail(RQ_E_IO, "parent link of leaf %u", k);
        seen[k] = 1;
        ++leaves;
        deepest = std::max(deepest, depth + 1);
      }
    }
    if (depth + 1 >= (uint32_t)TRACE_MAX_STACK)
      return why.fail(RQ_E_IO, "tree deeper than the traversal stack");
  }
  if (order.size() != ni || leaves != n)
    return why.fail(RQ_E_IO, "%zu of %zu nodes and %u of %u leaves reachable", order.size(), ni,
                    leaves, n);
  if (height != deepest) return why.fail(RQ_E_IO, "height %u, tree has %u", height, deepest);
  // boxes, children before parents: each stored child box is exactly the
  // union below it
  std::vector<float> nbox(6 * ni);
  for (size_t o = order.size(); o-- > 0;) {
    const uint32_t i = order[o];
    const NodeRec& N = nodes[i];
    const float cb[2][6] = {{N.lo0.x, N.lo0.y, N.lo0.z, N.hi0.x, N.hi0.y, N.hi0.z},
                            {N.lo1.x, N.lo1.y, N.lo1.z, N.hi1.x, N.hi1.y, N.hi1.z}};
    int32_t link[2];
    std::memcpy(&link[0], &N.lo0.w, 4);
    std::memcpy(&link[1], &N.hi0.w, 4);
    for (int side = 0; side < 2; ++side) {
      const float* below = link[side] >= 0 ? &nbox[6ull * link[side]] : &lbox[6ull * ~(uint32_t)link[side]];
      if (!same_bits(cb[side], below, 6))
        return why.fail(RQ_E_IO, "node %u: box of child %d is not the union below it", i, side);
    }
    float* u = &nbox[6ull * i];
    for (int q = 0; q < 3; ++q) {
      u[q] = std::fmin(cb[0][q], cb[1][q]);
      u[3 + q] = std::fmax(cb[0][3 + q], cb[1][3 + q]);
    }
  }
  if (!same_bits(root, nbox.data(), 6)) return why.fail(RQ_E_IO, "root box");
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
  HostScene hs;
  int rc = download_scene(c, b->scene, &hs);
  if (rc) return rc;
  h.scene_fp = prims_fingerprint(hs.view(b->scene));
  h.scene_elems = b->scene->nv;
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
  // words no kernel reads: the root's parent, the second child box's .w lanes
  if (n) {
    int32_t* par = reinterpret_cast<int32_t*>(p + 8ull * n);
    par[0] = -1;
    NodeRec* nd = reinterpret_cast<NodeRec*>(p + 8ull * n + 4 * np);
    for (size_t i = 0; i < ni; ++i) nd[i].lo1.w = 0.0f, nd[i].hi1.w = 0.0f;
  }
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
  HostScene hs;
  int rc = download_scene(c, s, &hs);
  if (rc) return rc;
  Header h;
  std::string why;
  rc = check_blob(blob, bytes, hs.view(s), &h, &why);
  if (rc) return set_err(c, rc, "%s", why.c_str());
  const uint8_t* p = static_cast<const uint8_t*>(blob) + sizeof(Header);
  hipSetDevice(c->device);
  auto* b = new rq_bvh();
  b->magic = BVH_MAGIC;
  b->scene = s;
  b->mode = h.mode;
  b->n = h.n;
  rc = bvh_alloc(c, b, b->n);
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

int rq_bvh_blob_validate(const void* blob, size_t bytes, int prim_kind, const float* vx,
                         const float* vy, const float* vz, uint32_t nv, const uint32_t* idx,
                         const float* radius, uint32_t nprim, char* msg, size_t msg_cap) {
  if (msg && msg_cap) msg[0] = 0;
  if ((prim_kind != RQ_PRIM_TRIANGLE && prim_kind != RQ_PRIM_SPHERE) ||
      (nv && (!vx || !vy || !vz)) ||
      (nprim && (prim_kind == RQ_PRIM_TRIANGLE ? !idx : !radius)) ||
      (prim_kind == RQ_PRIM_SPHERE && nv != nprim))
    return RQ_E_INVALID;
  if (prim_kind == RQ_PRIM_TRIANGLE)
    for (uint64_t k = 0; k < 3ull * nprim; ++k)
      if (idx[k] >= nv) return RQ_E_INVALID;
  HostPrims s;
  s.kind = prim_kind;
  s.vx = vx, s.vy = vy, s.vz = vz, s.nv = nv;
  s.idx = prim_kind == RQ_PRIM_TRIANGLE ? idx : nullptr;
  s.r = prim_kind == RQ_PRIM_SPHERE ? radius : nullptr;
  s.n = nprim;
  std::string why;
  const int rc = check_blob(blob, bytes, s, nullptr, &why);
  if (rc && msg && msg_cap) std::snprintf(msg, msg_cap, "%s", why.c_str());
  return rc;
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
