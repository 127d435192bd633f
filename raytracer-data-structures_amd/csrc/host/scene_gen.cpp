// Host-side scene generators and OBJ I/O (include/rq/rq_scene.h).
// The reference loads meshes in createScene_new (main.cpp:599-721), reading
// only `v` lines (it stops at the first other token, :694-697) and making a
// sphere per vertex: rq_gen_reference_scene rebuilds that sphere scene. The
// triangle scenes are generated here deterministically, or read from OBJ with
// both `v` and `f` records.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rq/rq.h"
#include "rq/rq_scene.h"

namespace {

struct Mesh {
  std::vector<float> x, y, z;
  std::vector<uint32_t> idx;
};

int export_mesh(Mesh&& m, rq_mesh* out) {
  std::memset(out, 0, sizeof *out);
  size_t nv = m.x.size(), nt = m.idx.size() / 3;
  out->vx = (float*)std::malloc(nv * 4 + 1);
  out->vy = (float*)std::malloc(nv * 4 + 1);
  out->vz = (float*)std::malloc(nv * 4 + 1);
  out->idx = (uint32_t*)std::malloc(nt * 12 + 1);
  if (!out->vx || !out->vy || !out->vz || !out->idx) {
    rq_mesh_free(out);
    return RQ_E_NOMEM;
  }
  std::memcpy(out->vx, m.x.data(), nv * 4);
  std::memcpy(out->vy, m.y.data(), nv * 4);
  std::memcpy(out->vz, m.z.data(), nv * 4);
  std::memcpy(out->idx, m.idx.data(), nt * 12);
  out->nv = (uint32_t)nv;
  out->ntri = (uint32_t)nt;
  return RQ_OK;
}

// lowbias32 integer hash (C. Wellons' hash-prospector constants).
inline uint32_t hash32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return h;
}

inline float lattice(int32_t i, int32_t j, int32_t k, uint32_t seed) {
  uint32_t h = hash32((uint32_t)i * 0x8DA6B343u ^ hash32((uint32_t)j * 0xD8163841u ^
                                                         hash32((uint32_t)k * 0xCB1AB31Fu ^ seed)));
  return (float)(h >> 8) * (2.0f / 16777216.0f) - 1.0f;  // [-1, 1)
}

inline float lerp(float a, float b, float t) { return a + (b - a) * t; }

float vnoise3(float x, float y, float z, uint32_t seed) {
  float fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
  int32_t ix = (int32_t)fx, iy = (int32_t)fy, iz = (int32_t)fz;
  float tx = x - fx, ty = y - fy, tz = z - fz;
  float sx = tx * tx * (3.0f - 2.0f * tx), sy = ty * ty * (3.0f - 2.0f * ty),
        sz = tz * tz * (3.0f - 2.0f * tz);
  float c[2][2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
  float x00 = lerp(c[0][0][0], c[1][0][0], sx), x10 = lerp(c[0][1][0], c[1][1][0], sx);
  float x01 = lerp(c[0][0][1], c[1][0][1], sx), x11 = lerp(c[0][1][1], c[1][1][1], sx);
  float y0 = lerp(x00, x10, sy), y1 = lerp(x01, x11, sy);
  return lerp(y0, y1, sz);
}

// Unit icosphere by midpoint subdivision; vertex order: the 12 icosahedron
// vertices, then midpoints in order of first use.
void icosphere_unit(uint32_t level, std::vector<double>& P, std::vector<uint32_t>& F) {
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  const double base[12][3] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0},
                              {0, -1, t}, {0, 1, t}, {0, -1, -t}, {0, 1, -t},
                              {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  const uint32_t faces[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                                 {1, 5, 9},  {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                 {3, 9, 4},  {3, 4, 2},  {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                                 {4, 9, 5},  {2, 4, 11}, {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};
  P.clear();
  F.clear();
  auto addv = [&](double x, double y, double z) {
    double n = std::sqrt(x * x + y * y + z * z);
    P.push_back(x / n), P.push_back(y / n), P.push_back(z / n);
    return (uint32_t)(P.size() / 3 - 1);
  };
  for (auto& v : base) addv(v[0], v[1], v[2]);
  for (auto& f : faces) F.insert(F.end(), {f[0], f[1], f[2]});
  for (uint32_t l = 0; l < level; ++l) {
    std::unordered_map<uint64_t, uint32_t> cache;
    cache.reserve(F.size());
    auto mid = [&](uint32_t a, uint32_t b) {
      uint64_t key = ((uint64_t)std::min(a, b) << 32) | std::max(a, b);
      auto it = cache.find(key);
      if (it != cache.end()) return it->second;
      uint32_t m = addv((P[3 * a] + P[3 * b]) * 0.5, (P[3 * a + 1] + P[3 * b + 1]) * 0.5,
                        (P[3 * a + 2] + P[3 * b + 2]) * 0.5);
      cache.emplace(key, m);
      return m;
    };
    std::vector<uint32_t> nf;
    nf.reserve(F.size() * 4);
    for (size_t i = 0; i < F.size(); i += 3) {
      uint32_t a = F[i], b = F[i + 1], c = F[i + 2];
      uint32_t ab = mid(a, b), bc = mid(b, c), ca = mid(c, a);
      nf.insert(nf.end(), {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca});
    }
    F.swap(nf);
  }
}

void displaced_icosphere(uint32_t level, uint32_t seed, float amp, float freq,
                         const float center[3], float scale, Mesh& m, uint32_t vbase) {
  std::vector<double> P;
  std::vector<uint32_t> F;
  icosphere_unit(level, P, F);
  size_t nv = P.size() / 3;
  for (size_t i = 0; i < nv; ++i) {
    float px = (float)P[3 * i], py = (float)P[3 * i + 1], pz = (float)P[3 * i + 2];
    float r = 1.0f + amp * vnoise3(freq * px, freq * py, freq * pz, seed);
    float s = scale * r;
    m.x.push_back(center[0] + s * px);
    m.y.push_back(center[1] + s * py);
    m.z.push_back(center[2] + s * pz);
  }
  for (uint32_t f : F) m.idx.push_back(vbase + f);
}

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {  // Box-Muller
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

}  // namespace

extern "C" {

int rq_gen_icosphere(uint32_t level, uint32_t seed, float amp, float freq, const float center[3],
                     float scale, rq_mesh* out) {
  if (!out || level > 12) return RQ_E_INVALID;
  const float zero[3] = {0, 0, 0};
  Mesh m;
  displaced_icosphere(level, seed, amp, freq, center ? center : zero, scale, m, 0);
  return export_mesh(std::move(m), out);
}

int rq_gen_icosphere_field(uint32_t count, uint32_t level, uint32_t seed0,
                           uint64_t placement_seed, float box, float rmin, float amp, float freq,
                           rq_mesh* out) {
  if (!out || level > 10) return RQ_E_INVALID;
  SplitMix64 rng{placement_seed};
  std::vector<float> C;
  uint64_t attempts = 0;
  while (C.size() / 3 < count) {
    if (++attempts > 1000000ull * (count + 1)) return RQ_E_INVALID;  // rmin too large for box
    float c[3] = {(float)(rng.uniform() * box), (float)(rng.uniform() * box),
                  (float)(rng.uniform() * box)};
    bool ok = true;
    for (size_t j = 0; j < C.size() && ok; j += 3) {
      float dx = C[j] - c[0], dy = C[j + 1] - c[1], dz = C[j + 2] - c[2];
      ok = dx * dx + dy * dy + dz * dz >= rmin * rmin;
    }
    if (ok) C.insert(C.end(), {c[0], c[1], c[2]});
  }
  Mesh m;
  for (uint32_t i = 0; i < count; ++i)
    displaced_icosphere(level, seed0 + i, amp, freq, &C[3 * i], 1.0f, m, (uint32_t)m.x.size());
  return export_mesh(std::move(m), out);
}

int rq_gen_soup(uint32_t ntri, uint64_t seed, float mean_edge, float sd_edge, float min_edge,
                rq_mesh* out) {
  if (!out) return RQ_E_INVALID;
  SplitMix64 rng{seed};
  Mesh m;
  m.x.reserve(3 * (size_t)ntri);
  for (uint32_t t = 0; t < ntri; ++t) {
    double cx = rng.uniform(), cy = rng.uniform(), cz = rng.uniform();
    double a[3], b[3];
    for (;;) {  // random unit vector by rejection
      a[0] = 2 * rng.uniform() - 1, a[1] = 2 * rng.uniform() - 1, a[2] = 2 * rng.uniform() - 1;
      double r2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
      if (r2 > 1e-6 && r2 <= 1) {
        double f = 1 / std::sqrt(r2);
        a[0] *= f, a[1] *= f, a[2] *= f;
        break;
      }
    }
    for (;;) {  // second direction, orthogonalised
      double w[3] = {2 * rng.uniform() - 1, 2 * rng.uniform() - 1, 2 * rng.uniform() - 1};
      double d = w[0] * a[0] + w[1] * a[1] + w[2] * a[2];
      b[0] = w[0] - d * a[0], b[1] = w[1] - d * a[1], b[2] = w[2] - d * a[2];
      double r2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
      if (r2 > 1e-6) {
        double f = 1 / std::sqrt(r2);
        b[0] *= f, b[1] *= f, b[2] *= f;
        break;
      }
    }
    double L = mean_edge + sd_edge * rng.normal();
    if (L < min_edge) L = min_edge;
    double R = L / std::sqrt(3.0), phi = 6.283185307179586 * rng.uniform();
    uint32_t base = (uint32_t)m.x.size();
    for (int k = 0; k < 3; ++k) {
      double th = phi + k * 2.0943951023931953;
      double cs = std::cos(th) * R, sn = std::sin(th) * R;
      m.x.push_back((float)(cx + cs * a[0] + sn * b[0]));
      m.y.push_back((float)(cy + cs * a[1] + sn * b[1]));
      m.z.push_back((float)(cz + cs * a[2] + sn * b[2]));
    }
    m.idx.insert(m.idx.end(), {base, base + 1, base + 2});
  }
  return export_mesh(std::move(m), out);
}

// OBJ parsing in parallel over a memory-mapped file. Pass 1 splits the file
// at line boundaries into chunks and counts each chunk's `v` records and
// triangles (fan of each `f` record); prefix sums give every chunk its first
// vertex and first triangle. Pass 2 parses the chunks into the final arrays:
// an `f` record's relative (negative) indices and range check use the number
// of vertices before that record, as a sequential reader would.
namespace {

struct ObjChunk {
  const char *b, *e;
  uint64_t nv = 0, nt = 0;    // pass 1 counts
  uint64_t v0 = 0, t0 = 0;    // prefix offsets
  bool bad = false;
};

inline const char* skip_ws(const char* p, const char* e) {
  while (p < e && (*p == ' ' || *p == '\t')) ++p;
  return p;
}

inline bool is_rec(const char* p, const char* eol, char k) {
  return p + 1 < eol && p[0] == k && (p[1] == ' ' || p[1] == '\t');
}

// vertices of an `f` record (tokens up to the end of line)
inline int64_t face_count(const char* q, const char* eol) {
  int64_t k = 0;
  while (q < eol) {
    while (q < eol && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
    if (q >= eol) break;
    ++k;
    while (q < eol && *q != ' ' && *q != '\t' && *q != '\r') ++q;
  }
  return k;
}

void count_chunk(ObjChunk& c) {
  for (const char* p = c.b; p < c.e;) {
    const char* eol = (const char*)std::memchr(p, '\n', (size_t)(c.e - p));
    if (!eol) eol = c.e;
    p = skip_ws(p, eol);
    if (is_rec(p, eol, 'v')) {
      ++c.nv;
    } else if (is_rec(p, eol, 'f')) {
      const int64_t k = face_count(p + 1, eol);
      if (k >= 3) c.nt += (uint64_t)(k - 2);
    }
    p = eol + 1;
  }
}

void parse_chunk(ObjChunk& c, const char* file_end, float* x, float* y, float* z, uint32_t* idx) {
  uint64_t nv = c.v0, nt = c.t0;
  std::vector<int64_t> poly;
  for (const char* p = c.b; p < c.e && !c.bad;) {
    const char* eol = (const char*)std::memchr(p, '\n', (size_t)(c.e - p));
    if (!eol) eol = c.e;
    p = skip_ws(p, eol);
    if (is_rec(p, eol, 'v')) {
      char* q;
      const float a = std::strtof(p + 1, &q);
      const float b = std::strtof(q, &q);
      const float d = std::strtof(q, &q);
      if (q > eol || q > file_end) {
        c.bad = true;
        break;
      }
      x[nv] = a, y[nv] = b, z[nv] = d;
      ++nv;
    } else if (is_rec(p, eol, 'f')) {
      poly.clear();
      const char* q = p + 1;
      while (q < eol) {
        while (q < eol && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
        if (q >= eol) break;
        char* r;
        const long long v = std::strtoll(q, &r, 10);
        const int64_t i0 = v > 0 ? v - 1 : (int64_t)nv + v;
        if (r == q || v == 0 || i0 < 0 || i0 >= (int64_t)nv) {
          c.bad = true;
          break;
        }
        poly.push_back(i0);
        q = r;
        while (q < eol && *q != ' ' && *q != '\t' && *q != '\r') ++q;  // skip /vt/vn
      }
      for (size_t k = 1; k + 1 < poly.size() && !c.bad; ++k, ++nt) {
        idx[3 * nt] = (uint32_t)poly[0];
        idx[3 * nt + 1] = (uint32_t)poly[k];
        idx[3 * nt + 2] = (uint32_t)poly[k + 1];
      }
    }
    p = eol + 1;
  }
}

}  // namespace

int rq_obj_read_ex(const char* path, int threads, rq_mesh* out) {
  if (!path || !out) return RQ_E_INVALID;
  std::memset(out, 0, sizeof *out);
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) return RQ_E_IO;
  struct stat sb;
  if (::fstat(fd, &sb) != 0) {
    ::close(fd);
    return RQ_E_IO;
  }
  const size_t len = (size_t)sb.st_size;
  const long page = ::sysconf(_SC_PAGESIZE);
  // strtof may look one byte past a last line without a newline: the mapping
  // gives zero bytes after the end unless the file ends on a page boundary,
  // in which case the file is read into a terminated buffer instead
  std::string copy;
  const char* data = nullptr;
  void* map = MAP_FAILED;
  if (len && page > 0 && len % (size_t)page != 0) {
    map = ::mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
    if (map != MAP_FAILED) data = (const char*)map;
  }
  if (!data && len) {
    copy.resize(len);
    size_t got = 0;
    while (got < len) {
      const ssize_t r = ::pread(fd, &copy[got], len - got, (off_t)got);
      if (r <= 0) break;
      got += (size_t)r;
    }
    if (got != len) {
      ::close(fd);
      return RQ_E_IO;
    }
    data = copy.c_str();
  }
  ::close(fd);
  const char* end = data + len;
  int nt = threads > 0 ? threads : (int)std::min<unsigned>(16, std::max(1u, std::thread::hardware_concurrency()));
  nt = (int)std::max<size_t>(1, std::min<size_t>((size_t)nt, len / (256 << 10) + 1));
  std::vector<ObjChunk> ch;
  for (int k = 0; k < nt && len; ++k) {
    const char* b = k == 0 ? data : ch.back().e;
    const char* e = k == nt - 1 ? end : data + len * (size_t)(k + 1) / (size_t)nt;
    if (e < b) e = b;
    while (e < end && e > data && e[-1] != '\n') ++e;  // end after a newline
    ObjChunk c;
    c.b = b, c.e = e;
    ch.push_back(c);
  }
  auto run = [&](auto&& f) {
    std::vector<std::thread> th;
    for (size_t k = 1; k < ch.size(); ++k) th.emplace_back([&, k] { f(ch[k]); });
    if (!ch.empty()) f(ch[0]);
    for (auto& t : th) t.join();
  };
  run([](ObjChunk& c) { count_chunk(c); });
  uint64_t nv = 0, ntri = 0;
  for (auto& c : ch) c.v0 = nv, c.t0 = ntri, nv += c.nv, ntri += c.nt;
  int rc = RQ_OK;
  if (nv > 0xFFFFFFFFull || ntri > 0xFFFFFFFFull / 3) rc = RQ_E_INVALID;
  Mesh m;
  if (!rc) {
    m.x.resize(nv), m.y.resize(nv), m.z.resize(nv), m.idx.resize(3 * ntri);
    run([&](ObjChunk& c) { parse_chunk(c, end, m.x.data(), m.y.data(), m.z.data(), m.idx.data()); });
    for (auto& c : ch)
      if (c.bad) rc = RQ_E_IO;
  }
  if (map != MAP_FAILED) ::munmap(map, len);
  if (rc) return rc;
  return export_mesh(std::move(m), out);
}

int rq_obj_read(const char* path, rq_mesh* out) { return rq_obj_read_ex(path, 0, out); }

int rq_obj_write(const char* path, const rq_mesh* m) {
  if (!path || !m) return RQ_E_INVALID;
  FILE* f = std::fopen(path, "wb");
  if (!f) return RQ_E_IO;
  for (uint32_t i = 0; i < m->nv; ++i)
    std::fprintf(f, "v %.9g %.9g %.9g\n", m->vx[i], m->vy[i], m->vz[i]);
  for (uint32_t t = 0; t < m->ntri; ++t)
    std::fprintf(f, "f %u %u %u\n", m->idx[3 * t] + 1, m->idx[3 * t + 1] + 1,
                 m->idx[3 * t + 2] + 1);
  return std::fclose(f) == 0 ? RQ_OK : RQ_E_IO;
}

void rq_mesh_bounds(const rq_mesh* m, float lo[3], float hi[3]) {
  lo[0] = lo[1] = lo[2] = INFINITY;
  hi[0] = hi[1] = hi[2] = -INFINITY;
  for (uint32_t i = 0; i < m->nv; ++i) {
    float v[3] = {m->vx[i], m->vy[i], m->vz[i]};
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::fmin(lo[k], v[k]);
      hi[k] = std::fmax(hi[k], v[k]);
    }
  }
}

int rq_gen_reference_scene(const char* path, uint32_t clones, uint32_t max_vertices,
                           rq_spheres* out) {
  if (!path || !out || clones == 0) return RQ_E_INVALID;
  std::memset(out, 0, sizeof *out);
  std::vector<float> v;  // x y z per vertex of one copy of the file
  {
    std::ifstream f(path);
    if (!f.good()) return RQ_E_IO;
    // main.cpp:665-703: whitespace-separated tokens; a `v` token is followed
    // by three floats; the first other token (or the end) stops the reading
    for (;;) {
      std::string tok;
      f >> tok;
      if (tok != "v" || (max_vertices && v.size() / 3 >= max_vertices)) break;
      float x = 0, y = 0, z = 0;
      f >> x >> y >> z;
      v.push_back(x), v.push_back(y), v.push_back(z);
    }
  }
  const size_t nv = v.size() / 3, n = nv * clones + 1;
  if (n > 0xFFFFFFFFull) return RQ_E_INVALID;
  out->cx = (float*)std::malloc(n * 4);
  out->cy = (float*)std::malloc(n * 4);
  out->cz = (float*)std::malloc(n * 4);
  out->r = (float*)std::malloc(n * 4);
  out->rgb = (float*)std::malloc(n * 12);
  if (!out->cx || !out->cy || !out->cz || !out->r || !out->rgb) {
    rq_spheres_free(out);
    return RQ_E_NOMEM;
  }
  size_t k = 0;
  const float radius = (float)(0.01 * 5);
  for (uint32_t clone = 0; clone < clones; ++clone) {
    const float shift = (float)(clone * 20);
    for (size_t i = 0; i < nv; ++i, ++k) {
      // s.center = vertex * 100 + shift; s.center.y += -10; s.center.z += -60
      out->cx[k] = v[3 * i] * 100.0f + shift;
      out->cy[k] = (v[3 * i + 1] * 100.0f + shift) + -10.0f;
      out->cz[k] = (v[3 * i + 2] * 100.0f + shift) + -60.0f;
      out->r[k] = radius;
      out->rgb[3 * k] = 0.8f, out->rgb[3 * k + 1] = 0.7f, out->rgb[3 * k + 2] = 0.0f;
    }
  }
  // the reference's double literals narrowed to float (Vec3f constructor)
  out->cx[k] = (float)0.93591022, out->cy[k] = (float)-105.47120094, out->cz[k] = (float)-43.2363205;
  out->r[k] = 100.0f;
  out->rgb[3 * k] = out->rgb[3 * k + 1] = out->rgb[3 * k + 2] = 0.0f;
  out->n = (uint32_t)n;
  return 0;
}

void rq_spheres_free(rq_spheres* s) {
  if (!s) return;
  std::free(s->cx), std::free(s->cy), std::free(s->cz), std::free(s->r), std::free(s->rgb);
  std::memset(s, 0, sizeof *s);
}

void rq_mesh_free(rq_mesh* m) {
  if (!m) return;
  std::free(m->vx), std::free(m->vy), std::free(m->vz), std::free(m->idx);
  std::memset(m, 0, sizeof *m);
}

}  // extern "C"
