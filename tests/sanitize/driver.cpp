// Host-code workout for sanitizer builds (AddressSanitizer + UndefinedBehavior
// Sanitizer, ThreadSanitizer): the threaded OBJ reader and the generators of
// librq's host part (csrc/host/scene_gen.cpp), and the threaded CPU oracle and
// reference-path port (oracle/oracle.cpp, oracle/port.cpp). Built from those
// sources by tests/test_sanitizers.py; exits 0 and prints "ok" when every
// result agrees across thread counts, nonzero otherwise.
//
//   driver <scratch dir>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "oracle.h"
#include "rq/rq_scene.h"

extern "C" {
void* port_bvh_build(const float*, const float*, const float*, const uint32_t*, uint32_t,
                     uint32_t*);
void* port_lbvh_build(const float*, const float*, const float*, const uint32_t*, uint32_t,
                      uint32_t*);
void port_bvh_free(void*);
void port_closest(const void*, const float*, const float*, const float*, const float*,
                  const float*, const float*, const float*, const float*, uint32_t, int32_t*,
                  float*, int);
}

namespace {

int failures = 0;
void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    ++failures;
  }
}

bool same_mesh(const rq_mesh& a, const rq_mesh& b) {
  return a.nv == b.nv && a.ntri == b.ntri &&
         !std::memcmp(a.vx, b.vx, 4 * (size_t)a.nv) && !std::memcmp(a.vy, b.vy, 4 * (size_t)a.nv) &&
         !std::memcmp(a.vz, b.vz, 4 * (size_t)a.nv) &&
         !std::memcmp(a.idx, b.idx, 12 * (size_t)a.ntri);
}

struct Rays {
  std::vector<float> a[8];
  explicit Rays(uint32_t n) {
    for (auto& v : a) v.resize(n);
  }
  or_rays view() const {
    return or_rays{a[0].data(), a[1].data(), a[2].data(), a[3].data(),
                   a[4].data(), a[5].data(), a[6].data(), a[7].data()};
  }
};

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) return 2;
  const std::string dir = argv[1];

  // ---- generators and the OBJ round trip at several thread counts
  rq_mesh ico{}, soup{}, field{};
  const float c[3] = {0.0f, 0.0f, -5.0f};
  expect(rq_gen_icosphere(4, 3, 0.1f, 4.0f, c, 1.0f, &ico) == 0, "icosphere");
  expect(rq_gen_soup(20000, 5, 0.02f, 0.005f, 0.002f, &soup) == 0, "soup");
  expect(rq_gen_icosphere_field(6, 2, 10, 77, 8.0f, 2.5f, 0.1f, 4.0f, &field) == 0, "field");
  const std::string obj = dir + "/mesh.obj";
  expect(rq_obj_write(obj.c_str(), &soup) == 0, "obj write");
  for (int threads : {1, 3, 8, 16}) {
    rq_mesh m{};
    expect(rq_obj_read_ex(obj.c_str(), threads, &m) == 0, "obj read");
    expect(m.ntri == soup.ntri && m.nv == soup.nv, "obj counts");
    rq_mesh m1{};
    expect(rq_obj_read_ex(obj.c_str(), 1, &m1) == 0, "obj read 1");
    expect(same_mesh(m, m1), "obj read equal across thread counts");
    rq_mesh_free(&m);
    rq_mesh_free(&m1);
  }
  // records the reader must accept or skip: negative and slashed indices,
  // polygons, comments, blank lines, CRLF ends, a final line without newline
  const std::string odd = dir + "/odd.obj";
  if (FILE* f = std::fopen(odd.c_str(), "wb")) {
    std::fputs("# comment\n\nv 0 0 0\r\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\n"
               "f 1/1/1 2//1 3\nf -4 -3 -2 -1\nusemtl x\nf 1 3 4", f);
    std::fclose(f);
  }
  for (int threads : {1, 2, 4}) {
    rq_mesh m{};
    expect(rq_obj_read_ex(odd.c_str(), threads, &m) == 0, "odd obj read");
    expect(m.nv == 4 && m.ntri == 4, "odd obj counts");
    rq_mesh_free(&m);
  }
  rq_spheres sph{};
  expect(rq_gen_reference_scene(obj.c_str(), 2, 500, &sph) == 0, "reference scene");
  expect(sph.n == 1001, "reference scene spheres");
  rq_spheres_free(&sph);

  // ---- oracle build and threaded queries
  const uint32_t n = 30000;
  Rays rays(n);
  float lo[3], hi[3];
  rq_mesh_bounds(&soup, lo, hi);
  or_uniform_rays(n, 9, lo, hi, rays.a[0].data(), rays.a[1].data(), rays.a[2].data(),
                  rays.a[3].data(), rays.a[4].data(), rays.a[5].data(), rays.a[6].data(),
                  rays.a[7].data());
  const or_rays rv = rays.view();
  or_bvh* b = or_bvh_build(soup.vx, soup.vy, soup.vz, soup.nv, soup.idx, soup.ntri, 1);
  std::vector<int32_t> p1(n), p8(n), pp(n), pl(n);
  std::vector<float> t1(n), t8(n), u(n), v(n), tp(n), tl(n);
  std::vector<uint8_t> o1(n), o8(n);
  or_hits h1{p1.data(), t1.data(), u.data(), v.data()};
  or_hits h8{p8.data(), t8.data(), u.data(), v.data()};
  or_closest_bvh(b, &rv, n, &h1, nullptr, nullptr, 1);
  or_closest_bvh(b, &rv, n, &h8, nullptr, nullptr, 8);
  or_any_bvh(b, &rv, n, o1.data(), nullptr, nullptr, 1);
  or_any_bvh(b, &rv, n, o8.data(), nullptr, nullptr, 8);
  expect(p1 == p8 && !std::memcmp(t1.data(), t8.data(), 4 * (size_t)n), "oracle closest threads");
  expect(o1 == o8, "oracle any threads");
  std::vector<uint32_t> st(n);
  or_closest_bvh_stack(b, &rv, n, st.data(), 8);
  expect(or_bvh_treelets(b) > 0, "oracle treelets");
  or_bvh_free(b);

  // ---- the reference-path port: median-split BVH and the 2-thread LBVH
  uint32_t nodes = 0;
  void* pb = port_bvh_build(soup.vx, soup.vy, soup.vz, soup.idx, soup.ntri, &nodes);
  void* plb = port_lbvh_build(soup.vx, soup.vy, soup.vz, soup.idx, soup.ntri, &nodes);
  port_closest(pb, rays.a[0].data(), rays.a[1].data(), rays.a[2].data(), rays.a[3].data(),
               rays.a[4].data(), rays.a[5].data(), rays.a[6].data(), rays.a[7].data(), n,
               pp.data(), tp.data(), 8);
  port_closest(plb, rays.a[0].data(), rays.a[1].data(), rays.a[2].data(), rays.a[3].data(),
               rays.a[4].data(), rays.a[5].data(), rays.a[6].data(), rays.a[7].data(), n,
               pl.data(), tl.data(), 4);
  expect(pp == p1 && pl == p1, "port closest equals the oracle");
  port_bvh_free(pb);
  port_bvh_free(plb);

  rq_mesh_free(&ico);
  rq_mesh_free(&soup);
  rq_mesh_free(&field);
  if (failures) return 1;
  std::printf("ok\n");
  return 0;
}
