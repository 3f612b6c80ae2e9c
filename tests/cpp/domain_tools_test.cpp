// DomainTools (include/ThunderEgg/DomainTools.h) the way the reference's
// test/DomainTools_MPI1.cpp drives it.
//   domain_tools_test         coordinates only (host code, no device)
//   domain_tools_test device  SetValues / SetValuesWithGhost / Integrate on the device
// Expected coordinates are written as start + (0.5 + i) * length / n, a different
// arithmetic path from the helpers' start + h / 2 + h * i.  Exit status 0 = pass.
#include <ThunderEgg/DomainTools.h>
#include <ThunderEgg/ForestDomainGenerator.h>

#include <cmath>
#include <cstring>

using namespace ThunderEgg;

static int failures = 0;
#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
      failures++;                                                       \
    }                                                                   \
  } while (0)

static bool near(double a, double b, double tol = 1e-12) { return std::fabs(a - b) <= tol * (1.0 + std::fabs(b)); }

static const double kStarts[] = { 0.0, -1.0, 1.0, 0.23, -0.23, 3.0, -3.0 };
static const double kLengths[] = { 1.0, 0.23, 3.0 };

// expected coordinate of index i with GetRealCoord's interface convention
static double expect_coord(int i, int n, double start, double length, bool interfaces)
{
  if (interfaces && i == -1) return start;
  if (interfaces && i == n) return start + length;
  return start + (0.5 + i) * length / n;
}

static void coordinates()
{
  for (int ny = 1; ny <= 3; ny++)
    for (double sy : kStarts)
      for (double ly : kLengths) {
        PatchInfo<2> p;
        const int nx = 2;
        p.ns = { nx, ny };
        p.spacings = { 1.0 / nx, ly / ny };
        p.starts = { 0.0, sy };
        for (int j = -1; j <= ny; j++)
          for (int i = -1; i <= nx; i++) {
            std::array<double, 2> r;
            DomainTools::GetRealCoord<2>(p, { i, j }, r);
            CHECK(near(r[0], expect_coord(i, nx, 0.0, 1.0, true)));
            CHECK(near(r[1], expect_coord(j, ny, sy, ly, true)));
            DomainTools::GetRealCoordGhost<2>(p, { i, j }, r);
            CHECK(near(r[0], expect_coord(i, nx, 0.0, 1.0, false)));
            CHECK(near(r[1], expect_coord(j, ny, sy, ly, false)));
          }
        for (int i = -1; i <= nx; i++) {
          std::array<double, 2> r;
          DomainTools::GetRealCoordBound<2>(p, { i }, Side<2>::south(), r);
          CHECK(near(r[0], expect_coord(i, nx, 0.0, 1.0, true)) && near(r[1], sy));
          DomainTools::GetRealCoordBound<2>(p, { i }, Side<2>::north(), r);
          CHECK(near(r[0], expect_coord(i, nx, 0.0, 1.0, true)) && near(r[1], sy + ly));
        }
        for (int j = -1; j <= ny; j++) {
          std::array<double, 2> r;
          DomainTools::GetRealCoordBound<2>(p, { j }, Side<2>::west(), r);
          CHECK(near(r[0], 0.0) && near(r[1], expect_coord(j, ny, sy, ly, true)));
          DomainTools::GetRealCoordBound<2>(p, { j }, Side<2>::east(), r);
          CHECK(near(r[0], 1.0) && near(r[1], expect_coord(j, ny, sy, ly, true)));
        }
      }
  for (int nz = 1; nz <= 3; nz++)
    for (double sz : kStarts)
      for (double lz : kLengths) {
        PatchInfo<3> p;
        p.ns = { 2, 3, nz };
        p.spacings = { 0.5, 2.0 / 3, lz / nz };
        p.starts = { 0.0, -1.0, sz };
        for (int k = -1; k <= nz; k++)
          for (int j = -1; j <= 3; j++)
            for (int i = -1; i <= 2; i++) {
              std::array<double, 3> r;
              DomainTools::GetRealCoord<3>(p, { i, j, k }, r);
              CHECK(near(r[0], expect_coord(i, 2, 0.0, 1.0, true)));
              CHECK(near(r[1], expect_coord(j, 3, -1.0, 2.0, true)));
              CHECK(near(r[2], expect_coord(k, nz, sz, lz, true)));
              DomainTools::GetRealCoordGhost<3>(p, { i, j, k }, r);
              CHECK(near(r[0], expect_coord(i, 2, 0.0, 1.0, false)));
              CHECK(near(r[1], expect_coord(j, 3, -1.0, 2.0, false)));
              CHECK(near(r[2], expect_coord(k, nz, sz, lz, false)));
            }
        // bottom / top: coord indexes x then y; south: x then z
        for (int j = -1; j <= 3; j++)
          for (int i = -1; i <= 2; i++) {
            std::array<double, 3> r;
            DomainTools::GetRealCoordBound<3>(p, { i, j }, Side<3>::top(), r);
            CHECK(near(r[0], expect_coord(i, 2, 0.0, 1.0, true)) && near(r[1], expect_coord(j, 3, -1.0, 2.0, true)) &&
                  near(r[2], sz + lz));
          }
        for (int k = -1; k <= nz; k++)
          for (int i = -1; i <= 2; i++) {
            std::array<double, 3> r;
            DomainTools::GetRealCoordBound<3>(p, { i, k }, Side<3>::south(), r);
            CHECK(near(r[0], expect_coord(i, 2, 0.0, 1.0, true)) && near(r[1], -1.0) &&
                  near(r[2], expect_coord(k, nz, sz, lz, true)));
          }
      }
}

// one patch of nx x ny cells, id 0, no neighbours
static Domain<2> one_patch(int nx, int ny, double sx, double sy, double hx, double hy)
{
  te_patch_info r;
  std::memset(&r, 0, sizeof(r));
  r.parent_id = r.parent_rank = r.orth_on_parent = -1;
  for (int k = 0; k < 8; k++) r.child_ids[k] = r.child_ranks[k] = -1;
  for (int k = 0; k < 26; k++) {
    r.nbr_type[k] = r.nbr_orth[k] = -1;
    for (int m = 0; m < 4; m++) r.nbr_ids[k][m] = r.nbr_ranks[k][m] = -1;
  }
  r.starts[0] = sx;
  r.starts[1] = sy;
  r.spacings[0] = hx;
  r.spacings[1] = hy;
  r.spacings[2] = 1.0;
  std::vector<te_patch_info> v{ r };
  return Domain<2>(Communicator(), 0, { nx, ny }, 1, v.begin(), v.end());
}

static void device()
{
  auto f = [](const std::array<double, 2>& c) { return c[0] + c[1]; };
  auto g = [](const std::array<double, 2>& c) { return c[0] * c[1]; };
  for (int ny : { 1, 2, 10, 13 })
    for (double sy : kStarts)
      for (double hy : { 0.01, 1.0, 3.14 }) {
        const int nx = 3, w = nx + 2, cells = w * (ny + 2);
        Domain<2> d = one_patch(nx, ny, 0.0, sy, 0.1, hy);
        Vector<2> v(d, 2), vg(d, 2);
        DomainTools::SetValues<2>(d, v, f, g);
        DomainTools::SetValuesWithGhost<2>(d, vg, f, g);
        std::vector<double> h = v.copyToHost(), hg = vg.copyToHost();
        for (int j = -1; j <= ny; j++)
          for (int i = -1; i <= nx; i++) {
            const double x = (0.5 + i) * 0.1, y = sy + (0.5 + j) * hy;
            const size_t at = (size_t)(j + 1) * w + (i + 1);
            const bool in = i >= 0 && i < nx && j >= 0 && j < ny;
            CHECK(near(h[at], in ? x + y : 0.0));
            CHECK(near(h[cells + at], in ? x * y : 0.0));
            CHECK(near(hg[at], x + y));
            CHECK(near(hg[cells + at], x * y));
          }
        // more functions than components: the first is set, then RuntimeError
        Vector<2> one(d, 1);
        bool threw = false;
        try {
          DomainTools::SetValues<2>(d, one, f, g);
        } catch (const RuntimeError&) {
          threw = true;
        }
        CHECK(threw);
        CHECK(near(one.copyToHost()[(size_t)w + 1], 0.05 + sy + 0.5 * hy));
        threw = false;
        try {
          DomainTools::SetValuesWithGhost<2>(d, one, f, g);
        } catch (const RuntimeError&) {
          threw = true;
        }
        CHECK(threw);
      }

  // the (x, y) overload and Integrate on a refined forest over the unit square:
  // the midpoint rule is exact for (1 + x)(1 - 2y) and for 1
  ForestDomainGenerator<2> gen(3, { 8, 6 });
  gen.refineWhere([](int level, std::array<int, 2> c) { return level == 3 && c[0] >= 4 && c[1] < 3; });
  Domain<2> d = gen.getFinestDomain();
  Vector<2> u(d, 1);
  DomainTools::SetValues(d, u, 0, [](double x, double y) { return (1 + x) * (1 - 2 * y); });
  double in = DomainTools::Integrate<2>(d, u);
  std::printf("integral of (1 + x)(1 - 2y) over %d patches: %.17g\n", d.getNumLocalPatches(), in);
  CHECK(near(in, 0.0, 1e-13));
  DomainTools::SetValues(d, u, 0, [](double x, double y) { return (1 + x) * (3 + 2 * y); });
  CHECK(near(DomainTools::Integrate<2>(d, u), 1.5 * 4.0, 1e-13));
  // a literal component index with an array-argument function selects the indexed overload
  DomainTools::SetValues<2>(d, u, 0, [](const std::array<double, 2>& c) { return c[0] * c[1]; });
  CHECK(near(DomainTools::Integrate<2>(d, u), 0.25, 1e-13));
  u.setWithGhost(7.0);
  u.set(1.0);  // ghosts stay 7: they must not enter
  CHECK(near(DomainTools::Integrate<2>(d, u), 1.0, 1e-13));
  // 3D: the x y z overload, integral of x y z over the unit cube = 1/8
  ForestDomainGenerator<3> g3(1, { 4, 4, 4 });
  Domain<3> d3 = g3.getFinestDomain();
  Vector<3> u3(d3, 2);
  DomainTools::SetValuesWithGhost(d3, u3, 1, [](double x, double y, double z) { return x * y * z; });
  CHECK(near(DomainTools::Integrate<3>(d3, u3), 0.125, 1e-13));
}

int main(int argc, char** argv)
{
  coordinates();
  if (argc > 1 && std::strcmp(argv[1], "device") == 0) device();
  std::printf(failures ? "domain_tools_test: %d failures\n" : "domain_tools_test: ok\n", failures);
  return failures ? 1 : 0;
}
