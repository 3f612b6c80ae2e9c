// The reference's per-patch plug-in points, driven the way user code written against the
// reference uses them: a user-written 5-point operator (PatchOperator's three per-patch
// virtuals on PatchViews), a user-written ghost filler (MPIGhostFiller's per-neighbour
// virtuals), the DFT and inner-BiCGStab patch solvers wrapping the user operator, a
// CycleBuilder cycle of user operators, the boundary-data helpers and the views.  Each is
// checked against the built-in (device) classes on the same input.  Exit status 0 = pass.
#include <ThunderEgg/DomainTools.h>
#include <ThunderEgg/GMG/CycleBuilder.h>
#include <ThunderEgg/Iterative/BiCGStab.h>
#include <ThunderEgg/Iterative/PatchSolver.h>
#include <ThunderEgg/Poisson/DFTPatchSolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>

using namespace ThunderEgg;

static int failures = 0;
#define CHECK(c)                                                          \
  do {                                                                    \
    if (!(c)) {                                                           \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
      failures++;                                                         \
    }                                                                     \
  } while (0)

// The reference's 5-point Poisson operator written as user code: physical ghosts -u, the
// patch's neighbour sides treated as boundaries in the patch solves.
class MyStar : public PatchOperator<2> {
 public:
  MyStar(const Domain<2>& d, const GhostFiller<2>& gf) : PatchOperator<2>(d, gf) {}
  MyStar* clone() const override { return new MyStar(*this); }
  void applySinglePatch(const PatchInfo<2>& pinfo, const PatchView<const double, 2>& u,
                        const PatchView<double, 2>& f) const override
  {
    stencil(pinfo, u, f, false);
  }
  void applySinglePatchWithInternalBoundaryConditions(const PatchInfo<2>& pinfo, const PatchView<const double, 2>& u,
                                                      const PatchView<double, 2>& f) const override
  {
    stencil(pinfo, u, f, true);
  }
  void modifyRHSForInternalBoundaryConditions(const PatchInfo<2>& pinfo, const PatchView<const double, 2>& u,
                                              const PatchView<double, 2>& f) const override
  {
    for (Side<2> s : Side<2>::getValues()) {
      if (!pinfo.hasNbr(s)) continue;
      const double h = pinfo.spacings[s.getAxisIndex()];
      View<double, 2> fi = f.getSliceOn(s, { 0 });
      View<const double, 2> ug = u.getSliceOn(s, { -1 }), ui = u.getSliceOn(s, { 0 });
      Loop::OverInteriorIndexes<2>(fi, [&](const std::array<int, 2>& c) { fi[c] -= (ug[c] + ui[c]) / (h * h); });
    }
  }

 private:
  void stencil(const PatchInfo<2>& pinfo, const PatchView<const double, 2>& u, const PatchView<double, 2>& f,
               bool internal) const
  {
    for (Side<2> s : Side<2>::getValues()) {
      if (pinfo.hasNbr(s) && !internal) continue;
      View<double, 2> g = u.getGhostSliceOn(s, { 0 });
      View<const double, 2> m = u.getSliceOn(s, { 0 });
      Loop::OverInteriorIndexes<2>(m, [&](const std::array<int, 2>& c) { g[c] = -m[c]; });
    }
    const double hx2 = pinfo.spacings[0] * pinfo.spacings[0], hy2 = pinfo.spacings[1] * pinfo.spacings[1];
    Loop::OverInteriorIndexes<3>(u, [&](const std::array<int, 3>& c) {
      const int x = c[0], y = c[1], k = c[2];
      f(x, y, k) = (u(x + 1, y, k) - 2 * u(x, y, k) + u(x - 1, y, k)) / hx2 +
                   (u(x, y + 1, k) - 2 * u(x, y, k) + u(x, y - 1, k)) / hy2;
    });
  }
};

// A user-written ghost filler for meshes whose neighbours are all on the same level: the
// ghost layer is the neighbour's boundary layer (faces and corners).
class MyFiller : public MPIGhostFiller<2> {
 public:
  MyFiller(const Domain<2>& d, GhostFillingType t) : MPIGhostFiller<2>(d, t) {}
  MyFiller* clone() const override { return new MyFiller(*this); }

 protected:
  void fillGhostCellsForNbrPatch(const PatchInfo<2>&, const PatchView<const double, 2>& local_view,
                                 const PatchView<const double, 2>& nbr_view, Side<2> side, NbrType type,
                                 Orthant<1>) const override
  {
    if (type != NbrType::Normal) throw RuntimeError("MyFiller: same-level neighbours only");
    View<double, 2> g = nbr_view.getGhostSliceOn(side.opposite(), { 0 });
    View<const double, 2> src = local_view.getSliceOn(side, { 0 });
    Loop::OverInteriorIndexes<2>(src, [&](const std::array<int, 2>& c) { g[c] = src[c]; });
  }
  void fillGhostCellsForEdgeNbrPatch(const PatchInfo<2>&, const PatchView<const double, 2>&,
                                     const PatchView<const double, 2>&, Edge, NbrType, Orthant<1>) const override
  {
  }
  void fillGhostCellsForCornerNbrPatch(const PatchInfo<2>&, const PatchView<const double, 2>& local_view,
                                       const PatchView<const double, 2>& nbr_view, Corner<2> corner,
                                       NbrType type) const override
  {
    if (type != NbrType::Normal) throw RuntimeError("MyFiller: same-level neighbours only");
    View<double, 1> g = nbr_view.getGhostSliceOn(corner.opposite(), { 0, 0 });
    View<const double, 1> src = local_view.getSliceOn(corner, { 0, 0 });
    Loop::OverInteriorIndexes<1>(src, [&](const std::array<int, 1>& c) { g[c] = src[c]; });
  }
  void fillGhostCellsForLocalPatch(const PatchInfo<2>&, const PatchView<const double, 2>&) const override {}
};

static double rel_diff(const Vector<2>& a, const Vector<2>& b)
{
  Vector<2> d(a);
  d.addScaled(-1.0, b);
  return d.twoNorm() / b.twoNorm();
}

static double max_diff_with_ghosts(const Vector<2>& a, const Vector<2>& b)
{
  std::vector<double> x = a.copyToHost(), y = b.copyToHost();
  double m = 0, s = 0;
  for (size_t i = 0; i < x.size(); i++) {
    m = std::max(m, std::fabs(x[i] - y[i]));
    s = std::max(s, std::fabs(y[i]));
  }
  return m / s;
}

static void random_fill(Vector<2>& v, unsigned seed)
{
  std::mt19937_64 rng(seed);
  std::normal_distribution<double> nd;
  std::vector<double> h = v.copyToHost();
  for (double& x : h) x = nd(rng);
  v.copyFromHost(h);
}

template <class Op>
static std::shared_ptr<GMG::Cycle<2>> build(const std::vector<Domain<2>>& doms, bool user)
{
  GMG::CycleOpts opts;
  GMG::CycleBuilder<2> builder(opts);
  for (size_t i = 0; i < doms.size(); i++) {
    BiLinearGhostFiller gf(doms[i], GhostFillingType::Faces);
    std::shared_ptr<PatchOperator<2>> op;
    if (user) op = std::make_shared<MyStar>(doms[i], gf);
    else op = std::make_shared<Poisson::StarPatchOperator<2>>(doms[i], gf);
    Poisson::DFTPatchSolver<2> sm(*op, std::bitset<4>());
    if (i == 0) {
      builder.addFinestLevel(*op, sm, GMG::LinearRestrictor<2>(doms[0], doms[1]));
    } else if (i + 1 < doms.size()) {
      builder.addIntermediateLevel(*op, sm, GMG::LinearRestrictor<2>(doms[i], doms[i + 1]),
                                   GMG::DirectInterpolator<2>(doms[i], doms[i - 1]));
    } else {
      builder.addCoarsestLevel(*op, sm, GMG::DirectInterpolator<2>(doms[i], doms[i - 1]));
    }
  }
  return builder.getCycle();
}

int main()
{
  ForestDomainGenerator<2> gen(2, { 16, 16 });  // a uniform 4 x 4 forest of 16 x 16 patches
  std::vector<Domain<2>> doms{ gen.getFinestDomain() };
  while (gen.hasCoarserDomain()) doms.push_back(gen.getCoarserDomain());
  const Domain<2>& d = doms[0];
  CHECK(d.getNumLocalPatches() == 16 && doms.size() == 3);

  // PatchInfo neighbour records: a corner patch has two neighbouring sides and one corner
  {
    const PatchInfo<2>& p0 = d.getPatchInfoVector()[0];
    int sides = 0;
    for (Side<2> s : Side<2>::getValues()) sides += p0.hasNbr(s) ? 1 : 0;
    CHECK(sides == 2 && p0.hasNbr());
    CHECK(p0.hasNbr(Side<2>::east()) && p0.getNbrType(Side<2>::east()) == NbrType::Normal);
    const int q = p0.getNormalNbrInfo(Side<2>::east()).local_index;
    CHECK(d.getPatchInfoVector()[(size_t)q].getNormalNbrInfo(Side<2>::west()).id == p0.id);
    CHECK(p0.hasNbr(Corner<2>::ne()) && !p0.hasNbr(Corner<2>::sw()));
  }

  Vector<2> u(d, 1);
  random_fill(u, 7);

  // views: a patch view, its component view and side slices address the same cells
  {
    std::vector<double> h = u.copyToHost();
    const int w = 18;
    PatchView<const double, 2> pv = u.getPatchView(3);
    CHECK(pv(0, 0, 0) == h[3 * w * w + 1 * w + 1]);
    CHECK(pv(15, 2, 0) == h[3 * w * w + 3 * w + 16]);
    CHECK(pv.getComponentView(0)(4, 5) == pv(4, 5, 0));
    View<const double, 2> east = pv.getSliceOn(Side<2>::east(), { 0 });
    CHECK(east(7, 0) == pv(15, 7, 0));
    View<double, 2> west_ghost = pv.getGhostSliceOn(Side<2>::west(), { 0 });
    CHECK(&west_ghost(7, 0) == &pv(-1, 7, 0));
    CHECK(pv.getSliceOn(Side<2>::north(), { -1 })(3, 0) == pv(3, 16, 0));
  }

  // the user filler against the built-in BiLinear filler (faces and corners, ghosts included)
  {
    Vector<2> a(u), b(u);
    MyFiller mine(d, GhostFillingType::Corners);
    BiLinearGhostFiller ref(d, GhostFillingType::Corners);
    mine.fillGhost(a);
    ref.fillGhost(b);
    const double e = max_diff_with_ghosts(a, b);
    std::printf("user filler vs BiLinearGhostFiller: %.3e\n", e);
    CHECK(e <= 1e-15);
  }

  // the user operator (with either filler) against the built-in Poisson operator
  BiLinearGhostFiller bil(d, GhostFillingType::Faces);
  Poisson::StarPatchOperator<2> A(d, bil);
  MyStar mineA(d, bil);
  MyStar mineB(d, MyFiller(d, GhostFillingType::Faces));
  {
    Vector<2> f0(d, 1), f1(d, 1), f2(d, 1);
    A.apply(u, f0);
    mineA.apply(u, f1);
    mineB.apply(u, f2);
    std::printf("user operator vs StarPatchOperator: %.3e %.3e\n", rel_diff(f1, f0), rel_diff(f2, f0));
    CHECK(rel_diff(f1, f0) <= 1e-13);
    CHECK(rel_diff(f2, f0) <= 1e-13);
    // the built-in operator's own per-patch virtuals give its device result too
    Vector<2> f3(d, 1);
    bil.fillGhost(u);
    for (const PatchInfo<2>& pinfo : d.getPatchInfoVector())
      A.applySinglePatch(pinfo, u.getPatchView(pinfo.local_index), f3.getPatchView(pinfo.local_index));
    CHECK(rel_diff(f3, f0) <= 1e-13);
  }

  // the DFT patch solver around the user operator (per-patch virtuals, host) against the
  // built-in one (device)
  Vector<2> f(d, 1);
  random_fill(f, 11);
  {
    Poisson::DFTPatchSolver<2> dev(A, std::bitset<4>());
    Poisson::DFTPatchSolver<2> host(mineA, std::bitset<4>());
    CHECK(dev.device() != nullptr && host.device() == nullptr);
    Vector<2> u1(u), u2(u);
    dev.smooth(f, u1);
    host.smooth(f, u2);
    std::printf("DFT smooth, user operator vs built-in: %.3e\n", rel_diff(u2, u1));
    CHECK(rel_diff(u2, u1) <= 1e-12);
  }

  // the inner-BiCGStab patch solver around the user operator against the built-in one
  {
    Iterative::BiCGStab<2> inner;
    inner.setMaxIterations(200);
    inner.setTolerance(1e-12);
    Iterative::PatchSolver<2> dev(inner, A), host(inner, mineA);
    CHECK(dev.onDevice() && !host.onDevice());
    Vector<2> u1(u), u2(u);
    dev.smooth(f, u1);
    host.smooth(f, u2);
    std::printf("BiCGStab patch smooth, user operator vs built-in: %.3e\n", rel_diff(u2, u1));
    CHECK(rel_diff(u2, u1) <= 1e-8);
  }

  // a V-cycle of user operators (generic cycle) against the device cycle, and the solves
  {
    auto cu = build<MyStar>(doms, true);
    auto cd = build<Poisson::StarPatchOperator<2>>(doms, false);
    CHECK(cu->device() == nullptr && cd->device() != nullptr);
    Vector<2> v1(d, 1), v2(d, 1);
    cd->apply(f, v1);
    cu->apply(f, v2);
    std::printf("V-cycle of user operators vs device cycle: %.3e\n", rel_diff(v2, v1));
    CHECK(rel_diff(v2, v1) <= 1e-11);
    Iterative::BiCGStab<2> bs;
    bs.setTolerance(1e-10);
    Vector<2> x1(d, 1), x2(d, 1);
    const int i1 = bs.solve(A, x1, f, cd.get());
    const int i2 = bs.solve(mineA, x2, f, cu.get());
    std::printf("solves: %d / %d iterations, %.3e apart\n", i1, i2, rel_diff(x2, x1));
    CHECK(i1 == i2);
    CHECK(rel_diff(x2, x1) <= 1e-9);
  }

  // addDrichletBCToRHS: for a linear g the operator gives exactly the boundary terms
  {
    Vector<2> g(d, 1), ag(d, 1), bc(d, 1);
    std::vector<double> h = g.copyToHost();
    const int w = 18;
    const auto& P = d.getPatchInfoVector();
    for (size_t p = 0; p < P.size(); p++)
      for (int j = -1; j <= 16; j++)
        for (int i = -1; i <= 16; i++) {
          const double x = P[p].starts[0] + (i + 0.5) * P[p].spacings[0];
          const double y = P[p].starts[1] + (j + 0.5) * P[p].spacings[1];
          h[p * w * w + (size_t)(j + 1) * w + (size_t)(i + 1)] = 1 + 5 * x + 13 * y;
        }
    g.copyFromHost(h);
    A.apply(g, ag);
    A.addDrichletBCToRHS(bc, [](const std::array<double, 2>& c) { return 1 + 5 * c[0] + 13 * c[1]; });
    std::printf("addDrichletBCToRHS linear exactness: %.3e\n", rel_diff(ag, bc));
    CHECK(rel_diff(ag, bc) <= 1e-9);
  }

  // LinearRestrictor with extrapolate_boundary_ghosts, as test/GMG/LinearRestrictor_MPI1.cpp
  // checks it: a linear function set with ghosts restricts to the same function on the coarse
  // interior and on the ghosts of every side without a neighbour
  {
    auto lin = [](const std::array<double, 2>& c) { return 1 + (c[0] * 0.3 + c[1]); };
    Vector<2> fine(doms[0], 1), expect(doms[1], 1);
    DomainTools::SetValuesWithGhost<2>(doms[0], fine, lin);
    DomainTools::SetValuesWithGhost<2>(doms[1], expect, lin);
    GMG::LinearRestrictor<2> r(doms[0], doms[1], true);
    Vector<2> got = r.restrict(fine);
    double worst = 0.0;
    int checked = 0;
    for (const PatchInfo<2>& pi : doms[1].getPatchInfoVector()) {
      auto g = got.getComponentView(0, pi.local_index);
      auto e = expect.getComponentView(0, pi.local_index);
      for (int j = 0; j < 8; j++)
        for (int i = 0; i < 8; i++) worst = std::max(worst, std::fabs(g(i, j) - e(i, j)));
      for (Side<2> sd : Side<2>::getValues()) {
        if (pi.hasNbr(sd)) continue;
        auto gs = g.getSliceOn(sd, { -1 });
        auto es = e.getSliceOn(sd, { -1 });
        for (int t = 0; t < 8; t++, checked++) worst = std::max(worst, std::fabs(gs(t) - es(t)));
      }
    }
    std::printf("LinearRestrictor extrapolation: %d boundary ghosts, max error %.3e\n", checked, worst);
    CHECK(checked > 0 && worst <= 1e-13);
  }

  if (failures) {
    std::printf("patch_api_test: %d failures\n", failures);
    return 1;
  }
  std::printf("patch_api_test: ok\n");
  return 0;
}
