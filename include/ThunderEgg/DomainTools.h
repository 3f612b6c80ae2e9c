/* DomainTools (DomainTools.h:36-537): cell-centre and interface coordinates of a patch,
 * vectors set from functions of the coordinates, and the integral of a vector over the
 * domain.
 *
 * SetValues / SetValuesWithGhost evaluate the caller's host function at every cell of
 * every patch and copy the vector's whole padded array back to the device (setup work,
 * as in the reference); Integrate is a device reduction (te_vec_integrate).  Ghost width
 * is one cell, as every vector of this library. */
#pragma once

#include <functional>
#include <type_traits>

#include "amd.h"

namespace ThunderEgg {

class DomainTools {
  /* coordinate of index i on axis a: the lower interface for -1, the upper interface for
   * ns[a], the cell centre otherwise (DomainTools.h:134-149) */
  template <int D>
  static double coordOnAxis(const PatchInfo<D>& p, int a, int i)
  {
    if (i == -1) return p.starts[a];
    if (i == p.ns[a]) return p.starts[a] + p.spacings[a] * p.ns[a];
    return p.starts[a] + p.spacings[a] / 2.0 + p.spacings[a] * i;
  }

  /* offset of cell (coord, ghost-relative) of component c of patch p in the padded layout
   * [patch][component][z][y][x] */
  template <int D>
  static size_t cellIndex(const std::array<int, D>& ns, int ncomp, int p, int c, const std::array<int, D>& coord)
  {
    size_t per = 1;
    for (int a = 0; a < D; a++) per *= (size_t)(ns[a] + 2);
    size_t off = 0;
    for (int a = D - 1; a >= 0; a--) off = off * (size_t)(ns[a] + 2) + (size_t)(coord[a] + 1);
    return ((size_t)p * ncomp + c) * per + off;
  }

  template <int D, class F>
  static void fill(const Domain<D>& domain, Vector<D>& vec, int component_index, bool ghost, F&& value)
  {
    if (component_index < 0 || component_index >= vec.getNumComponents())
      throw RuntimeError("Invalid component to set");
    std::vector<double> h = vec.copyToHost();
    const std::array<int, D>& ns = domain.getNs();
    const int g = ghost ? 1 : 0;
    std::array<double, D> real_coord;
    for (int p = 0; p < domain.getNumLocalPatches(); p++) {
      const PatchInfo<D> pinfo = domain.getPatchInfo(p);
      std::array<int, D> coord;
      for (int a = 0; a < D; a++) coord[a] = -g;
      while (true) {
        GetRealCoordGhost<D>(pinfo, coord, real_coord);
        h[cellIndex<D>(ns, vec.getNumComponents(), p, component_index, coord)] = value(real_coord);
        int a = 0;
        while (a < D && ++coord[a] == ns[a] + g) coord[a++] = -g;
        if (a == D) break;
      }
    }
    vec.copyFromHost(h);
  }

  template <int D>
  static void setAll(const Domain<D>&, Vector<D>&, int, bool)
  {
  }
  template <int D, class F, class... Args>
  static void setAll(const Domain<D>& domain, Vector<D>& vec, int c, bool ghost, F func, Args... args)
  {
    std::function<double(const std::array<double, D>&)> fn(func);
    fill<D>(domain, vec, c, ghost, fn);
    setAll<D>(domain, vec, c + 1, ghost, args...);
  }

 public:
  /* DomainTools::GetRealCoord (DomainTools.h:134-149): -1 and ns[axis] give the
   * interfaces, other indexes the cell centres. */
  template <int D>
  static void GetRealCoord(const PatchInfo<D>& pinfo, const std::array<int, D>& coord,
                           std::array<double, D>& real_coord)
  {
    for (int a = 0; a < D; a++) real_coord[a] = coordOnAxis<D>(pinfo, a, coord[a]);
  }
  /* DomainTools::GetRealCoordGhost (DomainTools.h:158-168): cell centres, ghost cells
   * included. */
  template <int D>
  static void GetRealCoordGhost(const PatchInfo<D>& pinfo, const std::array<int, D>& coord,
                                std::array<double, D>& real_coord)
  {
    for (int a = 0; a < D; a++)
      real_coord[a] = pinfo.starts[a] + pinfo.spacings[a] / 2.0 + pinfo.spacings[a] * coord[a];
  }
  /* DomainTools::GetRealCoordBound (DomainTools.h:178-210): a point of side s; coord
   * indexes the other axes in increasing order, as GetRealCoord. */
  template <int D>
  static void GetRealCoordBound(const PatchInfo<D>& pinfo, const std::array<int, D - 1>& coord, Side<D> s,
                                std::array<double, D>& real_coord)
  {
    const int axis = (int)s.getAxisIndex();
    for (int a = 0, k = 0; a < D; a++) {
      if (a == axis)
        real_coord[a] = s.isLowerOnAxis() ? pinfo.starts[a] : pinfo.starts[a] + pinfo.spacings[a] * pinfo.ns[a];
      else
        real_coord[a] = coordOnAxis<D>(pinfo, a, coord[k++]);
    }
  }

  /* DomainTools::SetValues (DomainTools.h:220-346): component component_index at the
   * cell centres; ghost values are left as they were. */
  template <int D>
  static void SetValues(const Domain<D>& domain, Vector<D>& vec, int component_index,
                        std::function<double(const std::array<double, D>&)> func)
  {
    fill<D>(domain, vec, component_index, false, func);
  }
  static void SetValues(const Domain<2>& domain, Vector<2>& vec, int component_index,
                        std::function<double(double, double)> func)
  {
    fill<2>(domain, vec, component_index, false, [&](const std::array<double, 2>& c) { return func(c[0], c[1]); });
  }
  static void SetValues(const Domain<3>& domain, Vector<3>& vec, int component_index,
                        std::function<double(double, double, double)> func)
  {
    fill<3>(domain, vec, component_index, false,
            [&](const std::array<double, 3>& c) { return func(c[0], c[1], c[2]); });
  }
  /* one function per component, from component 0; more functions than components
   * throw RuntimeError once the components run out (DomainTools.h:51-88, 340-346).  The
   * first function's type is a template parameter kept off integers, so that
   * SetValues<D>(domain, vec, 0, f) selects the component-index overload. */
  template <int D, typename F, typename... Args, typename = typename std::enable_if<!std::is_integral<F>::value>::type>
  static void SetValues(const Domain<D>& domain, Vector<D>& vec, F func, Args... args)
  {
    setAll<D>(domain, vec, 0, false, func, args...);
  }

  /* DomainTools::SetValuesWithGhost (DomainTools.h:357-506): as SetValues, ghost cells
   * included (their centres, GetRealCoordGhost). */
  template <int D>
  static void SetValuesWithGhost(const Domain<D>& domain, Vector<D>& vec, int component_index,
                                 std::function<double(const std::array<double, D>&)> func)
  {
    fill<D>(domain, vec, component_index, true, func);
  }
  static void SetValuesWithGhost(const Domain<2>& domain, Vector<2>& vec, int component_index,
                                 std::function<double(double, double)> func)
  {
    fill<2>(domain, vec, component_index, true, [&](const std::array<double, 2>& c) { return func(c[0], c[1]); });
  }
  static void SetValuesWithGhost(const Domain<3>& domain, Vector<3>& vec, int component_index,
                                 std::function<double(double, double, double)> func)
  {
    fill<3>(domain, vec, component_index, true,
            [&](const std::array<double, 3>& c) { return func(c[0], c[1], c[2]); });
  }
  template <int D, typename F, typename... Args, typename = typename std::enable_if<!std::is_integral<F>::value>::type>
  static void SetValuesWithGhost(const Domain<D>& domain, Vector<D>& vec, F func, Args... args)
  {
    setAll<D>(domain, vec, 0, true, func, args...);
  }

  /* DomainTools::Integrate (DomainTools.h:513-536): sum over patches and components of
   * the interior values times the cell volume, reduced on the device. */
  template <int D>
  static double Integrate(const Domain<D>& domain, const Vector<D>& u)
  {
    double r = 0.0;
    detail::check(te_vec_integrate(domain.handle(), TE_F64, u.data(), u.getNumComponents(), &r, nullptr));
    return r;
  }
};

}  // namespace ThunderEgg
