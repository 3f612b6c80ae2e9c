/*
 * ThunderEgg/amd.h -- C++ API of the MI355X GMG hot path, with the reference's class
 * names and method shapes (namespace ThunderEgg), over the C ABI of thunderegg_amd.h.
 *
 * A program written against the reference's hot-path classes (Domain, Vector,
 * BiLinear/BiQuadratic/TriLinearGhostFiller, Poisson/VarPoisson::StarPatchOperator,
 * Poisson::DFTPatchSolver, Iterative::PatchSolver, GMG::LinearRestrictor,
 * GMG::DirectInterpolator, GMG::CycleBuilder, Iterative::BiCGStab) compiles against these
 * headers with the include paths it already uses (<ThunderEgg/...>).  Differences:
 *  - Vector data lives in device memory (HBM).  Host access is by copy:
 *    Vector::copyToHost / copyFromHost (the reference's PatchView access has no cheap
 *    device equivalent).
 *  - One rank: Communicator is a stand-in with rank 0 of 1.  Patch tables are
 *    te_patch_info records (PatchInfo<D> flattened, see thunderegg_amd.h); meshes come
 *    from ForestDomainGenerator (in place of P4estDomainGenerator) or from a table.
 *  - GMG::CycleOpts gains `precision` ("double", or "mixed": FP32 V-cycle with an FP64
 *    coarsest level).  Cycles and BiCGStab solves whose components are all this
 *    library's built-in classes run as fused, graph-captured device code (te_gmg);
 *    any other (user-derived) components go through the generic virtual calls.
 * Errors throw RuntimeError, and Iterative::BreakdownError on a solver breakdown, as
 * the reference does (RuntimeError.h, Iterative/BreakdownError.h).
 * Link with -lthunderegg_amd (thunderegg_amd/libthunderegg_amd.so).
 */
#ifndef THUNDEREGG_AMD_HPP
#define THUNDEREGG_AMD_HPP

#include <array>
#include <bitset>
#include <cmath>
#include <cstdio>
#include <functional>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <typeinfo>
#include <type_traits>
#include <utility>
#include <vector>

#include "../thunderegg_amd.h"

namespace ThunderEgg {

class RuntimeError : public std::runtime_error {
 public:
  explicit RuntimeError(const std::string& msg) : std::runtime_error(msg) {}
};

namespace Iterative {
class BreakdownError : public RuntimeError {
 public:
  explicit BreakdownError(const std::string& msg) : RuntimeError(msg) {}
};
}  // namespace Iterative

namespace detail {
inline void check(int rc)
{
  if (rc == TE_OK) return;
  std::string msg = te_last_error();
  if (rc == TE_ERR_BREAKDOWN) throw Iterative::BreakdownError(msg);
  throw RuntimeError(msg);
}
template <typename T, int (*Destroy)(T*)>
struct Deleter {
  void operator()(T* p) const
  {
    if (p) Destroy(p);
  }
};
}  // namespace detail

/* Timer (Timer.h:28-204) over te_timer.  addDomain takes a Domain (the reference takes
 * its JSON); to_json becomes json(indent) returning the text.  The *On variants time
 * device work by events on a stream. */
class Timer {
 public:
  Timer()
  {
    te_timer* h = nullptr;
    detail::check(te_timer_create(&h));
    h_ = std::shared_ptr<te_timer>(h, detail::Deleter<te_timer, te_timer_destroy>());
  }
  template <class Comm>
  explicit Timer(const Comm&) : Timer()
  {
  }
  void start(const std::string& name) { detail::check(te_timer_start(h_.get(), name.c_str(), TE_TIMER_NONE, TE_TIMER_NONE)); }
  void stop(const std::string& name) { detail::check(te_timer_stop(h_.get(), name.c_str(), TE_TIMER_NONE, TE_TIMER_NONE)); }
  template <class Dom>
  void addDomain(int domain_id, const Dom& domain)
  {
    std::array<int, 3> n3{ 1, 1, 1 };
    for (size_t a = 0; a < domain.getNs().size(); a++) n3[a] = domain.getNs()[a];
    const auto& pv = domain.getPatchTable();
    detail::check(te_timer_add_domain(h_.get(), domain_id, (int)domain.getNs().size(), n3.data(), (int)pv.size(),
                                      pv.data()));
  }
  void startDomainTiming(int domain_id, const std::string& name)
  {
    detail::check(te_timer_start(h_.get(), name.c_str(), domain_id, TE_TIMER_NONE));
  }
  void stopDomainTiming(int domain_id, const std::string& name)
  {
    detail::check(te_timer_stop(h_.get(), name.c_str(), domain_id, TE_TIMER_NONE));
  }
  void startPatchTiming(int patch_id, int domain_id, const std::string& name)
  {
    detail::check(te_timer_start(h_.get(), name.c_str(), domain_id, patch_id));
  }
  void stopPatchTiming(int patch_id, int domain_id, const std::string& name)
  {
    detail::check(te_timer_stop(h_.get(), name.c_str(), domain_id, patch_id));
  }
  void startOn(const std::string& name, te_stream s, int domain_id = TE_TIMER_NONE, int patch_id = TE_TIMER_NONE)
  {
    detail::check(te_timer_start_on(h_.get(), name.c_str(), domain_id, patch_id, s));
  }
  void stopOn(const std::string& name, te_stream s, int domain_id = TE_TIMER_NONE, int patch_id = TE_TIMER_NONE)
  {
    detail::check(te_timer_stop_on(h_.get(), name.c_str(), domain_id, patch_id, s));
  }
  void addIntInfo(const std::string& name, int info) { detail::check(te_timer_add_int_info(h_.get(), name.c_str(), info)); }
  void addDoubleInfo(const std::string& name, double info)
  {
    detail::check(te_timer_add_double_info(h_.get(), name.c_str(), info));
  }
  /* to_json(json&, const Timer&) as text: compact (indent < 0) or indented */
  std::string json(int indent = -1) const
  {
    size_t n = 0;
    detail::check(te_timer_to_json(h_.get(), indent, nullptr, 0, &n));
    std::string out(n + 1, '\0');
    detail::check(te_timer_to_json(h_.get(), indent, &out[0], out.size(), &n));
    out.resize(n);
    return out;
  }
  void saveToFile(const std::string& filename) const { detail::check(te_timer_save(h_.get(), filename.c_str())); }
  friend std::ostream& operator<<(std::ostream& os, const Timer& t)
  {
    size_t n = 0;
    detail::check(te_timer_print(t.h_.get(), nullptr, 0, &n));
    std::string out(n + 1, '\0');
    detail::check(te_timer_print(t.h_.get(), &out[0], out.size(), &n));
    out.resize(n);
    return os << out;
  }
  te_timer* handle() const { return h_.get(); }

 private:
  std::shared_ptr<te_timer> h_;
};

/* Single-rank stand-in for the reference's MPI communicator (Communicator.h). */
class Communicator {
 public:
  int getRank() const { return 0; }
  int getSize() const { return 1; }
};

/* GhostFillingType.h */
enum class GhostFillingType { Faces, Edges, Corners };

/* Face<D, M> (Face.h:41-584): an M-dimensional face of a D-dimensional patch, with the
 * reference's encoding.  Side<D> = Face<D, D - 1> (value 2 axis + upper), Edge = Face<3, 1>
 * (value >> 2 = the axis parallel to the edge, low bits = upper on the other axes in
 * increasing order), Corner<D> = Face<D, 0> (bit a = upper on axis a). */
template <int D, int M>
class Face {
  static constexpr int choose(int n, int k) { return k == 0 ? 1 : choose(n - 1, k - 1) * n / k; }

 public:
  static constexpr int dimensionality = M;
  static constexpr size_t number_of = (size_t)(1 << (D - M)) * (size_t)choose(D, M);
  constexpr Face() : v_((unsigned char)number_of) {}
  constexpr explicit Face(unsigned char value) : v_(value) {}
  static constexpr Face null() { return Face((unsigned char)number_of); }
  size_t getIndex() const { return v_; }
  /* the D - M sides that meet in this face, in axis order (Face.h:496-524) */
  std::array<Face<D, D - 1>, D - M> getSides() const
  {
    std::array<Face<D, D - 1>, D - M> out;
    if (M == D - 1) {
      out[0] = Face<D, D - 1>(v_);
    } else if (M == 0) {
      for (int a = 0; a < D; a++) out[(size_t)a] = Face<D, D - 1>((unsigned char)(2 * a + ((v_ >> a) & 1)));
    } else {  // D == 3, M == 1
      const int par = v_ >> 2;
      for (int a = 0, k = 0; a < 3; a++)
        if (a != par) {
          out[(size_t)k] = Face<D, D - 1>((unsigned char)(2 * a + ((v_ >> k) & 1)));
          k++;
        }
    }
    return out;
  }
  Face opposite() const { return Face((unsigned char)(v_ ^ ~((~0u) << (D - M)))); }
  /* Side only */
  size_t getAxisIndex() const { return (size_t)(v_ >> 1); }
  bool isLowerOnAxis() const { return (v_ & 1) == 0; }
  bool isHigherOnAxis() const { return (v_ & 1) == 1; }
  /* Edge only: the axis parallel to the edge */
  size_t getAxisIndexOfEdge() const { return (size_t)(v_ >> 2); }
  bool operator==(const Face& o) const { return v_ == o.v_; }
  bool operator!=(const Face& o) const { return v_ != o.v_; }
  bool operator<(const Face& o) const { return v_ < o.v_; }
  /* all faces of this dimension, in value order */
  static std::vector<Face> getValues()
  {
    std::vector<Face> out;
    for (size_t k = 0; k < number_of; k++) out.push_back(Face((unsigned char)k));
    return out;
  }
  // named faces (Face.h:106-440)
  static constexpr Face west() { return Face(0); }
  static constexpr Face east() { return Face(1); }
  static constexpr Face south() { return Face(2); }
  static constexpr Face north() { return Face(3); }
  static constexpr Face bottom() { return Face(4); }
  static constexpr Face top() { return Face(5); }
  static constexpr Face sw() { return Face(0); }
  static constexpr Face se() { return Face(1); }
  static constexpr Face nw() { return Face(2); }
  static constexpr Face ne() { return Face(3); }
  static constexpr Face bsw() { return Face(0); }
  static constexpr Face bse() { return Face(1); }
  static constexpr Face bnw() { return Face(2); }
  static constexpr Face bne() { return Face(3); }
  static constexpr Face tsw() { return Face(4); }
  static constexpr Face tse() { return Face(5); }
  static constexpr Face tnw() { return Face(6); }
  static constexpr Face tne() { return Face(7); }
  static constexpr Face bs() { return Face(0); }
  static constexpr Face bn() { return Face(1); }
  static constexpr Face ts() { return Face(2); }
  static constexpr Face tn() { return Face(3); }
  static constexpr Face bw() { return Face(4); }
  static constexpr Face be() { return Face(5); }
  static constexpr Face tw() { return Face(6); }
  static constexpr Face te() { return Face(7); }
  static constexpr Face sw_edge() { return Face(8); }
  static constexpr Face se_edge() { return Face(9); }
  static constexpr Face nw_edge() { return Face(10); }
  static constexpr Face ne_edge() { return Face(11); }

 private:
  unsigned char v_;
};
template <int D>
using Side = Face<D, D - 1>;
using Edge = Face<3, 1>;
template <int D>
using Corner = Face<D, 0>;

/* Orthant<D> (Orthant.h:42-348): bit a = upper half on axis a. */
template <int D>
class Orthant {
 public:
  static constexpr size_t num_orthants = (size_t)1 << D;
  constexpr Orthant() : v_((unsigned char)num_orthants) {}
  constexpr explicit Orthant(unsigned char value) : v_(value) {}
  static constexpr Orthant null() { return Orthant((unsigned char)num_orthants); }
  static constexpr Orthant lower() { return Orthant(0); }
  static constexpr Orthant upper() { return Orthant(1); }
  static constexpr Orthant sw() { return Orthant(0); }
  static constexpr Orthant se() { return Orthant(1); }
  static constexpr Orthant nw() { return Orthant(2); }
  static constexpr Orthant ne() { return Orthant(3); }
  size_t getIndex() const { return v_; }
  bool isOnSide(Side<D> s) const { return (((v_ >> s.getAxisIndex()) & 1) != 0) == s.isHigherOnAxis(); }
  static std::vector<Orthant> getValues()
  {
    std::vector<Orthant> out;
    for (size_t k = 0; k < num_orthants; k++) out.push_back(Orthant((unsigned char)k));
    return out;
  }
  bool operator==(const Orthant& o) const { return v_ == o.v_; }
  bool operator!=(const Orthant& o) const { return v_ != o.v_; }

 private:
  unsigned char v_;
};

/* NbrType.h:33-48 */
enum class NbrType { Normal, Coarse, Fine };

/* View<T, D> (View.h:40-226): a strided D-dimensional view of patch data with a ghost
 * range; coordinates run from getGhostStart() (= -g on a padded axis) to getGhostEnd(). */
template <typename T, int D>
class View {
 public:
  using T_ptr = T*;
  View() = default;
  View(T* data, const std::array<int, D>& strides, const std::array<int, D>& ghost_start,
       const std::array<int, D>& start, const std::array<int, D>& end, const std::array<int, D>& ghost_end)
      : data_(data), strides_(strides), gs_(ghost_start), s_(start), e_(end), ge_(ghost_end)
  {
  }
  template <typename U, typename = typename std::enable_if<std::is_same<const U, T>::value>::type>
  View(const View<U, D>& o)
      : data_(o.getPointer()), strides_(o.getStrides()), gs_(o.getGhostStart()), s_(o.getStart()), e_(o.getEnd()),
        ge_(o.getGhostEnd())
  {
  }
  T& operator[](const std::array<int, D>& coord) const
  {
    long off = 0;
    for (int a = 0; a < D; a++) off += (long)coord[(size_t)a] * strides_[(size_t)a];
    return data_[off];
  }
  template <typename... I>
  T& operator()(I... coord) const
  {
    static_assert(sizeof...(I) == D, "one index per axis");
    return (*this)[std::array<int, D>{ { (int)coord... } }];
  }
  const std::array<int, D>& getStrides() const { return strides_; }
  const std::array<int, D>& getGhostStart() const { return gs_; }
  const std::array<int, D>& getStart() const { return s_; }
  const std::array<int, D>& getEnd() const { return e_; }
  const std::array<int, D>& getGhostEnd() const { return ge_; }
  /* pointer to the element at coordinate 0 */
  T* getPointer() const { return data_; }

 private:
  T* data_ = nullptr;
  std::array<int, D> strides_{}, gs_{}, s_{}, e_{}, ge_{};
};

namespace detail {
/* the view of face f (M-dimensional) of a view with D spatial axes (and, when C = 1, a
 * trailing component axis kept as the last axis of the result); offset[k] = layer
 * counted inward from the face's k-th side (0 = first interior layer, -1 = first ghost) */
template <typename T, int D, int M, int C>
View<T, M + C> face_slice(const View<T, D + C>& v, Face<D, M> f, const std::array<int, D - M>& offset)
{
  auto sides = f.getSides();
  T* p = v.getPointer();
  std::array<bool, D> fixed{};
  for (int k = 0; k < D - M; k++) {
    const Face<D, D - 1> sd = sides[(size_t)k];
    const int a = (int)sd.getAxisIndex();
    fixed[(size_t)a] = true;
    const int idx = sd.isLowerOnAxis() ? offset[(size_t)k] : v.getEnd()[(size_t)a] - offset[(size_t)k];
    p += (long)idx * v.getStrides()[(size_t)a];
  }
  std::array<int, M + C> st{}, gs{}, s{}, e{}, ge{};
  int j = 0;
  for (int a = 0; a < D + C; a++) {
    if (a < D && fixed[(size_t)a]) continue;
    st[(size_t)j] = v.getStrides()[(size_t)a];
    gs[(size_t)j] = v.getGhostStart()[(size_t)a];
    s[(size_t)j] = v.getStart()[(size_t)a];
    e[(size_t)j] = v.getEnd()[(size_t)a];
    ge[(size_t)j] = v.getGhostEnd()[(size_t)a];
    j++;
  }
  return View<T, M + C>(p, st, gs, s, e, ge);
}
}  // namespace detail

/* ComponentView<T, D> (ComponentView.h:38-266): one component of one patch */
template <typename T, int D>
class ComponentView : public View<T, D> {
 public:
  ComponentView() = default;
  explicit ComponentView(const View<T, D>& v) : View<T, D>(v) {}
  template <int M>
  View<T, M> getSliceOn(Face<D, M> f, const std::array<int, D - M>& offset) const
  {
    return detail::face_slice<T, D, M, 0>(*this, f, offset);
  }
  /* the ghost layers are writable through a view of const data (PatchView.h:261-262) */
  template <int M>
  View<typename std::remove_const<T>::type, M> getGhostSliceOn(Face<D, M> f, const std::array<size_t, D - M>& offset) const
  {
    std::array<int, D - M> o;
    for (size_t k = 0; k < (size_t)(D - M); k++) o[k] = -1 - (int)offset[k];
    View<typename std::remove_const<T>::type, D> w(const_cast<typename std::remove_const<T>::type*>(this->getPointer()),
                                                   this->getStrides(), this->getGhostStart(), this->getStart(),
                                                   this->getEnd(), this->getGhostEnd());
    return detail::face_slice<typename std::remove_const<T>::type, D, M, 0>(w, f, o);
  }
};

/* PatchView<T, D> (PatchView.h:38-306): all components of one patch; the last axis is the
 * component */
template <typename T, int D>
class PatchView : public View<T, D + 1> {
 public:
  PatchView() = default;
  explicit PatchView(const View<T, D + 1>& v) : View<T, D + 1>(v) {}
  template <typename U, typename = typename std::enable_if<std::is_same<const U, T>::value>::type>
  PatchView(const PatchView<U, D>& o) : View<T, D + 1>(o)
  {
  }
  template <int M>
  View<T, M + 1> getSliceOn(Face<D, M> f, const std::array<int, D - M>& offset) const
  {
    return detail::face_slice<T, D, M, 1>(*this, f, offset);
  }
  template <int M>
  View<typename std::remove_const<T>::type, M + 1> getGhostSliceOn(Face<D, M> f,
                                                                  const std::array<size_t, D - M>& offset) const
  {
    using U = typename std::remove_const<T>::type;
    std::array<int, D - M> o;
    for (size_t k = 0; k < (size_t)(D - M); k++) o[k] = -1 - (int)offset[k];
    View<U, D + 1> w(const_cast<U*>(this->getPointer()), this->getStrides(), this->getGhostStart(),
                     this->getStart(), this->getEnd(), this->getGhostEnd());
    return detail::face_slice<U, D, M, 1>(w, f, o);
  }
  ComponentView<T, D> getComponentView(int component) const
  {
    std::array<int, D> st, gs, s, e, ge;
    for (size_t a = 0; a < (size_t)D; a++) {
      st[a] = this->getStrides()[a];
      gs[a] = this->getGhostStart()[a];
      s[a] = this->getStart()[a];
      e[a] = this->getEnd()[a];
      ge[a] = this->getGhostEnd()[a];
    }
    return ComponentView<T, D>(
      View<T, D>(this->getPointer() + (long)component * this->getStrides()[(size_t)D], st, gs, s, e, ge));
  }
};

/* Loops.h:38-158 */
namespace Loop {
template <int D, class F>
void Nested(const std::array<int, D>& start, const std::array<int, D>& end, F f)
{
  std::array<int, D> c = start;
  if (D == 0) {
    f(c);
    return;
  }
  for (int a = 0; a < D; a++)
    if (end[(size_t)a] < start[(size_t)a]) return;
  while (true) {
    f(c);
    int a = 0;
    while (a < D) {
      if (c[(size_t)a] < end[(size_t)a]) {
        c[(size_t)a]++;
        break;
      }
      c[(size_t)a] = start[(size_t)a];
      a++;
    }
    if (a == D) return;
  }
}
template <int D, class V, class F>
void OverInteriorIndexes(const V& v, F f)
{
  Nested<D>(v.getStart(), v.getEnd(), f);
}
template <int D, class V, class F>
void OverAllIndexes(const V& v, F f)
{
  Nested<D>(v.getGhostStart(), v.getGhostEnd(), f);
}
}  // namespace Loop

/* Neighbour records of PatchInfo (NormalNbrInfo.h:38-100, CoarseNbrInfo.h:38-113,
 * FineNbrInfo.h:38-127); one rank, so every rank is 0 */
template <int M>
struct NormalNbrInfo {
  int id = -1, rank = 0, local_index = -1;
};
template <int M>
struct CoarseNbrInfo {
  int id = -1, rank = 0, local_index = -1;
  Orthant<M> orth_on_coarse;
};
template <int M>
struct FineNbrInfo {
  std::array<int, Orthant<M>::num_orthants> ids{}, ranks{}, local_indexes{};
};

/* PatchInfo<D> (PatchInfo.h:52-463): geometry, tree links and one neighbour slot per face,
 * edge and corner (slot = Face<D, M>::sum_of_faces + face value, PatchInfo.h:222-289).  The
 * default is one cell of unit size at the origin with no neighbours, as the reference's
 * default constructor. */
template <int D>
struct PatchInfo {
  int id = 0;
  int local_index = 0;
  int global_index = 0;
  int rank = 0;
  int refine_level = -1;
  int parent_id = -1;
  int parent_rank = -1;
  Orthant<D> orth_on_parent = Orthant<D>::null();
  int num_ghost_cells = 1;
  std::array<int, (1 << D)> child_ids;
  std::array<int, (1 << D)> child_ranks;
  std::array<int, D> ns;
  std::array<double, D> starts;
  std::array<double, D> spacings;
  PatchInfo()
  {
    ns.fill(1);
    starts.fill(0.0);
    spacings.fill(1.0);
    child_ids.fill(-1);
    child_ranks.fill(-1);
    nbr_type_.fill(-1);
    nbr_orth_.fill(-1);
    for (auto& a : nbr_id_) a.fill(-1);
    for (auto& a : nbr_local_) a.fill(-1);
  }
  template <int M>
  static constexpr int slot(Face<D, M> f)
  {
    return (D == 2 ? (M == 0 ? 0 : 4) : (M == 0 ? 0 : (M == 1 ? 8 : 20))) + (int)f.getIndex();
  }
  template <int M>
  bool hasNbr(Face<D, M> f) const
  {
    return nbr_type_[(size_t)slot(f)] >= 0;
  }
  /* any neighbour on a face, edge or corner */
  bool hasNbr() const
  {
    for (int t : nbr_type_)
      if (t >= 0) return true;
    return false;
  }
  template <int M>
  NbrType getNbrType(Face<D, M> f) const
  {
    const int t = nbr_type_[(size_t)slot(f)];
    if (t < 0) throw RuntimeError("PatchInfo: no neighbour on that face");
    return t == TE_NBR_NORMAL ? NbrType::Normal : (t == TE_NBR_COARSE ? NbrType::Coarse : NbrType::Fine);
  }
  template <int M>
  NormalNbrInfo<M> getNormalNbrInfo(Face<D, M> f) const
  {
    check_type(f, TE_NBR_NORMAL);
    NormalNbrInfo<M> r;
    r.id = nbr_id_[(size_t)slot(f)][0];
    r.local_index = nbr_local_[(size_t)slot(f)][0];
    return r;
  }
  template <int M>
  CoarseNbrInfo<M> getCoarseNbrInfo(Face<D, M> f) const
  {
    check_type(f, TE_NBR_COARSE);
    CoarseNbrInfo<M> r;
    r.id = nbr_id_[(size_t)slot(f)][0];
    r.local_index = nbr_local_[(size_t)slot(f)][0];
    r.orth_on_coarse = Orthant<M>((unsigned char)(nbr_orth_[(size_t)slot(f)] < 0 ? 0 : nbr_orth_[(size_t)slot(f)]));
    return r;
  }
  template <int M>
  FineNbrInfo<M> getFineNbrInfo(Face<D, M> f) const
  {
    check_type(f, TE_NBR_FINE);
    FineNbrInfo<M> r;
    for (size_t k = 0; k < Orthant<M>::num_orthants; k++) {
      r.ids[k] = nbr_id_[(size_t)slot(f)][k];
      r.local_indexes[k] = nbr_local_[(size_t)slot(f)][k];
      r.ranks[k] = 0;
    }
    return r;
  }
  bool operator<(const PatchInfo& o) const { return id < o.id; }
  bool operator==(const PatchInfo& o) const { return id == o.id; }
  /* the neighbour slots, filled by Domain */
  std::array<int, 26> nbr_type_;
  std::array<int, 26> nbr_orth_;
  std::array<std::array<int, 4>, 26> nbr_id_;
  std::array<std::array<int, 4>, 26> nbr_local_;

 private:
  template <int M>
  void check_type(Face<D, M> f, int t) const
  {
    if (nbr_type_[(size_t)slot(f)] != t) throw RuntimeError("PatchInfo: the neighbour is not of the requested type");
  }
};

/* Domain<D> (Domain.h:94-248): the patches of one level, local index = position. */
template <int D>
class Domain {
 public:
  Domain() = default;
  template <class It>
  Domain(Communicator, int id, std::array<int, D> ns, int num_ghost_cells, It first, It last)
      : id_(id), ns_(ns), num_ghost_(num_ghost_cells), patches_(first, last)
  {
    std::array<int, 3> n3{ 1, 1, 1 };
    for (int a = 0; a < D; a++) n3[a] = ns[a];
    te_domain* h = nullptr;
    detail::check(te_domain_create(D, n3.data(), num_ghost_cells, (int)patches_.size(), patches_.data(), &h));
    h_ = std::shared_ptr<te_domain>(h, detail::Deleter<te_domain, te_domain_destroy>());
    // the reference's PatchInfo objects, neighbour local indexes resolved (Domain.h:94-109)
    std::vector<int32_t> nl(patches_.size() * 26 * 4);
    if (!patches_.empty()) detail::check(te_domain_get_nbr_local(h, nl.data()));
    infos_ = std::make_shared<std::vector<PatchInfo<D>>>();
    infos_->reserve(patches_.size());
    for (size_t i = 0; i < patches_.size(); i++) {
      const te_patch_info& r = patches_[i];
      PatchInfo<D> pi;
      pi.id = r.id;
      pi.local_index = pi.global_index = (int)i;
      pi.rank = r.rank;
      pi.refine_level = r.refine_level;
      pi.parent_id = r.parent_id;
      pi.parent_rank = r.parent_rank;
      pi.orth_on_parent = r.orth_on_parent < 0 ? Orthant<D>::null() : Orthant<D>((unsigned char)r.orth_on_parent);
      pi.num_ghost_cells = num_ghost_;
      for (int k = 0; k < (1 << D); k++) {
        pi.child_ids[(size_t)k] = r.child_ids[k];
        pi.child_ranks[(size_t)k] = r.child_ranks[k];
      }
      for (int a = 0; a < D; a++) {
        pi.ns[(size_t)a] = ns_[(size_t)a];
        pi.starts[(size_t)a] = r.starts[a];
        pi.spacings[(size_t)a] = r.spacings[a];
      }
      for (int sl = 0; sl < 26; sl++) {
        pi.nbr_type_[(size_t)sl] = (D == 2 && sl >= 8) ? -1 : r.nbr_type[sl];
        pi.nbr_orth_[(size_t)sl] = r.nbr_orth[sl];
        for (int k = 0; k < 4; k++) {
          pi.nbr_id_[(size_t)sl][(size_t)k] = pi.nbr_type_[(size_t)sl] >= 0 ? r.nbr_ids[sl][k] : -1;
          pi.nbr_local_[(size_t)sl][(size_t)k] = nl[(i * 26 + (size_t)sl) * 4 + (size_t)k];
        }
      }
      infos_->push_back(pi);
    }
  }
  int getId() const { return id_; }
  const std::array<int, D>& getNs() const { return ns_; }
  int getNumGhostCells() const { return num_ghost_; }
  int getNumLocalPatches() const { return (int)patches_.size(); }
  int getNumGlobalPatches() const { return (int)patches_.size(); }
  long getNumLocalCells() const
  {
    long c = 1;
    for (int a = 0; a < D; a++) c *= ns_[a];
    return c * (long)patches_.size();
  }
  long getNumGlobalCells() const { return getNumLocalCells(); }
  /* Domain::getPatchInfoVector (Domain.h:257): the patches in local-index order */
  const std::vector<PatchInfo<D>>& getPatchInfoVector() const
  {
    static const std::vector<PatchInfo<D>> none;
    return infos_ ? *infos_ : none;
  }
  PatchInfo<D> getPatchInfo(int local_index) const { return getPatchInfoVector().at((size_t)local_index); }
  /* the flattened te_patch_info records this domain was made from */
  const std::vector<te_patch_info>& getPatchTable() const { return patches_; }
  Communicator getCommunicator() const { return {}; }
  te_domain* handle() const { return h_.get(); }
  bool operator==(const Domain& o) const { return h_ == o.h_; }
  /* Domain::setTimer / getTimer / hasTimer (Domain.h:316-332, json.cpp:357-362); the
   * timer is shared by every copy of this domain */
  void setTimer(std::shared_ptr<Timer> timer) const
  {
    *timer_ = timer;
    timer->addDomain(id_, *this);
  }
  std::shared_ptr<Timer> getTimer() const { return *timer_; }
  bool hasTimer() const { return *timer_ != nullptr; }

 private:
  std::shared_ptr<std::shared_ptr<Timer>> timer_{ std::make_shared<std::shared_ptr<Timer>>() };
  int id_ = -1;
  std::array<int, D> ns_{};
  int num_ghost_ = 1;
  std::vector<te_patch_info> patches_;
  std::shared_ptr<std::vector<PatchInfo<D>>> infos_;
  std::shared_ptr<te_domain> h_;
};

/* Vector<D> (Vector.h:112-611): device-resident, reference layout, FP64. */
template <int D>
class Vector {
 public:
  Vector() = default;
  Vector(const Domain<D>& domain, int num_components) : domain_(domain), ncomp_(num_components)
  {
    len_ = te_domain_vector_length(domain.handle(), num_components);
    void* p = nullptr;
    detail::check(te_device_alloc(sizeof(double) * (size_t)(len_ > 0 ? len_ : 1), &p));
    data_ = std::shared_ptr<double>(static_cast<double*>(p), [](double* q) { te_device_free(q); });
    detail::check(te_vec_set_with_ghost(domain_.handle(), TE_F64, data_.get(), 0.0, ncomp_, nullptr));
  }
  Vector(const Vector& o)
  {
    if (!o.data_) return;
    *this = Vector(o.domain_, o.ncomp_);
    copyWithGhost(o);
  }
  Vector(Vector&&) noexcept = default;
  Vector& operator=(const Vector& o)
  {
    if (this != &o) {
      if (!o.data_) {
        *this = Vector();
      } else {
        if (!data_ || len_ != o.len_) *this = Vector(o.domain_, o.ncomp_);
        copyWithGhost(o);
      }
    }
    return *this;
  }
  Vector& operator=(Vector&&) noexcept = default;

  Vector getZeroClone() const { return Vector(domain_, ncomp_); }
  int getNumComponents() const { return ncomp_; }
  int getNumLocalPatches() const { return domain_.getNumLocalPatches(); }
  long getNumLocalCells() const { return domain_.getNumLocalCells(); }
  const Domain<D>& getDomain() const { return domain_; }
  Communicator getCommunicator() const { return {}; }
  /* device pointer to the padded data.  Writes made through patch views (a host copy) are
   * copied back first, and the host copy is dropped: the caller may change the data on
   * the device. */
  double* data() const
  {
    to_device();
    host_->valid = false;
    return data_.get();
  }
  long size() const { return len_; }
  /* Vector::getPatchView / getComponentView (Vector.h:336-375).  The views address a host
   * copy of the vector (the data lives in device memory): it is fetched on the first view
   * and copied back before the next device operation on the vector (every other method,
   * and every operator, smoother or solver the vector is passed to), after which views taken
   * before that operation are stale.  Ghost cells are writable through views of a const
   * vector, as in the reference (PatchView.h:261-262). */
  PatchView<double, D> getPatchView(int local_index) { return PatchView<double, D>(host_view(local_index)); }
  PatchView<const double, D> getPatchView(int local_index) const
  {
    return PatchView<const double, D>(View<const double, D + 1>(host_view(local_index)));
  }
  ComponentView<double, D> getComponentView(int component, int local_index)
  {
    return getPatchView(local_index).getComponentView(component);
  }
  ComponentView<const double, D> getComponentView(int component, int local_index) const
  {
    return getPatchView(local_index).getComponentView(component);
  }

  void set(double alpha) { chk(te_vec_set(dh(), TE_F64, data(), alpha, ncomp_, nullptr)); }
  void setWithGhost(double alpha) { chk(te_vec_set_with_ghost(dh(), TE_F64, data(), alpha, ncomp_, nullptr)); }
  void scale(double alpha) { chk(te_vec_scale(dh(), TE_F64, data(), alpha, ncomp_, nullptr)); }
  void shift(double delta) { chk(te_vec_shift(dh(), TE_F64, data(), delta, ncomp_, nullptr)); }
  void copy(const Vector& b) { chk(te_vec_copy(dh(), TE_F64, data(), b.data(), ncomp_, nullptr)); }
  void add(const Vector& b) { chk(te_vec_add(dh(), TE_F64, data(), b.data(), ncomp_, nullptr)); }
  void addScaled(double alpha, const Vector& a)
  {
    chk(te_vec_add_scaled(dh(), TE_F64, data(), alpha, a.data(), ncomp_, nullptr));
  }
  void addScaled(double alpha, const Vector& a, double beta, const Vector& b)
  {
    chk(te_vec_add_scaled2(dh(), TE_F64, data(), alpha, a.data(), beta, b.data(), ncomp_, nullptr));
  }
  void scaleThenAdd(double alpha, const Vector& b)
  {
    chk(te_vec_scale_then_add(dh(), TE_F64, data(), alpha, b.data(), ncomp_, nullptr));
  }
  void scaleThenAddScaled(double alpha, double beta, const Vector& b)
  {
    chk(te_vec_scale_then_add_scaled(dh(), TE_F64, data(), alpha, beta, b.data(), ncomp_, nullptr));
  }
  void scaleThenAddScaled(double alpha, double beta, const Vector& b, double gamma, const Vector& c)
  {
    chk(te_vec_scale_then_add_scaled2(dh(), TE_F64, data(), alpha, beta, b.data(), gamma, c.data(), ncomp_,
                                      nullptr));
  }
  double dot(const Vector& b) const
  {
    double r = 0;
    chk(te_vec_dot(dh(), TE_F64, data(), b.data(), ncomp_, &r, nullptr));
    return r;
  }
  double twoNorm() const
  {
    double r = 0;
    chk(te_vec_two_norm(dh(), TE_F64, data(), ncomp_, &r, nullptr));
    return r;
  }
  /* max(0, max v), as the reference (Vector.h:568-579) */
  double infNorm() const
  {
    double r = 0;
    chk(te_vec_inf_norm(dh(), TE_F64, data(), ncomp_, &r, nullptr));
    return r;
  }
  /* host copies of the whole padded vector ([patch][comp][z][y][x]) */
  std::vector<double> copyToHost() const
  {
    std::vector<double> h((size_t)len_);
    chk(te_stream_synchronize(nullptr));
    chk(te_memcpy_d2h(h.data(), data(), sizeof(double) * (size_t)len_));
    return h;
  }
  void copyFromHost(const std::vector<double>& h)
  {
    if ((long)h.size() != len_) throw RuntimeError("copyFromHost: size mismatch");
    chk(te_stream_synchronize(nullptr));
    chk(te_memcpy_h2d(data(), h.data(), sizeof(double) * (size_t)len_));
  }

 private:
  te_domain* dh() const { return domain_.handle(); }
  static void chk(int rc) { detail::check(rc); }
  void copyWithGhost(const Vector& o)
  {
    chk(te_memcpy_d2d(data(), o.data(), sizeof(double) * (size_t)len_, nullptr));
  }
  // the host copy behind the patch views: valid = it holds the current values, dirty = it
  // may hold writes the device copy lacks
  struct HostCopy {
    std::vector<double> h;
    bool valid = false;
    bool dirty = false;
  };
  void to_device() const
  {
    if (!host_->dirty) return;
    chk(te_stream_synchronize(nullptr));
    chk(te_memcpy_h2d(data_.get(), host_->h.data(), sizeof(double) * (size_t)len_));
    host_->dirty = false;
  }
  View<double, D + 1> host_view(int local_index) const
  {
    if (local_index < 0 || local_index >= getNumLocalPatches()) throw RuntimeError("getPatchView: bad patch index");
    if (!host_->valid) {
      host_->h.resize((size_t)len_);
      chk(te_stream_synchronize(nullptr));
      chk(te_memcpy_d2h(host_->h.data(), data_.get(), sizeof(double) * (size_t)len_));
      host_->valid = true;
    }
    host_->dirty = true;
    const auto& ns = domain_.getNs();
    std::array<int, D + 1> st, gs, s, e, ge;
    long stride = 1, patch = 1;
    for (int a = 0; a < D; a++) {
      st[(size_t)a] = (int)stride;
      gs[(size_t)a] = -1;
      s[(size_t)a] = 0;
      e[(size_t)a] = ns[(size_t)a] - 1;
      ge[(size_t)a] = ns[(size_t)a];
      stride *= ns[(size_t)a] + 2;
    }
    st[(size_t)D] = (int)stride;  // one component's padded cells
    gs[(size_t)D] = s[(size_t)D] = 0;
    e[(size_t)D] = ge[(size_t)D] = ncomp_ - 1;
    patch = stride * ncomp_;
    // the element at coordinate 0: one ghost layer in on every axis
    long first = (long)local_index * patch;
    for (int a = 0; a < D; a++) first += st[(size_t)a];
    return View<double, D + 1>(host_->h.data() + first, st, gs, s, e, ge);
  }
  Domain<D> domain_;
  int ncomp_ = 0;
  long len_ = 0;
  std::shared_ptr<double> data_;
  mutable std::shared_ptr<HostCopy> host_ = std::make_shared<HostCopy>();
};

/* Operator.h:40-58 */
template <int D>
class Operator {
 public:
  virtual ~Operator() {}
  virtual Operator<D>* clone() const = 0;
  virtual void apply(const Vector<D>& x, Vector<D>& b) const = 0;
};

/* GhostFiller.h:37-53, MPIGhostFiller.h (fill types) */
template <int D>
class GhostFiller {
 public:
  virtual ~GhostFiller() {}
  virtual GhostFiller<D>* clone() const = 0;
  /* ghost cells are writable through the const reference, as in the reference */
  virtual void fillGhost(const Vector<D>& u) const = 0;
};

/* The three built-in fillers share one device implementation (te_ghost_filler). */
template <int D>
class DeviceGhostFiller : public GhostFiller<D> {
 public:
  DeviceGhostFiller(const Domain<D>& domain, GhostFillingType fill_type, int kind)
      : domain_(domain), fill_type_(fill_type), kind_(kind)
  {
    te_ghost_filler* h = nullptr;
    detail::check(te_ghost_filler_create(domain.handle(), kind, (int)fill_type, &h));
    h_ = std::shared_ptr<te_ghost_filler>(h, detail::Deleter<te_ghost_filler, te_ghost_filler_destroy>());
  }
  void fillGhost(const Vector<D>& u) const override
  {
    detail::check(te_ghost_fill(h_.get(), TE_F64, u.data(), u.getNumComponents(), nullptr));
  }
  const Domain<D>& getDomain() const { return domain_; }
  GhostFillingType getFillType() const { return fill_type_; }
  int kind() const { return kind_; }
  te_ghost_filler* handle() const { return h_.get(); }

 private:
  Domain<D> domain_;
  GhostFillingType fill_type_;
  int kind_;
  std::shared_ptr<te_ghost_filler> h_;
};

/* BiLinearGhostFiller.h / BiQuadraticGhostFiller.h / TriLinearGhostFiller.h */
class BiLinearGhostFiller : public DeviceGhostFiller<2> {
 public:
  BiLinearGhostFiller(const Domain<2>& domain, GhostFillingType fill_type)
      : DeviceGhostFiller<2>(domain, fill_type, TE_BILINEAR) {}
  BiLinearGhostFiller* clone() const override { return new BiLinearGhostFiller(*this); }
};
class BiQuadraticGhostFiller : public DeviceGhostFiller<2> {
 public:
  BiQuadraticGhostFiller(const Domain<2>& domain, GhostFillingType fill_type)
      : DeviceGhostFiller<2>(domain, fill_type, TE_BIQUADRATIC) {}
  BiQuadraticGhostFiller* clone() const override { return new BiQuadraticGhostFiller(*this); }
};
class TriLinearGhostFiller : public DeviceGhostFiller<3> {
 public:
  TriLinearGhostFiller(const Domain<3>& domain, GhostFillingType fill_type)
      : DeviceGhostFiller<3>(domain, fill_type, TE_TRILINEAR) {}
  TriLinearGhostFiller* clone() const override { return new TriLinearGhostFiller(*this); }
};

namespace detail {
/* A function sampled at the face points of every side of every patch (the points of
 * DomainTools::GetRealCoordBound, DomainTools.h:178-210), in the boundary-data layout of
 * te_domain_boundary_length: per patch, per side, the side's face cells with the first of the
 * other axes fastest.  func(side axis, point). */
template <int D, class F>
std::vector<double> boundary_samples(const Domain<D>& domain, F func)
{
  std::vector<double> out;
  out.reserve((size_t)te_domain_boundary_length(domain.handle()));
  const auto& ns = domain.getNs();
  for (const auto& r : domain.getPatchInfoVector()) {
    for (int s = 0; s < 2 * D; s++) {
      const int a = s >> 1;
      int fr[2] = { 0, 0 }, nf = 0;
      for (int b = 0; b < D; b++)
        if (b != a) fr[nf++] = b;
      const int n0 = nf > 0 ? ns[fr[0]] : 1, n1 = nf > 1 ? ns[fr[1]] : 1;
      for (int j = 0; j < n1; j++)
        for (int i = 0; i < n0; i++) {
          std::array<double, D> x;
          x[a] = (s & 1) ? r.starts[a] + r.spacings[a] * ns[a] : r.starts[a];
          const int idx[2] = { i, j };
          for (int k = 0; k < nf; k++)
            x[fr[k]] = r.starts[fr[k]] + r.spacings[fr[k]] / 2 + r.spacings[fr[k]] * idx[k];
          out.push_back(func(a, x));
        }
    }
  }
  return out;
}

/* a host array copied to a temporary device buffer */
struct DeviceCopy {
  explicit DeviceCopy(const std::vector<double>& h)
  {
    check(te_device_alloc(sizeof(double) * (h.empty() ? 1 : h.size()), &p));
    if (!h.empty()) check(te_memcpy_h2d(p, h.data(), sizeof(double) * h.size()));
  }
  ~DeviceCopy() { te_device_free(p); }
  DeviceCopy(const DeviceCopy&) = delete;
  DeviceCopy& operator=(const DeviceCopy&) = delete;
  const double* get() const { return static_cast<const double*>(p); }
  void* p = nullptr;
};
}  // namespace detail

/* PatchOperator.h:45-143.  The per-patch virtuals take host views of one patch (see
 * Vector::getPatchView).  apply() fills the ghosts of u and calls applySinglePatch on every
 * patch; the built-in operators override it with their device kernels (when they are used
 * as they are, not through a subclass), so a user-derived operator runs through the
 * virtuals and a built-in one on the device. */
template <int D>
class PatchOperator : public Operator<D> {
 public:
  PatchOperator(const Domain<D>& domain, const GhostFiller<D>& ghost_filler)
      : domain_(domain), filler_(ghost_filler.clone())
  {
    dev_ = dynamic_cast<const DeviceGhostFiller<D>*>(filler_.get());
  }
  PatchOperator<D>* clone() const override = 0;
  /* f = A u on one patch; u's ghost values are filled (PatchOperator.h:85-93) */
  virtual void applySinglePatch(const PatchInfo<D>& pinfo, const PatchView<const double, D>& u_view,
                                const PatchView<double, D>& f_view) const = 0;
  /* the patch's internal boundaries treated as domain boundaries (:96-107) */
  virtual void modifyRHSForInternalBoundaryConditions(const PatchInfo<D>& pinfo,
                                                      const PatchView<const double, D>& u_view,
                                                      const PatchView<double, D>& f_view) const = 0;
  virtual void applySinglePatchWithInternalBoundaryConditions(const PatchInfo<D>& pinfo,
                                                              const PatchView<const double, D>& u_view,
                                                              const PatchView<double, D>& f_view) const = 0;
  /* PatchOperator::apply (:126-143): f := 0 with ghosts, fill u's ghosts, every patch */
  void apply(const Vector<D>& u, Vector<D>& f) const override
  {
    f.setWithGhost(0.0);
    filler_->fillGhost(u);
    for (const PatchInfo<D>& pinfo : domain_.getPatchInfoVector())
      applySinglePatch(pinfo, u.getPatchView(pinfo.local_index), f.getPatchView(pinfo.local_index));
  }
  const Domain<D>& getDomain() const { return domain_; }
  const GhostFiller<D>& getGhostFiller() const { return *filler_; }
  bool hasDeviceFiller() const { return dev_ != nullptr; }
  const DeviceGhostFiller<D>& deviceFiller() const
  {
    if (!dev_) throw RuntimeError("PatchOperator: the device path needs one of the built-in ghost fillers");
    return *dev_;
  }

 private:
  Domain<D> domain_;
  std::shared_ptr<const GhostFiller<D>> filler_;
  const DeviceGhostFiller<D>* dev_ = nullptr;
};

namespace detail {
/* the physical-side (internal = false) or neighbour-side (internal = true) ghost values
 * of u set to sign * the first interior layer (StarPatchOperator::enforceBoundaryConditions
 * / enforceInternalBoundaryConditions, Poisson/StarPatchOperator.h:115-164) */
template <int D>
void set_side_ghosts(const PatchInfo<D>& pinfo, const PatchView<const double, D>& u_view, bool internal,
                     double sign)
{
  for (Side<D> s : Side<D>::getValues()) {
    if (pinfo.hasNbr(s) != internal) continue;
    View<double, D> g = u_view.getGhostSliceOn(s, { 0 });
    View<const double, D> mid = u_view.getSliceOn(s, { 0 });
    Loop::OverInteriorIndexes<D>(mid, [&](const std::array<int, D>& c) { g[c] = sign * mid[c]; });
  }
}
}  // namespace detail

namespace Poisson {
/* Poisson/StarPatchOperator.h:39-243: f = sum_a (u+ - 2u + u-)/h_a^2 */
template <int D>
class StarPatchOperator : public PatchOperator<D> {
 public:
  StarPatchOperator(const Domain<D>& domain, const GhostFiller<D>& ghost_filler, bool neumann = false)
      : PatchOperator<D>(domain, ghost_filler), neumann_(neumann)
  {
    if (domain.getNumGhostCells() < 1) throw RuntimeError("StarPatchOperator needs at least one set of ghost cells");
  }
  StarPatchOperator* clone() const override { return new StarPatchOperator(*this); }
  void apply(const Vector<D>& x, Vector<D>& b) const override
  {
    if (typeid(*this) != typeid(StarPatchOperator) || !this->hasDeviceFiller()) {
      PatchOperator<D>::apply(x, b);  // a subclass or a user filler: the per-patch virtuals
      return;
    }
    detail::check(te_star_apply(this->deviceFiller().handle(), TE_F64, neumann_, x.data(), b.data(),
                                x.getNumComponents(), nullptr));
  }
  bool neumann() const { return neumann_; }
  void applySinglePatch(const PatchInfo<D>& pinfo, const PatchView<const double, D>& u_view,
                        const PatchView<double, D>& f_view) const override
  {
    applySinglePatch(pinfo, u_view, f_view, false);
  }
  void applySinglePatchWithInternalBoundaryConditions(const PatchInfo<D>& pinfo,
                                                      const PatchView<const double, D>& u_view,
                                                      const PatchView<double, D>& f_view) const override
  {
    applySinglePatch(pinfo, u_view, f_view, true);
  }
  /* physical ghosts: -u (Dirichlet) or +u (Neumann) (:115-144) */
  void enforceBoundaryConditions(const PatchInfo<D>& pinfo, const PatchView<const double, D>& u_view) const
  {
    detail::set_side_ghosts<D>(pinfo, u_view, false, neumann_ ? 1.0 : -1.0);
  }
  /* neighbour-side ghosts: -u (:145-164) */
  void enforceInternalBoundaryConditions(const PatchInfo<D>& pinfo, const PatchView<const double, D>& u_view) const
  {
    detail::set_side_ghosts<D>(pinfo, u_view, true, -1.0);
  }
  /* f_b -= (u_g + u_b) / h^2 on sides with a neighbour (:165-180) */
  void modifyRHSForInternalBoundaryConditions(const PatchInfo<D>& pinfo, const PatchView<const double, D>& u_view,
                                              const PatchView<double, D>& f_view) const override
  {
    for (Side<D> s : Side<D>::getValues()) {
      if (!pinfo.hasNbr(s)) continue;
      const double h2 = pinfo.spacings[s.getAxisIndex()] * pinfo.spacings[s.getAxisIndex()];
      View<double, D> f_inner = f_view.getSliceOn(s, { 0 });
      View<const double, D> u_ghost = u_view.getSliceOn(s, { -1 });
      View<const double, D> u_inner = u_view.getSliceOn(s, { 0 });
      Loop::OverInteriorIndexes<D>(f_inner, [&](const std::array<int, D>& c) {
        f_inner[c] -= (u_ghost[c] + u_inner[c]) / h2;
      });
    }
  }
  /* addDrichletBCToRHS (Poisson/StarPatchOperator.h:187-204): on every side without a
   * neighbour, f_b += -2 g(face point) / h^2 (gfunc sampled on the host, added on the device) */
  void addDrichletBCToRHS(Vector<D>& f, std::function<double(const std::array<double, D>&)> gfunc) const
  {
    const detail::DeviceCopy g(
      detail::boundary_samples<D>(this->getDomain(), [&](int, const std::array<double, D>& x) { return gfunc(x); }));
    chk_sync(te_star_add_dirichlet_bc(this->getDomain().handle(), TE_F64, f.data(), g.get(), nullptr, nullptr));
  }
  /* addNeumannBCToRHS (:206-242): f_b += dg/dx_a / h on lower sides, -= on upper sides, with
   * dg/dx_a = gfunc_grad[a] of the side's axis a */
  void addNeumannBCToRHS(Vector<D>& f, std::function<double(const std::array<double, D>&)> gfunc,
                         std::array<std::function<double(const std::array<double, D>&)>, D> gfunc_grad) const
  {
    (void)gfunc;  // the reference takes it too and uses only the gradient
    const detail::DeviceCopy dg(detail::boundary_samples<D>(
      this->getDomain(), [&](int a, const std::array<double, D>& x) { return gfunc_grad[(size_t)a](x); }));
    chk_sync(te_star_add_neumann_bc(this->getDomain().handle(), TE_F64, f.data(), dg.get(), nullptr));
  }

 private:
  /* the stencil on one patch (:75-101), axis by axis: f = (axis 0 ? 0 : f) + (u+ - 2u + u-)/h^2 */
  void applySinglePatch(const PatchInfo<D>& pinfo, const PatchView<const double, D>& u_view,
                        const PatchView<double, D>& f_view, bool internal) const
  {
    if (internal) enforceInternalBoundaryConditions(pinfo, u_view);
    enforceBoundaryConditions(pinfo, u_view);
    for (int a = 0; a < D; a++) {
      const double h2 = pinfo.spacings[(size_t)a] * pinfo.spacings[(size_t)a];
      const int st = u_view.getStrides()[(size_t)a];
      Loop::OverInteriorIndexes<D + 1>(u_view, [&](const std::array<int, D + 1>& c) {
        const double* q = &u_view[c];
        f_view[c] = (a == 0 ? 0.0 : 1.0) * f_view[c] + (q[st] - 2 * q[0] + q[-st]) / h2;
      });
    }
  }
  static void chk_sync(int rc)
  {
    detail::check(rc);
    detail::check(te_stream_synchronize(nullptr));  // before the temporary boundary data is freed
  }
  bool neumann_;
};
}  // namespace Poisson

namespace VarPoisson {
/* VarPoisson/StarPatchOperator.h:40-220; the coefficients' ghosts are filled once here (:70) */
template <int D>
class StarPatchOperator : public PatchOperator<D> {
 public:
  StarPatchOperator(const Vector<D>& coeffs, const Domain<D>& domain, const GhostFiller<D>& ghost_filler)
      : PatchOperator<D>(domain, ghost_filler), coeffs_(coeffs)
  {
    if (domain.getNumGhostCells() < 1) throw RuntimeError("StarPatchOperator needs at least one set of ghost cells");
    this->getGhostFiller().fillGhost(coeffs_);
  }
  StarPatchOperator* clone() const override { return new StarPatchOperator(*this); }
  void apply(const Vector<D>& x, Vector<D>& b) const override
  {
    if (typeid(*this) != typeid(StarPatchOperator) || !this->hasDeviceFiller()) {
      PatchOperator<D>::apply(x, b);
      return;
    }
    detail::check(te_varstar_apply(this->deviceFiller().handle(), TE_F64, coeffs_.data(), x.data(), b.data(),
                                   x.getNumComponents(), nullptr));
  }
  const Vector<D>& getCoefficients() const { return coeffs_; }
  void applySinglePatch(const PatchInfo<D>& pinfo, const PatchView<const double, D>& u_view,
                        const PatchView<double, D>& f_view) const override
  {
    applySinglePatch(pinfo, u_view, f_view, false);
  }
  void applySinglePatchWithInternalBoundaryConditions(const PatchInfo<D>& pinfo,
                                                      const PatchView<const double, D>& u_view,
                                                      const PatchView<double, D>& f_view) const override
  {
    applySinglePatch(pinfo, u_view, f_view, true);
  }
  /* physical ghosts -u (:132-150) */
  void enforceBoundaryConditions(const PatchInfo<D>& pinfo, const PatchView<const double, D>& u_view) const
  {
    detail::set_side_ghosts<D>(pinfo, u_view, false, -1.0);
  }
  void enforceInternalBoundaryConditions(const PatchInfo<D>& pinfo, const PatchView<const double, D>& u_view) const
  {
    detail::set_side_ghosts<D>(pinfo, u_view, true, -1.0);
  }
  /* f_b -= (u_g + u_b)(c_b + c_g) / (2 h^2) on sides with a neighbour (:168-186) */
  void modifyRHSForInternalBoundaryConditions(const PatchInfo<D>& pinfo, const PatchView<const double, D>& u_view,
                                              const PatchView<double, D>& f_view) const override
  {
    ComponentView<const double, D> c = coeffs_.getComponentView(0, pinfo.local_index);
    for (Side<D> s : Side<D>::getValues()) {
      if (!pinfo.hasNbr(s)) continue;
      const double h2 = pinfo.spacings[s.getAxisIndex()] * pinfo.spacings[s.getAxisIndex()];
      View<double, D> f_inner = f_view.getSliceOn(s, { 0 });
      View<double, D> u_ghost = u_view.getGhostSliceOn(s, { 0 });
      View<const double, D> u_inner = u_view.getSliceOn(s, { 0 });
      View<const double, D - 1> c_ghost = c.getSliceOn(s, { -1 });
      View<const double, D - 1> c_inner = c.getSliceOn(s, { 0 });
      Loop::OverInteriorIndexes<D>(f_inner, [&](const std::array<int, D>& k) {
        std::array<int, D - 1> kc;
        for (int a = 0; a < D - 1; a++) kc[(size_t)a] = k[(size_t)a];
        f_inner[k] -= (u_ghost[k] + u_inner[k]) * (c_inner[kc] + c_ghost[kc]) / (2 * h2);
      });
    }
  }
  /* addDrichletBCToRHS (VarPoisson/StarPatchOperator.h:194-220): on every side without a
   * neighbour, f_b -= 2 g c / h^2 with g = gfunc and c = hfunc at the face point */
  void addDrichletBCToRHS(Vector<D>& f, std::function<double(const std::array<double, D>&)> gfunc,
                          std::function<double(const std::array<double, D>&)> hfunc) const
  {
    const detail::DeviceCopy g(
      detail::boundary_samples<D>(this->getDomain(), [&](int, const std::array<double, D>& x) { return gfunc(x); }));
    const detail::DeviceCopy c(
      detail::boundary_samples<D>(this->getDomain(), [&](int, const std::array<double, D>& x) { return hfunc(x); }));
    detail::check(te_star_add_dirichlet_bc(this->getDomain().handle(), TE_F64, f.data(), g.get(), c.get(), nullptr));
    detail::check(te_stream_synchronize(nullptr));
  }
  /* shared by this operator and its clones (each clone deep-copies the coefficients, so
   * their buffers differ while their values are the same) */
  const void* coefficientIdentity() const { return ident_.get(); }

 private:
  /* :61-109: f = sum_a [(c+ + c)(u+ - u) - (c- + c)(u - u-)] / (2 h_a^2) */
  void applySinglePatch(const PatchInfo<D>& pinfo, const PatchView<const double, D>& u_view,
                        const PatchView<double, D>& f_view, bool internal) const
  {
    if (internal) enforceInternalBoundaryConditions(pinfo, u_view);
    enforceBoundaryConditions(pinfo, u_view);
    ComponentView<const double, D> c = coeffs_.getComponentView(0, pinfo.local_index);
    for (int a = 0; a < D; a++) {
      const double h2 = pinfo.spacings[(size_t)a] * pinfo.spacings[(size_t)a];
      const int st = u_view.getStrides()[(size_t)a], cst = c.getStrides()[(size_t)a];
      Loop::OverInteriorIndexes<D + 1>(u_view, [&](const std::array<int, D + 1>& k) {
        std::array<int, D> kc;
        for (int b = 0; b < D; b++) kc[(size_t)b] = k[(size_t)b];
        const double* q = &u_view[k];
        const double* cq = &c[kc];
        f_view[k] = (a == 0 ? 0.0 : 1.0) * f_view[k] +
                    ((cq[cst] + cq[0]) * (q[st] - q[0]) - (cq[-cst] + cq[0]) * (q[0] - q[-st])) / (2 * h2);
      });
    }
  }
  Vector<D> coeffs_;
  std::shared_ptr<const char> ident_{ std::make_shared<const char>(0) };
};
}  // namespace VarPoisson

/* MPIGhostFiller<D> (MPIGhostFiller.h:44-1107) on one rank, for user-written fillers: the
 * four per-neighbour virtuals on host views, called in the reference's order (zero the
 * ghosts facing neighbours, the local fill of every patch, then per face dimension -- sides,
 * edges, corners as the fill type includes them -- per patch, per face in value order, per
 * neighbour: the patch pushes into that neighbour's ghosts, :1055-1092, :821-895). */
template <int D>
class MPIGhostFiller : public GhostFiller<D> {
 public:
  MPIGhostFiller(const Domain<D>& domain, GhostFillingType fill_type) : domain_(domain), fill_type_(fill_type) {}
  void fillGhost(const Vector<D>& u) const override
  {
    const auto& infos = domain_.getPatchInfoVector();
    for (const PatchInfo<D>& pinfo : infos) zero_ghosts(pinfo, u.getPatchView(pinfo.local_index));
    for (const PatchInfo<D>& pinfo : infos) fillGhostCellsForLocalPatch(pinfo, u.getPatchView(pinfo.local_index));
    calls<D - 1>(u);
    if (D == 3 && fill_type_ != GhostFillingType::Faces) calls<(D == 3 ? 1 : 0)>(u);
    if (fill_type_ == GhostFillingType::Corners) calls<0>(u);
  }
  GhostFillingType getFillType() const { return fill_type_; }
  const Domain<D>& getDomain() const { return domain_; }

 protected:
  virtual void fillGhostCellsForNbrPatch(const PatchInfo<D>& pinfo, const PatchView<const double, D>& local_view,
                                         const PatchView<const double, D>& nbr_view, Side<D> side, NbrType nbr_type,
                                         Orthant<D - 1> orthant_on_coarse) const = 0;
  virtual void fillGhostCellsForEdgeNbrPatch(const PatchInfo<D>& pinfo, const PatchView<const double, D>& local_view,
                                             const PatchView<const double, D>& nbr_view, Edge edge, NbrType nbr_type,
                                             Orthant<1> orthant_on_coarse) const = 0;
  virtual void fillGhostCellsForCornerNbrPatch(const PatchInfo<D>& pinfo,
                                               const PatchView<const double, D>& local_view,
                                               const PatchView<const double, D>& nbr_view, Corner<D> corner,
                                               NbrType nbr_type) const = 0;
  virtual void fillGhostCellsForLocalPatch(const PatchInfo<D>& pinfo, const PatchView<const double, D>& view) const = 0;

 private:
  template <int M>
  void zero_faces(const PatchInfo<D>& pinfo, const PatchView<const double, D>& v) const
  {
    for (Face<D, M> f : Face<D, M>::getValues()) {
      if (!pinfo.hasNbr(f)) continue;
      std::array<size_t, D - M> off{};
      View<double, M + 1> g = v.getGhostSliceOn(f, off);
      Loop::OverInteriorIndexes<M + 1>(g, [&](const std::array<int, M + 1>& c) { g[c] = 0.0; });
    }
  }
  /* zeroGhostCells (:904-947) */
  void zero_ghosts(const PatchInfo<D>& pinfo, const PatchView<const double, D>& v) const
  {
    if (fill_type_ == GhostFillingType::Corners) zero_faces<0>(pinfo, v);
    if (D == 3 && fill_type_ != GhostFillingType::Faces) zero_faces<(D == 3 ? 1 : 0)>(pinfo, v);
    zero_faces<D - 1>(pinfo, v);
  }
  template <int M>
  void call(const PatchInfo<D>& pinfo, const PatchView<const double, D>& lv, const PatchView<const double, D>& nv,
            Face<D, M> f, NbrType t, int orth) const
  {
    if constexpr (M == D - 1) {
      fillGhostCellsForNbrPatch(pinfo, lv, nv, f, t, Orthant<D - 1>((unsigned char)orth));
    } else if constexpr (M == 0) {
      fillGhostCellsForCornerNbrPatch(pinfo, lv, nv, f, t);
    } else {
      fillGhostCellsForEdgeNbrPatch(pinfo, lv, nv, Edge((unsigned char)f.getIndex()), t,
                                    Orthant<1>((unsigned char)orth));
    }
  }
  /* the local calls of face dimension M (enumerateCalls, :821-895; processLocalFills,
   * :704-717) */
  template <int M>
  void calls(const Vector<D>& u) const
  {
    for (const PatchInfo<D>& pinfo : domain_.getPatchInfoVector()) {
      PatchView<const double, D> lv = u.getPatchView(pinfo.local_index);
      for (Face<D, M> f : Face<D, M>::getValues()) {
        if (!pinfo.hasNbr(f)) continue;
        switch (pinfo.getNbrType(f)) {
          case NbrType::Normal:
            call<M>(pinfo, lv, u.getPatchView(pinfo.getNormalNbrInfo(f).local_index), f, NbrType::Normal, 0);
            break;
          case NbrType::Coarse: {
            auto ni = pinfo.getCoarseNbrInfo(f);
            call<M>(pinfo, lv, u.getPatchView(ni.local_index), f, NbrType::Coarse,
                    (int)ni.orth_on_coarse.getIndex());
            break;
          }
          case NbrType::Fine: {
            auto ni = pinfo.getFineNbrInfo(f);
            for (size_t k = 0; k < Orthant<M>::num_orthants; k++)
              call<M>(pinfo, lv, u.getPatchView(ni.local_indexes[k]), f, NbrType::Fine, (int)k);
            break;
          }
        }
      }
    }
  }
  Domain<D> domain_;
  GhostFillingType fill_type_;
};

namespace GMG {
/* GMG/Smoother.h, GMG/Restrictor.h, GMG/Interpolator.h */
template <int D>
class Smoother {
 public:
  virtual ~Smoother() {}
  virtual Smoother<D>* clone() const = 0;
  virtual void smooth(const Vector<D>& f, Vector<D>& u) const = 0;
};
template <int D>
class Restrictor {
 public:
  virtual ~Restrictor() {}
  virtual Restrictor<D>* clone() const = 0;
  virtual Vector<D> restrict(const Vector<D>& fine) const = 0;
};
template <int D>
class Interpolator {
 public:
  virtual ~Interpolator() {}
  virtual Interpolator<D>* clone() const = 0;
  virtual void interpolate(const Vector<D>& coarse, Vector<D>& fine) const = 0;
};
}  // namespace GMG

namespace detail {
/* PatchSolver::smooth's "Total Patch Smooth" domain timing (PatchSolver.h:153-168): by
 * events on the null stream for the device launches, by the host clock for the per-patch
 * virtuals; stop() ends it */
template <int D>
struct SmoothTiming {
  SmoothTiming(const Domain<D>& d, bool host = false) : d_(d), host_(host)
  {
    if (!d_.hasTimer()) return;
    if (host_) d_.getTimer()->startDomainTiming(d_.getId(), "Total Patch Smooth");
    else d_.getTimer()->startOn("Total Patch Smooth", nullptr, d_.getId());
  }
  void stop()
  {
    if (!d_.hasTimer()) return;
    if (host_) d_.getTimer()->stopDomainTiming(d_.getId(), "Total Patch Smooth");
    else d_.getTimer()->stopOn("Total Patch Smooth", nullptr, d_.getId());
  }
  const Domain<D>& d_;
  bool host_;
};

/* one padded patch of one component in host memory, with the views of Vector::getPatchView */
template <int D>
struct PatchBuf {
  explicit PatchBuf(const std::array<int, D>& ns) : ns_(ns)
  {
    size_t n = 1;
    for (int a = 0; a < D; a++) n *= (size_t)(ns[(size_t)a] + 2);
    v.assign(n, 0.0);
  }
  PatchView<double, D> view()
  {
    std::array<int, D + 1> st, gs, s, e, ge;
    long stride = 1, first = 0;
    for (int a = 0; a < D; a++) {
      st[(size_t)a] = (int)stride;
      gs[(size_t)a] = -1;
      s[(size_t)a] = 0;
      e[(size_t)a] = ns_[(size_t)a] - 1;
      ge[(size_t)a] = ns_[(size_t)a];
      first += stride;
      stride *= ns_[(size_t)a] + 2;
    }
    st[(size_t)D] = (int)stride;
    gs[(size_t)D] = s[(size_t)D] = e[(size_t)D] = ge[(size_t)D] = 0;
    return PatchView<double, D>(View<double, D + 1>(v.data() + first, st, gs, s, e, ge));
  }
  std::array<int, D> ns_;
  std::vector<double> v;
};

/* the interior cells of a patch of n cells per axis, x fastest, with their running index */
template <int D, class F>
void each_cell(const std::array<int, D>& n, F fn)
{
  std::array<int, D> e;
  for (int a = 0; a < D; a++) e[(size_t)a] = n[(size_t)a] - 1;
  long i = 0;
  Loop::Nested<D>(std::array<int, D>{}, e, [&](const std::array<int, D>& c) { fn(c, i++); });
}
}  // namespace detail

/* PatchSolver.h:40-171.  solveSinglePatch takes host views of one patch; smooth() fills
 * u's ghosts and solves every patch through it.  The built-in solvers override smooth()
 * with their device kernels when their operator is a built-in one. */
template <int D>
class PatchSolver : public virtual Operator<D>, public virtual GMG::Smoother<D> {
 public:
  PatchSolver(const Domain<D>& domain, const GhostFiller<D>& filler) : domain_(domain), filler_(filler.clone()) {}
  PatchSolver* clone() const override = 0;
  /* u = the patch solution for f, u's ghosts filled (PatchSolver.h:98) */
  virtual void solveSinglePatch(const PatchInfo<D>& pinfo, const PatchView<const double, D>& f_view,
                                const PatchView<double, D>& u_view) const = 0;
  /* PatchSolver::apply (:107-134): u := 0, then smooth */
  void apply(const Vector<D>& f, Vector<D>& u) const override
  {
    u.setWithGhost(0.0);
    this->smooth(f, u);
  }
  /* PatchSolver::smooth (:142-170) */
  void smooth(const Vector<D>& f, Vector<D>& u) const override
  {
    detail::SmoothTiming<D> t(domain_, true);
    filler_->fillGhost(u);
    for (const PatchInfo<D>& pinfo : domain_.getPatchInfoVector())
      solveSinglePatch(pinfo, f.getPatchView(pinfo.local_index), u.getPatchView(pinfo.local_index));
    t.stop();
  }
  const Domain<D>& getDomain() const { return domain_; }
  const GhostFiller<D>& getGhostFiller() const { return *filler_; }

 private:
  Domain<D> domain_;
  std::shared_ptr<const GhostFiller<D>> filler_;
};

namespace detail {
/* DFTPatchSolver::getTransformArray (Poisson/DFTPatchSolver.h:131-196): row-major n x n,
 * y_i = sum_j M[i][j] x_j; type 0 DCT-II, 1 DCT-III, 2 DCT-IV, 3 DST-II, 4 DST-III, 5 DST-IV */
inline std::vector<double> dft_matrix(int type, int n)
{
  const double pi = 3.14159265358979323846;
  std::vector<double> m((size_t)n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double v;
      switch (type) {
        case 0: v = std::cos(pi / n * (i * (j + 0.5))); break;
        case 1: v = j == 0 ? 0.5 : std::cos(pi / n * ((i + 0.5) * j)); break;
        case 2: v = std::cos(pi / n * ((i + 0.5) * (j + 0.5))); break;
        case 3: v = std::sin(pi / n * ((i + 1) * (j + 0.5))); break;
        case 4: v = j == n - 1 ? ((i % 2 == 0) ? 0.5 : -0.5) : std::sin(pi / n * ((i + 0.5) * (j + 1))); break;
        default: v = std::sin(pi / n * ((i + 0.5) * (j + 0.5))); break;
      }
      m[(size_t)i * n + j] = v;
    }
  return m;
}

/* the transform along axis a of an interior block (x fastest): y = M x on every line */
template <int D>
void transform_axis(const std::array<int, D>& n, int a, const std::vector<double>& M, std::vector<double>& blk)
{
  long st = 1;
  for (int b = 0; b < a; b++) st *= n[(size_t)b];
  const int na = n[(size_t)a];
  std::vector<double> line((size_t)na);
  for (long base = 0; base < (long)blk.size(); base++) {
    if ((base / st) % na != 0) continue;
    for (int i = 0; i < na; i++) {
      double acc = 0;
      for (int j = 0; j < na; j++) acc += M[(size_t)i * na + j] * blk[(size_t)(base + j * st)];
      line[(size_t)i] = acc;
    }
    for (int i = 0; i < na; i++) blk[(size_t)(base + i * st)] = line[(size_t)i];
  }
}
}  // namespace detail

namespace Poisson {
/* Poisson/DFTPatchSolver.h:47-451: on the device for the built-in operators; for any other
 * PatchOperator the per-patch solve below (host views, the operator's own
 * modifyRHSForInternalBoundaryConditions) */
template <int D>
class DFTPatchSolver : public PatchSolver<D> {
 public:
  DFTPatchSolver(const PatchOperator<D>& op, std::bitset<2 * D> neumann)
      : PatchSolver<D>(op.getDomain(), op.getGhostFiller()), op_(op.clone()), neumann_(neumann)
  {
    int op_kind = -1;
    const double* coeffs = nullptr;  // copied by te_dft_plan_create
    if (typeid(*op_) == typeid(VarPoisson::StarPatchOperator<D>)) {
      op_kind = 1;
      coeffs = static_cast<const VarPoisson::StarPatchOperator<D>*>(op_.get())->getCoefficients().data();
    } else if (typeid(*op_) == typeid(Poisson::StarPatchOperator<D>)) {
      op_kind = 0;
    }
    if (op_kind >= 0 && op_->hasDeviceFiller()) {
      te_dft_plan* h = nullptr;
      detail::check(
        te_dft_plan_create(op_->deviceFiller().handle(), (unsigned)neumann.to_ulong(), op_kind, coeffs, &h));
      h_ = std::shared_ptr<te_dft_plan>(h, detail::Deleter<te_dft_plan, te_dft_plan_destroy>());
    }
  }
  DFTPatchSolver* clone() const override { return new DFTPatchSolver(*this); }
  void smooth(const Vector<D>& f, Vector<D>& u) const override
  {
    if (!h_) {
      PatchSolver<D>::smooth(f, u);
      return;
    }
    detail::SmoothTiming<D> t(this->getDomain());
    detail::check(te_dft_smooth(h_.get(), TE_F64, f.data(), u.data(), nullptr));
    t.stop();
  }
  /* solveSinglePatch (:420-450): f_copy = f with the interface terms, transforms along each
   * axis, division by the eigenvalues, inverse transforms, scaled by prod 2/n */
  void solveSinglePatch(const PatchInfo<D>& pinfo, const PatchView<const double, D>& f_view,
                        const PatchView<double, D>& u_view) const override
  {
    const auto& ns = this->getDomain().getNs();
    detail::PatchBuf<D> fc(ns);
    PatchView<double, D> fv = fc.view();
    ComponentView<double, D> fcv = fv.getComponentView(0);
    ComponentView<const double, D> fin = f_view.getComponentView(0);
    Loop::OverInteriorIndexes<D>(fcv, [&](const std::array<int, D>& c) { fcv[c] = fin[c]; });
    op_->modifyRHSForInternalBoundaryConditions(pinfo, u_view, fv);
    std::array<int, D> n;
    long N = 1;
    for (int a = 0; a < D; a++) N *= (n[(size_t)a] = ns[(size_t)a]);
    std::vector<double> blk((size_t)N);
    detail::each_cell<D>(n, [&](const std::array<int, D>& c, long i) { blk[(size_t)i] = fcv[c]; });
    std::array<int, D> tf, ti;
    std::array<double, D> eshift;  // the eigenvalue index shift: 0 both Neumann, 1/2 one, 1 none
    for (int a = 0; a < D; a++) {
      // getTransformsForPatch / getInverseTransformsForPatch (:272-313)
      const bool lo = !pinfo.hasNbr(Side<D>((unsigned char)(2 * a))) && neumann_[(size_t)(2 * a)];
      const bool hi = !pinfo.hasNbr(Side<D>((unsigned char)(2 * a + 1))) && neumann_[(size_t)(2 * a + 1)];
      tf[(size_t)a] = (lo && hi) ? 0 : lo ? 2 : hi ? 5 : 3;
      ti[(size_t)a] = (lo && hi) ? 1 : lo ? 2 : hi ? 5 : 4;
      eshift[(size_t)a] = (lo && hi) ? 0.0 : (lo || hi) ? 0.5 : 1.0;
    }
    for (int a = 0; a < D; a++) detail::transform_axis<D>(n, a, detail::dft_matrix(tf[(size_t)a], n[(size_t)a]), blk);
    // getEigenValues (:320-361): -sum_a 4/h_a^2 sin^2((k + shift) pi / (2 n_a))
    const double pi = 3.14159265358979323846;
    detail::each_cell<D>(n, [&](const std::array<int, D>& c, long i) {
      double eig = 0;
      for (int a = 0; a < D; a++) {
        const double h = pinfo.spacings[(size_t)a];
        eig -= 4 / (h * h) * std::pow(std::sin((c[(size_t)a] + eshift[(size_t)a]) * pi / (2 * n[(size_t)a])), 2);
      }
      blk[(size_t)i] /= eig;
    });
    if (neumann_.all() && !pinfo.hasNbr()) blk[0] = 0;  // (:438-440)
    for (int a = 0; a < D; a++) detail::transform_axis<D>(n, a, detail::dft_matrix(ti[(size_t)a], n[(size_t)a]), blk);
    double scale = 1;
    for (int a = 0; a < D; a++) scale *= 2.0 / n[(size_t)a];
    ComponentView<double, D> uc = u_view.getComponentView(0);
    detail::each_cell<D>(n, [&](const std::array<int, D>& c, long i) { uc[c] = blk[(size_t)i] * scale; });
  }
  const PatchOperator<D>& getOperator() const { return *op_; }
  std::bitset<2 * D> getNeumann() const { return neumann_; }
  /* the device plan, or NULL when the per-patch virtuals are used */
  te_dft_plan* device() const { return h_.get(); }

 private:
  std::shared_ptr<const PatchOperator<D>> op_;
  std::bitset<2 * D> neumann_;
  std::shared_ptr<te_dft_plan> h_;
};
}  // namespace Poisson

namespace Iterative {
/* Iterative/Solver.h:40-69 */
template <int D>
class Solver {
 public:
  virtual ~Solver() {}
  virtual Solver<D>* clone() const = 0;
  virtual int solve(const Operator<D>& A, Vector<D>& x, const Vector<D>& b, const Operator<D>* Mr = nullptr,
                    bool output = false, std::ostream& os = std::cout) const = 0;
};
}  // namespace Iterative

namespace GMG {
/* GMG/LinearRestrictor.h:101-191 */
template <int D>
class LinearRestrictor : public Restrictor<D> {
 public:
  LinearRestrictor(const Domain<D>& fine_domain, const Domain<D>& coarse_domain,
                   bool extrapolate_boundary_ghosts = false)
      : fine_(fine_domain), coarse_(coarse_domain), extrapolate_(extrapolate_boundary_ghosts)
  {
    te_level_link* h = nullptr;
    detail::check(te_level_link_create(fine_domain.handle(), coarse_domain.handle(), &h));
    h_ = std::shared_ptr<te_level_link>(h, detail::Deleter<te_level_link, te_level_link_destroy>());
  }
  LinearRestrictor* clone() const override { return new LinearRestrictor(*this); }
  Vector<D> restrict(const Vector<D>& fine) const override
  {
    Vector<D> coarse(coarse_, fine.getNumComponents());
    if (extrapolate_)
      detail::check(
        te_restrict_extrapolate(h_.get(), TE_F64, fine.data(), coarse.data(), fine.getNumComponents(), nullptr));
    else
      detail::check(te_restrict(h_.get(), TE_F64, fine.data(), coarse.data(), fine.getNumComponents(), nullptr));
    return coarse;
  }
  const Domain<D>& fineDomain() const { return fine_; }
  const Domain<D>& coarseDomain() const { return coarse_; }
  bool extrapolatesBoundaryGhosts() const { return extrapolate_; }

 private:
  Domain<D> fine_, coarse_;
  bool extrapolate_;
  std::shared_ptr<te_level_link> h_;
};

/* GMG/DirectInterpolator.h:36-91 */
template <int D>
class DirectInterpolator : public Interpolator<D> {
 public:
  DirectInterpolator(const Domain<D>& coarse_domain, const Domain<D>& fine_domain)
      : fine_(fine_domain), coarse_(coarse_domain)
  {
    te_level_link* h = nullptr;
    detail::check(te_level_link_create(fine_domain.handle(), coarse_domain.handle(), &h));
    h_ = std::shared_ptr<te_level_link>(h, detail::Deleter<te_level_link, te_level_link_destroy>());
  }
  DirectInterpolator* clone() const override { return new DirectInterpolator(*this); }
  void interpolate(const Vector<D>& coarse, Vector<D>& fine) const override
  {
    detail::check(
      te_interpolate(h_.get(), TE_F64, coarse.data(), fine.data(), fine.getNumComponents(), nullptr));
  }

 private:
  Domain<D> fine_, coarse_;
  std::shared_ptr<te_level_link> h_;
};
}  // namespace GMG

namespace Iterative {
/* Iterative/PatchSolver.h:53-175: an inner BiCGStab per patch (the solver must be an
 * Iterative::BiCGStab); on the device for the built-in operators, else on host views
 * through the operator's per-patch virtuals */
template <int D>
class BiCGStab;
template <int D>
class PatchSolver : public ThunderEgg::PatchSolver<D> {
 public:
  PatchSolver(const Solver<D>& solver, const PatchOperator<D>& op, bool continue_on_breakdown = false);
  PatchSolver* clone() const override { return new PatchSolver(*this); }
  void smooth(const Vector<D>& f, Vector<D>& u) const override
  {
    if (!device_) {
      ThunderEgg::PatchSolver<D>::smooth(f, u);
      return;
    }
    ThunderEgg::detail::SmoothTiming<D> t(this->getDomain());
    detail::check(te_patch_bicgstab_smooth(op_->deviceFiller().handle(), TE_F64, op_kind_, neumann_, coeffs_,
                                           max_it_, tol_, continue_ ? 1 : 0, f.data(), u.data(), nullptr,
                                           nullptr));
    t.stop();
  }
  /* solveSinglePatch (Iterative/PatchSolver.h:134-175): f' = f with the interface terms, then
   * BiCGStab (Iterative/BiCGStab.h:111-203) on the patch's operator with internal boundary
   * conditions, started from u */
  void solveSinglePatch(const PatchInfo<D>& pinfo, const PatchView<const double, D>& f_view,
                        const PatchView<double, D>& u_view) const override;
  int maxIterations() const { return max_it_; }
  double tolerance() const { return tol_; }
  bool continueOnBreakdown() const { return continue_; }
  /* the per-patch solves run on the device (a built-in operator) */
  bool onDevice() const { return device_; }

 private:
  std::shared_ptr<const PatchOperator<D>> op_;
  bool device_ = false;
  int op_kind_ = 0, neumann_ = 0;
  const double* coeffs_ = nullptr;
  int max_it_ = 1000;
  double tol_ = 1e-12;
  bool continue_;
};

/* Iterative/CG.h:40-194 */
template <int D>
class CG : public Solver<D> {
 public:
  void setMaxIterations(int n) { max_iterations_ = n; }
  int getMaxIterations() const { return max_iterations_; }
  void setTolerance(double t) { tolerance_ = t; }
  double getTolerance() const { return tolerance_; }
  CG* clone() const override { return new CG(*this); }
  /* "Iteration" timings (Iterative/CG.h) */
  void setTimer(std::shared_ptr<Timer> timer) { timer_ = timer; }
  std::shared_ptr<Timer> getTimer() const { return timer_; }
  int solve(const Operator<D>& A, Vector<D>& x, const Vector<D>& b, const Operator<D>* Mr = nullptr,
            bool output = false, std::ostream& os = std::cout) const override;

 private:
  int max_iterations_ = 1000;
  double tolerance_ = 1e-12;
  std::shared_ptr<Timer> timer_;
};

/* Iterative/BiCGStab.h:40-203 */
template <int D>
class BiCGStab : public Solver<D> {
 public:
  void setMaxIterations(int n) { max_iterations_ = n; }
  int getMaxIterations() const { return max_iterations_; }
  void setTolerance(double t) { tolerance_ = t; }
  double getTolerance() const { return tolerance_; }
  BiCGStab* clone() const override { return new BiCGStab(*this); }
  /* "Iteration" timings (Iterative/BiCGStab.h) */
  void setTimer(std::shared_ptr<Timer> timer) { timer_ = timer; }
  std::shared_ptr<Timer> getTimer() const { return timer_; }
  int solve(const Operator<D>& A, Vector<D>& x, const Vector<D>& b, const Operator<D>* Mr = nullptr,
            bool output = false, std::ostream& os = std::cout) const override;

 private:
  int max_iterations_ = 1000;
  double tolerance_ = 1e-12;
  std::shared_ptr<Timer> timer_;
};

template <int D>
PatchSolver<D>::PatchSolver(const Solver<D>& solver, const PatchOperator<D>& op, bool continue_on_breakdown)
    : ThunderEgg::PatchSolver<D>(op.getDomain(), op.getGhostFiller()), op_(op.clone()), continue_(continue_on_breakdown)
{
  auto bs = dynamic_cast<const BiCGStab<D>*>(&solver);
  if (!bs) throw RuntimeError("Iterative::PatchSolver: the inner solver must be Iterative::BiCGStab");
  max_it_ = bs->getMaxIterations();
  tol_ = bs->getTolerance();
  // the coefficients come from the owned clone: the caller's operator may go away first
  if (typeid(*op_) == typeid(VarPoisson::StarPatchOperator<D>)) {
    op_kind_ = 1;
    coeffs_ = static_cast<const VarPoisson::StarPatchOperator<D>*>(op_.get())->getCoefficients().data();
    device_ = op_->hasDeviceFiller();
  } else if (typeid(*op_) == typeid(Poisson::StarPatchOperator<D>)) {
    neumann_ = static_cast<const Poisson::StarPatchOperator<D>*>(op_.get())->neumann() ? 1 : 0;
    device_ = op_->hasDeviceFiller();
  }
}

template <int D>
void PatchSolver<D>::solveSinglePatch(const PatchInfo<D>& pinfo, const PatchView<const double, D>& f_view,
                                      const PatchView<double, D>& u_view) const
{
  const auto& ns = this->getDomain().getNs();
  std::array<int, D> n;
  for (int a = 0; a < D; a++) n[(size_t)a] = ns[(size_t)a];
  using Buf = ThunderEgg::detail::PatchBuf<D>;
  Buf fb(ns), xb(ns), resid(ns), rhat(ns), p(ns), ap(ns), sv(ns), as(ns);
  auto cv = [](Buf& b) { return b.view().getComponentView(0); };
  // f_copy = f, modifyRHSForInternalBoundaryConditions (:160-163); x = u
  {
    ComponentView<double, D> f = cv(fb), x = cv(xb);
    ComponentView<const double, D> fi = f_view.getComponentView(0);
    ComponentView<double, D> ui = u_view.getComponentView(0);
    ThunderEgg::detail::each_cell<D>(n, [&](const std::array<int, D>& c, long) {
      f[c] = fi[c];
      x[c] = ui[c];
    });
  }
  op_->modifyRHSForInternalBoundaryConditions(pinfo, u_view, fb.view());
  auto A = [&](Buf& in, Buf& out) {  // SinglePatchOp::apply (:76-81)
    std::fill(out.v.begin(), out.v.end(), 0.0);
    op_->applySinglePatchWithInternalBoundaryConditions(pinfo, in.view(), out.view());
  };
  auto dot = [&](Buf& a, Buf& b) {
    double r = 0;
    ComponentView<double, D> av = cv(a), bv = cv(b);
    ThunderEgg::detail::each_cell<D>(n, [&](const std::array<int, D>& c, long) { r += av[c] * bv[c]; });
    return r;
  };
  auto axpby = [&](Buf& y, double alpha, Buf& x, double beta, Buf* z, double gamma) {  // y = alpha y + beta x + gamma z
    ComponentView<double, D> yv = cv(y), xv = cv(x);
    ThunderEgg::detail::each_cell<D>(n, [&](const std::array<int, D>& c, long) {
      double v = yv[c] * alpha + xv[c] * beta;
      if (z) v += cv(*z)[c] * gamma;
      yv[c] = v;
    });
  };
  // Iterative::BiCGStab::solve (Iterative/BiCGStab.h:111-203), unpreconditioned
  A(xb, resid);
  axpby(resid, -1.0, fb, 1.0, nullptr, 0.0);
  const double r0 = std::sqrt(dot(fb, fb));
  rhat.v = resid.v;
  p.v = resid.v;
  double rho = dot(rhat, resid);
  int its = 0;
  double residual = std::sqrt(dot(resid, resid)) / r0;
  if (r0 != 0) {
    while (residual > tol_ && its < max_it_) {
      if (rho == 0) {
        if (continue_) break;
        throw BreakdownError("BiCGStab broke down, rho was 0 on iteration " + std::to_string(its));
      }
      A(p, ap);
      const double alpha = rho / dot(rhat, ap);
      sv.v = resid.v;
      axpby(sv, 1.0, ap, -alpha, nullptr, 0.0);
      if (std::sqrt(dot(sv, sv)) / r0 <= tol_) {
        axpby(xb, 1.0, p, alpha, nullptr, 0.0);
        break;
      }
      A(sv, as);
      const double omega = dot(as, sv) / dot(as, as);
      axpby(xb, 1.0, p, alpha, &sv, omega);
      axpby(resid, 1.0, ap, -alpha, &as, -omega);
      const double rho_new = dot(resid, rhat);
      const double beta = rho_new * alpha / (rho * omega);
      axpby(p, 1.0, ap, -omega, nullptr, 0.0);
      axpby(p, beta, resid, 1.0, nullptr, 0.0);
      its++;
      rho = rho_new;
      residual = std::sqrt(dot(resid, resid)) / r0;
    }
  }
  ComponentView<double, D> ui = u_view.getComponentView(0), x = cv(xb);
  ThunderEgg::detail::each_cell<D>(n, [&](const std::array<int, D>& c, long) { ui[c] = x[c]; });
}
}  // namespace Iterative

namespace GMG {
/* GMG/CycleOpts.h:33-55, plus `precision` ("double" or "mixed") */
struct CycleOpts {
  int pre_sweeps = 1;
  int post_sweeps = 1;
  int mid_sweeps = 1;
  int coarse_sweeps = 1;
  std::string cycle_type = "V";
  std::string precision = "double";
};

template <int D>
struct Level {
  std::shared_ptr<const Operator<D>> op;
  std::shared_ptr<const Smoother<D>> smoother;
  std::shared_ptr<const Restrictor<D>> restrictor;      // to the next coarser level
  std::shared_ptr<const Interpolator<D>> interpolator;  // from the next coarser level
};

/* GMG/Cycle.h:40-98 + GMG/VCycle.h:46-70, GMG/WCycle.h:50-84, GMG/FMGCycle.h:43-104
 * (CycleOpts::cycle_type "V", "W", "F", as CycleBuilder.h:182-189 dispatches) */
template <int D>
class Cycle : public Operator<D> {
 public:
  Cycle(const CycleOpts& opts, std::vector<Level<D>> levels) : opts_(opts), levels_(std::move(levels))
  {
    build_device_cycle();
  }
  Cycle* clone() const override { return new Cycle(*this); }
  /* Cycle::apply (GMG/Cycle.h:94-98): u := cycle(f) from u = 0 */
  void apply(const Vector<D>& f, Vector<D>& u) const override
  {
    if (g_) {
      attachTimers(nullptr);
      int rc = te_gmg_vcycle(g_.get(), f.data(), u.data(), nullptr);
      detachTimers();
      detail::check(rc);
      return;
    }
    u.setWithGhost(0.0);
    visit(0, f, u, false);
  }
  /* the fused device cycle, or NULL when a component is user-defined */
  te_gmg* device() const { return g_.get(); }
  /* Attach to the device cycle the timer of a solver (its "Iteration" timings; may be
   * null) and, for the levels whose domain carries the same timer (or, without a solver
   * timer, the first timer found on a level), the levels' sweep timings; detachTimers
   * undoes it. */
  void attachTimers(const std::shared_ptr<Timer>& solver_timer) const
  {
    std::shared_ptr<Timer> t = solver_timer;
    for (size_t i = 0; !t && i < domains_.size(); i++) t = domains_[i].getTimer();
    if (!t || !g_) return;
    std::vector<int> ids;
    for (const Domain<D>& d : domains_) ids.push_back(d.getTimer() == t ? d.getId() : TE_TIMER_NONE);
    detail::check(te_gmg_set_timer(g_.get(), t->handle(), solver_timer == t ? 1 : 0, ids.data()));
  }
  void detachTimers() const
  {
    if (g_) te_gmg_set_timer(g_.get(), nullptr, 0, nullptr);
  }
  const Operator<D>& finestOperator() const { return *levels_[0].op; }
  int numLevels() const { return (int)levels_.size(); }

 private:
  /* vonly: FMGCycle::v_visit (one coarse visit).  The W- and F-cycles' second coarse
   * visit starts from the first one's coarse u, as in the reference. */
  void visit(size_t i, const Vector<D>& f, Vector<D>& u, bool vonly) const
  {
    const Level<D>& L = levels_[i];
    if (i + 1 == levels_.size()) {
      for (int k = 0; k < opts_.coarse_sweeps; k++) L.smoother->smooth(f, u);
      return;
    }
    for (int k = 0; k < opts_.pre_sweeps; k++) L.smoother->smooth(f, u);
    auto restrict_residual = [&]() {  // Cycle::restrict: r = f - A u, restricted
      Vector<D> r = f.getZeroClone();
      L.op->apply(u, r);
      r.scaleThenAdd(-1.0, f);
      return L.restrictor->restrict(r);
    };
    Vector<D> fc = restrict_residual();
    Vector<D> uc = fc.getZeroClone();
    visit(i + 1, fc, uc, false);
    levels_[i + 1].interpolator->interpolate(uc, u);
    if (opts_.cycle_type != "V" && !vonly) {
      for (int k = 0; k < opts_.mid_sweeps; k++) L.smoother->smooth(f, u);
      fc = restrict_residual();
      visit(i + 1, fc, uc, opts_.cycle_type == "F");
      levels_[i + 1].interpolator->interpolate(uc, u);
    }
    for (int k = 0; k < opts_.post_sweeps; k++) L.smoother->smooth(f, u);
  }
  void build_device_cycle();

  CycleOpts opts_;
  std::vector<Level<D>> levels_;
  std::shared_ptr<te_gmg> g_;
  std::vector<Vector<D>> coeffs_;  // keeps VarPoisson coefficients alive
  std::vector<Domain<D>> domains_;  // the device cycle's levels (for their timers)
};

template <int D>
void Cycle<D>::build_device_cycle()
{
  if (opts_.cycle_type != "V" && opts_.cycle_type != "W" && opts_.cycle_type != "F")
    throw RuntimeError("Unsupported Cycle type: " + opts_.cycle_type);
  if (opts_.precision != "double" && opts_.precision != "mixed")
    throw RuntimeError("CycleOpts::precision must be \"double\" or \"mixed\"");
  te_gmg_options o{};
  std::vector<te_domain*> doms;
  std::vector<const double*> coeffs;
  int op_kind = -1, filler_kind = -1, fill_type = -1, neumann = 0, cont = -1;
  unsigned mask = 0;
  bool any_dft = false;
  // one smoother per level (CycleBuilder takes one per level): kind and inner settings
  std::vector<int> lsm, lmax;
  std::vector<double> ltol;
  for (size_t i = 0; i < levels_.size(); i++) {
    const Level<D>& L = levels_[i];
    // only the built-in classes themselves (not subclasses, which may override the per-patch
    // virtuals) run as the device cycle
    const PatchOperator<D>* po = dynamic_cast<const PatchOperator<D>*>(L.op.get());
    if (!po || !po->hasDeviceFiller()) return;
    int ok = -1, nm = 0;
    const double* c = nullptr;
    if (typeid(*po) == typeid(VarPoisson::StarPatchOperator<D>)) {
      ok = 1;
      c = static_cast<const VarPoisson::StarPatchOperator<D>*>(po)->getCoefficients().data();
    } else if (typeid(*po) == typeid(Poisson::StarPatchOperator<D>)) {
      ok = 0;
      nm = static_cast<const Poisson::StarPatchOperator<D>*>(po)->neumann() ? 1 : 0;
    } else {
      return;
    }
    int sk = -1, im = 1000, ct = -1;
    double it = 1e-12;
    const GMG::Smoother<D>& sref = *L.smoother;
    if (typeid(sref) == typeid(Poisson::DFTPatchSolver<D>)) {
      auto s = dynamic_cast<const Poisson::DFTPatchSolver<D>*>(L.smoother.get());
      if (!s->device()) return;
      sk = 0;
      const unsigned mk = (unsigned)s->getNeumann().to_ulong();
      if (any_dft && mk != mask) return;  // one Neumann mask for every DFT level
      mask = mk;
      any_dft = true;
    } else if (typeid(sref) == typeid(Iterative::PatchSolver<D>)) {
      auto s = dynamic_cast<const Iterative::PatchSolver<D>*>(L.smoother.get());
      if (!s->onDevice()) return;
      sk = 1;
      im = s->maxIterations();
      it = s->tolerance();
      ct = s->continueOnBreakdown() ? 1 : 0;
    } else {
      return;
    }
    if (ct >= 0) {
      if (cont >= 0 && ct != cont) return;  // one breakdown policy for every level
      cont = ct;
    }
    if (i + 1 < levels_.size() && typeid(*L.restrictor) != typeid(LinearRestrictor<D>)) return;
    // the device cycle's restriction leaves the coarse ghosts at zero: a restrictor that
    // extrapolates them runs in the generic cycle
    if (i + 1 < levels_.size() &&
        static_cast<const LinearRestrictor<D>&>(*L.restrictor).extrapolatesBoundaryGhosts())
      return;
    if (i > 0 && typeid(*L.interpolator) != typeid(DirectInterpolator<D>)) return;
    const DeviceGhostFiller<D>& gf = po->deviceFiller();
    if (i == 0) {
      op_kind = ok, filler_kind = gf.kind(), fill_type = (int)gf.getFillType(), neumann = nm;
    } else if (ok != op_kind || gf.kind() != filler_kind || (int)gf.getFillType() != fill_type || nm != neumann) {
      return;  // levels with different operators or fillers: generic cycle
    }
    lsm.push_back(sk);
    lmax.push_back(im);
    ltol.push_back(it);
    doms.push_back(po->getDomain().handle());
    coeffs.push_back(c);
  }
  for (const Level<D>& L : levels_) domains_.push_back(dynamic_cast<const PatchOperator<D>*>(L.op.get())->getDomain());
  o.filler_kind = filler_kind;
  o.fill_type = fill_type;
  o.op_kind = op_kind;
  o.neumann = neumann;
  o.smoother = lsm[0];
  o.neumann_mask = mask;
  o.inner_max_it = lmax[0];
  o.inner_tol = ltol[0];
  o.continue_on_breakdown = cont > 0 ? 1 : 0;
  o.level_smoother = lsm.data();
  o.level_inner_max_it = lmax.data();
  o.level_inner_tol = ltol.data();
  o.pre_sweeps = opts_.pre_sweeps;
  o.post_sweeps = opts_.post_sweeps;
  o.coarse_sweeps = opts_.coarse_sweeps;
  o.mid_sweeps = opts_.mid_sweeps;
  o.cycle_type = opts_.cycle_type == "W" ? TE_CYCLE_W : (opts_.cycle_type == "F" ? TE_CYCLE_F : TE_CYCLE_V);
  o.precision = opts_.precision == "mixed" ? TE_F32 : TE_F64;
  o.use_graph = 1;
  te_gmg* h = nullptr;
  detail::check(te_gmg_create(doms.data(), (int)doms.size(), &o, op_kind == 1 ? coeffs.data() : nullptr, &h));
  g_ = std::shared_ptr<te_gmg>(h, detail::Deleter<te_gmg, te_gmg_destroy>());
}

/* GMG/CycleBuilder.h:40-192 */
template <int D>
class CycleBuilder {
 public:
  explicit CycleBuilder(const CycleOpts& opts) : opts_(opts) {}
  void addFinestLevel(const Operator<D>& op, const Smoother<D>& smoother, const Restrictor<D>& restrictor)
  {
    if (!levels_.empty()) throw RuntimeError("addFinestLevel was already called");
    levels_.push_back({ std::shared_ptr<const Operator<D>>(op.clone()),
                        std::shared_ptr<const Smoother<D>>(smoother.clone()),
                        std::shared_ptr<const Restrictor<D>>(restrictor.clone()), nullptr });
  }
  void addIntermediateLevel(const Operator<D>& op, const Smoother<D>& smoother, const Restrictor<D>& restrictor,
                            const Interpolator<D>& interpolator)
  {
    if (levels_.empty()) throw RuntimeError("addFinestLevel has not been called yet");
    if (coarsest_) throw RuntimeError("addCoarsestLevel has already been called");
    levels_.push_back({ std::shared_ptr<const Operator<D>>(op.clone()),
                        std::shared_ptr<const Smoother<D>>(smoother.clone()),
                        std::shared_ptr<const Restrictor<D>>(restrictor.clone()),
                        std::shared_ptr<const Interpolator<D>>(interpolator.clone()) });
  }
  void addCoarsestLevel(const Operator<D>& op, const Smoother<D>& smoother, const Interpolator<D>& interpolator)
  {
    if (levels_.empty()) throw RuntimeError("addFinestLevel has not been called yet");
    if (coarsest_) throw RuntimeError("addCoarsestLevel has already been called");
    coarsest_ = true;
    levels_.push_back({ std::shared_ptr<const Operator<D>>(op.clone()),
                        std::shared_ptr<const Smoother<D>>(smoother.clone()), nullptr,
                        std::shared_ptr<const Interpolator<D>>(interpolator.clone()) });
  }
  std::shared_ptr<Cycle<D>> getCycle() const
  {
    if (!coarsest_) throw RuntimeError("Coarsest level was not added");
    return std::make_shared<Cycle<D>>(opts_, levels_);
  }

 private:
  CycleOpts opts_;
  std::vector<Level<D>> levels_;
  bool coarsest_ = false;
};
}  // namespace GMG

namespace Iterative {
namespace detail_bicg {
inline void log_line(std::ostream& os, int it, double res)
{
  char buf[100];
  std::snprintf(buf, sizeof(buf), "%5d %16.8e\n", it, res);
  os << buf;
}
}  // namespace detail_bicg

/* A and B are the same built-in operator (kind, domain, boundary flag, coefficients) */
template <int D>
bool same_operator(const Operator<D>& A, const Operator<D>& B)
{
  if (typeid(A) != typeid(B)) return false;
  if (auto a = dynamic_cast<const Poisson::StarPatchOperator<D>*>(&A)) {
    auto b = dynamic_cast<const Poisson::StarPatchOperator<D>*>(&B);
    return a->getDomain() == b->getDomain() && a->neumann() == b->neumann() &&
           a->deviceFiller().kind() == b->deviceFiller().kind() &&
           a->deviceFiller().getFillType() == b->deviceFiller().getFillType();
  }
  if (auto a = dynamic_cast<const VarPoisson::StarPatchOperator<D>*>(&A)) {
    auto b = dynamic_cast<const VarPoisson::StarPatchOperator<D>*>(&B);
    return a->getDomain() == b->getDomain() && a->coefficientIdentity() == b->coefficientIdentity() &&
           a->deviceFiller().kind() == b->deviceFiller().kind() &&
           a->deviceFiller().getFillType() == b->deviceFiller().getFillType();
  }
  return false;
}

template <int D>
int BiCGStab<D>::solve(const Operator<D>& A, Vector<D>& x, const Vector<D>& b, const Operator<D>* Mr, bool output,
                       std::ostream& os) const
{
  // Fused device solve when Mr is a device cycle whose finest operator is A.
  auto cyc = dynamic_cast<const GMG::Cycle<D>*>(Mr);
  if (cyc && cyc->device()) {
    if (same_operator(A, cyc->finestOperator()) && b.getNumComponents() == 1) {
      std::vector<double> hist((size_t)max_iterations_ + 2, 0.0);
      int its = 0;
      cyc->attachTimers(timer_);
      int rc = te_gmg_bicgstab(cyc->device(), b.data(), x.data(), 1, max_iterations_, tolerance_, &its, hist.data(),
                               (int)hist.size(), nullptr);
      cyc->detachTimers();
      detail::check(rc);
      if (output)
        for (int k = 0; k <= its && k < (int)hist.size(); k++) detail_bicg::log_line(os, k, hist[(size_t)k]);
      return its;
    }
  }
  // Generic right-preconditioned BiCGStab over the virtual interfaces (same recurrences).
  Vector<D> resid = b.getZeroClone();
  A.apply(x, resid);
  resid.scaleThenAdd(-1.0, b);
  const double r0_norm = b.twoNorm();
  Vector<D> rhat(resid), p(resid), ap = b.getZeroClone(), as = b.getZeroClone(), s = b.getZeroClone();
  Vector<D> ms, mp;
  if (Mr) {
    ms = b.getZeroClone();
    mp = b.getZeroClone();
  }
  double rho = rhat.dot(resid);
  int num_its = 0;
  double residual = resid.twoNorm() / r0_norm;
  if (output) detail_bicg::log_line(os, num_its, residual);
  if (r0_norm == 0) return num_its;
  while (residual > tolerance_ && num_its < max_iterations_) {
    if (timer_) timer_->start("Iteration");
    if (rho == 0) throw BreakdownError("BiCGStab broke down, rho was 0 on iteration " + std::to_string(num_its));
    if (Mr) {
      Mr->apply(p, mp);
      A.apply(mp, ap);
    } else {
      A.apply(p, ap);
    }
    const double alpha = rho / rhat.dot(ap);
    s.copy(resid);
    s.addScaled(-alpha, ap);
    if (s.twoNorm() / r0_norm <= tolerance_) {
      x.addScaled(alpha, p);  // the reference's early exit (Iterative/BiCGStab.h:162-167)
      if (timer_) timer_->stop("Iteration");
      break;
    }
    if (Mr) {
      Mr->apply(s, ms);
      A.apply(ms, as);
    } else {
      A.apply(s, as);
    }
    const double omega = as.dot(s) / as.dot(as);
    if (Mr) x.addScaled(alpha, mp, omega, ms);
    else x.addScaled(alpha, p, omega, s);
    resid.addScaled(-alpha, ap, -omega, as);
    const double rho_new = resid.dot(rhat);
    const double beta = rho_new * alpha / (rho * omega);
    p.addScaled(-omega, ap);
    p.scaleThenAdd(beta, resid);
    num_its++;
    rho = rho_new;
    residual = resid.twoNorm() / r0_norm;
    if (output) detail_bicg::log_line(os, num_its, residual);
    if (timer_) timer_->stop("Iteration");
  }
  return num_its;
}

template <int D>
int CG<D>::solve(const Operator<D>& A, Vector<D>& x, const Vector<D>& b, const Operator<D>* Mr, bool output,
                 std::ostream& os) const
{
  // Device solve when Mr is a device cycle whose finest operator is A.
  auto cyc = dynamic_cast<const GMG::Cycle<D>*>(Mr);
  if (cyc && cyc->device() && same_operator(A, cyc->finestOperator()) && b.getNumComponents() == 1) {
    std::vector<double> hist((size_t)max_iterations_ + 2, 0.0);
    int its = 0;
    cyc->attachTimers(timer_);
    int rc = te_gmg_cg(cyc->device(), b.data(), x.data(), 1, max_iterations_, tolerance_, &its, hist.data(),
                       (int)hist.size(), nullptr);
    cyc->detachTimers();
    detail::check(rc);
    if (output)
      for (int k = 0; k <= its && k < (int)hist.size(); k++) detail_bicg::log_line(os, k, hist[(size_t)k]);
    return its;
  }
  // Generic right-preconditioned CG over the virtual interfaces (Iterative/CG.h:126-194).
  Vector<D> resid = b.getZeroClone();
  A.apply(x, resid);
  resid.scaleThenAdd(-1.0, b);
  Vector<D> initial_guess(x);
  x.set(0.0);
  const double r0_norm = b.twoNorm();
  Vector<D> p(resid), ap = b.getZeroClone(), tmp = b.getZeroClone();
  double rho = resid.dot(resid);
  int num_its = 0;
  if (r0_norm == 0) return num_its;
  double residual = resid.twoNorm() / r0_norm;
  if (output) detail_bicg::log_line(os, num_its, residual);
  while (residual > tolerance_ && num_its < max_iterations_) {
    if (timer_) timer_->start("Iteration");
    if (rho == 0) throw BreakdownError("CG broke down, rho was 0 on iteration " + std::to_string(num_its));
    if (Mr) {
      Mr->apply(p, tmp);
      A.apply(tmp, ap);
    } else {
      A.apply(p, ap);
    }
    const double alpha = rho / p.dot(ap);
    x.addScaled(alpha, p);
    resid.addScaled(-alpha, ap);
    const double rho_new = resid.dot(resid);
    const double beta = rho_new / rho;
    p.scaleThenAdd(beta, resid);
    num_its++;
    rho = rho_new;
    residual = resid.twoNorm() / r0_norm;
    if (output) detail_bicg::log_line(os, num_its, residual);
    if (timer_) timer_->stop("Iteration");
  }
  if (Mr) {
    Mr->apply(x, resid);
    x.copy(resid);
  }
  x.add(initial_guess);
  return num_its;
}
}  // namespace Iterative

/* Forest-based domain generation in place of P4estDomainGenerator / P8estDomainGenerator
 * (same getters: getFinestDomain, hasCoarserDomain, getCoarserDomain). */
template <int D>
class ForestDomainGenerator {
 public:
  /* uniform forest at base_level, refined where refine(level, coords) says so (one pass per
   * call of refineWhere), 2:1 balanced */
  ForestDomainGenerator(int base_level, std::array<int, D> ns, int num_ghost_cells = 1)
      : ns_(ns), num_ghost_(num_ghost_cells)
  {
    te_forest* f = nullptr;
    detail::check(te_forest_create(D, base_level, &f));
    f_ = std::shared_ptr<te_forest>(f, detail::Deleter<te_forest, te_forest_destroy>());
  }
  template <class Pred>
  void refineWhere(Pred refine)
  {
    int n = te_forest_num_leaves(f_.get());
    std::vector<int32_t> lv((size_t)n), co((size_t)n * D);
    detail::check(te_forest_get_leaves(f_.get(), lv.data(), co.data()));
    std::vector<uint8_t> flags((size_t)n);
    for (int i = 0; i < n; i++) {
      std::array<int, D> c;
      for (int a = 0; a < D; a++) c[a] = co[(size_t)i * D + a];
      flags[(size_t)i] = refine(lv[(size_t)i], c) ? 1 : 0;
    }
    detail::check(te_forest_refine(f_.get(), flags.data()));
    detail::check(te_forest_balance(f_.get()));
    built_ = false;
  }
  Domain<D> getFinestDomain() { return domain(0); }
  bool hasCoarserDomain()
  {
    build();
    return next_ < num_levels_;
  }
  Domain<D> getCoarserDomain()
  {
    build();
    if (next_ >= num_levels_) throw RuntimeError("no coarser domain");
    return domain(next_++);
  }

 private:
  void build()
  {
    if (built_) return;
    std::array<int, 3> n3{ 1, 1, 1 };
    for (int a = 0; a < D; a++) n3[a] = ns_[a];
    detail::check(te_forest_build_domains(f_.get(), n3.data(), &num_levels_));
    built_ = true;
    next_ = 1;
  }
  Domain<D> domain(int k)
  {
    build();
    int n = 0;
    detail::check(te_forest_domain_size(f_.get(), k, &n));
    std::vector<te_patch_info> p((size_t)n);
    detail::check(te_forest_domain_patches(f_.get(), k, p.data()));
    return Domain<D>(Communicator(), num_levels_ - 1 - k, ns_, num_ghost_, p.begin(), p.end());
  }
  std::array<int, D> ns_;
  int num_ghost_;
  std::shared_ptr<te_forest> f_;
  bool built_ = false;
  int num_levels_ = 0, next_ = 1;
};

}  // namespace ThunderEgg

#endif
