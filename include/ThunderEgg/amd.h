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
#include <cstdio>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <typeinfo>
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
    const auto& pv = domain.getPatchInfoVector();
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

/* Side<D> = Face<D, D - 1> (Face.h:41-584): 0 west, 1 east, 2 south, 3 north, 4 bottom,
 * 5 top; axis = value / 2, lower side = even value. */
template <int D>
class Side {
 public:
  constexpr explicit Side(int value = 0) : v_(value) {}
  static constexpr Side west() { return Side(0); }
  static constexpr Side east() { return Side(1); }
  static constexpr Side south() { return Side(2); }
  static constexpr Side north() { return Side(3); }
  static constexpr Side bottom() { return Side(4); }
  static constexpr Side top() { return Side(5); }
  size_t getIndex() const { return (size_t)v_; }
  size_t getAxisIndex() const { return (size_t)(v_ >> 1); }
  bool isLowerOnAxis() const { return (v_ & 1) == 0; }
  bool isHigherOnAxis() const { return (v_ & 1) == 1; }
  bool operator==(const Side& o) const { return v_ == o.v_; }

 private:
  int v_;
};

/* The PatchInfo<D> fields that locate a patch's cells (PatchInfo.h:50-140); the default
 * is one cell of unit size at the origin, as the reference's default constructor. */
template <int D>
struct PatchInfo {
  int id = 0;
  int local_index = 0;
  int refine_level = -1;
  int num_ghost_cells = 1;
  std::array<int, D> ns;
  std::array<double, D> starts;
  std::array<double, D> spacings;
  PatchInfo()
  {
    ns.fill(1);
    starts.fill(0.0);
    spacings.fill(1.0);
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
  const std::vector<te_patch_info>& getPatchInfoVector() const { return patches_; }
  /* the cell-locating fields of one patch, with the domain's ns and ghost width */
  PatchInfo<D> getPatchInfo(int local_index) const
  {
    const te_patch_info& r = patches_.at((size_t)local_index);
    PatchInfo<D> pi;
    pi.id = r.id;
    pi.local_index = local_index;
    pi.refine_level = r.refine_level;
    pi.num_ghost_cells = num_ghost_;
    for (int a = 0; a < D; a++) {
      pi.ns[a] = ns_[a];
      pi.starts[a] = r.starts[a];
      pi.spacings[a] = r.spacings[a];
    }
    return pi;
  }
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
  /* device pointer to the padded data */
  double* data() const { return data_.get(); }
  long size() const { return len_; }

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
  Domain<D> domain_;
  int ncomp_ = 0;
  long len_ = 0;
  std::shared_ptr<double> data_;
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

/* PatchOperator.h:45-143 */
template <int D>
class PatchOperator : public Operator<D> {
 public:
  PatchOperator(const Domain<D>& domain, const GhostFiller<D>& ghost_filler)
      : domain_(domain), filler_(ghost_filler.clone())
  {
    dev_ = dynamic_cast<const DeviceGhostFiller<D>*>(filler_.get());
    if (!dev_) throw RuntimeError("PatchOperator: the ghost filler must be one of the built-in fillers");
  }
  PatchOperator<D>* clone() const override = 0;
  const Domain<D>& getDomain() const { return domain_; }
  const GhostFiller<D>& getGhostFiller() const { return *filler_; }
  const DeviceGhostFiller<D>& deviceFiller() const { return *dev_; }

 private:
  Domain<D> domain_;
  std::shared_ptr<const GhostFiller<D>> filler_;
  const DeviceGhostFiller<D>* dev_ = nullptr;
};

namespace Poisson {
/* Poisson/StarPatchOperator.h:39-204: f = sum_a (u+ - 2u + u-)/h_a^2 */
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
    detail::check(te_star_apply(this->deviceFiller().handle(), TE_F64, neumann_, x.data(), b.data(),
                                x.getNumComponents(), nullptr));
  }
  bool neumann() const { return neumann_; }

 private:
  bool neumann_;
};
}  // namespace Poisson

namespace VarPoisson {
/* VarPoisson/StarPatchOperator.h:40-186; the coefficients' ghosts are filled once here (:70) */
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
    detail::check(te_varstar_apply(this->deviceFiller().handle(), TE_F64, coeffs_.data(), x.data(), b.data(),
                                   x.getNumComponents(), nullptr));
  }
  const Vector<D>& getCoefficients() const { return coeffs_; }
  /* shared by this operator and its clones (each clone deep-copies the coefficients, so
   * their buffers differ while their values are the same) */
  const void* coefficientIdentity() const { return ident_.get(); }

 private:
  Vector<D> coeffs_;
  std::shared_ptr<const char> ident_{ std::make_shared<const char>(0) };
};
}  // namespace VarPoisson

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

/* PatchSolver.h:40-170: apply = smooth from u = 0 */
template <int D>
class PatchSolver : public virtual Operator<D>, public virtual GMG::Smoother<D> {
 public:
  PatchSolver(const Domain<D>& domain, const GhostFiller<D>& filler) : domain_(domain), filler_(filler.clone()) {}
  PatchSolver* clone() const override = 0;
  void apply(const Vector<D>& f, Vector<D>& u) const override
  {
    u.setWithGhost(0.0);
    this->smooth(f, u);
  }
  const Domain<D>& getDomain() const { return domain_; }
  const GhostFiller<D>& getGhostFiller() const { return *filler_; }

 private:
  Domain<D> domain_;
  std::shared_ptr<const GhostFiller<D>> filler_;
};

namespace detail {
/* PatchSolver::smooth's "Total Patch Smooth" domain timing (PatchSolver.h:153-168), by
 * events on the null stream the smoother's launches use; stop() ends it */
template <int D>
struct SmoothTiming {
  explicit SmoothTiming(const Domain<D>& d) : d_(d)
  {
    if (d_.hasTimer()) d_.getTimer()->startOn("Total Patch Smooth", nullptr, d_.getId());
  }
  void stop()
  {
    if (d_.hasTimer()) d_.getTimer()->stopOn("Total Patch Smooth", nullptr, d_.getId());
  }
  const Domain<D>& d_;
};
}  // namespace detail

namespace Poisson {
/* Poisson/DFTPatchSolver.h:369-450 */
template <int D>
class DFTPatchSolver : public PatchSolver<D> {
 public:
  DFTPatchSolver(const PatchOperator<D>& op, std::bitset<2 * D> neumann)
      : PatchSolver<D>(op.getDomain(), op.getGhostFiller()), op_(op.clone()), neumann_(neumann)
  {
    int op_kind = 0;
    const double* coeffs = nullptr;  // copied by te_dft_plan_create
    if (auto v = dynamic_cast<const VarPoisson::StarPatchOperator<D>*>(op_.get())) {
      op_kind = 1;
      coeffs = v->getCoefficients().data();
    } else if (!dynamic_cast<const Poisson::StarPatchOperator<D>*>(op_.get())) {
      throw RuntimeError("DFTPatchSolver: the operator must be a Poisson or VarPoisson StarPatchOperator");
    }
    te_dft_plan* h = nullptr;
    detail::check(te_dft_plan_create(op.deviceFiller().handle(), (unsigned)neumann.to_ulong(), op_kind, coeffs, &h));
    h_ = std::shared_ptr<te_dft_plan>(h, detail::Deleter<te_dft_plan, te_dft_plan_destroy>());
  }
  DFTPatchSolver* clone() const override { return new DFTPatchSolver(*this); }
  void smooth(const Vector<D>& f, Vector<D>& u) const override
  {
    detail::SmoothTiming<D> t(this->getDomain());
    detail::check(te_dft_smooth(h_.get(), TE_F64, f.data(), u.data(), nullptr));
    t.stop();
  }
  const PatchOperator<D>& getOperator() const { return *op_; }
  std::bitset<2 * D> getNeumann() const { return neumann_; }

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
      : fine_(fine_domain), coarse_(coarse_domain)
  {
    if (extrapolate_boundary_ghosts)
      throw RuntimeError("LinearRestrictor: boundary-ghost extrapolation is not provided by this build");
    te_level_link* h = nullptr;
    detail::check(te_level_link_create(fine_domain.handle(), coarse_domain.handle(), &h));
    h_ = std::shared_ptr<te_level_link>(h, detail::Deleter<te_level_link, te_level_link_destroy>());
  }
  LinearRestrictor* clone() const override { return new LinearRestrictor(*this); }
  Vector<D> restrict(const Vector<D>& fine) const override
  {
    Vector<D> coarse(coarse_, fine.getNumComponents());
    detail::check(te_restrict(h_.get(), TE_F64, fine.data(), coarse.data(), fine.getNumComponents(), nullptr));
    return coarse;
  }
  const Domain<D>& fineDomain() const { return fine_; }
  const Domain<D>& coarseDomain() const { return coarse_; }

 private:
  Domain<D> fine_, coarse_;
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
/* Iterative/PatchSolver.h:53-175: inner solver per patch (BiCGStab only) */
template <int D>
class BiCGStab;
template <int D>
class PatchSolver : public ThunderEgg::PatchSolver<D> {
 public:
  PatchSolver(const Solver<D>& solver, const PatchOperator<D>& op, bool continue_on_breakdown = false);
  PatchSolver* clone() const override { return new PatchSolver(*this); }
  void smooth(const Vector<D>& f, Vector<D>& u) const override
  {
    ThunderEgg::detail::SmoothTiming<D> t(this->getDomain());
    detail::check(te_patch_bicgstab_smooth(op_->deviceFiller().handle(), TE_F64, op_kind_, neumann_, coeffs_,
                                           max_it_, tol_, continue_ ? 1 : 0, f.data(), u.data(), nullptr,
                                           nullptr));
    t.stop();
  }
  int maxIterations() const { return max_it_; }
  double tolerance() const { return tol_; }
  bool continueOnBreakdown() const { return continue_; }

 private:
  std::shared_ptr<const PatchOperator<D>> op_;
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
  if (auto v = dynamic_cast<const VarPoisson::StarPatchOperator<D>*>(op_.get())) {
    op_kind_ = 1;
    coeffs_ = v->getCoefficients().data();
  } else if (auto p = dynamic_cast<const Poisson::StarPatchOperator<D>*>(op_.get())) {
    neumann_ = p->neumann() ? 1 : 0;
  } else {
    throw RuntimeError("Iterative::PatchSolver: the operator must be a Poisson or VarPoisson StarPatchOperator");
  }
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
    const PatchOperator<D>* po = dynamic_cast<const PatchOperator<D>*>(L.op.get());
    if (!po) return;
    int ok = -1, nm = 0;
    const double* c = nullptr;
    if (auto v = dynamic_cast<const VarPoisson::StarPatchOperator<D>*>(po)) {
      ok = 1;
      c = v->getCoefficients().data();
    } else if (auto p = dynamic_cast<const Poisson::StarPatchOperator<D>*>(po)) {
      ok = 0;
      nm = p->neumann() ? 1 : 0;
    } else {
      return;
    }
    int sk = -1, im = 1000, ct = -1;
    double it = 1e-12;
    if (auto s = dynamic_cast<const Poisson::DFTPatchSolver<D>*>(L.smoother.get())) {
      sk = 0;
      const unsigned mk = (unsigned)s->getNeumann().to_ulong();
      if (any_dft && mk != mask) return;  // one Neumann mask for every DFT level
      mask = mk;
      any_dft = true;
    } else if (auto s = dynamic_cast<const Iterative::PatchSolver<D>*>(L.smoother.get())) {
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
    if (i + 1 < levels_.size() && !dynamic_cast<const LinearRestrictor<D>*>(L.restrictor.get())) return;
    if (i > 0 && !dynamic_cast<const DirectInterpolator<D>*>(L.interpolator.get())) return;
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
