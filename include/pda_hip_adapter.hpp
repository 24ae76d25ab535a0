/*
 * pda_hip_adapter.hpp -- C++ drop-in facade of the MI355X engine with the
 * reference's problem API (include/pressiodemoapps/adapter_cpp.hpp:60-264,
 * the create_problem_eigen factories of euler1d.hpp:82-99, euler2d.hpp:88-173,
 * euler3d.hpp:80-122, swe2d.hpp:143-195, advection_diffusion2d.hpp:80-179,
 * advection_diffusion_reaction2d.hpp:111-171, and
 * load_cellcentered_uniform_mesh_eigen, mesh.hpp:87-91).
 *
 * Code written against the reference compiles against this header after a
 * namespace change (pressiodemoapps -> pressiodemoapps::hip): the factory and
 * method names, argument order and error behaviour (std::runtime_error) are
 * the reference's.  What differs is where the data lives: state, velocity
 * and Jacobian values are HIP device arrays (DeviceVector, DeviceMatrix,
 * DeviceCsrMatrix below) instead of Eigen containers, and every evaluation
 * is ordered on the problem's stream (setStream; default: the null stream)
 * without an implicit device synchronization.
 *
 * Header-only over the C ABI of pda_c_api.h; host code, built with any C++17
 * compiler: add <repo>/include and /opt/rocm/include to the include path,
 * define __HIP_PLATFORM_AMD__, link -lpda -lamdhip64.
 */
#ifndef PDA_HIP_ADAPTER_HPP
#define PDA_HIP_ADAPTER_HPP

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "pda_c_api.h"

namespace pressiodemoapps {
namespace hip {

inline void check(int rc) {
  if (rc != PDA_OK) throw std::runtime_error(pda_last_error());
}

inline void check_hip(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// enums with the reference's names and values (schemes_info.hpp:61-63,
// euler1d.hpp:65-70, euler2d.hpp:67-77, euler3d.hpp:65-68, swe2d.hpp:66-69,
// advection_diffusion2d.hpp:65-68, advection_diffusion_reaction2d.hpp:61-62)
enum class InviscidFluxReconstruction { FirstOrder = PDA_FIRSTORDER, Weno3 = PDA_WENO3, Weno5 = PDA_WENO5 };
enum class Euler1d { PeriodicSmooth = 0, Sod, Lax, ShuOsher };
enum class Euler2d {
  PeriodicSmooth = 0, KelvinHelmholtz, SedovFull, SedovSymmetry, Riemann, NormalShock,
  DoubleMachReflection, CrossShock, testingonlyneumann
};
enum class Euler3d { PeriodicSmooth = 0, SedovSymmetry };
enum class Swe2d { SlipWall = 0, CustomBCs };
enum class AdvectionDiffusion2d { BurgersPeriodic = 0, BurgersOutflow };
enum class AdvectionDiffusionReaction2d { ProblemA = 0 };

namespace detail {
inline int family(Euler1d) { return PDA_EULER1D; }
inline int family(Euler2d) { return PDA_EULER2D; }
inline int family(Euler3d) { return PDA_EULER3D; }
inline int family(Swe2d) { return PDA_SWE2D; }
inline int family(AdvectionDiffusion2d) { return PDA_BURGERS2D; }
inline int family(AdvectionDiffusionReaction2d) { return PDA_ADR2D; }
}  // namespace detail

// ------------------------------------------------------------ containers

// Owned array of doubles in device memory (move-only; clone() copies).
class DeviceVector {
 public:
  DeviceVector() = default;
  explicit DeviceVector(int64_t n) : n_(n) {
    if (n_ > 0) {
      check_hip(hipMalloc(reinterpret_cast<void **>(&p_), sizeof(double) * size_t(n_)), "hipMalloc");
      check_hip(hipMemset(p_, 0, sizeof(double) * size_t(n_)), "hipMemset");
    }
  }
  ~DeviceVector() { release(); }
  DeviceVector(const DeviceVector &) = delete;
  DeviceVector &operator=(const DeviceVector &) = delete;
  DeviceVector(DeviceVector &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DeviceVector &operator=(DeviceVector &&o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_; n_ = o.n_;
      o.p_ = nullptr; o.n_ = 0;
    }
    return *this;
  }
  double *data() { return p_; }
  const double *data() const { return p_; }
  int64_t size() const { return n_; }
  DeviceVector clone() const {
    DeviceVector c(n_);
    if (n_ > 0) check_hip(hipMemcpy(c.p_, p_, sizeof(double) * size_t(n_), hipMemcpyDeviceToDevice), "hipMemcpy");
    return c;
  }
  void setZero() {
    if (n_ > 0) check_hip(hipMemset(p_, 0, sizeof(double) * size_t(n_)), "hipMemset");
  }
  // synchronous host transfers
  void copyFromHost(const double *h) {
    if (n_ > 0) check_hip(hipMemcpy(p_, h, sizeof(double) * size_t(n_), hipMemcpyHostToDevice), "hipMemcpy");
  }
  void copyToHost(double *h) const {
    if (n_ > 0) check_hip(hipMemcpy(h, p_, sizeof(double) * size_t(n_), hipMemcpyDeviceToHost), "hipMemcpy");
  }
  std::vector<double> toHost() const {
    std::vector<double> h(static_cast<size_t>(n_));
    copyToHost(h.data());
    return h;
  }

 private:
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  double *p_ = nullptr;
  int64_t n_ = 0;
};

enum class Layout { ColMajor = 0, RowMajor = 1 };

// Dense rows x cols matrix of doubles in device memory, column- or
// row-major (the two operand types of applyJacobian, adapter_cpp.hpp:76-77).
class DeviceMatrix {
 public:
  DeviceMatrix() = default;
  DeviceMatrix(int64_t rows, int64_t cols, Layout layout = Layout::ColMajor)
      : v_(rows * cols), rows_(rows), cols_(cols), layout_(layout) {}
  int64_t rows() const { return rows_; }
  int64_t cols() const { return cols_; }
  Layout layout() const { return layout_; }
  double *data() { return v_.data(); }
  const double *data() const { return v_.data(); }
  DeviceVector &storage() { return v_; }
  const DeviceVector &storage() const { return v_; }

 private:
  DeviceVector v_;
  int64_t rows_ = 0, cols_ = 0;
  Layout layout_ = Layout::ColMajor;
};

// The Jacobian: the reference's CSR (row = cell*m + dof, sorted columns,
// impl/euler_2d_prob_class.hpp:223-237) with the pattern shared with the
// problem (valid while it lives) and owned values, named as Eigen's
// compressed-storage accessors.
class DeviceCsrMatrix {
 public:
  DeviceCsrMatrix() = default;
  DeviceCsrMatrix(const int64_t *rowptr, const int32_t *colidx, int64_t rows, int64_t cols, int64_t nnz)
      : rowptr_(rowptr), colidx_(colidx), values_(nnz), rows_(rows), cols_(cols) {}
  int64_t rows() const { return rows_; }
  int64_t cols() const { return cols_; }
  int64_t nonZeros() const { return values_.size(); }
  const int64_t *outerIndexPtr() const { return rowptr_; }  // device
  const int32_t *innerIndexPtr() const { return colidx_; }  // device
  double *valuePtr() { return values_.data(); }             // device
  const double *valuePtr() const { return values_.data(); }
  DeviceVector &values() { return values_; }
  const DeviceVector &values() const { return values_; }
  DeviceCsrMatrix clone() const {
    DeviceCsrMatrix c;
    c.rowptr_ = rowptr_;
    c.colidx_ = colidx_;
    c.values_ = values_.clone();
    c.rows_ = rows_;
    c.cols_ = cols_;
    return c;
  }

 private:
  const int64_t *rowptr_ = nullptr;
  const int32_t *colidx_ = nullptr;
  DeviceVector values_;
  int64_t rows_ = 0, cols_ = 0;
};

// ------------------------------------------------------------------ mesh

// Cell-centred uniform mesh (impl/mesh_ccu.hpp:60-473), read from the
// reference's text files or built as meshing_scripts/create_full_mesh.py
// writes it.  Move-only; a problem keeps its own reference on the mesh.
class CellCenteredUniformMesh {
 public:
  explicit CellCenteredUniformMesh(pda_mesh m) : m_(m) { query(); }
  ~CellCenteredUniformMesh() {
    if (m_) pda_mesh_destroy(m_);
  }
  CellCenteredUniformMesh(const CellCenteredUniformMesh &) = delete;
  CellCenteredUniformMesh &operator=(const CellCenteredUniformMesh &) = delete;
  CellCenteredUniformMesh(CellCenteredUniformMesh &&o) noexcept { *this = std::move(o); }
  CellCenteredUniformMesh &operator=(CellCenteredUniformMesh &&o) noexcept {
    if (this != &o) {
      if (m_) pda_mesh_destroy(m_);
      m_ = o.m_; dim_ = o.dim_; st_ = o.st_; ns_ = o.ns_; nst_ = o.nst_;
      d_[0] = o.d_[0]; d_[1] = o.d_[1]; d_[2] = o.d_[2];
      o.m_ = nullptr;
    }
    return *this;
  }
  pda_mesh handle() const { return m_; }
  int dimensionality() const { return dim_; }
  int stencilSize() const { return st_; }
  int64_t sampleMeshSize() const { return ns_; }
  int64_t stencilMeshSize() const { return nst_; }
  double dx() const { return d_[0]; }
  double dy() const { return d_[1]; }
  double dz() const { return d_[2]; }

 private:
  void query() { check(pda_mesh_info(m_, &dim_, &st_, &ns_, &nst_, d_)); }
  pda_mesh m_ = nullptr;
  int dim_ = 0, st_ = 0;
  int64_t ns_ = 0, nst_ = 0;
  double d_[3] = {0, 0, 0};
};

// Reads info.dat / coordinates.dat / connectivity.dat; the reference calls
// exit() on a missing file, this throws.
inline CellCenteredUniformMesh load_cellcentered_uniform_mesh_eigen(const std::string &dir) {
  pda_mesh m = nullptr;
  check(pda_mesh_load(dir.c_str(), &m));
  return CellCenteredUniformMesh(m);
}
inline CellCenteredUniformMesh load_cellcentered_uniform_mesh(const std::string &dir) {
  return load_cellcentered_uniform_mesh_eigen(dir);
}

// The full mesh meshing_scripts/create_full_mesh.py would write: n cells per
// axis (n.size() = dimension), bounds {xmin, xmax[, ymin, ymax[, zmin, zmax]]},
// periodic bit a = axis a.
inline CellCenteredUniformMesh create_full_mesh(const std::vector<int64_t> &n, const std::vector<double> &bounds,
                                                int stencil, int periodic_mask = 0) {
  int64_t nn[3] = {1, 1, 1};
  double b[6] = {0, 0, 0, 0, 0, 0};
  for (size_t a = 0; a < n.size() && a < 3; ++a) nn[a] = n[a];
  for (size_t a = 0; a < bounds.size() && a < 6; ++a) b[a] = bounds[a];
  pda_mesh m = nullptr;
  check(pda_mesh_create_structured(int(n.size()), nn, b, stencil, periodic_mask, &m));
  return CellCenteredUniformMesh(m);
}

// --------------------------------------------------------------- problem

// PublicProblemEigenMixinCpp (adapter_cpp.hpp:60-264) on device arrays.
class Problem {
 public:
  using scalar_type = double;
  using independent_variable_type = double;
  using time_type = double;
  using state_type = DeviceVector;
  using rhs_type = DeviceVector;
  using right_hand_side_type = DeviceVector;
  using jacobian_type = DeviceCsrMatrix;

  Problem(const CellCenteredUniformMesh &mesh, int family, int prob, InviscidFluxReconstruction rec, int icFlag,
          const std::unordered_map<std::string, double> &params) {
    std::vector<const char *> keys;
    std::vector<double> vals;
    for (const auto &kv : params) {
      keys.push_back(kv.first.c_str());
      vals.push_back(kv.second);
    }
    check(pda_problem_create(mesh.handle(), family, prob, int(rec), icFlag, keys.data(), vals.data(),
                             int(keys.size()), &p_));
    check(pda_problem_sizes(p_, &m_, &ns_, &nst_, &nnz_));
  }
  ~Problem() {
    if (p_) pda_problem_destroy(p_);
  }
  // the reference deletes both assignments (adapter_cpp.hpp:90-91); the
  // factories return by value, so the facade keeps the move constructor
  Problem(const Problem &) = delete;
  Problem &operator=(const Problem &) = delete;
  Problem &operator=(Problem &&) = delete;
  Problem(Problem &&o) noexcept
      : p_(o.p_), m_(o.m_), ns_(o.ns_), nst_(o.nst_), nnz_(o.nnz_), stream_(o.stream_) {
    o.p_ = nullptr;
  }

  pda_problem handle() const { return p_; }
  // hipStream_t the evaluations are ordered on (nullptr: the null stream)
  void setStream(void *stream) { stream_ = stream; }
  void *stream() const { return stream_; }

  int numDofPerCell() const { return m_; }
  int64_t totalDofSampleMesh() const { return ns_; }
  int64_t totalDofStencilMesh() const { return nst_; }
  double queryParameter(const std::string &name) const {
    double v = 0.;
    check(pda_problem_parameter(p_, name.c_str(), &v));
    return v;
  }

  state_type initialCondition() const {
    state_type U(nst_);
    check(pda_initial_condition(p_, U.data(), stream_));
    return U;
  }
  state_type createState() const { return state_type(nst_); }  // zeros, as adapter_cpp.hpp:105-109
  right_hand_side_type createRightHandSide() const { return right_hand_side_type(ns_); }
  rhs_type createRhs() const { return rhs_type(ns_); }
  jacobian_type createJacobian() const {
    int64_t rows = 0, nnz = 0;
    const int64_t *rp = nullptr;
    const int32_t *ci = nullptr;
    check(pda_jacobian_pattern(p_, &rows, &nnz, &rp, &ci));
    return jacobian_type(rp, ci, rows, nst_, nnz);
  }
  // the pattern on the host (rows+1 row pointers, nnz column indices)
  void jacobianPatternHost(std::vector<int64_t> &rowptr, std::vector<int32_t> &colidx) const {
    rowptr.resize(size_t(ns_ + 1));
    colidx.resize(size_t(nnz_));
    check(pda_jacobian_pattern_host(p_, rowptr.data(), colidx.data()));
  }

  // evaluation, adapter_cpp.hpp:162-229
  void operator()(const state_type &U, double t, right_hand_side_type &V) const { rightHandSide(U, t, V); }
  void operator()(const state_type &U, double t, right_hand_side_type &V, jacobian_type &J,
                  bool computeJacobian) const {
    if (computeJacobian) rightHandSideAndJacobian(U, t, V, J);
    else rightHandSide(U, t, V);
  }
  void rhs(const state_type &U, double t, right_hand_side_type &V) const { rightHandSide(U, t, V); }
  void rightHandSide(const state_type &U, double t, right_hand_side_type &V) const {
    sizes(U, &V, nullptr);
    check(pda_velocity(p_, U.data(), t, V.data(), stream_));
  }
  void rightHandSideAndJacobian(const state_type &U, double t, right_hand_side_type &V, jacobian_type &J) const {
    sizes(U, &V, &J);
    check(pda_jacobian(p_, U.data(), t, J.valuePtr(), V.data(), stream_));
  }
  void rhsAndJacobian(const state_type &U, double t, rhs_type &V, std::optional<jacobian_type *> J) const {
    if (J) rightHandSideAndJacobian(U, t, V, *J.value());
    else rightHandSide(U, t, V);
  }
  void jacobian(const state_type &U, double t, jacobian_type &J) const {
    sizes(U, nullptr, &J);
    check(pda_jacobian(p_, U.data(), t, J.valuePtr(), nullptr, stream_));
  }

  // Jacobian action, adapter_cpp.hpp:125-157,231-265
  DeviceVector createApplyJacobianResult(const DeviceVector &) const { return DeviceVector(ns_); }
  DeviceMatrix createApplyJacobianResult(const DeviceMatrix &B) const {
    return DeviceMatrix(ns_, B.cols(), B.layout());
  }
  DeviceVector createResultOfJacobianActionOn(const DeviceVector &B) const { return createApplyJacobianResult(B); }
  DeviceMatrix createResultOfJacobianActionOn(const DeviceMatrix &B) const { return createApplyJacobianResult(B); }
  void applyJacobian(const state_type &U, const DeviceVector &B, double t, DeviceVector &AJ) const {
    if (U.size() != nst_ || B.size() != nst_ || AJ.size() != ns_)
      throw std::runtime_error("applyJacobian: operand or result has the wrong size");
    check(pda_apply_jacobian(p_, U.data(), t, B.data(), 1, 0, AJ.data(), stream_));
  }
  void applyJacobian(const state_type &U, const DeviceMatrix &B, double t, DeviceMatrix &AJ) const {
    if (U.size() != nst_ || B.rows() != nst_ || AJ.rows() != ns_ || AJ.cols() != B.cols() ||
        AJ.layout() != B.layout())
      throw std::runtime_error("applyJacobian: operand or result has the wrong shape or layout");
    check(pda_apply_jacobian(p_, U.data(), t, B.data(), B.cols(), int(B.layout()), AJ.data(), stream_));
  }

 private:
  void sizes(const state_type &U, const right_hand_side_type *V, const jacobian_type *J) const {
    if (U.size() != nst_) throw std::runtime_error("state has the wrong size");
    if (V && V->size() != ns_) throw std::runtime_error("right-hand side has the wrong size");
    if (J && J->nonZeros() != nnz_) throw std::runtime_error("Jacobian has the wrong number of nonzeros");
  }

  pda_problem p_ = nullptr;
  int m_ = 0;
  int64_t ns_ = 0, nst_ = 0, nnz_ = 0;
  void *stream_ = nullptr;
};

// ------------------------------------------------------------- factories

// create_problem_eigen(mesh, ProblemEnum, InviscidFluxReconstruction[, icFlag]
// [, parameter map]); icFlag <= 0 selects the family's default initial
// condition (the reference's 1 for SWE and Burgers, the problem's own
// default otherwise)
template <class ProblemEnum>
Problem create_problem_eigen(const CellCenteredUniformMesh &mesh, ProblemEnum e, InviscidFluxReconstruction rec,
                             int icFlag = 0, const std::unordered_map<std::string, double> &params = {}) {
  return Problem(mesh, detail::family(e), int(e), rec, icFlag, params);
}
template <class ProblemEnum>
Problem create_problem(const CellCenteredUniformMesh &mesh, ProblemEnum e, InviscidFluxReconstruction rec,
                       int icFlag = 0, const std::unordered_map<std::string, double> &params = {}) {
  return create_problem_eigen(mesh, e, rec, icFlag, params);
}

// advection_diffusion_reaction2d.hpp:148-171
inline Problem create_advdiffreac_2d_problem_A_eigen(const CellCenteredUniformMesh &mesh,
                                                     InviscidFluxReconstruction rec, double ux, double uy,
                                                     double diffusion, double sigma) {
  return Problem(mesh, PDA_ADR2D, int(AdvectionDiffusionReaction2d::ProblemA), rec, 0,
                 {{"ux", ux}, {"uy", uy}, {"diffusion", diffusion}, {"sigma", sigma}});
}

}  // namespace hip
}  // namespace pressiodemoapps

#endif
