// ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A thin C entry layer over the reference pressio-demoapps headers, compiled
// (by oracle/Makefile) from the reference sources where they lie into
// oracle/_ref/.  It is used to pin the CPU oracle (oracle/pda_oracle.c) and,
// with OpenMP, as the reference CPU baseline in bench.py.  Nothing in the
// product links it.
//
// Entry points take a mesh directory in the reference text format
// (info.dat / coordinates.dat / connectivity.dat), a problem family and the
// reference enum values, and expose the reference's own
// rightHandSide / rightHandSideAndJacobian / initialCondition.
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "pressiodemoapps/euler1d.hpp"
#include "pressiodemoapps/euler2d.hpp"
#include "pressiodemoapps/euler3d.hpp"
#include "pressiodemoapps/swe2d.hpp"
#include "pressiodemoapps/advection_diffusion2d.hpp"
// The reference ADR class does not compile with PRESSIODEMOAPPS_ENABLE_OPENMP
// (impl/advection_diffusion_reaction_2d_prob_class.hpp:493: a declaration
// follows "#pragma omp for"), so the OpenMP build leaves ADR out.
#ifndef PRESSIODEMOAPPS_ENABLE_OPENMP
#define PDA_REF_HAS_ADR 1
#include "pressiodemoapps/advection_diffusion_reaction2d.hpp"
#endif

namespace pda = pressiodemoapps;
using mesh_t = pda::cellcentered_uniform_mesh_eigen_type;

namespace {

struct Base {
  virtual ~Base() = default;
  virtual int ndpc() const = 0;
  virtual int64_t rows() const = 0;
  virtual int64_t cols() const = 0;
  virtual int64_t nnz() const = 0;
  virtual void ic(double *) const = 0;
  virtual void velocity(const double *, double, double *) const = 0;
  virtual void jacobian(const double *, double, double *, int64_t *, int32_t *, double *) const = 0;
  virtual void time_calls(const double *, double, int, int, double *) const = 0;
};

template <class P>
struct Holder : Base {
  std::unique_ptr<mesh_t> mesh;
  P p;
  template <class F>
  Holder(std::unique_ptr<mesh_t> m, F &&make) : mesh(std::move(m)), p(make(*mesh)) {}
  int ndpc() const override { return p.numDofPerCell(); }
  int64_t rows() const override { return p.totalDofSampleMesh(); }
  int64_t cols() const override { return p.totalDofStencilMesh(); }
  int64_t nnz() const override { return p.createJacobian().nonZeros(); }
  void ic(double *out) const override {
    auto s = p.initialCondition();
    std::memcpy(out, s.data(), sizeof(double) * s.size());
  }
  void velocity(const double *U, double t, double *V) const override {
    typename P::state_type u(cols());
    std::memcpy(u.data(), U, sizeof(double) * cols());
    auto v = p.createRightHandSide();
    p.rightHandSide(u, t, v);
    std::memcpy(V, v.data(), sizeof(double) * rows());
  }
  void jacobian(const double *U, double t, double *V, int64_t *rowptr,
                int32_t *colidx, double *vals) const override {
    typename P::state_type u(cols());
    std::memcpy(u.data(), U, sizeof(double) * cols());
    auto v = p.createRightHandSide();
    auto J = p.createJacobian();
    p.rightHandSideAndJacobian(u, t, v, J);
    if (V) std::memcpy(V, v.data(), sizeof(double) * rows());
    for (int64_t i = 0; i <= J.rows(); ++i) rowptr[i] = J.outerIndexPtr()[i];
    for (int64_t i = 0; i < J.nonZeros(); ++i) {
      colidx[i] = J.innerIndexPtr()[i];
      vals[i] = J.valuePtr()[i];
    }
  }
  void time_calls(const double *U, double t, int which, int reps, double *secs) const override;
};

template <class P>
void Holder<P>::time_calls(const double *U, double t, int which, int reps, double *secs) const {
  // the caller's side of adapter_cpp.hpp:105-229: state, velocity and
  // Jacobian created once, then the evaluation called on them repeatedly
  typename P::state_type u(cols());
  std::memcpy(u.data(), U, sizeof(double) * cols());
  auto v = p.createRightHandSide();
  // (no Jacobian copy when only the velocity is timed)
  auto J = which == 0 ? typename P::jacobian_type() : p.createJacobian();
  auto call = [&] {
    if (which == 0) p.rightHandSide(u, t, v);
    else if (which == 1) p.jacobian(u, t, J);
    else p.rightHandSideAndJacobian(u, t, v, J);
  };
  call();  // warm-up
  for (int r = 0; r < reps; ++r) {
    const auto t0 = std::chrono::steady_clock::now();
    call();
    secs[r] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
}

template <class P, class F>
Base *make(std::unique_ptr<mesh_t> m, F &&f) {
  return new Holder<P>(std::move(m), std::forward<F>(f));
}

pda::InviscidFluxReconstruction rec(int s) {
  return s == 0 ? pda::InviscidFluxReconstruction::FirstOrder
       : s == 1 ? pda::InviscidFluxReconstruction::Weno3
                : pda::InviscidFluxReconstruction::Weno5;
}

thread_local std::string g_err;

}  // namespace

extern "C" {

const char *ref_last_error() { return g_err.c_str(); }

// family: 1 Euler1d, 2 Euler2d, 3 Euler3d, 4 Swe2d, 5 AdvectionDiffusion2d,
// 6 AdvectionDiffusionReaction2d.  icflag <= 0 selects the default factory.
void *ref_create(const char *meshdir, int family, int prob, int scheme,
                 int icflag, const char *const *keys, const double *vals,
                 int nparams) {
  try {
    auto mesh = std::make_unique<mesh_t>(std::string(meshdir));
    std::unordered_map<std::string, double> params;
    for (int i = 0; i < nparams; ++i) params[keys[i]] = vals[i];
    const auto r = rec(scheme);
    switch (family) {
      case 1: {
        using P = decltype(pda::create_problem_eigen(*mesh, pda::Euler1d::Sod, r));
        const auto e = static_cast<pda::Euler1d>(prob);
        return make<P>(std::move(mesh), [&](const mesh_t &m) { return pda::create_problem_eigen(m, e, r); });
      }
      case 2: {
        using P = decltype(pda::create_problem_eigen(*mesh, pda::Euler2d::Riemann, r));
        const auto e = static_cast<pda::Euler2d>(prob);
        if (nparams > 0)
          return make<P>(std::move(mesh), [&](const mesh_t &m) { return pda::create_problem_eigen(m, e, r, icflag, params); });
        if (icflag > 0)
          return make<P>(std::move(mesh), [&](const mesh_t &m) { return pda::create_problem_eigen(m, e, r, icflag); });
        return make<P>(std::move(mesh), [&](const mesh_t &m) { return pda::create_problem_eigen(m, e, r); });
      }
      case 3: {
        using P = decltype(pda::create_problem_eigen(*mesh, pda::Euler3d::SedovSymmetry, r));
        const auto e = static_cast<pda::Euler3d>(prob);
        return make<P>(std::move(mesh), [&](const mesh_t &m) { return pda::create_problem_eigen(m, e, r); });
      }
      case 4: {
        using P = decltype(pda::create_problem_eigen(*mesh, pda::Swe2d::SlipWall, r));
        const auto e = static_cast<pda::Swe2d>(prob);
        const int ic = icflag > 0 ? icflag : 1;
        return make<P>(std::move(mesh), [&](const mesh_t &m) { return pda::create_problem_eigen(m, e, r, ic, params); });
      }
      case 5: {
        using P = decltype(pda::create_problem_eigen(*mesh, pda::AdvectionDiffusion2d::BurgersPeriodic, r, 1, params));
        const auto e = static_cast<pda::AdvectionDiffusion2d>(prob);
        const int ic = icflag > 0 ? icflag : 1;
        return make<P>(std::move(mesh), [&](const mesh_t &m) { return pda::create_problem_eigen(m, e, r, ic, params); });
      }
#ifdef PDA_REF_HAS_ADR
      case 6: {
        using P = decltype(pda::create_problem_eigen(*mesh, pda::AdvectionDiffusionReaction2d::ProblemA, r));
        const auto e = static_cast<pda::AdvectionDiffusionReaction2d>(prob);
        if (nparams == 4) {
          const double ux = vals[0], uy = vals[1], D = vals[2], sg = vals[3];
          return make<P>(std::move(mesh), [&](const mesh_t &m) {
            return pda::create_advdiffreac_2d_problem_A_eigen(m, r, ux, uy, D, sg);
          });
        }
        return make<P>(std::move(mesh), [&](const mesh_t &m) { return pda::create_problem_eigen(m, e, r); });
      }
#endif
      default:
        g_err = "unknown family";
        return nullptr;
    }
  } catch (const std::exception &e) {
    g_err = e.what();
    return nullptr;
  }
}

void ref_destroy(void *h) { delete static_cast<Base *>(h); }

int ref_sizes(void *h, int *ndpc, int64_t *rows, int64_t *cols, int64_t *nnz) {
  auto *b = static_cast<Base *>(h);
  *ndpc = b->ndpc();
  *rows = b->rows();
  *cols = b->cols();
  *nnz = b->nnz();
  return 0;
}

int ref_initial_condition(void *h, double *U) {
  try {
    static_cast<Base *>(h)->ic(U);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

int ref_velocity(void *h, const double *U, double t, double *V) {
  try {
    static_cast<Base *>(h)->velocity(U, t, V);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// Times the reference's own evaluation calls as a caller of
// adapter_cpp.hpp runs them: state, velocity and Jacobian created once,
// then which = 0 rightHandSide (:193-199), 1 jacobian (:223-229) or
// 2 rightHandSideAndJacobian (:201-208) called once to warm up and `reps`
// more times; secs[r] receives the wall time of call r.
int ref_time_calls(void *h, const double *U, double t, int which, int reps, double *secs) {
  try {
    static_cast<Base *>(h)->time_calls(U, t, which, reps, secs);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

int ref_velocity_and_jacobian(void *h, const double *U, double t, double *V,
                              int64_t *rowptr, int32_t *cols, double *vals) {
  try {
    static_cast<Base *>(h)->jacobian(U, t, V, rowptr, cols, vals);
    return 0;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

}  // extern "C"
