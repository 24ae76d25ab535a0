// pda_stepper.hip -- explicit time integration of dU/dt = velocity(U) on the
// device: forward Euler, SSP Runge-Kutta 3 and classical Runge-Kutta 4 with
// pressio's stage formulas
// (tests_cpp/pressio/include/pressio/ode/impl/ode_explicit_stepper_without_mass_matrix.hpp:
//  forward Euler :168-187, SSPRK3 :230-281, RK4 :284-352; the Python
// module's advanceRK2/advanceRK4/advanceSSP3, pressiodemoapps/__init__.py:90-185).
//
// A step is a fixed sequence of velocity evaluations and stage updates on
// the stepper's own buffers.  With graphs enabled it is captured once into a
// hipGraph (on an internal stream) and replayed per step on the caller's
// stream; the capture is redone when the state pointer or dt changes.  The
// evaluation kernels use no memsets, so the captured step has only kernel
// nodes.
#include "pda_rk.hpp"

namespace pda {
namespace {

// out[i] = c0*x0[i] + c1*x1[i] + ... (left to right, up to five terms)
struct Lin {
  const double *x[5];
  double c[5];
  int n;
};

// out may alias an input (element-wise read before write)
__global__ void __launch_bounds__(256) k_lincomb(double *out, Lin L, int64_t len) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += stride) {
    double v = L.c[0] * L.x[0][i];
    for (int k = 1; k < L.n; ++k) v += L.c[k] * L.x[k][i];
    out[i] = v;
  }
}

// The Python module's stage updates (pressiodemoapps/__init__.py:90-185)
// as numpy evaluates them: one rounding per product and per sum, in the
// module's order, no fused multiply-add.  v is the velocity just
// evaluated; k_i = dt * v_i as the module forms them.
#pragma clang fp contract(off)

// k = dt * v (in place, kept for the final combination); out = x + a * k
__global__ void __launch_bounds__(256) k_py_stage(double *kv, const double *x, double *out, double dt, double a,
                                                  int64_t n) {
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double k = dt * kv[i];
    kv[i] = k;
    out[i] = x[i] + a * k;
  }
}

// advanceSSP3, second stage: out = 3/4 y + 1/4 y1 + (1/4 dt) v
__global__ void __launch_bounds__(256) k_py_ssp3_second(double *out, const double *y, const double *y1,
                                                        const double *v, double qdt, int64_t n) {
  const double threeOverFour = 3. / 4., oneOverFour = 1. / 4.;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double a = threeOverFour * y[i];
    const double b = oneOverFour * y1[i];
    const double c = qdt * v[i];
    out[i] = (a + b) + c;
  }
}

// advanceSSP3, last stage: y = 1/3 (y + 2 y2 + (2 dt) v)
__global__ void __launch_bounds__(256) k_py_ssp3_final(double *y, const double *y2, const double *v, double tdt,
                                                       int64_t n) {
  const double oneOverThree = 1. / 3., two = 2.;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double a = y[i] + two * y2[i];
    const double b = tdt * v[i];
    y[i] = oneOverThree * (a + b);
  }
}

// advanceRK4, last stage: k4 = dt v; y = y + (k1 + 2 k2 + 2 k3 + k4) / 6
__global__ void __launch_bounds__(256) k_py_rk4_final(double *y, const double *k1, const double *k2,
                                                      const double *k3, const double *v, double dt, int64_t n) {
  const double two = 2., oneOverSix = 1. / 6.;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double k4 = dt * v[i];
    double sum = k1[i] + two * k2[i];
    sum = sum + two * k3[i];
    sum = sum + k4;
    y[i] = y[i] + sum * oneOverSix;
  }
}

// advanceRK2, last stage: k2 = dt v; y = y + k2/2 + k1/2
__global__ void __launch_bounds__(256) k_py_rk2_final(double *y, const double *k1, const double *v, double dt,
                                                      int64_t n) {
  const double half = 0.5;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double k2 = dt * v[i];
    const double a = y[i] + k2 * half;
    y[i] = a + k1[i] * half;
  }
}

#pragma clang fp contract(on)

dim3 vec_grid(int64_t len) {
  return dim3(unsigned(std::max<int64_t>(std::min<int64_t>((len + 255) / 256, 4096), 1)));
}

}  // namespace

int py_stage(double *kv, const double *x, double *out, double dt, double a, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_py_stage, vec_grid(n), dim3(256), 0, s, kv, x, out, dt, a, n);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}
int py_ssp3_second(double *out, const double *y, const double *y1, const double *v, double dt, int64_t n,
                   hipStream_t s) {
  const double oneOverFour = 1. / 4.;
  hipLaunchKernelGGL(k_py_ssp3_second, vec_grid(n), dim3(256), 0, s, out, y, y1, v, oneOverFour * dt, n);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}
int py_ssp3_final(double *y, const double *y2, const double *v, double dt, int64_t n, hipStream_t s) {
  const double two = 2.;
  hipLaunchKernelGGL(k_py_ssp3_final, vec_grid(n), dim3(256), 0, s, y, y2, v, two * dt, n);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}
int py_rk4_final(double *y, const double *k1, const double *k2, const double *k3, const double *v, double dt,
                 int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_py_rk4_final, vec_grid(n), dim3(256), 0, s, y, k1, k2, k3, v, dt, n);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}
int py_rk2_final(double *y, const double *k1, const double *v, double dt, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_py_rk2_final, vec_grid(n), dim3(256), 0, s, y, k1, v, dt, n);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

int lincomb(double *out, std::initializer_list<std::pair<double, const double *>> terms, int64_t len,
            hipStream_t s) {
  Lin L{};
  L.n = 0;
  for (auto &t : terms) {
    L.c[L.n] = t.first;
    L.x[L.n] = t.second;
    ++L.n;
  }
  const int64_t blocks = std::min<int64_t>((len + 255) / 256, 4096);
  hipLaunchKernelGGL(k_lincomb, dim3(unsigned(std::max<int64_t>(blocks, 1))), dim3(256), 0, s, out, L,
                     len);
  PDA_HIP_CHECK(hipGetLastError());
  return PDA_OK;
}

}  // namespace pda

struct pda_stepper_s {
  pda_problem_s *prob = nullptr;
  int scheme = 0;
  int64_t len = 0;
  pda::DevBuf aux, k[4];
  hipStream_t cap = nullptr;         // capture stream
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const double *g_U = nullptr;       // arguments the captured step was recorded with
  double g_dt = 0., g_t = 0.;
  bool fused1d = false, periodic1d = false; // 1D problems: several steps per launch
  ~pda_stepper_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

namespace {
using namespace pda;

// one step U <- step(U) at time t; buffers of the stepper
RkStage stage(double *kout, double *out, std::initializer_list<std::pair<double, const double *>> terms,
              double cv) {
  RkStage S;
  S.kout = kout;
  S.out = out;
  for (auto &x : terms) {
    S.c[S.nx] = x.first;
    S.x[S.nx] = x.second;
    ++S.nx;
  }
  S.cv = cv;
  return S;
}

int step_once(pda_stepper_s &st, double *U, double t, double dt, hipStream_t s) {
  const pda_problem_s &p = *st.prob;
  double *k[4] = {st.k[0].as<double>(), st.k[1].as<double>(), st.k[2].as<double>(), st.k[3].as<double>()};
  double *aux = st.aux.as<double>();
  // SSPRK3 and RK4 on a problem evaluated by the generic row kernel alone:
  // each velocity launch also forms the next stage state (the stage
  // formulas of rk_step with each combination's terms in the same order,
  // the velocity term last), half the launches of a step; a stage never
  // writes the vector its velocity reads (ping-pong through k[1] / k[3])
  if (velocity_stage_fits(p) && st.scheme == PDA_SSPRK3) {
    const double third = 1. / 3., twoThird = 2. / 3., quarter = 1. / 4., threeQuarter = 3. / 4.;
    double *aux2 = k[1];
    int rc = launch_velocity_stage(p, U, stage(nullptr, aux, {{1., U}}, dt), s);
    if (!rc) rc = launch_velocity_stage(p, aux, stage(nullptr, aux2, {{quarter, aux}, {threeQuarter, U}}, quarter * dt), s);
    if (!rc) rc = launch_velocity_stage(p, aux2, stage(nullptr, U, {{third, U}, {twoThird, aux2}}, twoThird * dt), s);
    return rc;
  }
  if (velocity_stage_fits(p) && st.scheme == PDA_RK4) {
    const double h2 = dt / 2., h6 = dt / 6., h3 = dt / 3.;
    double *aux2 = k[3];
    int rc = launch_velocity_stage(p, U, stage(k[0], aux, {{1., U}}, h2), s);
    if (!rc) rc = launch_velocity_stage(p, aux, stage(k[1], aux2, {{1., U}}, h2), s);
    if (!rc) rc = launch_velocity_stage(p, aux2, stage(k[2], aux, {{1., U}}, dt), s);
    if (!rc) rc = launch_velocity_stage(p, aux, stage(nullptr, U, {{1., U}, {h6, k[0]}, {h3, k[1]}, {h3, k[2]}}, h6), s);
    return rc;
  }
  return rk_step(
      st.scheme, [&](const double *y, double tt, double *f) { return launch_velocity(p, y, tt, f, s); }, U,
      aux, k, st.len, t, dt, s);
}

}  // namespace

extern "C" {

int pda_stepper_create(pda_problem p, int scheme, pda_stepper *out) {
  if (!p || !out || scheme < PDA_FORWARD_EULER || scheme > PDA_MODULE_SSP3) {
    set_error("pda_stepper_create: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  if (p->mesh->n_sample != p->mesh->n_stencil) {
    set_error("pda_stepper_create: explicit stepping needs a full mesh (sample == stencil mesh)");
    return PDA_E_INVALID_ARG;
  }
  int rc = ensure_device(*p);
  if (rc) return rc;
  auto *st = new pda_stepper_s();
  st->prob = p;
  st->scheme = scheme;
  st->len = p->mesh->n_stencil * p->m;
  st->fused1d = steps1d_fits(*p, scheme, &st->periodic1d);
  const size_t bytes = sizeof(double) * size_t(st->len);
  hipError_t e = st->aux.alloc(bytes);
  for (int i = 0; i < 4 && e == hipSuccess; ++i) e = st->k[i].alloc(bytes);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st->cap, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error(std::string("pda_stepper_create: ") + hipGetErrorString(e));
    delete st;
    return PDA_E_HIP;
  }
  *out = st;
  return PDA_OK;
}

void pda_stepper_destroy(pda_stepper st) { delete st; }

int pda_stepper_advance(pda_stepper st, double *d_U, double t0, double dt, int64_t nsteps,
                        int use_graph, pda_stream stream) {
  if (!st || !d_U || nsteps < 0) {
    set_error("pda_stepper_advance: invalid argument");
    return PDA_E_INVALID_ARG;
  }
  hipStream_t s = hipStream_t(stream);
  if (st->fused1d) {
    // several steps per launch, the state alternating between U and the
    // stepper's buffer (a launch's workgroups read cells that other
    // workgroups of it write); one copy back when the count of launches is
    // odd.  use_graph does not change this path.
    const int kmax = steps1d_per_launch(*st->prob, st->scheme);
    double *a = d_U, *b = st->aux.as<double>();
    for (int64_t done = 0; done < nsteps;) {
      const int k = int(std::min<int64_t>(nsteps - done, kmax));
      const int rc = launch_steps1d(*st->prob, st->scheme, a, b, k, dt, st->periodic1d, s);
      if (rc) return rc;
      std::swap(a, b);
      done += k;
    }
    if (a != d_U) return lincomb(d_U, {{1., a}}, st->len, s);
    return PDA_OK;
  }
  if (!use_graph) {
    for (int64_t i = 0; i < nsteps; ++i) {
      const int rc = step_once(*st, d_U, t0 + double(i) * dt, dt, s);
      if (rc) return rc;
    }
    return PDA_OK;
  }
  // the problems' velocity does not depend on t, so one captured step
  // serves every step; it is re-recorded when U or dt change
  if (!st->exec || st->g_U != d_U || st->g_dt != dt) {
    if (st->exec) PDA_HIP_CHECK(hipGraphExecDestroy(st->exec));
    if (st->graph) PDA_HIP_CHECK(hipGraphDestroy(st->graph));
    st->exec = nullptr;
    st->graph = nullptr;
    PDA_HIP_CHECK(hipStreamBeginCapture(st->cap, hipStreamCaptureModeThreadLocal));
    const int rc = step_once(*st, d_U, t0, dt, st->cap);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st->cap, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) {
      set_error(std::string("pda_stepper_advance: capture: ") + hipGetErrorString(e));
      return PDA_E_HIP;
    }
    st->graph = g;
    PDA_HIP_CHECK(hipGraphInstantiate(&st->exec, st->graph, nullptr, nullptr, 0));
    st->g_U = d_U;
    st->g_dt = dt;
  }
  for (int64_t i = 0; i < nsteps; ++i) PDA_HIP_CHECK(hipGraphLaunch(st->exec, s));
  return PDA_OK;
}

}  // extern "C"
