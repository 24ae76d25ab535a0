// pda_rk.hpp -- explicit Runge-Kutta steps over a velocity callable, with
// pressio's stage formulas
// (tests_cpp/pressio/include/pressio/ode/impl/ode_explicit_stepper_without_mass_matrix.hpp:
//  forward Euler :168-187, SSPRK3 :230-281, RK4 :284-352) and the Python
// module's RK2 (pressiodemoapps/__init__.py:90-114).
#pragma once
#include <initializer_list>
#include <utility>

#include "pda_internal.hpp"

namespace pda {

// out[i] = sum of c_k * x_k[i] over up to five terms (out may alias an x_k)
int lincomb(double *out, std::initializer_list<std::pair<double, const double *>> terms, int64_t len,
            hipStream_t s);

// the Python module's stage updates, rounded as numpy rounds them
// (pressiodemoapps/__init__.py:90-185); v is the velocity just evaluated:
//   py_stage:       kv <- dt * kv (kept as k_i); out = x + a * k_i
//   py_ssp3_second: out = 3/4 y + 1/4 y1 + (1/4 dt) v
//   py_ssp3_final:  y = 1/3 (y + 2 y2 + (2 dt) v)
//   py_rk4_final:   y = y + (k1 + 2 k2 + 2 k3 + dt v) / 6
//   py_rk2_final:   y = y + (dt v) / 2 + k1 / 2
int py_stage(double *kv, const double *x, double *out, double dt, double a, int64_t n, hipStream_t s);
int py_ssp3_second(double *out, const double *y, const double *y1, const double *v, double dt, int64_t n,
                   hipStream_t s);
int py_ssp3_final(double *y, const double *y2, const double *v, double dt, int64_t n, hipStream_t s);
int py_rk4_final(double *y, const double *k1, const double *k2, const double *k3, const double *v, double dt,
                 int64_t n, hipStream_t s);
int py_rk2_final(double *y, const double *k1, const double *v, double dt, int64_t n, hipStream_t s);

// vel(y, t, f) evaluates f = velocity(y) at time t on stream s
template <class Vel>
int rk_step(int scheme, Vel &&vel, double *U, double *aux, double *const k[4], int64_t n, double t, double dt,
            hipStream_t s) {
  int rc;
#define PDA_TRY(x) \
  do {             \
    rc = (x);      \
    if (rc) return rc; \
  } while (0)
  double *k1 = k[0], *k2 = k[1], *k3 = k[2], *k4 = k[3];
  switch (scheme) {
    case PDA_FORWARD_EULER:
      // y_{n+1} = y_n + dt f(y_n)
      PDA_TRY(vel(U, t, k1));
      PDA_TRY(lincomb(U, {{1., U}, {dt, k1}}, n, s));
      break;
    case PDA_SSPRK3: {
      const double third = 1. / 3., twoThird = 2. / 3., quarter = 1. / 4., threeQuarter = 3. / 4.;
      PDA_TRY(vel(U, t, k1));
      PDA_TRY(lincomb(aux, {{1., U}, {dt, k1}}, n, s)); // u1 = u + dt f(u)
      PDA_TRY(vel(aux, t + dt, k1));
      PDA_TRY(lincomb(aux, {{quarter, aux}, {threeQuarter, U}, {quarter * dt, k1}}, n, s));
      PDA_TRY(vel(aux, t + dt / 2., k1));
      PDA_TRY(lincomb(U, {{third, U}, {twoThird, aux}, {twoThird * dt, k1}}, n, s));
      break;
    }
    case PDA_RK4: {
      const double h2 = dt / 2., h6 = dt / 6., h3 = dt / 3.;
      PDA_TRY(vel(U, t, k1));
      PDA_TRY(lincomb(aux, {{1., U}, {h2, k1}}, n, s));
      PDA_TRY(vel(aux, t + h2, k2));
      PDA_TRY(lincomb(aux, {{1., U}, {h2, k2}}, n, s));
      PDA_TRY(vel(aux, t + h2, k3));
      PDA_TRY(lincomb(aux, {{1., U}, {dt, k3}}, n, s));
      PDA_TRY(vel(aux, t + dt, k4));
      PDA_TRY(lincomb(U, {{1., U}, {h6, k1}, {h3, k2}, {h3, k3}, {h6, k4}}, n, s));
      break;
    }
    case PDA_RK2: {
      // pressiodemoapps/__init__.py:90-114: k1 = dt f(y), k2 = dt f(y + k1),
      // y = y + k2/2 + k1/2
      PDA_TRY(vel(U, t, k1));
      PDA_TRY(py_stage(k1, U, aux, dt, 1., n, s));
      PDA_TRY(vel(aux, t + dt, k2));
      PDA_TRY(py_rk2_final(U, k1, k2, dt, n, s));
      break;
    }
    case PDA_MODULE_RK4: {
      // pressiodemoapps/__init__.py:117-152: k_i = dt v_i, stages at
      // y + k1/2, y + k2/2, y + k3, then y + (k1 + 2 k2 + 2 k3 + k4) / 6
      const double half = 0.5;
      PDA_TRY(vel(U, t, k1));
      PDA_TRY(py_stage(k1, U, aux, dt, half, n, s));
      PDA_TRY(vel(aux, t + half * dt, k2));
      PDA_TRY(py_stage(k2, U, aux, dt, half, n, s));
      PDA_TRY(vel(aux, t + half * dt, k3));
      PDA_TRY(py_stage(k3, U, aux, dt, 1., n, s));
      PDA_TRY(vel(aux, t + dt, k4));
      PDA_TRY(py_rk4_final(U, k1, k2, k3, k4, dt, n, s));
      break;
    }
    case PDA_MODULE_SSP3: {
      // pressiodemoapps/__init__.py:155-185: y1 = y + dt v1;
      // y2 = 3/4 y + 1/4 y1 + 1/4 dt v2; y = 1/3 (y + 2 y2 + 2 dt v3),
      // velocities at t, t + dt, t + dt/2
      double *y2 = k2;
      PDA_TRY(vel(U, t, k1));
      PDA_TRY(py_stage(k1, U, aux, dt, 1., n, s));
      PDA_TRY(vel(aux, t + dt, k1));
      PDA_TRY(py_ssp3_second(y2, U, aux, k1, dt, n, s));
      PDA_TRY(vel(y2, t + dt * 0.5, k1));
      PDA_TRY(py_ssp3_final(U, y2, k1, dt, n, s));
      break;
    }
    default:
      set_error("stepper: unknown scheme");
      return PDA_E_INVALID_ARG;
  }
#undef PDA_TRY
  return PDA_OK;
}

}  // namespace pda
