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
      // k1 = dt f(y), k2 = dt f(y + k1), y += k2/2 + k1/2
      PDA_TRY(vel(U, t, k1));
      PDA_TRY(lincomb(aux, {{1., U}, {dt, k1}}, n, s));
      PDA_TRY(vel(aux, t + dt, k2));
      PDA_TRY(lincomb(U, {{1., U}, {0.5 * dt, k2}, {0.5 * dt, k1}}, n, s));
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
