// pda_physics.hpp -- device-side numerics of the finite-volume hot path:
// WENO3/WENO5 reconstruction (values and stencil gradients) and the face
// fluxes with their exact Jacobians for every family.
//
// Formulas follow the reference files cited per function; evaluation order
// is kept close to them, and FP64 throughout.
#pragma once
#include <hip/hip_runtime.h>

namespace pda {
namespace phys {

constexpr double kEs = 1.e-30;

// n / d from the hardware reciprocal estimate refined by two Newton steps
// and one residual correction: within an ulp of the rounded quotient for d
// and n / d in the normal range (the WENO weight sums, the densities and the
// Roe factors of the velocity's fluxes), without the IEEE division
// sequence's scaling for denormal and infinite operands.
__device__ __forceinline__ double fdiv(double n, double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = __fma_rn(r, __fma_rn(-d, r, 1.), r);
  r = __fma_rn(r, __fma_rn(-d, r, 1.), r);
  const double q = n * r;
  return __fma_rn(r, __fma_rn(-d, q, n), q);
}

// ---------------------------------------------------------------- WENO3
// impl/weno3.hpp:56-132.  q = {q_{i-1}, q_i, q_{i+1}, q_{i+2}}
__device__ __forceinline__ void weno3(const double *q, double &un, double &up) {
  constexpr double eps = 1.e-6, c13 = 1. / 3., c23 = 2. / 3.;
  {
    const double p0 = (-q[0] + 3. * q[1]) * 0.5;
    const double p1 = (q[1] + q[2]) * 0.5;
    const double B0 = (q[0] - q[1]) * (q[0] - q[1]);
    const double B1 = (q[1] - q[2]) * (q[1] - q[2]);
    const double a0 = c13 / (eps * eps + 2. * eps * B0 + B0 * B0);
    const double a1 = c23 / (eps * eps + 2. * eps * B1 + B1 * B1);
    const double s = 1. / (a0 + a1);
    un = (a0 * s) * p0 + (a1 * s) * p1;
  }
  {
    const double p0 = (q[1] + q[2]) * 0.5;
    const double p1 = (3. * q[2] - q[3]) * 0.5;
    const double B0 = (q[1] - q[2]) * (q[1] - q[2]);
    const double B1 = (q[2] - q[3]) * (q[2] - q[3]);
    const double a0 = c23 / (eps * eps + 2. * eps * B0 + B0 * B0);
    const double a1 = c13 / (eps * eps + 2. * eps * B1 + B1 * B1);
    const double s = 1. / (a0 + a1);
    up = (a0 * s) * p0 + (a1 * s) * p1;
  }
}

// impl/weno3.hpp:134-246: values plus d(face)/d(q_j), j = 0..3
__device__ __forceinline__ void weno3_grad(const double *q, double &un, double &up,
                                           double *gn, double *gp) {
#pragma clang fp contract(off)
  constexpr double eps = 1.e-6, c13 = 1. / 3., c23 = 2. / 3.;
  {
    const double qim1 = q[0], qi = q[1], qip1 = q[2];
    const double p0 = (-qim1 + 3. * qi) * 0.5, p1 = (qi + qip1) * 0.5;
    const double d0 = qim1 - qi, d1 = qi - qip1;
    const double B0 = d0 * d0, B1 = d1 * d1;
    const double a0 = c13 / (eps * eps + 2. * eps * B0 + B0 * B0);
    const double a1 = c23 / (eps * eps + 2. * eps * B1 + B1 * B1);
    const double s = 1. / (a0 + a1);
    const double e0 = B0 + eps, e1 = B1 + eps;
    const double e03 = e0 * e0 * e0, e13 = e1 * e1 * e1;
    const double S = 2. / (3. * (e1 * e1)) + 1. / (3. * (e0 * e0));
    const double S2 = S * S;
    const double da0[3] = {-4. * d0 / (3. * e03), 4. * d0 / (3. * e03), 0.};
    const double da1[3] = {0., -8. * d1 / (3. * e13), 8. * d1 / (3. * e13)};
    const double ds[3] = {(4. * d0) / (3. * e03 * S2),
                          -((4. * d0) / (3. * e03) - (8. * d1) / (3. * e13)) / S2,
                          -(8. * d1) / (3. * S2 * e13)};
    const double dp0[3] = {-0.5, 1.5, 0.}, dp1[3] = {0., 0.5, 0.5};
    const double w0 = a0 * s, w1 = a1 * s;
    for (int i = 0; i < 3; ++i) {
      double v = (da0[i] * s + ds[i] * a0) * p0 + dp0[i] * w0;
      v += (da1[i] * s + ds[i] * a1) * p1 + dp1[i] * w1;
      gn[i] = v;
    }
    gn[3] = 0.;
    un = w0 * p0 + w1 * p1;
  }
  {
    const double qi = q[1], qip1 = q[2], qip2 = q[3];
    const double p0 = (qi + qip1) * 0.5, p1 = (3. * qip1 - qip2) * 0.5;
    const double d0 = qi - qip1, d1 = qip1 - qip2;
    const double B0 = d0 * d0, B1 = d1 * d1;
    const double a0 = c23 / (eps * eps + 2. * eps * B0 + B0 * B0);
    const double a1 = c13 / (eps * eps + 2. * eps * B1 + B1 * B1);
    const double s = 1. / (a0 + a1);
    const double w0 = a0 * s, w1 = a1 * s;
    const double e0 = B0 + eps, e1 = B1 + eps;
    const double e03 = e0 * e0 * e0, e13 = e1 * e1 * e1;
    const double S = 1. / (3. * (e1 * e1)) + 2. / (3. * (e0 * e0));
    const double S2 = S * S;
    const double da0[3] = {-8. * d0 / (3. * e03), 8. * d0 / (3. * e03), 0.};
    const double da1[3] = {0., -4. * d1 / (3. * e13), 4. * d1 / (3. * e13)};
    const double ds[3] = {8. * d0 / (3. * e03 * S2),
                          -(8. * d0 / (3. * e03) - 4. * d1 / (3. * e13)) / S2,
                          -4. * d1 / (3. * S2 * e13)};
    const double dp0[3] = {0.5, 0.5, 0.}, dp1[3] = {0., 1.5, -0.5};
    for (int i = 0; i < 3; ++i) {
      double v = (da0[i] * s + ds[i] * a0) * p0 + dp0[i] * w0;
      v += (da1[i] * s + ds[i] * a1) * p1 + dp1[i] * w1;
      gp[i + 1] = v;
    }
    gp[0] = 0.;
    up = w0 * p0 + w1 * p1;
  }
}

// ---------------------------------------------------------------- WENO5
// impl/weno5.hpp:56-178.  q = {q_{i-2}, ..., q_{i+3}}
__device__ __forceinline__ void weno5(const double *q, double &un, double &up) {
  constexpr double eps = 1.e-6, c1312 = 13. / 12., c16 = 1. / 6.;
  {
    const double p0 = (2. * q[0] - 7. * q[1] + 11. * q[2]) * c16;
    const double p1 = (-q[1] + 5. * q[2] + 2. * q[3]) * c16;
    const double p2 = (2. * q[2] + 5. * q[3] - q[4]) * c16;
    const double t0 = q[0] - 2. * q[1] + q[2], s0 = q[0] - 4. * q[1] + 3. * q[2];
    const double t1 = q[1] - 2. * q[2] + q[3], s1 = q[1] - q[3];
    const double t2 = q[2] - 2. * q[3] + q[4], s2 = 3. * q[2] - 4. * q[3] + q[4];
    const double B0 = c1312 * t0 * t0 + 0.25 * s0 * s0;
    const double B1 = c1312 * t1 * t1 + 0.25 * s1 * s1;
    const double B2 = c1312 * t2 * t2 + 0.25 * s2 * s2;
    const double a0 = 0.1 / (eps * eps + 2. * eps * B0 + B0 * B0);
    const double a1 = 0.6 / (eps * eps + 2. * eps * B1 + B1 * B1);
    const double a2 = 0.3 / (eps * eps + 2. * eps * B2 + B2 * B2);
    const double s = 1. / (a0 + a1 + a2);
    un = (a0 * s) * p0 + (a1 * s) * p1 + (a2 * s) * p2;
  }
  {
    const double p0 = (-q[1] + 5. * q[2] + 2. * q[3]) * c16;
    const double p1 = (2. * q[2] + 5. * q[3] - q[4]) * c16;
    const double p2 = (11. * q[3] - 7. * q[4] + 2. * q[5]) * c16;
    const double t0 = q[1] - 2. * q[2] + q[3], s0 = q[1] - 4. * q[2] + 3. * q[3];
    const double t1 = q[2] - 2. * q[3] + q[4], s1 = q[2] - q[4];
    const double t2 = q[3] - 2. * q[4] + q[5], s2 = 3. * q[3] - 4. * q[4] + q[5];
    const double B0 = c1312 * t0 * t0 + 0.25 * s0 * s0;
    const double B1 = c1312 * t1 * t1 + 0.25 * s1 * s1;
    const double B2 = c1312 * t2 * t2 + 0.25 * s2 * s2;
    const double a0 = 0.3 / (eps * eps + 2. * eps * B0 + B0 * B0);
    const double a1 = 0.6 / (eps * eps + 2. * eps * B1 + B1 * B1);
    const double a2 = 0.1 / (eps * eps + 2. * eps * B2 + B2 * B2);
    const double s = 1. / (a0 + a1 + a2);
    up = (a0 * s) * p0 + (a1 * s) * p1 + (a2 * s) * p2;
  }
}

// Same reconstruction with the normalised weights formed with one division:
// alpha_k / sum(alpha) = N_k / sum(N) with N_k = d_k * prod_{j != k} D_j and
// D_k = eps^2 + 2 eps B_k + B_k^2 (the reference's denominator).  Equal to
// weno5() in exact arithmetic; the results differ by a few ulps.
//
// D of a side window (far, mid, near), near = the reconstructed cell:
// B = 13/12 (far - 2 mid + near)^2 + 1/4 (far - 4 mid + 3 near)^2.  The
// operations are fixed (explicit FMAs, far + near summed first), so that a
// window gives the same D whichever of its two cells' reconstructions reads
// it: the low-side and high-side values of a cell share their windows.
__device__ __forceinline__ double w5_dside(double far, double mid, double near) {
  constexpr double eps = 1.e-6, c1312 = 13. / 12.;
  const double t = __fma_rn(-2., mid, far + near);
  const double s = __fma_rn(-4., mid, __fma_rn(3., near, far));
  const double B = __fma_rn(c1312 * t, t, (0.25 * s) * s);
  return __fma_rn(B, B, __fma_rn(2. * eps, B, eps * eps));
}
// D of the centred window (lo, mid, hi): B = 13/12 (lo - 2 mid + hi)^2 +
// 1/4 (lo - hi)^2, the same for the window read in either direction
__device__ __forceinline__ double w5_dmid(double lo, double mid, double hi) {
  constexpr double eps = 1.e-6, c1312 = 13. / 12.;
  const double t = __fma_rn(-2., mid, lo + hi);
  const double s = lo - hi;
  const double B = __fma_rn(c1312 * t, t, (0.25 * s) * s);
  return __fma_rn(B, B, __fma_rn(2. * eps, B, eps * eps));
}
// reconstruction at the face of cell q2 towards q3 from the windows' D
__device__ __forceinline__ double w5_value(double q0, double q1, double q2, double q3, double q4,
                                           double D0, double D1, double D2) {
  constexpr double c16 = 1. / 6.;
  const double p0 = (2. * q0 - 7. * q1 + 11. * q2) * c16;
  const double p1 = (-q1 + 5. * q2 + 2. * q3) * c16;
  const double p2 = (2. * q2 + 5. * q3 - q4) * c16;
  const double n0 = 0.1 * (D1 * D2), n1 = 0.6 * (D0 * D2), n2 = 0.3 * (D0 * D1);
  return fdiv(n0 * p0 + n1 * p1 + n2 * p2, n0 + n1 + n2);
}
__device__ __forceinline__ double weno5_side_1div(double q0, double q1, double q2, double q3,
                                                  double q4) {
  return w5_value(q0, q1, q2, q3, q4, w5_dside(q0, q1, q2), w5_dmid(q1, q2, q3),
                  w5_dside(q4, q3, q2));
}

// q = {q_{i-2}, ..., q_{i+3}}; un, up: left and right states of face i+1/2
__device__ __forceinline__ void weno5_1div(const double *q, double &un, double &up) {
  un = weno5_side_1div(q[0], q[1], q[2], q[3], q[4]);
  // mirrored stencil: the right state of the face is the left-biased
  // reconstruction of the reversed sequence
  up = weno5_side_1div(q[5], q[4], q[3], q[2], q[1]);
}

// Both face values of one cell from q = cells i-2 .. i+2: hi = the left
// state of face i+1/2, lo = the right state of face i-1/2.  Bit-identical to
// the two weno5_side_1div calls of weno5_1div at those faces; the three
// windows' D are formed once.
__device__ __forceinline__ void weno5_cell_1div(double q0, double q1, double q2, double q3,
                                                double q4, double &hi, double &lo) {
  const double D0 = w5_dside(q0, q1, q2), D1 = w5_dmid(q1, q2, q3), D2 = w5_dside(q4, q3, q2);
  hi = w5_value(q0, q1, q2, q3, q4, D0, D1, D2);
  lo = w5_value(q4, q3, q2, q1, q0, D2, D1, D0);
}

// WENO3 counterpart of weno5_1div (impl/weno3.hpp:56-132).  D of the
// two-cell window (a, b): B = (a - b)^2, the same in either direction.
__device__ __forceinline__ double w3_d(double a, double b) {
  constexpr double eps = 1.e-6;
  const double B = (a - b) * (a - b);
  return __fma_rn(B, B, __fma_rn(2. * eps, B, eps * eps));
}
// the face values of cell b (neighbours a below, c above) from the windows'
// D: hi at the face towards c, lo at the face towards a
__device__ __forceinline__ double w3_hi(double a, double b, double c, double Dab, double Dbc) {
  constexpr double c13 = 1. / 3., c23 = 2. / 3.;
  const double p0 = (-a + 3. * b) * 0.5, p1 = (b + c) * 0.5;
  const double n0 = c13 * Dbc, n1 = c23 * Dab;
  return fdiv(n0 * p0 + n1 * p1, n0 + n1);
}
__device__ __forceinline__ double w3_lo(double a, double b, double c, double Dab, double Dbc) {
  constexpr double c13 = 1. / 3., c23 = 2. / 3.;
  const double p0 = (a + b) * 0.5, p1 = (3. * b - c) * 0.5;
  const double n0 = c23 * Dbc, n1 = c13 * Dab;
  return fdiv(n0 * p0 + n1 * p1, n0 + n1);
}
__device__ __forceinline__ void weno3_1div(const double *q, double &un, double &up) {
  un = w3_hi(q[0], q[1], q[2], w3_d(q[0], q[1]), w3_d(q[1], q[2]));
  up = w3_lo(q[1], q[2], q[3], w3_d(q[1], q[2]), w3_d(q[2], q[3]));
}
// both face values of cell b, bit-identical to weno3_1div's at its faces
__device__ __forceinline__ void weno3_cell_1div(double a, double b, double c, double &hi, double &lo) {
  const double Dab = w3_d(a, b), Dbc = w3_d(b, c);
  hi = w3_hi(a, b, c, Dab, Dbc);
  lo = w3_lo(a, b, c, Dab, Dbc);
}

// one side of impl/weno5.hpp:180-434.  c = linear weights {d0,d1,d2};
// P[k][i] = d p_k / d q_i; T/S = the two smoothness building blocks with
// their gradients; returns value, writes grad[0..4]
// cube with one rounding in almost all cases (the reference uses pow(x, 3.))
__device__ __forceinline__ double cube(double e) {
  const double s = e * e;
  const double s_lo = __fma_rn(e, e, -s);
  return __fma_rn(s, e, s_lo * e);
}

// The gradient terms cancel strongly where the smoothness indicators are
// small (d alpha / d beta ~ 1/(eps+beta)^3), so this path keeps the
// reference's expression forms and is compiled without FMA contraction.
// dd = {1, 6, 3} or {3, 6, 1}: alpha_k = (dd_k/10) / (eps+beta_k)^2 and
// d alpha_k / d beta_k = -dd_k / (5 (eps+beta_k)^3), impl/weno5.hpp:313-323.
__device__ __forceinline__ double weno5_side_grad(const double p[3], const double B[3],
                                                  const double dB[3][5], const double dp[3][5],
                                                  const double c[3], const double dd[3],
                                                  double *grad) {
#pragma clang fp contract(off)
  constexpr double eps = 1.e-6;
  double a[3], e3[3];
  for (int k = 0; k < 3; ++k) {
    const double e = eps + B[k];
    a[k] = c[k] / (e * e);
    e3[k] = cube(B[k] + eps);
  }
  const double s = 1. / (a[0] + a[1] + a[2]);
  const double S = a[2] + a[1] + a[0];
  const double mS2 = -1. / (S * S);
  const double w0 = a[0] * s, w1 = a[1] * s, w2 = a[2] * s;
  for (int i = 0; i < 5; ++i) {
    const double da0 = -dd[0] / (5. * e3[0]) * dB[0][i];
    const double da1 = -dd[1] / (5. * e3[1]) * dB[1][i];
    const double da2 = -dd[2] / (5. * e3[2]) * dB[2][i];
    const double ds = mS2 * (da0 + da1 + da2);
    double v = (da0 * s + ds * a[0]) * p[0] + dp[0][i] * w0;
    v += (da1 * s + ds * a[1]) * p[1] + dp[1][i] * w1;
    v += (da2 * s + ds * a[2]) * p[2] + dp[2][i] * w2;
    grad[i] = v;
  }
  return w0 * p[0] + w1 * p[1] + w2 * p[2];
}

// values plus d(face)/d(q_j), j = 0..5 (gn[5] = gp[0] = 0)
__device__ __forceinline__ void weno5_grad(const double *q, double &un, double &up,
                                           double *gn, double *gp) {
#pragma clang fp contract(off)
  constexpr double c1312 = 13. / 12., c16 = 1. / 6.;
  {
    const double qim2 = q[0], qim1 = q[1], qi = q[2], qip1 = q[3], qip2 = q[4];
    const double p[3] = {(2. * qim2 - 7. * qim1 + 11. * qi) * c16,
                         (-qim1 + 5. * qi + 2. * qip1) * c16,
                         (2. * qi + 5. * qip1 - qip2) * c16};
    const double t0 = qim2 - 2. * qim1 + qi, s0 = qim2 - 4. * qim1 + 3. * qi;
    const double t1 = qim1 - 2. * qi + qip1, s1 = qim1 - qip1;
    const double t2 = qi - 2. * qip1 + qip2, s2 = 3. * qi - 4. * qip1 + qip2;
    const double B[3] = {c1312 * (t0 * t0) + 0.25 * (s0 * s0),
                         c1312 * (t1 * t1) + 0.25 * (s1 * s1),
                         c1312 * (t2 * t2) + 0.25 * (s2 * s2)};
    const double dB[3][5] = {
        {13. * t0 / 6. + s0 / 2., -13. * t0 / 3. - 2. * s0, 13. * t0 / 6. + 3. * s0 / 2., 0., 0.},
        {0., 13. * t1 / 6. + s1 / 2., -13. * t1 / 3., 13. * t1 / 6. - s1 / 2., 0.},
        {0., 0., 13. * t2 / 6. + 3. * s2 / 2., -13. * t2 / 3. - 2. * s2, 13. * t2 / 6. + s2 / 2.}};
    const double dp[3][5] = {{1. / 3., -7. / 6., 11. / 6., 0., 0.},
                             {0., -1. / 6., 5. / 6., 1. / 3., 0.},
                             {0., 0., 1. / 3., 5. / 6., -1. / 6.}};
    const double c[3] = {0.1, 0.6, 0.3}, dd[3] = {1., 6., 3.};
    un = weno5_side_grad(p, B, dB, dp, c, dd, gn);
    gn[5] = 0.;
  }
  {
    const double qim1 = q[1], qi = q[2], qip1 = q[3], qip2 = q[4], qip3 = q[5];
    const double p[3] = {(-qim1 + 5. * qi + 2. * qip1) * c16,
                         (2. * qi + 5. * qip1 - qip2) * c16,
                         (11. * qip1 - 7. * qip2 + 2. * qip3) * c16};
    const double t0 = qim1 - 2. * qi + qip1, s0 = qim1 - 4. * qi + 3. * qip1;
    const double t1 = qi - 2. * qip1 + qip2, s1 = qi - qip2;
    const double t2 = qip1 - 2. * qip2 + qip3, s2 = 3. * qip1 - 4. * qip2 + qip3;
    const double B[3] = {c1312 * (t0 * t0) + 0.25 * (s0 * s0),
                         c1312 * (t1 * t1) + 0.25 * (s1 * s1),
                         c1312 * (t2 * t2) + 0.25 * (s2 * s2)};
    const double dB[3][5] = {
        {s0 / 2. + 13. * t0 / 6., -2. * s0 - 13. * t0 / 3., 3. * s0 / 2. + 13. * t0 / 6., 0., 0.},
        {0., 13. * t1 / 6. + s1 / 2., -13. * t1 / 3., 13. * t1 / 6. - s1 / 2., 0.},
        {0., 0., 13. * t2 / 6. + 3. * s2 / 2., -13. * t2 / 3. - 2. * s2, 13. * t2 / 6. + s2 / 2.}};
    const double dp[3][5] = {{-1. / 6., 5. / 6., 1. / 3., 0., 0.},
                             {0., 1. / 3., 5. / 6., -1. / 6., 0.},
                             {0., 0., 11. / 6., -7. / 6., 1. / 3.}};
    const double c[3] = {0.3, 0.6, 0.1}, dd[3] = {3., 6., 1.};
    up = weno5_side_grad(p, B, dB, dp, c, dd, gp + 1);
    gp[0] = 0.;
  }
}

// ------------------------------------------------------------ Euler
// Rusanov flux for 1, 2, 3 dimensions (M = DIM+2 conserved variables),
// impl/euler_rusanov_flux_values_function.hpp:54-208.  A = face normal axis.
template <int DIM, int A>
__device__ __forceinline__ void euler_flux(const double *qL, const double *qR, double g,
                                           double *F) {
  constexpr int M = DIM + 2;
  const double half = 0.5;
  double velL[DIM], velR[DIM];
  const double rL = qL[0], rR = qR[0];
  // one reciprocal per side (the reference divides per component)
  const double iL = fdiv(1., rL + kEs), iR = fdiv(1., rR + kEs);
  double v2L = 0., v2R = 0.;
  for (int d = 0; d < DIM; ++d) {
    velL[d] = qL[1 + d] * iL;
    velR[d] = qR[1 + d] * iR;
    v2L += velL[d] * velL[d];
    v2R += velR[d] * velR[d];
  }
  const double unL = velL[A], unR = velR[A];
  const double pL = (g - 1.) * (qL[M - 1] - half * rL * v2L);
  const double pR = (g - 1.) * (qR[M - 1] - half * rR * v2R);
  const double HL = (qL[M - 1] + pL) * iL, HR = (qR[M - 1] + pR) * iR;
  double FL[M], FR[M];
  FL[0] = rL * unL;
  FR[0] = rR * unR;
  for (int d = 0; d < DIM; ++d) {
    FL[1 + d] = rL * unL * velL[d] + (d == A ? pL : 0.);
    FR[1 + d] = rR * unR * velR[d] + (d == A ? pR : 0.);
  }
  FL[M - 1] = rL * unL * HL;
  FR[M - 1] = rR * unR * HR;
  const double RT = sqrt(fdiv(rR, rL));
  const double inv = fdiv(1., 1. + RT);
  double vr2 = 0.;
  for (int d = 0; d < DIM; ++d) {
    const double v = (velL[d] + RT * velR[d]) * inv;
    vr2 += v * v;
  }
  const double H = (HL + RT * HR) * inv;
  const double a = sqrt((g - 1.) * (H - half * vr2));
  const double smax = (DIM == 1 ? fabs((velL[0] + RT * velR[0]) * inv) : sqrt(vr2)) + a;
  for (int i = 0; i < M; ++i) F[i] = half * (FL[i] + FR[i] + smax * (qL[i] - qR[i]));
}

// exact Jacobians of the Rusanov flux w.r.t. the left and right states,
// including the derivative of the Roe-averaged wave speed,
// impl/euler_rusanov_flux_jacobian_function.hpp:54-406.
//
// Each side's M x M block has the form
//   J_s = 1/2 A(q_s) +- 1/2 smax I + 1/2 grad_s (qL - qR)^T-outer,
// i.e. J_s(i,j) = 1/2 A_s(i,j) + [i==j] dsm + 1/2 grad_s[j] dq[i], where A_s
// is the physical flux Jacobian along axis A at the side's state.  EulerSide
// keeps the factors (side velocity, |v|^2, total enthalpy, gradient of smax,
// dq and the signed diagonal term); entries are formed on demand, so a
// kernel can stream a block to memory without holding all M*M values.
template <int DIM> struct EulerSide {
  double vel[DIM], V2, H, grad[DIM + 2], dq[DIM + 2], dsm;
};

template <int DIM, int A>
__device__ __forceinline__ void euler_sides(const double *qL, const double *qR, double gamma,
                                            EulerSide<DIM> &L, EulerSide<DIM> &R) {
  constexpr int M = DIM + 2;
  const double half = 0.5, gm1 = gamma - 1.;
  const double igm1 = 1. / gm1;
  double aa[2];
  {
    const double iL = 1. / (qL[0] + kEs), iR = 1. / (qR[0] + kEs);
    L.V2 = 0.;
    R.V2 = 0.;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      L.vel[d] = qL[1 + d] * iL;
      R.vel[d] = qR[1 + d] * iR;
      L.V2 += L.vel[d] * L.vel[d];
      R.V2 += R.vel[d] * R.vel[d];
    }
    const double pL = gm1 * (qL[M - 1] - half * qL[0] * L.V2);
    const double pR = gm1 * (qR[M - 1] - half * qR[0] * R.V2);
    L.H = (qL[M - 1] + pL) * iL;
    R.H = (qR[M - 1] + pR) * iR;
    aa[0] = gm1 * (L.H - half * L.V2);
    aa[1] = gm1 * (R.H - half * R.V2);
  }
  const double rL = qL[0], rR = qR[0];
  const double r = sqrt(rR * rL), RT = sqrt(rR / rL);
  const double inv = 1. / (1. + RT);
  double vR[DIM];
  double vr2 = 0.;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    vR[d] = (L.vel[d] + RT * R.vel[d]) * inv;
    vr2 += vR[d] * vR[d];
  }
  const double Hr = (L.H + RT * R.H) * inv;
  const double a = sqrt(gm1 * (Hr - half * vr2));
  const double ia = 1. / a;
  const double smax = (DIM == 1 ? fabs(vR[0]) : sqrt(vr2)) + a;
  const double VMagSqrRoe = vr2 + kEs;
  const double rs = 1. / sqrt(VMagSqrRoe);
  double rel[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) rel[d] = vR[d] * rs;
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    EulerSide<DIM> &S = s ? R : L;
    const double rX = s ? rR : rL;
    double VV = 0.;
#pragma unroll
    for (int d = 0; d < DIM; ++d) VV += S.vel[d] * vR[d];
    const double f = 1. / (rX + r);
    double g0 = 0.;
#pragma unroll
    for (int d = 0; d < DIM; ++d) g0 -= half * (S.vel[d] + vR[d]) * rel[d];
    g0 += half * gm1 * ia *
          (half * (VMagSqrRoe + VV) + half * (S.H - Hr) - aa[s] * igm1 + half * (gamma - 2.) * S.V2);
    S.grad[0] = f * g0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) S.grad[1 + d] = f * (rel[d] - half * (gm1 * (vR[d] + gm1 * S.vel[d])) * ia);
    S.grad[M - 1] = half * f * gamma * gm1 * ia;
#pragma unroll
    for (int i = 0; i < M; ++i) S.dq[i] = qL[i] - qR[i];
    S.dsm = s ? -half * smax : half * smax;
  }
}

// entry (i, j) of one side's block; i and j are compile-time after unrolling
template <int DIM, int A>
__device__ __forceinline__ double euler_side_entry(const EulerSide<DIM> &S, double gamma, int i,
                                                   int j) {
  constexpr int M = DIM + 2;
  const double half = 0.5, gm1 = gamma - 1.;
  const double uX = S.vel[A];
  double v;
  if (i == 0) {
    v = (j == 1 + A) ? half : 0.;
  } else if (i <= DIM) {
    const int ii = i - 1;
    if (j == 0) {
      v = half * ((ii == A ? half * gm1 * S.V2 : 0.) - S.vel[ii] * uX);
    } else if (j <= DIM) {
      const int d = j - 1;
      double w = (d == A ? S.vel[ii] : 0.) - (ii == A ? gm1 * S.vel[d] : 0.);
      if (d == ii) w += uX;
      v = half * w;
    } else {
      v = (ii == A ? half * gm1 : 0.);
    }
  } else {
    if (j == 0) v = half * ((half * gm1 * S.V2 - S.H) * uX);
    else if (j <= DIM) v = half * ((j - 1 == A ? S.H : 0.) - gm1 * S.vel[j - 1] * uX);
    else v = half * gamma * uX;
  }
  if (i == j) v += S.dsm;
  return v + half * S.grad[j] * S.dq[i];
}

// JL, JR row-major M x M
template <int DIM, int A>
__device__ __forceinline__ void euler_fjac(const double *qL, const double *qR, double gamma,
                                           double *JL, double *JR) {
  constexpr int M = DIM + 2;
  EulerSide<DIM> L, R;
  euler_sides<DIM, A>(qL, qR, gamma, L, R);
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int j = 0; j < M; ++j) {
      JL[i * M + j] = euler_side_entry<DIM, A>(L, gamma, i, j);
      JR[i * M + j] = euler_side_entry<DIM, A>(R, gamma, i, j);
    }
}

// ------------------------------------------------------- shallow water
// impl/swe_rusanov_flux_values_function.hpp:54-97
template <int A>
__device__ __forceinline__ void swe_flux(const double *qL, const double *qR, double g, double *F) {
  const double hL = qL[0], uL = fdiv(qL[1], hL + kEs), vL = fdiv(qL[2], hL + kEs);
  const double hR = qR[0], uR = fdiv(qR[1], hR + kEs), vR = fdiv(qR[2], hR + kEs);
  const double unL = A == 0 ? uL : vL, unR = A == 0 ? uR : vR;
  const double pL = 0.5 * g * hL * hL, pR = 0.5 * g * hR * hR;
  double FL[3], FR[3];
  FL[0] = hL * unL; FL[1] = hL * unL * uL + (A == 0 ? pL : 0.); FL[2] = hL * unL * vL + (A == 1 ? pL : 0.);
  FR[0] = hR * unR; FR[1] = hR * unR * uR + (A == 0 ? pR : 0.); FR[2] = hR * unR * vR + (A == 1 ? pR : 0.);
  const double hm = 0.5 * (hL + hR);
  const double sL = sqrt(hL), sR = sqrt(hR);
  const double um = fdiv(unL * sL + unR * sR, sL + sR + kEs);
  const double c = sqrt(g * hm);
  // sqrt(x * x) is |x| exactly in IEEE arithmetic (no overflow here)
  const double smax = fabs(um) + fabs(c);
  for (int i = 0; i < 3; ++i) F[i] = 0.5 * (FL[i] + FR[i] + smax * (qL[i] - qR[i]));
}

// impl/swe_rusanov_flux_jacobian_function.hpp:54-137
template <int A>
__device__ __forceinline__ void swe_fjac(const double *qL, const double *qR, double g,
                                         double *JL, double *JR) {
  const double n0 = A == 0 ? 1. : 0., n1 = A == 1 ? 1. : 0.;
  const double hL = qL[0], uL = qL[1] / (hL + kEs), vL = qL[2] / (hL + kEs);
  const double hR = qR[0], uR = qR[1] / (hR + kEs), vR = qR[2] / (hR + kEs);
  const double unL = A == 0 ? uL : vL, unR = A == 0 ? uR : vR;
  const double hm = 0.5 * (hL + hR);
  const double sL = sqrt(hL), sR = sqrt(hR);
  const double um = (unL * sL + unR * sR) / (sL + sR + kEs);
  const double smax = fabs(um) + fabs(sqrt(g * hm));
  const double termL = (n0 * qL[1] + n1 * qL[2]) / (qL[0] * qL[0]);
  const double termR = (n0 * qR[1] + n1 * qR[2]) / (qR[0] * qR[0]);
  const double hsun = sL * unL + sR * unR + kEs;
  const double ahs = fabs(hsun), ssum = sL + sR;
  const double gterm = g / (2.8284271247461903 * sqrt(g * (hL + hR)));
  double dL[3], dR[3];
  dL[0] = -ahs / (2. * sL * (ssum * ssum)) + (0.5 * unL / sL - sL * termL) * hsun / (ssum * ahs) + gterm;
  dL[1] = n0 * hsun / (sL * ssum * ahs);
  dL[2] = n1 * hsun / (sL * ssum * ahs);
  dR[0] = -ahs / (2. * sR * (ssum * ssum)) + (0.5 * unR / sR - sR * termR) * hsun / (ssum * ahs) + gterm;
  dR[1] = n0 * hsun / (sR * ssum * ahs);
  dR[2] = n1 * hsun / (sR * ssum * ahs);
  const double dq0 = qR[0] - qL[0], dq1 = qR[1] - qL[1], dq2 = qR[2] - qL[2];
  JL[0] = -0.5 * dL[0] * dq0 + 0.5 * (n0 * uL + n1 * vL - qL[0] * termL) + 0.5 * smax;
  JL[1] = 0.5 * n0 - 0.5 * dL[1] * dq0;
  JL[2] = 0.5 * n1 - 0.5 * dL[2] * dq0;
  JL[3] = 0.5 * (g * n0 * qL[0] - qL[1] * termL) - 0.5 * dL[0] * dq1;
  JL[4] = n0 * uL + 0.5 * n1 * vL + 0.5 * smax - 0.5 * dL[1] * dq1;
  JL[5] = 0.5 * n1 * uL - 0.5 * dL[2] * dq1;
  JL[6] = 0.5 * (g * n1 * qL[0] - qL[2] * termL) - 0.5 * dL[0] * dq2;
  JL[7] = 0.5 * n0 * vL - 0.5 * dL[1] * dq2;
  JL[8] = n1 * vL + 0.5 * n0 * uL + 0.5 * smax - 0.5 * dL[2] * dq2;
  JR[0] = -0.5 * dR[0] * dq0 + 0.5 * (n0 * uR + n1 * vR - qR[0] * termR) - 0.5 * smax;
  JR[1] = 0.5 * n0 - 0.5 * dR[1] * dq0;
  JR[2] = 0.5 * n1 - 0.5 * dR[2] * dq0;
  JR[3] = 0.5 * (g * n0 * qR[0] - qR[1] * termR) - 0.5 * dR[0] * dq1;
  JR[4] = n0 * uR + 0.5 * n1 * vR - 0.5 * smax - 0.5 * dR[1] * dq1;
  JR[5] = 0.5 * n1 * uR - 0.5 * dR[2] * dq1;
  JR[6] = 0.5 * (g * n1 * qR[0] - qR[2] * termR) - 0.5 * dR[0] * dq2;
  JR[7] = 0.5 * n0 * vR - 0.5 * dR[1] * dq2;
  JR[8] = n1 * vR + 0.5 * n0 * uR - 0.5 * smax - 0.5 * dR[2] * dq2;
}

// ------------------------------------------------------------ Burgers
// impl/advection_diffusion_2d_flux_functions.hpp:54-114
template <int A>
__device__ __forceinline__ void burgers_flux(const double *qL, const double *qR, double *F) {
  const double a0 = fmax(fabs(qL[0]), fabs(qR[0]));
  const double a1 = fmax(fabs(qL[1]), fabs(qR[1]));
  const double cross = qL[0] * qL[1] + qR[0] * qR[1];
  double f0 = a0 * (qL[0] - qR[0]);
  double f1 = a1 * (qL[1] - qR[1]);
  if (A == 0) {
    f0 += qL[0] * qL[0] + qR[0] * qR[0];
    f1 += cross;
  } else {
    f0 += cross;
    f1 += qL[1] * qL[1] + qR[1] * qR[1];
  }
  F[0] = f0 * 0.25;
  F[1] = f1 * 0.25;
}

template <int A>
__device__ __forceinline__ void burgers_fjac(const double *qL, const double *qR, double *JL,
                                             double *JR) {
  const double n0 = A == 0 ? 1. : 0., n1 = A == 1 ? 1. : 0.;
  if (fabs(qL[0]) > fabs(qR[0])) {
    JL[0] = (2. * qL[0] - qR[0]) * copysign(1., qL[0]) + n0 * 2. * qL[0] + n1 * qL[1];
    JR[0] = n0 * 2. * qR[0] + n1 * qR[1] - fabs(qL[0]);
  } else {
    JL[0] = n0 * 2. * qL[0] + n1 * qL[1] + fabs(qR[0]);
    JR[0] = (qL[0] - 2. * qR[0]) * copysign(1., qR[0]) + n0 * 2. * qR[0] + n1 * qR[1];
  }
  if (fabs(qL[1]) > fabs(qR[1])) {
    JL[3] = (2. * qL[1] - qR[1]) * copysign(1., qL[1]) + n0 * qL[0] + n1 * 2. * qL[1];
    JR[3] = n0 * qR[0] + n1 * 2. * qR[1] - fabs(qL[1]);
  } else {
    JL[3] = n0 * qL[0] + n1 * 2. * qL[1] + fabs(qR[1]);
    JR[3] = (qL[1] - 2. * qR[1]) * copysign(1., qR[1]) + n0 * qR[0] + n1 * 2. * qR[1];
  }
  JL[0] *= 0.25; JR[0] *= 0.25; JL[3] *= 0.25; JR[3] *= 0.25;
  JL[1] = n1 * qL[0] * 0.25;
  JL[2] = n0 * qL[1] * 0.25;
  JR[1] = n1 * qR[0] * 0.25;
  JR[2] = n0 * qR[1] * 0.25;
}

}  // namespace phys
}  // namespace pda
