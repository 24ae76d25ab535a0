/*
 * pda_oracle.c -- CPU restatement of pressio-demoapps' finite-volume hot path.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the checker for the HIP product.
 * It is compiled into oracle/liboracle.so and loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; nothing in
 * the product links or calls it.
 *
 * It mirrors the reference loop structure on purpose (one row at a time,
 * per-row ghost arrays, per-scalar CSR search on every Jacobian update) so
 * that it is also a faithful CPU proxy of the reference cost.
 *
 * Every function cites the reference file it restates; paths are relative
 * to the pressio-demoapps source tree (include/pressiodemoapps/...).
 * Compile with -ffp-contract=off so that evaluation order matches the
 * reference built for a plain x86-64 target.
 */
#include "pda_oracle.h"
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>

#define MAXM 5

/* ------------------------------------------------------------------ */
/* WENO3: impl/weno3.hpp:56-132 (values) and 134-246 (with gradient)   */
/* ------------------------------------------------------------------ */
static void weno3_core(double *uNeg, double *uPos, double qim1, double qi,
                       double qip1, double qip2)
{
  const double epsilon = 1.e-6, one = 1., two = 2., three = 3.;
  const double oneOvtwo = one / two, oneOvthree = one / three, twoOvthree = two / three;
  {
    const double p0 = (-qim1 + three * qi) * oneOvtwo;
    const double p1 = (qi + qip1) * oneOvtwo;
    const double B0 = (qim1 - qi) * (qim1 - qi);
    const double B1 = (qi - qip1) * (qi - qip1);
    const double alpha0 = oneOvthree / (epsilon * epsilon + two * epsilon * B0 + B0 * B0);
    const double alpha1 = twoOvthree / (epsilon * epsilon + two * epsilon * B1 + B1 * B1);
    const double alphaSInv = one / (alpha0 + alpha1);
    const double w0 = alpha0 * alphaSInv, w1 = alpha1 * alphaSInv;
    *uNeg = w0 * p0 + w1 * p1;
  }
  {
    const double p0 = (qi + qip1) * oneOvtwo;
    const double p1 = (three * qip1 - qip2) * oneOvtwo;
    const double B0 = (qi - qip1) * (qi - qip1);
    const double B1 = (qip1 - qip2) * (qip1 - qip2);
    const double alpha0 = twoOvthree / (epsilon * epsilon + two * epsilon * B0 + B0 * B0);
    const double alpha1 = oneOvthree / (epsilon * epsilon + two * epsilon * B1 + B1 * B1);
    const double alphaSInv = one / (alpha0 + alpha1);
    const double w0 = alpha0 * alphaSInv, w1 = alpha1 * alphaSInv;
    *uPos = w0 * p0 + w1 * p1;
  }
}

void orc_weno3(double *uNeg, double *uPos, const double q[4])
{
  weno3_core(uNeg, uPos, q[0], q[1], q[2], q[3]);
}

static double P(double x, double e) { return pow(x, e); }

void orc_weno3_grad(double *uNeg, double *uPos, double duNeg_dq[4],
                    double duPos_dq[4], const double q[4])
{
  const double qim1 = q[0], qi = q[1], qip1 = q[2], qip2 = q[3];
  const double epsilon = 1e-6, one = 1, two = 2, three = 3;
  const double oneOvtwo = one / two, oneOvthree = one / three, twoOvthree = two / three;
  {
    double dp0_dq[3], dp1_dq[3], dalpha0_dq[3], dalpha1_dq[3], dalphaSInv_dq[3];
    const double p0 = (-qim1 + three * qi) * oneOvtwo;
    const double p1 = (qi + qip1) * oneOvtwo;
    dp0_dq[0] = -1. / 2.; dp0_dq[1] = 3. / 2.; dp0_dq[2] = 0.;
    dp1_dq[0] = 0.; dp1_dq[1] = 1. / 2.; dp1_dq[2] = 1. / 2.;
    const double B0 = (qim1 - qi) * (qim1 - qi);
    const double B1 = (qi - qip1) * (qi - qip1);
    const double alpha0 = oneOvthree / (epsilon * epsilon + two * epsilon * B0 + B0 * B0);
    const double alpha1 = twoOvthree / (epsilon * epsilon + two * epsilon * B1 + B1 * B1);
    const double alphaSInv = one / (alpha0 + alpha1);
    dalpha0_dq[0] = -4. * (qim1 - qi) / (3. * P(P(qim1 - qi, 2.) + epsilon, 3));
    dalpha0_dq[1] = 4. * (qim1 - qi) / (3. * P(P(qim1 - qi, 2.) + epsilon, 3.));
    dalpha0_dq[2] = 0.;
    dalpha1_dq[0] = 0.;
    dalpha1_dq[1] = -8. * (qi - qip1) / (3. * P(P(qi - qip1, 2.) + epsilon, 3.));
    dalpha1_dq[2] = 8. * (qi - qip1) / (3. * P(P(qi - qip1, 2.) + epsilon, 3.));
    dalphaSInv_dq[0] = (4. * (qim1 - qi)) / (3. * P(P(qim1 - qi, 2.) + epsilon, 3.) *
        P(2. / (3. * P(P(qi - qip1, 2.) + epsilon, 2.)) + 1. / (3. * P(P(qim1 - qi, 2.) + epsilon, 2.)), 2.));
    dalphaSInv_dq[1] = -((4. * (qim1 - qi)) / (3. * P(P(qim1 - qi, 2.) + epsilon, 3.)) -
                         (8. * (qi - qip1)) / (3. * P(P(qi - qip1, 2.) + epsilon, 3.))) /
        P(2. / (3. * P(P(qi - qip1, 2.) + epsilon, 2.)) + 1. / (3. * P(P(qim1 - qi, 2.) + epsilon, 2.)), 2.);
    dalphaSInv_dq[2] = -(8. * (qi - qip1)) / (3. *
        P(2. / (3. * P(P(qi - qip1, 2.) + epsilon, 2.)) + 1. / (3. * P(P(qim1 - qi, 2.) + epsilon, 2.)), 2.) *
        P(P(qi - qip1, 2.) + epsilon, 3.));
    const double w0 = alpha0 * alphaSInv, w1 = alpha1 * alphaSInv;
    for (int i = 0; i < 3; i++) {
      duNeg_dq[i] = (dalpha0_dq[i] * alphaSInv + dalphaSInv_dq[i] * alpha0) * p0 + dp0_dq[i] * w0;
      duNeg_dq[i] += (dalpha1_dq[i] * alphaSInv + dalphaSInv_dq[i] * alpha1) * p1 + dp1_dq[i] * w1;
    }
    duNeg_dq[3] = 0.;
    *uNeg = w0 * p0 + w1 * p1;
  }
  {
    double dp0_dq[3], dp1_dq[3], dalpha0_dq[3], dalpha1_dq[3], dalphaSInv_dq[3];
    const double p0 = (qi + qip1) * oneOvtwo;
    const double p1 = (three * qip1 - qip2) * oneOvtwo;
    const double B0 = (qi - qip1) * (qi - qip1);
    const double B1 = (qip1 - qip2) * (qip1 - qip2);
    const double alpha0 = twoOvthree / (epsilon * epsilon + two * epsilon * B0 + B0 * B0);
    const double alpha1 = oneOvthree / (epsilon * epsilon + two * epsilon * B1 + B1 * B1);
    const double alphaSInv = one / (alpha0 + alpha1);
    const double w0 = alpha0 * alphaSInv, w1 = alpha1 * alphaSInv;
    dp0_dq[0] = 1. / 2.; dp0_dq[1] = 1 / 2.; dp0_dq[2] = 0.;
    dp1_dq[0] = 0.; dp1_dq[1] = 3. / 2.; dp1_dq[2] = -1. / 2.;
    dalpha0_dq[0] = -8. * (qi - qip1) / (3. * P(P(qi - qip1, 2.) + epsilon, 3.));
    dalpha0_dq[1] = 8. * (qi - qip1) / (3. * P(P(qi - qip1, 2.) + epsilon, 3.));
    dalpha0_dq[2] = 0.;
    dalpha1_dq[0] = 0.;
    dalpha1_dq[1] = -4. * (qip1 - qip2) / (3. * P(P(qip1 - qip2, 2.) + epsilon, 3.));
    dalpha1_dq[2] = 4. * (qip1 - qip2) / (3. * P(P(qip1 - qip2, 2) + epsilon, 3.));
    dalphaSInv_dq[0] = 8. * (qi - qip1) / (3. * P(P(qi - qip1, 2.) + epsilon, 3.) *
        P(1. / (3. * P(P(qip1 - qip2, 2.) + epsilon, 2.)) + 2. / (3. * P(P(qi - qip1, 2.) + epsilon, 2.)), 2.));
    dalphaSInv_dq[1] = -(8. * (qi - qip1) / (3 * P(P(qi - qip1, 2.) + epsilon, 3.)) -
                         4. * (qip1 - qip2) / (3. * P(P(qip1 - qip2, 2.) + epsilon, 3.))) /
        P(1. / (3. * P(P(qip1 - qip2, 2.) + epsilon, 2.)) + 2. / (3. * P(P(qi - qip1, 2.) + epsilon, 2.)), 2.);
    dalphaSInv_dq[2] = -4. * (qip1 - qip2) / (3. *
        P(1. / (3. * P(P(qip1 - qip2, 2.) + epsilon, 2.)) + 2. / (3. * P(P(qi - qip1, 2.) + epsilon, 2.)), 2.) *
        P(P(qip1 - qip2, 2.) + epsilon, 3.));
    for (int i = 0; i < 3; i++) {
      duPos_dq[i + 1] = (dalpha0_dq[i] * alphaSInv + dalphaSInv_dq[i] * alpha0) * p0 + dp0_dq[i] * w0;
      duPos_dq[i + 1] += (dalpha1_dq[i] * alphaSInv + dalphaSInv_dq[i] * alpha1) * p1 + dp1_dq[i] * w1;
    }
    duPos_dq[0] = 0.;
    *uPos = w0 * p0 + w1 * p1;
  }
}

/* ------------------------------------------------------------------ */
/* WENO5: impl/weno5.hpp:56-178 (values) and 180-434 (with gradient)   */
/* ------------------------------------------------------------------ */
void orc_weno5(double *uNeg, double *uPos, const double q[6])
{
  const double qim2 = q[0], qim1 = q[1], qi = q[2], qip1 = q[3], qip2 = q[4], qip3 = q[5];
  const double epsilon = 1e-6, one = 1, two = 2, three = 3;
  const double four = two * two, five = three + two, six = three * two, seven = four + three;
  const double ten = five * two, eleven = five + six, twelve = six * two, thirteen = six + seven;
  const double oneOvfour = one / four, oneOvsix = one / six, oneOvten = one / ten;
  const double threeOvten = three / ten, sixOvten = six / ten;
  const double thirteenOvtwelve = thirteen / twelve;
  {
    const double p0 = (two * qim2 - seven * qim1 + eleven * qi) * oneOvsix;
    const double p1 = (-qim1 + five * qi + two * qip1) * oneOvsix;
    const double p2 = (two * qi + five * qip1 - qip2) * oneOvsix;
    const double B0 = thirteenOvtwelve * (qim2 - two * qim1 + qi) * (qim2 - two * qim1 + qi) +
                      oneOvfour * (qim2 - four * qim1 + three * qi) * (qim2 - four * qim1 + three * qi);
    const double B1 = thirteenOvtwelve * (qim1 - two * qi + qip1) * (qim1 - two * qi + qip1) +
                      oneOvfour * (qim1 - qip1) * (qim1 - qip1);
    const double B2 = thirteenOvtwelve * (qi - two * qip1 + qip2) * (qi - two * qip1 + qip2) +
                      oneOvfour * (three * qi - four * qip1 + qip2) * (three * qi - four * qip1 + qip2);
    const double alpha0 = oneOvten / (epsilon * epsilon + 2. * epsilon * B0 + B0 * B0);
    const double alpha1 = sixOvten / (epsilon * epsilon + 2. * epsilon * B1 + B1 * B1);
    const double alpha2 = threeOvten / (epsilon * epsilon + 2. * epsilon * B2 + B2 * B2);
    const double alphaSInv = one / (alpha0 + alpha1 + alpha2);
    const double w0 = alpha0 * alphaSInv, w1 = alpha1 * alphaSInv, w2 = alpha2 * alphaSInv;
    *uNeg = w0 * p0 + w1 * p1 + w2 * p2;
  }
  {
    const double p0 = (-qim1 + five * qi + two * qip1) * oneOvsix;
    const double p1 = (two * qi + five * qip1 - qip2) * oneOvsix;
    const double p2 = (eleven * qip1 - seven * qip2 + two * qip3) * oneOvsix;
    const double B0 = thirteenOvtwelve * (qim1 - two * qi + qip1) * (qim1 - two * qi + qip1) +
                      oneOvfour * (qim1 - four * qi + three * qip1) * (qim1 - four * qi + three * qip1);
    const double B1 = thirteenOvtwelve * (qi - two * qip1 + qip2) * (qi - two * qip1 + qip2) +
                      oneOvfour * (qi - qip2) * (qi - qip2);
    const double B2 = thirteenOvtwelve * (qip1 - two * qip2 + qip3) * (qip1 - two * qip2 + qip3) +
                      oneOvfour * (three * qip1 - four * qip2 + qip3) * (three * qip1 - four * qip2 + qip3);
    const double alpha0 = threeOvten / (epsilon * epsilon + 2. * epsilon * B0 + B0 * B0);
    const double alpha1 = sixOvten / (epsilon * epsilon + 2. * epsilon * B1 + B1 * B1);
    const double alpha2 = oneOvten / (epsilon * epsilon + 2. * epsilon * B2 + B2 * B2);
    const double alphaSInv = one / (alpha0 + alpha1 + alpha2);
    const double w0 = alpha0 * alphaSInv, w1 = alpha1 * alphaSInv, w2 = alpha2 * alphaSInv;
    *uPos = w0 * p0 + w1 * p1 + w2 * p2;
  }
}

void orc_weno5_grad(double *uNeg, double *uPos, double duNeg_dq[6],
                    double duPos_dq[6], const double q[6])
{
  const double qim2 = q[0], qim1 = q[1], qi = q[2], qip1 = q[3], qip2 = q[4], qip3 = q[5];
  const double epsilon = 1e-6, one = 1, two = 2, three = 3;
  const double four = two * two, five = three + two, six = three * two, seven = four + three;
  const double ten = five * two, eleven = five + six, twelve = six * two, thirteen = six + seven;
  const double oneOvfour = one / four, oneOvsix = one / six, oneOvten = one / ten;
  const double threeOvten = three / ten, sixOvten = six / ten;
  const double thirteenOvtwelve = thirteen / twelve;
  {
    const double p0 = (two * qim2 - seven * qim1 + eleven * qi) * oneOvsix;
    const double p1 = (-qim1 + five * qi + two * qip1) * oneOvsix;
    const double p2 = (two * qi + five * qip1 - qip2) * oneOvsix;
    const double dp0_dq[5] = {1. / 3., -7. / 6., 11. / 6., 0., 0.};
    const double dp1_dq[5] = {0., -1. / 6., 5. / 6., 1. / 3., 0.};
    const double dp2_dq[5] = {0., 0., 1. / 3., 5. / 6., -1. / 6.};
    const double a0 = qim2 - two * qim1 + qi, b0 = qim2 - four * qim1 + three * qi;
    const double a1 = qim1 - two * qi + qip1, b1 = qim1 - qip1;
    const double a2 = qi - two * qip1 + qip2, b2 = three * qi - four * qip1 + qip2;
    const double B0 = thirteenOvtwelve * (a0 * a0) + oneOvfour * (b0 * b0);
    const double B1 = thirteenOvtwelve * (a1 * a1) + oneOvfour * (b1 * b1);
    const double B2 = thirteenOvtwelve * (a2 * a2) + oneOvfour * (b2 * b2);
    double dB0_dq[5], dB1_dq[5], dB2_dq[5];
    dB0_dq[0] = (13. * (qim2 - 2. * qim1 + qi)) / 6. + (qim2 - 4. * qim1 + 3. * qi) / 2.;
    dB0_dq[1] = -(13. * (qim2 - 2. * qim1 + qi)) / 3. - 2. * (qim2 - 4. * qim1 + 3. * qi);
    dB0_dq[2] = (13. * (qim2 - 2. * qim1 + qi)) / 6. + (3. * (qim2 - 4. * qim1 + 3. * qi)) / 2.;
    dB0_dq[3] = 0.; dB0_dq[4] = 0.;
    dB1_dq[0] = 0.;
    dB1_dq[1] = (13. * (qip1 + qim1 - 2. * qi)) / 6. + (qim1 - qip1) / 2.;
    dB1_dq[2] = -(13. * (qip1 + qim1 - 2. * qi)) / 3.;
    dB1_dq[3] = (13. * (qip1 + qim1 - 2. * qi)) / 6. - (qim1 - qip1) / 2.;
    dB1_dq[4] = 0.;
    dB2_dq[0] = 0.; dB2_dq[1] = 0.;
    dB2_dq[2] = (13. * (qip2 - 2. * qip1 + qi)) / 6. + (3. * (qip2 - 4. * qip1 + 3. * qi)) / 2.;
    dB2_dq[3] = -(13. * (qip2 - 2. * qip1 + qi)) / 3. - 2. * (qip2 - 4. * qip1 + 3. * qi);
    dB2_dq[4] = (13. * (qip2 - 2. * qip1 + qi)) / 6. + (qip2 - 4. * qip1 + 3. * qi) / 2.;
    const double e0 = epsilon + B0, e1 = epsilon + B1, e2 = epsilon + B2;
    const double alpha0 = oneOvten / (e0 * e0);
    const double alpha1 = sixOvten / (e1 * e1);
    const double alpha2 = threeOvten / (e2 * e2);
    const double alphaSInv = one / (alpha0 + alpha1 + alpha2);
    double dalpha0_dq[5], dalpha1_dq[5], dalpha2_dq[5], dalphaSInv_dq[5];
    for (int i = 0; i < 5; i++) {
      dalpha0_dq[i] = -1. / (5. * P(B0 + epsilon, 3.)) * dB0_dq[i];
      dalpha1_dq[i] = -6. / (5. * P(B1 + epsilon, 3.)) * dB1_dq[i];
      dalpha2_dq[i] = -3. / (5. * P(B2 + epsilon, 3.)) * dB2_dq[i];
      const double S = alpha2 + alpha1 + alpha0;
      dalphaSInv_dq[i] = -1. / (S * S) * (dalpha0_dq[i] + dalpha1_dq[i] + dalpha2_dq[i]);
    }
    const double w0 = alpha0 * alphaSInv, w1 = alpha1 * alphaSInv, w2 = alpha2 * alphaSInv;
    for (int i = 0; i < 5; i++) {
      duNeg_dq[i] = (dalpha0_dq[i] * alphaSInv + dalphaSInv_dq[i] * alpha0) * p0 + dp0_dq[i] * w0;
      duNeg_dq[i] += (dalpha1_dq[i] * alphaSInv + dalphaSInv_dq[i] * alpha1) * p1 + dp1_dq[i] * w1;
      duNeg_dq[i] += (dalpha2_dq[i] * alphaSInv + dalphaSInv_dq[i] * alpha2) * p2 + dp2_dq[i] * w2;
    }
    duNeg_dq[5] = 0.;
    *uNeg = w0 * p0 + w1 * p1 + w2 * p2;
  }
  {
    const double p0 = (-qim1 + five * qi + two * qip1) * oneOvsix;
    const double p1 = (two * qi + five * qip1 - qip2) * oneOvsix;
    const double p2 = (eleven * qip1 - seven * qip2 + two * qip3) * oneOvsix;
    const double dp0_dq[5] = {-1. / 6., 5. / 6., 1. / 3., 0., 0.};
    const double dp1_dq[5] = {0., 1. / 3., 5. / 6., -1. / 6., 0.};
    const double dp2_dq[5] = {0., 0., 11. / 6., -7. / 6., 1. / 3.};
    const double a0 = qim1 - two * qi + qip1, b0 = qim1 - four * qi + three * qip1;
    const double a1 = qi - two * qip1 + qip2, b1 = qi - qip2;
    const double a2 = qip1 - two * qip2 + qip3, b2 = three * qip1 - four * qip2 + qip3;
    const double B0 = thirteenOvtwelve * (a0 * a0) + oneOvfour * (b0 * b0);
    const double B1 = thirteenOvtwelve * (a1 * a1) + oneOvfour * (b1 * b1);
    const double B2 = thirteenOvtwelve * (a2 * a2) + oneOvfour * (b2 * b2);
    double dB0_dq[5], dB1_dq[5], dB2_dq[5];
    dB0_dq[0] = (3. * qip1 + qim1 - 4. * qi) / 2. + (13. * (qip1 + qim1 - 2. * qi)) / 6.;
    dB0_dq[1] = -2. * (3. * qip1 + qim1 - 4. * qi) - (13. * (qip1 + qim1 - 2 * qi)) / 3.;
    dB0_dq[2] = (3. * (3. * qip1 + qim1 - 4. * qi)) / 2. + (13. * (qip1 + qim1 - 2. * qi)) / 6.;
    dB0_dq[3] = 0.; dB0_dq[4] = 0.;
    dB1_dq[0] = 0.;
    dB1_dq[1] = (13. * (qip2 - 2 * qip1 + qi)) / 6. + (qi - qip2) / 2.;
    dB1_dq[2] = -(13. * (qip2 - 2. * qip1 + qi)) / 3.;
    dB1_dq[3] = (13. * (qip2 - 2. * qip1 + qi)) / 6. - (qi - qip2) / 2.;
    dB1_dq[4] = 0.;
    dB2_dq[0] = 0.; dB2_dq[1] = 0.;
    dB2_dq[2] = (13. * (qip3 - 2. * qip2 + qip1)) / 6. + (3. * (qip3 - 4. * qip2 + 3. * qip1)) / 2.;
    dB2_dq[3] = -(13. * (qip3 - 2. * qip2 + qip1)) / 3. - 2. * (qip3 - 4. * qip2 + 3. * qip1);
    dB2_dq[4] = (13. * (qip3 - 2. * qip2 + qip1)) / 6. + (qip3 - 4. * qip2 + 3. * qip1) / 2.;
    const double e0 = epsilon + B0, e1 = epsilon + B1, e2 = epsilon + B2;
    const double alpha0 = threeOvten / (e0 * e0);
    const double alpha1 = sixOvten / (e1 * e1);
    const double alpha2 = oneOvten / (e2 * e2);
    const double alphaSInv = one / (alpha0 + alpha1 + alpha2);
    double dalpha0_dq[5], dalpha1_dq[5], dalpha2_dq[5], dalphaSInv_dq[5];
    for (int i = 0; i < 5; i++) {
      dalpha0_dq[i] = -3. / (5. * P(B0 + epsilon, 3.)) * dB0_dq[i];
      dalpha1_dq[i] = -6. / (5. * P(B1 + epsilon, 3.)) * dB1_dq[i];
      dalpha2_dq[i] = -1. / (5. * P(B2 + epsilon, 3.)) * dB2_dq[i];
      const double S = alpha2 + alpha1 + alpha0;
      dalphaSInv_dq[i] = -1. / (S * S) * (dalpha0_dq[i] + dalpha1_dq[i] + dalpha2_dq[i]);
    }
    const double w0 = alpha0 * alphaSInv, w1 = alpha1 * alphaSInv, w2 = alpha2 * alphaSInv;
    for (int i = 0; i < 5; i++) {
      duPos_dq[i + 1] = (dalpha0_dq[i] * alphaSInv + dalphaSInv_dq[i] * alpha0) * p0 + dp0_dq[i] * w0;
      duPos_dq[i + 1] += (dalpha1_dq[i] * alphaSInv + dalphaSInv_dq[i] * alpha1) * p1 + dp1_dq[i] * w1;
      duPos_dq[i + 1] += (dalpha2_dq[i] * alphaSInv + dalphaSInv_dq[i] * alpha2) * p2 + dp2_dq[i] * w2;
    }
    duPos_dq[0] = 0.;
    *uPos = w0 * p0 + w1 * p1 + w2 * p2;
  }
}

/* ------------------------------------------------------------------ */
/* Euler Rusanov flux: impl/euler_rusanov_flux_values_function.hpp     */
/* ------------------------------------------------------------------ */
static void euler_flux3(double *F, const double *qL, const double *qR, double gamma)
{ /* :54-97 */
  const double one = 1, two = 2, half = one / two, es = 1.e-30;
  double FL[3], FR[3];
  const double rL = qL[0], uL = qL[1] / (rL + es);
  const double pL = (gamma - 1) * (qL[2] - half * rL * (uL * uL));
  const double HL = (qL[2] + pL) / rL;
  FL[0] = rL * uL; FL[1] = rL * uL * uL + pL; FL[2] = rL * uL * HL;
  const double rR = qR[0], uR = qR[1] / (rR + es);
  const double pR = (gamma - 1) * (qR[2] - half * rR * (uR * uR));
  const double HR = (qR[2] + pR) / rR;
  FR[0] = rR * uR; FR[1] = rR * uR * uR + pR; FR[2] = rR * uR * HR;
  const double RT = sqrt(rR / (rL));
  const double u = (uL + RT * uR) / (1. + RT);
  const double H = (HL + RT * HR) / (1. + RT);
  const double a = sqrt((gamma - 1.) * (H - half * (u * u)));
  const double smax = fabs(u) + a;
  for (int i = 0; i < 3; ++i) F[i] = half * (FL[i] + FR[i] + smax * (qL[i] - qR[i]));
}

static void euler_flux4(double *F, const double *qL, const double *qR, const double *n, double gamma)
{ /* :99-149 */
  const double one = 1, two = 2, half = one / two, es = 1.e-30;
  double FL[4], FR[4];
  const double rL = qL[0], uL = qL[1] / (rL + es), vL = qL[2] / (rL + es);
  const double unL = uL * n[0] + vL * n[1];
  const double pL = (gamma - 1) * (qL[3] - half * rL * (uL * uL + vL * vL));
  const double HL = (qL[3] + pL) / rL;
  FL[0] = rL * unL; FL[1] = rL * unL * uL + pL * n[0];
  FL[2] = rL * unL * vL + pL * n[1]; FL[3] = rL * unL * HL;
  const double rR = qR[0], uR = qR[1] / (rR + es), vR = qR[2] / (rR + es);
  const double unR = uR * n[0] + vR * n[1];
  const double pR = (gamma - 1) * (qR[3] - half * rR * (uR * uR + vR * vR));
  const double HR = (qR[3] + pR) / rR;
  FR[0] = rR * unR; FR[1] = rR * unR * uR + pR * n[0];
  FR[2] = rR * unR * vR + pR * n[1]; FR[3] = rR * unR * HR;
  const double RT = sqrt(rR / (rL));
  const double u = (uL + RT * uR) / (1. + RT);
  const double v = (vL + RT * vR) / (1. + RT);
  const double H = (HL + RT * HR) / (1. + RT);
  const double a = sqrt((gamma - 1.) * (H - half * (u * u + v * v)));
  const double smax = sqrt(u * u + v * v) + a;
  for (int i = 0; i < 4; ++i) F[i] = half * (FL[i] + FR[i] + smax * (qL[i] - qR[i]));
}

static void euler_flux5(double *F, const double *qL, const double *qR, const double *n, double gamma)
{ /* :151-208 */
  const double one = 1, two = 2, half = one / two, es = 1.e-30;
  double FL[5], FR[5];
  const double rL = qL[0], uL = qL[1] / (rL + es), vL = qL[2] / (rL + es), wL = qL[3] / (rL + es);
  const double unL = uL * n[0] + vL * n[1] + wL * n[2];
  const double pL = (gamma - 1) * (qL[4] - half * rL * (uL * uL + vL * vL + wL * wL));
  const double HL = (qL[4] + pL) / rL;
  FL[0] = rL * unL; FL[1] = rL * unL * uL + pL * n[0]; FL[2] = rL * unL * vL + pL * n[1];
  FL[3] = rL * unL * wL + pL * n[2]; FL[4] = rL * unL * HL;
  const double rR = qR[0], uR = qR[1] / (rR + es), vR = qR[2] / (rR + es), wR = qR[3] / (rR + es);
  const double unR = uR * n[0] + vR * n[1] + wR * n[2];
  const double pR = (gamma - 1) * (qR[4] - half * rR * (uR * uR + vR * vR + wR * wR));
  const double HR = (qR[4] + pR) / rR;
  FR[0] = rR * unR; FR[1] = rR * unR * uR + pR * n[0]; FR[2] = rR * unR * vR + pR * n[1];
  FR[3] = rR * unR * wR + pR * n[2]; FR[4] = rR * unR * HR;
  const double RT = sqrt(rR / (rL));
  const double u = (uL + RT * uR) / (1. + RT);
  const double v = (vL + RT * vR) / (1. + RT);
  const double w = (wL + RT * wR) / (1. + RT);
  const double H = (HL + RT * HR) / (1. + RT);
  const double a = sqrt((gamma - 1.) * (H - half * (u * u + v * v + w * w)));
  const double smax = sqrt(u * u + v * v + w * w) + a;
  for (int i = 0; i < 5; ++i) F[i] = half * (FL[i] + FR[i] + smax * (qL[i] - qR[i]));
}

/* ------------------------------------------------------------------ */
/* Euler flux Jacobians: impl/euler_rusanov_flux_jacobian_function.hpp */
/* J arrays are row-major m x m.                                        */
/* ------------------------------------------------------------------ */
static void euler_fjac3(double *JL, double *JR, const double *qL, const double *qR, double gamma)
{ /* :54-143 */
  const double one = 1, two = 2, three = 3, half = one / two, gm1 = gamma - one, es = 1.e-30;
  const double rL = qL[0], uL = qL[1] / (rL + es);
  const double pL = gm1 * (qL[2] - half * rL * (uL * uL));
  const double HL = (qL[2] + pL) / rL;
  const double aL = sqrt(gm1 * (HL - half * (uL * uL)));
  const double rR = qR[0], uR = qR[1] / (rR + es);
  const double pR = gm1 * (qR[2] - half * rR * (uR * uR));
  const double HR = (qR[2] + pR) / rR;
  const double aR = sqrt(gm1 * (HR - half * (uR * uR)));
  const double r = sqrt(rR * rL), RT = sqrt(rR / (rL));
  const double u = (uL + RT * uR) / (one + RT);
  const double H = (HL + RT * HR) / (one + RT);
  const double a = sqrt(gm1 * (H - half * (u * u)));
  const double smax = fabs(u) + a;
  double gradL[3], gradR[3];
  const double VMagSqrRoe = u * u + es;
  const double VMagSqrL = uL * uL, VVRoeL = uL * u;
  const double VMagSqrR = uR * uR, VVRoeR = uR * u;
  const double uRelRoe = u / sqrt(VMagSqrRoe);
  gradL[0] = one / (rL + r) * (-half * (uL + u) * uRelRoe + half * gm1 / a * (half * (VMagSqrRoe + VVRoeL) + half * (HL - H) - aL * aL / gm1 + half * (gamma - two) * VMagSqrL));
  gradL[1] = one / (rL + r) * (uRelRoe - half * (gm1 * (u + gm1 * uL)) / (a));
  gradL[2] = half / (rL + r) * gamma * gm1 / (a);
  gradR[0] = one / (rR + r) * (-half * (uR + u) * uRelRoe + half * gm1 / a * (half * (VMagSqrRoe + VVRoeR) + half * (HR - H) - aR * aR / gm1 + half * (gamma - two) * VMagSqrR));
  gradR[1] = one / (rR + r) * (uRelRoe - half * (gm1 * (u + gm1 * uR)) / (a));
  gradR[2] = half / (rR + r) * gamma * gm1 / (a);
#define JLm(i, j) JL[(i) * 3 + (j)]
#define JRm(i, j) JR[(i) * 3 + (j)]
  JLm(0, 0) = 0; JLm(0, 1) = half; JLm(0, 2) = 0;
  JLm(1, 0) = half * (half * gm1 * uL * uL - uL * uL);
  JLm(1, 1) = half * ((three - gamma) * uL);
  JLm(1, 2) = half * gm1;
  JLm(2, 0) = half * ((half * gm1 * uL * uL - HL) * uL);
  JLm(2, 1) = half * (HL - gm1 * uL * uL);
  JLm(2, 2) = half * gamma * uL;
  for (int i = 0; i < 3; i++) {
    JLm(i, i) += half * smax;
    for (int j = 0; j < 3; j++) JLm(i, j) += half * gradL[j] * (qL[i] - qR[i]);
  }
  JRm(0, 0) = 0; JRm(0, 1) = half; JRm(0, 2) = 0;
  JRm(1, 0) = half * (half * gm1 * uR * uR - uR * uR);
  JRm(1, 1) = half * ((three - gamma) * uR);
  JRm(1, 2) = half * gm1;
  JRm(2, 0) = half * ((half * gm1 * uR * uR - HR) * uR);
  JRm(2, 1) = half * (HR - gm1 * uR * uR);
  JRm(2, 2) = half * gamma * uR;
  for (int i = 0; i < 3; i++) {
    JRm(i, i) -= half * smax;
    for (int j = 0; j < 3; j++) JRm(i, j) += half * gradR[j] * (qL[i] - qR[i]);
  }
#undef JLm
#undef JRm
}

static void euler_fjac4(double *JL, double *JR, const double *qL, const double *qR, const double *n, double gamma)
{ /* :145-262 */
  const double one = 1, two = 2, half = one / two, gm1 = gamma - one, es = 1.e-30;
  const double rL = qL[0], uL = qL[1] / (rL + es), vL = qL[2] / (rL + es);
  const double unL = uL * n[0] + vL * n[1];
  const double pL = gm1 * (qL[3] - half * rL * (uL * uL + vL * vL));
  const double HL = (qL[3] + pL) / rL;
  const double aL = sqrt(gm1 * (HL - half * (uL * uL + vL * vL)));
  const double rR = qR[0], uR = qR[1] / (rR + es), vR = qR[2] / (rR + es);
  const double unR = uR * n[0] + vR * n[1];
  const double pR = gm1 * (qR[3] - half * rR * (uR * uR + vR * vR));
  const double HR = (qR[3] + pR) / rR;
  const double aR = sqrt(gm1 * (HR - half * (uR * uR + vR * vR)));
  const double r = sqrt(rR * rL), RT = sqrt(rR / (rL));
  const double u = (uL + RT * uR) / (one + RT);
  const double v = (vL + RT * vR) / (one + RT);
  const double H = (HL + RT * HR) / (one + RT);
  const double a = sqrt(gm1 * (H - half * (u * u + v * v)));
  const double smax = sqrt(u * u + v * v) + a;
  double gradL[4], gradR[4];
  const double VMagSqrRoe = u * u + v * v + es;
  const double VMagSqrL = uL * uL + vL * vL, VVRoeL = uL * u + vL * v;
  const double VMagSqrR = uR * uR + vR * vR, VVRoeR = uR * u + vR * v;
  const double uRelRoe = u / sqrt(VMagSqrRoe), vRelRoe = v / sqrt(VMagSqrRoe);
  gradL[0] = one / (rL + r) * (-half * (uL + u) * uRelRoe - half * (vL + v) * vRelRoe + half * gm1 / a * (half * (VMagSqrRoe + VVRoeL) + half * (HL - H) - aL * aL / gm1 + half * (gamma - two) * VMagSqrL));
  gradL[1] = one / (rL + r) * (uRelRoe - half * (gm1 * (u + gm1 * uL)) / (a));
  gradL[2] = one / (rL + r) * (vRelRoe - half * (gm1 * (v + gm1 * vL)) / (a));
  gradL[3] = half / (rL + r) * gamma * gm1 / (a);
  gradR[0] = one / (rR + r) * (-half * (uR + u) * uRelRoe - half * (vR + v) * vRelRoe + half * gm1 / a * (half * (VMagSqrRoe + VVRoeR) + half * (HR - H) - aR * aR / gm1 + half * (gamma - two) * VMagSqrR));
  gradR[1] = one / (rR + r) * (uRelRoe - half * (gm1 * (u + gm1 * uR)) / (a));
  gradR[2] = one / (rR + r) * (vRelRoe - half * (gm1 * (v + gm1 * vR)) / (a));
  gradR[3] = half / (rR + r) * gamma * gm1 / (a);
#define JLm(i, j) JL[(i) * 4 + (j)]
#define JRm(i, j) JR[(i) * 4 + (j)]
  JLm(0, 0) = 0; JLm(0, 1) = half * n[0]; JLm(0, 2) = half * n[1]; JLm(0, 3) = 0;
  JLm(1, 0) = half * (half * gm1 * VMagSqrL * n[0] - uL * unL);
  JLm(1, 1) = half * (uL * n[0] - gm1 * uL * n[0] + unL);
  JLm(1, 2) = half * (uL * n[1] - gm1 * vL * n[0]);
  JLm(1, 3) = half * gm1 * n[0];
  JLm(2, 0) = half * (half * gm1 * VMagSqrL * n[1] - vL * unL);
  JLm(2, 1) = half * (vL * n[0] - gm1 * uL * n[1]);
  JLm(2, 2) = half * (vL * n[1] - gm1 * vL * n[1] + unL);
  JLm(2, 3) = half * gm1 * n[1];
  JLm(3, 0) = half * ((half * gm1 * VMagSqrL - HL) * unL);
  JLm(3, 1) = half * (HL * n[0] - gm1 * uL * unL);
  JLm(3, 2) = half * (HL * n[1] - gm1 * vL * unL);
  JLm(3, 3) = half * gamma * unL;
  for (int i = 0; i < 4; i++) {
    JLm(i, i) += half * smax;
    for (int j = 0; j < 4; j++) JLm(i, j) += half * gradL[j] * (qL[i] - qR[i]);
  }
  JRm(0, 0) = 0; JRm(0, 1) = half * n[0]; JRm(0, 2) = half * n[1]; JRm(0, 3) = 0;
  JRm(1, 0) = half * (half * gm1 * VMagSqrR * n[0] - uR * unR);
  JRm(1, 1) = half * (uR * n[0] - gm1 * uR * n[0] + unR);
  JRm(1, 2) = half * (uR * n[1] - gm1 * vR * n[0]);
  JRm(1, 3) = half * gm1 * n[0];
  JRm(2, 0) = half * (half * gm1 * VMagSqrR * n[1] - vR * unR);
  JRm(2, 1) = half * (vR * n[0] - gm1 * uR * n[1]);
  JRm(2, 2) = half * (vR * n[1] - gm1 * vR * n[1] + unR);
  JRm(2, 3) = half * gm1 * n[1];
  JRm(3, 0) = half * ((half * gm1 * VMagSqrR - HR) * unR);
  JRm(3, 1) = half * (HR * n[0] - gm1 * uR * unR);
  JRm(3, 2) = half * (HR * n[1] - gm1 * vR * unR);
  JRm(3, 3) = half * gamma * unR;
  for (int i = 0; i < 4; i++) {
    JRm(i, i) -= half * smax;
    for (int j = 0; j < 4; j++) JRm(i, j) += half * gradR[j] * (qL[i] - qR[i]);
  }
#undef JLm
#undef JRm
}

static void euler_fjac5(double *JL, double *JR, const double *qL, const double *qR, const double *n, double gamma)
{ /* :264-406 */
  const double one = 1, two = 2, half = one / two, gm1 = gamma - one, es = 1.e-30;
  const double rL = qL[0], uL = qL[1] / (rL + es), vL = qL[2] / (rL + es), wL = qL[3] / (rL + es);
  const double unL = uL * n[0] + vL * n[1] + wL * n[2];
  const double pL = gm1 * (qL[4] - half * rL * (uL * uL + vL * vL + wL * wL));
  const double HL = (qL[4] + pL) / rL;
  const double aL = sqrt(gm1 * (HL - half * (uL * uL + vL * vL + wL * wL)));
  const double rR = qR[0], uR = qR[1] / (rR + es), vR = qR[2] / (rR + es), wR = qR[3] / (rR + es);
  const double unR = uR * n[0] + vR * n[1] + wR * n[2];
  const double pR = gm1 * (qR[4] - half * rR * (uR * uR + vR * vR + wR * wR));
  const double HR = (qR[4] + pR) / rR;
  const double aR = sqrt(gm1 * (HR - half * (uR * uR + vR * vR + wR * wR)));
  const double r = sqrt(rR * rL), RT = sqrt(rR / (rL));
  const double u = (uL + RT * uR) / (one + RT);
  const double v = (vL + RT * vR) / (one + RT);
  const double w = (wL + RT * wR) / (one + RT);
  const double H = (HL + RT * HR) / (one + RT);
  const double a = sqrt(gm1 * (H - half * (u * u + v * v + w * w)));
  const double smax = sqrt(u * u + v * v + w * w) + a;
  double gradL[5], gradR[5];
  const double VMagSqrRoe = u * u + v * v + w * w + es;
  const double VMagSqrL = uL * uL + vL * vL + wL * wL, VVRoeL = uL * u + vL * v + wL * w;
  const double VMagSqrR = uR * uR + vR * vR + wR * wR, VVRoeR = uR * u + vR * v + wR * w;
  const double uRelRoe = u / sqrt(VMagSqrRoe), vRelRoe = v / sqrt(VMagSqrRoe), wRelRoe = w / sqrt(VMagSqrRoe);
  gradL[0] = one / (rL + r) * (-half * (uL + u) * uRelRoe - half * (vL + v) * vRelRoe - half * (wL + w) * wRelRoe + half * gm1 / a * (half * (VMagSqrRoe + VVRoeL) + half * (HL - H) - aL * aL / gm1 + half * (gamma - two) * VMagSqrL));
  gradL[1] = one / (rL + r) * (uRelRoe - half * (gm1 * (u + gm1 * uL)) / (a));
  gradL[2] = one / (rL + r) * (vRelRoe - half * (gm1 * (v + gm1 * vL)) / (a));
  gradL[3] = one / (rL + r) * (wRelRoe - half * (gm1 * (w + gm1 * wL)) / (a));
  gradL[4] = half / (rL + r) * gamma * gm1 / (a);
  gradR[0] = one / (rR + r) * (-half * (uR + u) * uRelRoe - half * (vR + v) * vRelRoe - half * (wR + w) * wRelRoe + half * gm1 / a * (half * (VMagSqrRoe + VVRoeR) + half * (HR - H) - aR * aR / gm1 + half * (gamma - two) * VMagSqrR));
  gradR[1] = one / (rR + r) * (uRelRoe - half * (gm1 * (u + gm1 * uR)) / (a));
  gradR[2] = one / (rR + r) * (vRelRoe - half * (gm1 * (v + gm1 * vR)) / (a));
  gradR[3] = one / (rR + r) * (wRelRoe - half * (gm1 * (w + gm1 * wR)) / (a));
  gradR[4] = half / (rR + r) * gamma * gm1 / (a);
  const double *Q[2] = {qL, qR};
  double *JJ[2] = {JL, JR};
  const double un[2] = {unL, unR}, uu[2] = {uL, uR}, vv[2] = {vL, vR}, ww[2] = {wL, wR};
  const double HH[2] = {HL, HR}, VM[2] = {VMagSqrL, VMagSqrR};
  for (int s = 0; s < 2; ++s) {
    double *J = JJ[s];
    const double uX = uu[s], vX = vv[s], wX = ww[s], unX = un[s], HX = HH[s], VX = VM[s];
#define Jm(i, j) J[(i) * 5 + (j)]
    Jm(0, 0) = 0; Jm(0, 1) = half * n[0]; Jm(0, 2) = half * n[1]; Jm(0, 3) = half * n[2]; Jm(0, 4) = 0;
    Jm(1, 0) = half * (half * gm1 * VX * n[0] - uX * unX);
    Jm(1, 1) = half * (uX * n[0] - gm1 * uX * n[0] + unX);
    Jm(1, 2) = half * (uX * n[1] - gm1 * vX * n[0]);
    Jm(1, 3) = half * (uX * n[2] - gm1 * wX * n[0]);
    Jm(1, 4) = half * gm1 * n[0];
    Jm(2, 0) = half * (half * gm1 * VX * n[1] - vX * unX);
    Jm(2, 1) = half * (vX * n[0] - gm1 * uX * n[1]);
    Jm(2, 2) = half * (vX * n[1] - gm1 * vX * n[1] + unX);
    Jm(2, 3) = half * (vX * n[2] - gm1 * wX * n[1]);
    Jm(2, 4) = half * gm1 * n[1];
    Jm(3, 0) = half * (half * gm1 * VX * n[2] - wX * unX);
    Jm(3, 1) = half * (wX * n[0] - gm1 * uX * n[2]);
    Jm(3, 2) = half * (wX * n[1] - gm1 * vX * n[2]);
    Jm(3, 3) = half * (wX * n[2] - gm1 * wX * n[2] + unX);
    Jm(3, 4) = half * gm1 * n[2];
    Jm(4, 0) = half * ((half * gm1 * VX - HX) * unX);
    Jm(4, 1) = half * (HX * n[0] - gm1 * uX * unX);
    Jm(4, 2) = half * (HX * n[1] - gm1 * vX * unX);
    Jm(4, 3) = half * (HX * n[2] - gm1 * wX * unX);
    Jm(4, 4) = half * gamma * unX;
    const double *grad = s == 0 ? gradL : gradR;
    for (int i = 0; i < 5; i++) {
      if (s == 0) Jm(i, i) += half * smax; else Jm(i, i) -= half * smax;
      for (int j = 0; j < 5; j++) Jm(i, j) += half * grad[j] * (Q[0][i] - Q[1][i]);
    }
#undef Jm
  }
}

/* ------------------------------------------------------------------ */
/* Shallow water: impl/swe_rusanov_flux_values_function.hpp:54-97 and  */
/* impl/swe_rusanov_flux_jacobian_function.hpp:54-137                  */
/* ------------------------------------------------------------------ */
static void swe_flux(double *F, const double *qL, const double *qR, const double *n, double gravity)
{
  const double one = 1, two = 2, half = one / two, es = 1.e-30;
  double FL[3], FR[3];
  const double hL = qL[0], uL = qL[1] / (hL + es), vL = qL[2] / (hL + es);
  const double unL = uL * n[0] + vL * n[1];
  const double pL = 0.5 * gravity * hL * hL;
  FL[0] = hL * unL; FL[1] = hL * unL * uL + pL * n[0]; FL[2] = hL * unL * vL + pL * n[1];
  const double hR = qR[0], uR = qR[1] / (hR + es), vR = qR[2] / (hR + es);
  const double unR = uR * n[0] + vR * n[1];
  const double pR = 0.5 * gravity * hR * hR;
  FR[0] = hR * unR; FR[1] = hR * unR * uR + pR * n[0]; FR[2] = hR * unR * vR + pR * n[1];
  const double hm = 0.5 * (hL + hR);
  const double um = (unL * pow(hL, 0.5) + unR * pow(hR, 0.5)) / (pow(hL, 0.5) + pow(hR, 0.5) + es);
  const double smax = sqrt(um * um) + sqrt(pow(sqrt(gravity * hm), 2.));
  for (int i = 0; i < 3; ++i) F[i] = half * (FL[i] + FR[i] + smax * (qL[i] - qR[i]));
}

static void swe_fjac(double *JL, double *JR, const double *qL, const double *qR, const double *n, double g)
{
  const double es = 1.e-30;
  const double hL = qL[0], uL = qL[1] / (hL + es), vL = qL[2] / (hL + es), unL = uL * n[0] + vL * n[1];
  const double hR = qR[0], uR = qR[1] / (hR + es), vR = qR[2] / (hR + es), unR = uR * n[0] + vR * n[1];
  const double hm = 0.5 * (hL + hR);
  const double um = (unL * pow(hL, 0.5) + unR * pow(hR, 0.5)) / (pow(hL, 0.5) + pow(hR, 0.5) + es);
  const double smax = fabs(um) + fabs(pow(g * hm, 0.5));
  const double termL = (n[0] * qL[1] + n[1] * qL[2]) / pow(qL[0], 2.);
  const double termR = (n[0] * qR[1] + n[1] * qR[2]) / pow(qR[0], 2.);
  const double hL_sqrt = pow(hL, 0.5), hR_sqrt = pow(hR, 0.5);
  const double hsqrt_un = hL_sqrt * unL + hR_sqrt * unR + es;
  double dsmaxL[3], dsmaxR[3];
  dsmaxL[0] = -fabs(hsqrt_un) / (2. * hL_sqrt * pow(hL_sqrt + hR_sqrt, 2.)) +
              (0.5 * unL / hL_sqrt - hL_sqrt * termL) * hsqrt_un / ((hL_sqrt + hR_sqrt) * fabs(hsqrt_un)) +
              g / (pow(2., 3. / 2.) * pow(g * (hL + hR), 0.5));
  dsmaxL[1] = n[0] * hsqrt_un / (hL_sqrt * (hL_sqrt + hR_sqrt) * fabs(hsqrt_un));
  dsmaxL[2] = n[1] * hsqrt_un / (hL_sqrt * (hL_sqrt + hR_sqrt) * fabs(hsqrt_un));
  dsmaxR[0] = -fabs(hsqrt_un) / (2. * hR_sqrt * pow(hL_sqrt + hR_sqrt, 2.)) +
              (0.5 * unR / hR_sqrt - hR_sqrt * termR) * hsqrt_un / ((hL_sqrt + hR_sqrt) * fabs(hsqrt_un)) +
              g / (pow(2., 3. / 2.) * pow(g * (hL + hR), 0.5));
  dsmaxR[1] = n[0] * hsqrt_un / (hR_sqrt * (hL_sqrt + hR_sqrt) * fabs(hsqrt_un));
  dsmaxR[2] = n[1] * hsqrt_un / (hR_sqrt * (hL_sqrt + hR_sqrt) * fabs(hsqrt_un));
#define JLm(i, j) JL[(i) * 3 + (j)]
#define JRm(i, j) JR[(i) * 3 + (j)]
  JLm(0, 0) = -0.5 * dsmaxL[0] * (qR[0] - qL[0]) + 0.5 * (n[0] * uL + n[1] * vL - qL[0] * termL) + 0.5 * smax;
  JLm(0, 1) = 0.5 * n[0] - 0.5 * dsmaxL[1] * (qR[0] - qL[0]);
  JLm(0, 2) = 0.5 * n[1] - 0.5 * dsmaxL[2] * (qR[0] - qL[0]);
  JLm(1, 0) = 0.5 * (g * n[0] * qL[0] - qL[1] * termL) - 0.5 * dsmaxL[0] * (qR[1] - qL[1]);
  JLm(1, 1) = n[0] * uL + 0.5 * n[1] * vL + 0.5 * smax - 0.5 * dsmaxL[1] * (qR[1] - qL[1]);
  JLm(1, 2) = 0.5 * n[1] * uL - 0.5 * dsmaxL[2] * (qR[1] - qL[1]);
  JLm(2, 0) = 0.5 * (g * n[1] * qL[0] - qL[2] * termL) - 0.5 * dsmaxL[0] * (qR[2] - qL[2]);
  JLm(2, 1) = 0.5 * n[0] * vL - 0.5 * dsmaxL[1] * (qR[2] - qL[2]);
  JLm(2, 2) = n[1] * vL + 0.5 * n[0] * uL + 0.5 * smax - 0.5 * dsmaxL[2] * (qR[2] - qL[2]);
  JRm(0, 0) = -0.5 * dsmaxR[0] * (qR[0] - qL[0]) + 0.5 * (n[0] * uR + n[1] * vR - qR[0] * termR) - 0.5 * smax;
  JRm(0, 1) = 0.5 * n[0] - 0.5 * dsmaxR[1] * (qR[0] - qL[0]);
  JRm(0, 2) = 0.5 * n[1] - 0.5 * dsmaxR[2] * (qR[0] - qL[0]);
  JRm(1, 0) = 0.5 * (g * n[0] * qR[0] - qR[1] * termR) - 0.5 * dsmaxR[0] * (qR[1] - qL[1]);
  JRm(1, 1) = n[0] * uR + 0.5 * n[1] * vR - 0.5 * smax - 0.5 * dsmaxR[1] * (qR[1] - qL[1]);
  JRm(1, 2) = 0.5 * n[1] * uR - 0.5 * dsmaxR[2] * (qR[1] - qL[1]);
  JRm(2, 0) = 0.5 * (g * n[1] * qR[0] - qR[2] * termR) - 0.5 * dsmaxR[0] * (qR[2] - qL[2]);
  JRm(2, 1) = 0.5 * n[0] * vR - 0.5 * dsmaxR[1] * (qR[2] - qL[2]);
  JRm(2, 2) = n[1] * vR + 0.5 * n[0] * uR - 0.5 * smax - 0.5 * dsmaxR[2] * (qR[2] - qL[2]);
#undef JLm
#undef JRm
}

/* ------------------------------------------------------------------ */
/* Burgers 2D: impl/advection_diffusion_2d_flux_functions.hpp:54-114   */
/* ------------------------------------------------------------------ */
static void burgers_flux(double *F, const double *qL, const double *qR, const double *n)
{
  const double fourInv = 1. / 4.;
  const double alpha_0 = fmax(fabs(qL[0]), fabs(qR[0]));
  const double alpha_1 = fmax(fabs(qL[1]), fabs(qR[1]));
  F[0] = alpha_0 * (qL[0] - qR[0]);
  F[0] += n[0] * (qL[0] * qL[0] + qR[0] * qR[0]);
  F[0] += n[1] * (qL[0] * qL[1] + qR[0] * qR[1]);
  F[0] *= fourInv;
  F[1] = alpha_1 * (qL[1] - qR[1]);
  F[1] += n[0] * (qL[0] * qL[1] + qR[0] * qR[1]);
  F[1] += n[1] * (qL[1] * qL[1] + qR[1] * qR[1]);
  F[1] *= fourInv;
}

static void burgers_fjac(double *JL, double *JR, const double *qL, const double *qR, const double *n)
{
  const double two = 2, fourInv = 1. / 4.;
  if (fabs(qL[0]) > fabs(qR[0])) {
    JL[0] = (two * qL[0] - qR[0]) * copysign(1., qL[0]) + n[0] * two * qL[0] + n[1] * qL[1];
    JR[0] = n[0] * two * qR[0] + n[1] * qR[1] - fabs(qL[0]);
  } else {
    JL[0] = n[0] * two * qL[0] + n[1] * qL[1] + fabs(qR[0]);
    JR[0] = (qL[0] - two * qR[0]) * copysign(1., qR[0]) + n[0] * two * qR[0] + n[1] * qR[1];
  }
  JL[0] *= fourInv; JR[0] *= fourInv;
  if (fabs(qL[1]) > fabs(qR[1])) {
    JL[3] = (two * qL[1] - qR[1]) * copysign(1., qL[1]) + n[0] * qL[0] + n[1] * two * qL[1];
    JR[3] = n[0] * qR[0] + n[1] * two * qR[1] - fabs(qL[1]);
  } else {
    JL[3] = n[0] * qL[0] + n[1] * two * qL[1] + fabs(qR[1]);
    JR[3] = (qL[1] - two * qR[1]) * copysign(1., qR[1]) + n[0] * qR[0] + n[1] * two * qR[1];
  }
  JL[3] *= fourInv; JR[3] *= fourInv;
  JL[1] = n[1] * qL[0] * fourInv;
  JL[2] = n[0] * qL[1] * fourInv;
  JR[1] = n[1] * qR[0] * fourInv;
  JR[2] = n[0] * qR[1] * fourInv;
}

/* ------------------------------------------------------------------ */
/* family dispatch                                                      */
/* ------------------------------------------------------------------ */
int orc_ndpc(int family)
{
  switch (family) {
  case ORC_EULER1D: return 3;
  case ORC_EULER2D: return 4;
  case ORC_EULER3D: return 5;
  case ORC_SWE2D: return 3;
  case ORC_BURGERS2D: return 2;
  case ORC_ADR2D: return 1;
  default: return -1;
  }
}

static int family_dim(int family)
{
  switch (family) {
  case ORC_EULER1D: return 1;
  case ORC_EULER3D: return 3;
  case ORC_EULER2D: case ORC_SWE2D: case ORC_BURGERS2D: case ORC_ADR2D: return 2;
  default: return -1;
  }
}

/* advection velocity for ADR (impl/advection_diffusion_reaction_2d_flux_mixin.hpp:87-105):
   the flux is the "Neg" state times the axis velocity */
static double adr_velocity(const orc_problem *pb, const double *n)
{
  return n[0] != 0. ? pb->p[0] : pb->p[1];
}

int orc_face_flux(const orc_problem *pb, int dim, const double *qL,
                  const double *qR, const double *n, double *F)
{
  switch (pb->family) {
  case ORC_EULER1D: euler_flux3(F, qL, qR, pb->p[0]); return 0;
  case ORC_EULER2D: euler_flux4(F, qL, qR, n, pb->p[0]); return 0;
  case ORC_EULER3D: euler_flux5(F, qL, qR, n, pb->p[0]); return 0;
  case ORC_SWE2D: swe_flux(F, qL, qR, n, pb->p[0]); return 0;
  case ORC_BURGERS2D: burgers_flux(F, qL, qR, n); return 0;
  case ORC_ADR2D: F[0] = qL[0] * adr_velocity(pb, n); return 0;
  }
  (void)dim;
  return -1;
}

int orc_face_flux_jac(const orc_problem *pb, int dim, const double *qL,
                      const double *qR, const double *n, double *JL, double *JR)
{
  switch (pb->family) {
  case ORC_EULER1D: euler_fjac3(JL, JR, qL, qR, pb->p[0]); return 0;
  case ORC_EULER2D: euler_fjac4(JL, JR, qL, qR, n, pb->p[0]); return 0;
  case ORC_EULER3D: euler_fjac5(JL, JR, qL, qR, n, pb->p[0]); return 0;
  case ORC_SWE2D: swe_fjac(JL, JR, qL, qR, n, pb->p[0]); return 0;
  case ORC_BURGERS2D: burgers_fjac(JL, JR, qL, qR, n); return 0;
  case ORC_ADR2D: JL[0] = adr_velocity(pb, n); JR[0] = 0; return 0;
  }
  (void)dim;
  return -1;
}

/* ------------------------------------------------------------------ */
/* mesh helpers: impl/mesh_ccu.hpp                                     */
/* ------------------------------------------------------------------ */
static int ncols_of(const orc_mesh *m) { return 1 + (m->stencil - 1) * m->dim; }

static int32_t G(const orc_mesh *m, int64_t r, int c) { return m->graph[r * ncols_of(m) + c]; }

/* graph column of the neighbour at layer L on the minus (side=0) or plus
   (side=1) side of axis a; layouts from impl/functor_fill_stencil.hpp */
static int col_of(int dim, int a, int side, int L)
{
  if (dim == 1) return 1 + 2 * L + side;
  if (dim == 2) {
    const int base = 1 + 4 * L;
    if (a == 0) return base + (side ? 2 : 0);  /* left / right */
    return base + (side ? 1 : 3);               /* back / front */
  }
  const int base = 1 + 6 * L;
  if (a == 0) return base + (side ? 2 : 0);
  if (a == 1) return base + (side ? 1 : 3);
  return base + (side ? 5 : 4);                 /* bottom / top */
}

/* "has boundary within the mesh stencil width" predicates,
   impl/mesh_ccu.hpp:162-296; 3D stencil-7 meshes (not supported by the
   reference) check all three layers like 1D/2D do. */
static int has_bd(const orc_mesh *m, int64_t r, int a, int side)
{
  const int layers = (m->stencil - 1) / 2;
  for (int L = 0; L < layers; ++L)
    if (G(m, r, col_of(m->dim, a, side, L)) == -1) return 1;
  return 0;
}

static int is_near_bd(const orc_mesh *m, int64_t r)
{
  for (int a = 0; a < m->dim; ++a)
    for (int s = 0; s < 2; ++s)
      if (has_bd(m, r, a, s)) return 1;
  return 0;
}

/* ------------------------------------------------------------------ */
/* evaluation context                                                   */
/* ------------------------------------------------------------------ */
typedef struct {
  const orc_mesh *mesh;
  const orc_problem *pb;
  int m, dim, w;
  int64_t *ghostRow;   /* per graph row: index in the near-boundary list or -1 */
  int64_t nNear;
  double *ghost[3][2]; /* [axis][side] -> nNear x (w*m) */
  double hInv[3];
  /* CSR */
  const int64_t *rowptr;
  const int32_t *cols;
  double *vals;
} ctx_t;

static int scheme_w(int scheme) { return scheme == ORC_WENO5 ? 3 : scheme == ORC_WENO3 ? 2 : 1; }

static int setup(ctx_t *c, const orc_mesh *mesh, const orc_problem *pb)
{
  memset(c, 0, sizeof(*c));
  c->mesh = mesh; c->pb = pb;
  c->m = orc_ndpc(pb->family);
  c->dim = family_dim(pb->family);
  c->w = scheme_w(pb->scheme);
  if (c->m < 0 || c->dim != mesh->dim) return -1;
  if (mesh->stencil < 2 * c->w + 1) return -2;
  for (int a = 0; a < c->dim; ++a) c->hInv[a] = 1. / mesh->dx[a];
  c->ghostRow = (int64_t *)malloc(sizeof(int64_t) * (size_t)(mesh->nrows > 0 ? mesh->nrows : 1));
  int64_t nn = 0;
  for (int64_t r = 0; r < mesh->nrows; ++r) c->ghostRow[r] = is_near_bd(mesh, r) ? nn++ : -1;
  c->nNear = nn;
  const size_t gsz = (size_t)(nn > 0 ? nn : 1) * (size_t)(c->w * c->m);
  for (int a = 0; a < c->dim; ++a)
    for (int s = 0; s < 2; ++s) {
      c->ghost[a][s] = (double *)malloc(sizeof(double) * gsz);
      for (size_t i = 0; i < gsz; ++i) c->ghost[a][s][i] = DBL_MIN;
    }
  return 0;
}

static void teardown(ctx_t *c)
{
  free(c->ghostRow);
  for (int a = 0; a < 3; ++a)
    for (int s = 0; s < 2; ++s) free(c->ghost[a][s]);
}

static double *GH(ctx_t *c, int a, int side, int64_t g) { return c->ghost[a][side] + g * (c->w * c->m); }

/* ------------------------------------------------------------------ */
/* ghost filling                                                        */
/* ------------------------------------------------------------------ */
static void copy_cell(double *dst, const double *U, int64_t cell, int m, const double *sgn)
{
  for (int k = 0; k < m; ++k) dst[k] = sgn[k] * U[cell * m + k];
}

/* Euler 1D Neumann, impl/euler_1d_ghost_filler.hpp:77-143 */
static void ghosts_euler1d(ctx_t *c, const double *U)
{
  const orc_mesh *M = c->mesh;
  const double one[3] = {1, 1, 1};
  for (int64_t r = 0; r < M->nrows; ++r) {
    const int64_t g = c->ghostRow[r];
    if (g < 0) continue;
    const int32_t self = G(M, r, 0), l0 = G(M, r, 1), r0 = G(M, r, 2);
    if (l0 == -1) copy_cell(GH(c, 0, 0, g), U, self, 3, one);
    if (r0 == -1) copy_cell(GH(c, 0, 1, g), U, self, 3, one);
    if (c->w >= 2) {
      const int32_t l1 = G(M, r, 3), r1 = G(M, r, 4);
      if (l1 == -1) copy_cell(GH(c, 0, 0, g) + 3, U, r0, 3, one);
      if (r1 == -1) copy_cell(GH(c, 0, 1, g) + 3, U, l0, 3, one);
    }
    if (c->w == 3) {
      const int32_t l1 = G(M, r, 3), r1 = G(M, r, 4), l2 = G(M, r, 5), r2 = G(M, r, 6);
      if (l2 == -1) copy_cell(GH(c, 0, 0, g) + 6, U, r1, 3, one);
      if (r2 == -1) copy_cell(GH(c, 0, 1, g) + 6, U, l1, 3, one);
    }
  }
}

/* 2D mirror-type fillers: Euler Neumann impl/euler_2d_ghost_filler_neumann.hpp:95-279,
   shallow-water slip wall impl/swe_2d_ghost_filler_inviscid_wall.hpp:95-270 (same cells,
   sign flip of the wall-normal momentum). */
static void ghosts_mirror2d(ctx_t *c, const double *U, int swe)
{
  const orc_mesh *M = c->mesh;
  const int m = c->m;
  for (int64_t r = 0; r < M->nrows; ++r) {
    const int64_t g = c->ghostRow[r];
    if (g < 0) continue;
    const int32_t self = G(M, r, 0);
    for (int a = 0; a < 2; ++a)
      for (int side = 0; side < 2; ++side) {
        double sgn[MAXM] = {1, 1, 1, 1, 1};
        if (swe) sgn[a == 0 ? 1 : 2] = -1;
        const int opp = 1 - side;
        const int32_t n0 = G(M, r, col_of(2, a, side, 0));
        const int32_t o0 = G(M, r, col_of(2, a, opp, 0));
        double *dst = GH(c, a, side, g);
        if (n0 == -1) copy_cell(dst, U, self, m, sgn);
        if (c->w >= 2) {
          const int32_t n1 = G(M, r, col_of(2, a, side, 1));
          if (n1 == -1) copy_cell(dst + m, U, (n0 == -1) ? o0 : n0, m, sgn);
        }
        if (c->w == 3) {
          const int32_t n1 = G(M, r, col_of(2, a, side, 1));
          const int32_t n2 = G(M, r, col_of(2, a, side, 2));
          const int32_t o1 = G(M, r, col_of(2, a, opp, 1));
          if (n2 == -1) {
            int32_t src = self;
            if (n1 != -1 && n0 != -1) src = n1;
            if (n1 == -1 && n0 != -1) src = self;
            if (n1 == -1 && n0 == -1) src = o1;
            copy_cell(dst + 2 * m, U, src, m, sgn);
          }
        }
      }
  }
}

/* Euler 3D Sedov symmetry, impl/euler_3d_ghost_filler_sedov.hpp:79-236:
   reflective at the minus sides (x,y,z momentum flipped), Neumann at the
   plus sides; layer L>0 copies the opposite layer-(L-1) neighbour.  The
   layer-2 rule (WENO5 in 3D) extends the same pattern; the reference has
   no 3D WENO5. */
static void ghosts_sedov3d(ctx_t *c, const double *U)
{
  const orc_mesh *M = c->mesh;
  for (int64_t r = 0; r < M->nrows; ++r) {
    const int64_t g = c->ghostRow[r];
    if (g < 0) continue;
    const int32_t self = G(M, r, 0);
    for (int a = 0; a < 3; ++a)
      for (int side = 0; side < 2; ++side) {
        double sgn[MAXM] = {1, 1, 1, 1, 1};
        if (side == 0) sgn[1 + a] = -1;
        double *dst = GH(c, a, side, g);
        for (int L = 0; L < c->w; ++L) {
          const int32_t nL = G(M, r, col_of(3, a, side, L));
          if (nL != -1) continue;
          const int32_t src = (L == 0) ? self : G(M, r, col_of(3, a, 1 - side, L - 1));
          copy_cell(dst + L * 5, U, src, 5, sgn);
        }
      }
  }
}

/* ADR problem A homogeneous Dirichlet,
   impl/advection_diffusion_reaction_2d_ghost_filler_problemA.hpp:85-133 */
static void ghosts_adr(ctx_t *c, const double *U)
{
  const orc_mesh *M = c->mesh;
  const double neg[1] = {-1};
  for (int64_t r = 0; r < M->nrows; ++r) {
    const int64_t g = c->ghostRow[r];
    if (g < 0) continue;
    const int32_t self = G(M, r, 0);
    for (int a = 0; a < 2; ++a)
      for (int side = 0; side < 2; ++side) {
        double *dst = GH(c, a, side, g);
        for (int L = 0; L < c->w; ++L) {
          const int32_t nL = G(M, r, col_of(2, a, side, L));
          if (nL != -1) continue;
          const int32_t src = (L == 0) ? self : G(M, r, col_of(2, a, 1 - side, L - 1));
          copy_cell(dst + L, U, src, 1, neg);
        }
      }
  }
}

/* Burgers 2D outflow, impl/advection_diffusion_2d_ghost_filler_outflow.hpp:
   homogeneous Dirichlet (every ghost 0) at the left and back sides, the
   mirror images about the boundary (homogeneous Neumann) at the right and
   front sides, with the layer rules of the 2D Euler Neumann filler */
static void ghosts_burgers_outflow(ctx_t *c, const double *U)
{
  const orc_mesh *M = c->mesh;
  for (int64_t r = 0; r < M->nrows; ++r) {
    const int64_t g = c->ghostRow[r];
    if (g < 0) continue;
    const int32_t self = G(M, r, 0);
    for (int a = 0; a < 2; ++a)
      for (int side = 0; side < 2; ++side) {
        const double v = side == 0 ? 0. : 1.;
        const double sgn[MAXM] = {v, v, v, v, v};
        const int opp = 1 - side;
        const int32_t n0 = G(M, r, col_of(2, a, side, 0));
        const int32_t o0 = G(M, r, col_of(2, a, opp, 0));
        double *dst = GH(c, a, side, g);
        if (n0 == -1) copy_cell(dst, U, self, 2, sgn);
        if (c->w >= 2) {
          const int32_t n1 = G(M, r, col_of(2, a, side, 1));
          if (n1 == -1) copy_cell(dst + 2, U, (n0 == -1) ? o0 : n0, 2, sgn);
        }
        if (c->w == 3) {
          const int32_t n1 = G(M, r, col_of(2, a, side, 1));
          const int32_t n2 = G(M, r, col_of(2, a, side, 2));
          const int32_t o1 = G(M, r, col_of(2, a, opp, 1));
          if (n2 == -1) {
            int32_t src = self;
            if (n1 != -1 && n0 != -1) src = n1;
            if (n1 == -1 && n0 != -1) src = self;
            if (n1 == -1 && n0 == -1) src = o1;
            copy_cell(dst + 4, U, src, 2, sgn);
          }
        }
      }
  }
}

static int fill_ghosts(ctx_t *c, const double *U)
{
  const orc_problem *pb = c->pb;
  switch (pb->family) {
  case ORC_EULER1D:
    if (pb->prob >= 1 && pb->prob <= 3) ghosts_euler1d(c, U);
    return 0;
  case ORC_EULER2D:
    if (pb->prob == 2 || pb->prob == 4 || pb->prob == 8) ghosts_mirror2d(c, U, 0);
    else if (pb->prob != 0 && pb->prob != 1) return -3; /* not restated */
    return 0;
  case ORC_EULER3D:
    if (pb->prob == 1) ghosts_sedov3d(c, U);
    return 0;
  case ORC_SWE2D:
    if (pb->prob != 0) return -3;
    ghosts_mirror2d(c, U, 1);
    return 0;
  case ORC_BURGERS2D:
    if (pb->prob == 1) ghosts_burgers_outflow(c, U);
    else if (pb->prob != 0) return -3;
    return 0;
  case ORC_ADR2D:
    ghosts_adr(c, U);
    return 0;
  }
  return -1;
}

/* Jacobian factors for the ghost dependence of first-order boundary blocks:
   Euler 1D impl/euler_1d_prob_class.hpp:608-610; Euler 2D
   impl/euler_2d_prob_class.hpp:1115-1224; Euler 3D
   impl/euler_3d_prob_class.hpp:1017-1044; SWE impl/swe_2d_prob_class.hpp:839-854;
   ADR impl/advection_diffusion_reaction_2d_prob_class.hpp (Dirichlet: -1);
   Burgers outflow impl/advection_diffusion_2d_prob_class.hpp:1131-1156
   (0 at the left and back boundaries, 1 at the right and front ones). */
static void bc_factors(ctx_t *c, int64_t r, int a, double *f)
{
  for (int k = 0; k < MAXM; ++k) f[k] = 1.;
  const orc_problem *pb = c->pb;
  switch (pb->family) {
  case ORC_SWE2D: f[a == 0 ? 1 : 2] = -1.; break;
  case ORC_EULER3D:
    if (has_bd(c->mesh, r, a, 0)) f[1 + a] = -1.;
    break;
  case ORC_ADR2D: f[0] = -1.; break;
  case ORC_BURGERS2D:
    if (has_bd(c->mesh, r, a, 0)) f[0] = f[1] = 0.;
    break;
  default: break;
  }
}

/* ------------------------------------------------------------------ */
/* stencil gather (impl/functor_fill_stencil.hpp) and reconstruction    */
/* ------------------------------------------------------------------ */
/* st[p*m+k], p = 0..2w, centre at p=w */
static void gather(ctx_t *c, int64_t r, int a, int w, const double *U, double *st)
{
  const orc_mesh *M = c->mesh;
  const int m = c->m;
  const int64_t g = c->ghostRow[r];
  const int32_t self = G(M, r, 0);
  for (int k = 0; k < m; ++k) st[w * m + k] = U[(int64_t)self * m + k];
  for (int side = 0; side < 2; ++side)
    for (int L = 0; L < w; ++L) {
      const int32_t nb = G(M, r, col_of(M->dim, a, side, L));
      const int p = side ? (w + 1 + L) : (w - 1 - L);
      for (int k = 0; k < m; ++k)
        st[p * m + k] = (nb == -1) ? GH(c, a, side, g)[L * m + k] : U[(int64_t)nb * m + k];
    }
}

/* face states; the left face of the cell uses stencil points 0..2w-1,
   the right face 1..2w (weno.hpp:63-87, 212-236; first order
   impl/functor_reconstruct_from_state.hpp:262-268) */
typedef struct {
  double lm[MAXM], lp[MAXM], rm[MAXM], rp[MAXM];
  /* gradients d(face)/d(stencil point), per dof: [k][pos] with pos 0..2w-1 */
  double glm[MAXM][6], glp[MAXM][6], grm[MAXM][6], grp[MAXM][6];
} faces_t;

static void reconstruct(int scheme, int m, int w, const double *st, faces_t *f, int withGrad)
{
  for (int k = 0; k < m; ++k) {
    double q[7];
    for (int p = 0; p <= 2 * w; ++p) q[p] = st[p * m + k];
    if (scheme == ORC_FIRSTORDER) {
      f->lm[k] = q[0]; f->lp[k] = q[1]; f->rm[k] = q[1]; f->rp[k] = q[2];
    } else if (scheme == ORC_WENO3) {
      if (withGrad) {
        orc_weno3_grad(&f->lm[k], &f->lp[k], f->glm[k], f->glp[k], q);
        orc_weno3_grad(&f->rm[k], &f->rp[k], f->grm[k], f->grp[k], q + 1);
      } else {
        orc_weno3(&f->lm[k], &f->lp[k], q);
        orc_weno3(&f->rm[k], &f->rp[k], q + 1);
      }
    } else {
      if (withGrad) {
        orc_weno5_grad(&f->lm[k], &f->lp[k], f->glm[k], f->glp[k], q);
        orc_weno5_grad(&f->rm[k], &f->rp[k], f->grm[k], f->grp[k], q + 1);
      } else {
        orc_weno5(&f->lm[k], &f->lp[k], q);
        orc_weno5(&f->rm[k], &f->rp[k], q + 1);
      }
    }
  }
}

static void normal_of(int dim, int a, double *n)
{
  n[0] = n[1] = n[2] = 0.;
  if (dim > 1) n[a] = 1.;
  else n[0] = 1.;
}

/* ------------------------------------------------------------------ */
/* CSR update, the equivalent of Eigen's coeffRef on a compressed row   */
/* ------------------------------------------------------------------ */
static void jadd(ctx_t *c, int64_t row, int64_t col, double v)
{
  int64_t lo = c->rowptr[row], hi = c->rowptr[row + 1] - 1;
  while (lo <= hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (c->cols[mid] == col) { c->vals[mid] += v; return; }
    if (c->cols[mid] < col) lo = mid + 1; else hi = mid - 1;
  }
  abort(); /* pattern mismatch: never expected */
}

/* ------------------------------------------------------------------ */
/* per-row evaluation                                                   */
/* ------------------------------------------------------------------ */
/* interior Jacobian, impl/mixin_directional_flux_balance_jacobian.hpp:142-284 */
static void jac_interior(ctx_t *c, int64_t r, int a, const faces_t *f,
                         const double *JLn, const double *JLp,
                         const double *JRn, const double *JRp)
{
  const orc_mesh *M = c->mesh;
  const int m = c->m, w = c->w;
  const double h = c->hInv[a];
  const int64_t row = r * m;
  int64_t colp[7]; /* column of stencil point p (p=0..2w) */
  colp[w] = (int64_t)G(M, r, 0) * m;
  for (int L = 0; L < w; ++L) {
    colp[w - 1 - L] = (int64_t)G(M, r, col_of(M->dim, a, 0, L)) * m;
    colp[w + 1 + L] = (int64_t)G(M, r, col_of(M->dim, a, 1, L)) * m;
  }
#define A(X, k, j) X[(k) * m + (j)]
  if (c->pb->scheme == ORC_FIRSTORDER) {
    for (int k = 0; k < m; ++k)
      for (int j = 0; j < m; ++j) {
        jadd(c, row + k, colp[0] + j, A(JLn, k, j) * h);
        jadd(c, row + k, colp[1] + j, (A(JLp, k, j) - A(JRn, k, j)) * h);
        jadd(c, row + k, colp[2] + j, -A(JRp, k, j) * h);
      }
    return;
  }
  const int np = 2 * w; /* grads per face */
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < m; ++j) {
      for (int p = 0; p < np; ++p) {
        jadd(c, row + k, colp[p] + j, A(JLn, k, j) * f->glm[j][p] * h);
        jadd(c, row + k, colp[p] + j, A(JLp, k, j) * f->glp[j][p] * h);
      }
      for (int p = 0; p < np; ++p) {
        jadd(c, row + k, colp[p + 1] + j, -(A(JRn, k, j) * f->grm[j][p] * h));
        jadd(c, row + k, colp[p + 1] + j, -(A(JRp, k, j) * f->grp[j][p] * h));
      }
    }
#undef A
}

/* first-order boundary Jacobian,
   impl/mixin_directional_flux_balance_jacobian.hpp:311-371 */
static void jac_boundary(ctx_t *c, int64_t r, int a,
                         const double *JLn, const double *JLp,
                         const double *JRn, const double *JRp)
{
  const orc_mesh *M = c->mesh;
  const int m = c->m;
  const double h = c->hInv[a];
  const int64_t row = r * m;
  const int32_t l0 = G(M, r, col_of(M->dim, a, 0, 0));
  const int32_t r0 = G(M, r, col_of(M->dim, a, 1, 0));
  const int64_t col_i = (int64_t)G(M, r, 0) * m;
  double fac[MAXM];
  bc_factors(c, r, a, fac);
#define A(X, k, j) X[(k) * m + (j)]
  for (int k = 0; k < m; ++k)
    for (int j = 0; j < m; ++j) jadd(c, row + k, col_i + j, (A(JLp, k, j) - A(JRn, k, j)) * h);
  if (l0 != -1)
    for (int k = 0; k < m; ++k)
      for (int j = 0; j < m; ++j) jadd(c, row + k, (int64_t)l0 * m + j, A(JLn, k, j) * h);
  if (r0 != -1)
    for (int k = 0; k < m; ++k)
      for (int j = 0; j < m; ++j) jadd(c, row + k, (int64_t)r0 * m + j, -A(JRp, k, j) * h);
  if (l0 == -1)
    for (int k = 0; k < m; ++k)
      for (int j = 0; j < m; ++j) jadd(c, row + k, col_i + j, (fac[j] * A(JLn, k, j)) * h);
  if (r0 == -1)
    for (int k = 0; k < m; ++k)
      for (int j = 0; j < m; ++j) jadd(c, row + k, col_i + j, (fac[j] * -A(JRp, k, j)) * h);
#undef A
}

/* sources after the axis sweeps */
static void sources(ctx_t *c, int64_t r, const double *U, double t, double *Vr, int withJac)
{
  const orc_mesh *M = c->mesh;
  const orc_problem *pb = c->pb;
  const int m = c->m;
  (void)t;
  if (pb->family == ORC_SWE2D) {
    /* Coriolis, impl/swe_2d_prob_class.hpp:984-1012 */
    const double f = pb->p[1];
    const int64_t ci = (int64_t)G(M, r, 0) * m;
    Vr[1] -= f * U[ci + 2] / U[ci];
    Vr[2] += f * U[ci + 1] / U[ci];
    if (withJac) {
      const int64_t row = r * m;
      jadd(c, row + 1, ci, f * U[ci + 2] / (U[ci] * U[ci]));
      jadd(c, row + 1, ci + 2, -f / U[ci]);
      jadd(c, row + 2, ci + 1, f / U[ci]);
      jadd(c, row + 2, ci, -f * U[ci + 1] / (U[ci] * U[ci]));
    }
  } else if (pb->family == ORC_BURGERS2D && c->ghostRow[r] < 0) {
    /* diffusion on inner cells, impl/advection_diffusion_2d_prob_class.hpp:1167-1204 */
    const double two = 2;
    const double dxInvSq = c->hInv[0] * c->hInv[0], dyInvSq = c->hInv[1] * c->hInv[1];
    const double dX = pb->p[0] * dxInvSq, dY = pb->p[0] * dyInvSq;
    const int64_t ui = (int64_t)G(M, r, 0) * m, uL = (int64_t)G(M, r, 1) * m, uF = (int64_t)G(M, r, 2) * m;
    const int64_t uR = (int64_t)G(M, r, 3) * m, uB = (int64_t)G(M, r, 4) * m;
    Vr[0] += dX * (U[uR] - two * U[ui] + U[uL]);
    Vr[0] += dY * (U[uF] - two * U[ui] + U[uB]);
    Vr[1] += dX * (U[uR + 1] - two * U[ui + 1] + U[uL + 1]);
    Vr[1] += dY * (U[uF + 1] - two * U[ui + 1] + U[uB + 1]);
    if (withJac) {
      const int64_t row = r * m;
      for (int k = 0; k < 2; ++k) {
        jadd(c, row + k, ui + k, -two * dX - two * dY);
        jadd(c, row + k, uL + k, dX);
        jadd(c, row + k, uF + k, dY);
        jadd(c, row + k, uR + k, dX);
        jadd(c, row + k, uB + k, dY);
      }
    }
  } else if (pb->family == ORC_BURGERS2D && withJac) {
    /* diffusion Jacobian of a near-boundary cell (its velocity was added
       after each axis, burgers_diffusion_near): the diagonal, then per
       first neighbour (left, front, right, back) its weight, or minus the
       weight on the diagonal where the neighbour is missing,
       impl/advection_diffusion_2d_prob_class.hpp:752-790 */
    const double two = 2;
    const double dX = pb->p[0] * (c->hInv[0] * c->hInv[0]), dY = pb->p[0] * (c->hInv[1] * c->hInv[1]);
    const int64_t row = r * m, ui = (int64_t)G(M, r, 0) * m;
    const int32_t nb[4] = {G(M, r, 1), G(M, r, 2), G(M, r, 3), G(M, r, 4)};
    const double wt[4] = {dX, dY, dX, dY};
    for (int k = 0; k < 2; ++k) jadd(c, row + k, ui + k, -two * dX - two * dY);
    for (int q = 0; q < 4; ++q)
      for (int k = 0; k < 2; ++k) {
        if (nb[q] > -1) jadd(c, row + k, (int64_t)nb[q] * m + k, wt[q]);
        else jadd(c, row + k, ui + k, -wt[q]);
      }
  } else if (pb->family == ORC_ADR2D) {
    /* ADR problem A, impl/advection_diffusion_reaction_2d_prob_class.hpp:
       inner rows add both axes' diffusion here (:1044-1049 with
       addDiffusionToVelocityAndOptionalJacobianInnerCells :1057-1083),
       near-boundary rows added theirs after each axis (adr_diffusion_near);
       then the source f = 1 (advection_diffusion_reaction2d.hpp:87-99) and
       the reaction -sigma*u (:1051-1052, Jacobian :509-511 and :840-872) */
    const double two = 2;
    const double dX = pb->p[2] * (c->hInv[0] * c->hInv[0]);
    const double dY = pb->p[2] * (c->hInv[1] * c->hInv[1]);
    const double sigma = pb->p[3];
    const int64_t row = r;
    const int32_t ui = G(M, r, 0), uL = G(M, r, 1), uF = G(M, r, 2), uR = G(M, r, 3), uB = G(M, r, 4);
    const int near = c->ghostRow[r] >= 0;
    if (!near) {
      Vr[0] += dX * (U[uR] - two * U[ui] + U[uL]);
      Vr[0] += dY * (U[uF] - two * U[ui] + U[uB]);
      if (withJac) {
        jadd(c, row, ui, -two * dX - two * dY);
        jadd(c, row, uL, dX);
        jadd(c, row, uF, dY);
        jadd(c, row, uR, dX);
        jadd(c, row, uB, dY);
      }
    }
    Vr[0] += 1.0;
    Vr[0] -= sigma * U[ui];
    if (withJac) {
      if (!near) {
        jadd(c, row, ui, -sigma);
      } else {
        /* the missing neighbour is the ghost -u of the cell itself */
        double self = -two * dX - two * dY - sigma;
        if (uL != -1) jadd(c, row, uL, dX); else self += -dX;
        if (uF != -1) jadd(c, row, uF, dY); else self += -dY;
        if (uR != -1) jadd(c, row, uR, dX); else self += -dX;
        if (uB != -1) jadd(c, row, uB, dY); else self += -dY;
        jadd(c, row, ui, self);
      }
    }
  }
}

/* Burgers diffusion of a near-boundary row along axis a from its stencil,
   a missing neighbour read from the ghost layer,
   impl/advection_diffusion_2d_prob_class.hpp:721-729 and :892-894 */
static void burgers_diffusion_near(ctx_t *c, int64_t r, int a, const double *U, double *Vr)
{
  double st[3 * 2];
  gather(c, r, a, 1, U, st);
  const double two = 2;
  const double d = c->pb->p[0] * (c->hInv[a] * c->hInv[a]);
  Vr[0] += d * (st[4] - two * st[2] + st[0]);
  Vr[1] += d * (st[5] - two * st[3] + st[1]);
}

/* ADR diffusion of a near-boundary row along axis a from its stencil, the
   missing neighbour read from the ghost layer,
   impl/advection_diffusion_reaction_2d_prob_class.hpp:946-976 and :658-676 */
static void adr_diffusion_near(ctx_t *c, int64_t r, int a, const double *U, double *Vr)
{
  double st[3];
  gather(c, r, a, 1, U, st);
  const double two = 2;
  const double d = c->pb->p[2] * (c->hInv[a] * c->hInv[a]);
  Vr[0] += d * (st[2] - two * st[1] + st[0]);
}

static void eval_row(ctx_t *c, int64_t r, const double *U, double t, double *V, int withJac)
{
  const orc_problem *pb = c->pb;
  const int m = c->m, dim = c->dim, w = c->w;
  const int near = c->ghostRow[r] >= 0;
  double Vr[MAXM] = {0, 0, 0, 0, 0};
  double st[7 * MAXM], n[3];
  faces_t f;
  double FL[MAXM], FR[MAXM];
  double JLn[25], JLp[25], JRn[25], JRp[25];
  for (int a = 0; a < dim; ++a) {
    normal_of(dim, a, n);
    const double h = c->hInv[a];
    if (!withJac || (near && pb->scheme != ORC_FIRSTORDER)) {
      /* values only: plain WENO from the stencil */
      gather(c, r, a, w, U, st);
      reconstruct(pb->scheme, m, w, st, &f, 0);
      orc_face_flux(pb, dim, f.lm, f.lp, n, FL);
      orc_face_flux(pb, dim, f.rm, f.rp, n, FR);
      for (int k = 0; k < m; ++k) Vr[k] += h * (FL[k] - FR[k]);
      if (withJac) {
        /* first-order Jacobian of a near-boundary cell */
        gather(c, r, a, 1, U, st);
        faces_t f1;
        reconstruct(ORC_FIRSTORDER, m, 1, st, &f1, 0);
        orc_face_flux_jac(pb, dim, f1.lm, f1.lp, n, JLn, JLp);
        orc_face_flux_jac(pb, dim, f1.rm, f1.rp, n, JRn, JRp);
        jac_boundary(c, r, a, JLn, JLp, JRn, JRp);
      }
    } else {
      gather(c, r, a, w, U, st);
      reconstruct(pb->scheme, m, w, st, &f, 1);
      orc_face_flux(pb, dim, f.lm, f.lp, n, FL);
      orc_face_flux(pb, dim, f.rm, f.rp, n, FR);
      orc_face_flux_jac(pb, dim, f.lm, f.lp, n, JLn, JLp);
      orc_face_flux_jac(pb, dim, f.rm, f.rp, n, JRn, JRp);
      for (int k = 0; k < m; ++k) Vr[k] += h * (FL[k] - FR[k]);
      if (near) jac_boundary(c, r, a, JLn, JLp, JRn, JRp);
      else jac_interior(c, r, a, &f, JLn, JLp, JRn, JRp);
    }
    if (pb->family == ORC_ADR2D && near) adr_diffusion_near(c, r, a, U, Vr);
    if (pb->family == ORC_BURGERS2D && near) burgers_diffusion_near(c, r, a, U, Vr);
  }
  sources(c, r, U, t, Vr, withJac);
  for (int k = 0; k < m; ++k) V[r * m + k] = Vr[k];
}

int orc_velocity(const orc_mesh *mesh, const orc_problem *pb,
                 const double *U, double t, double *V)
{
  ctx_t c;
  int rc = setup(&c, mesh, pb);
  if (rc) return rc;
  rc = fill_ghosts(&c, U);
  if (!rc)
    for (int64_t r = 0; r < mesh->nrows; ++r) eval_row(&c, r, U, t, V, 0);
  teardown(&c);
  return rc;
}

static int cmp_i32(const void *a, const void *b)
{
  const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

/* sparsity, e.g. impl/euler_2d_prob_class.hpp:223-237 and 314-387 */
int orc_jacobian_pattern(const orc_mesh *mesh, const orc_problem *pb,
                         int64_t *nnz, int64_t *rowptr, int32_t *cols)
{
  ctx_t c;
  int rc = setup(&c, mesh, pb);
  if (rc) return rc;
  const int m = c.m, dim = c.dim;
  int64_t total = 0;
  if (rowptr) rowptr[0] = 0;
  for (int64_t r = 0; r < mesh->nrows; ++r) {
    int32_t cells[1 + 6 * 3];
    int nc = 0;
    cells[nc++] = G(mesh, r, 0);
    const int near = c.ghostRow[r] >= 0;
    const int nnb = near ? 2 * dim : 2 * dim * c.w;
    for (int i = 1; i <= nnb; ++i) {
      const int32_t nb = G(mesh, r, i);
      if (nb != -1) cells[nc++] = nb;
    }
    qsort(cells, (size_t)nc, sizeof(int32_t), cmp_i32);
    int nu = 0;
    for (int i = 0; i < nc; ++i)
      if (i == 0 || cells[i] != cells[i - 1]) cells[nu++] = cells[i];
    for (int k = 0; k < m; ++k) {
      if (cols)
        for (int i = 0; i < nu; ++i)
          for (int j = 0; j < m; ++j) cols[total + i * m + j] = cells[i] * m + j;
      total += (int64_t)nu * m;
      if (rowptr) rowptr[r * m + k + 1] = total;
    }
  }
  *nnz = total;
  teardown(&c);
  return 0;
}

int orc_velocity_and_jacobian(const orc_mesh *mesh, const orc_problem *pb,
                              const double *U, double t, double *V,
                              const int64_t *rowptr, const int32_t *cols, double *vals)
{
  ctx_t c;
  int rc = setup(&c, mesh, pb);
  if (rc) return rc;
  c.rowptr = rowptr; c.cols = cols; c.vals = vals;
  memset(vals, 0, sizeof(double) * (size_t)rowptr[mesh->nrows * c.m]);
  rc = fill_ghosts(&c, U);
  if (!rc)
    for (int64_t r = 0; r < mesh->nrows; ++r) eval_row(&c, r, U, t, V, 1);
  teardown(&c);
  return rc;
}
