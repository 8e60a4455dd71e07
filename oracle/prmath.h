/* CPU oracle: exp / log / pow used by the reactor arithmetic (test infrastructure only).
 *
 * The HIP kernels evaluate the same functions with the same operation sequence (their own
 * device copy), so the oracle and the GPU reactor produce bit-identical results on the same
 * inputs: every step of both is an IEEE-754 double operation (+ - * / sqrt, fma, rint, frexp,
 * ldexp are correctly rounded or exact on both targets) with no implicit contraction.
 *
 * exp: Cody & Waite argument reduction x = k ln2 + r, |r| <= ln2/2, with ln2 split in two parts,
 *      then the degree-13 Taylor polynomial of e^r in Horner form (truncation < 2^-58 relative),
 *      scaled by 2^k.
 * log: x = 2^e m with m in [sqrt(1/2), sqrt(2)), f = m - 1, s = f / (2 + f),
 *      log m = 2 atanh(s) = 2 s (1 + s^2/3 + s^4/5 + ... + s^20/21)  (|s| <= 0.1716),
 *      log x = e ln2 + log m.
 * Both are within a few ulp of the correctly rounded result (checked against libm in tests).
 */
#ifndef PR_ORACLE_PRMATH_H
#define PR_ORACLE_PRMATH_H
#include <math.h>

#define PRM_LN2_HI 6.93147180559945286227e-01
#define PRM_LN2_LO 2.31904681384629955842e-17
#define PRM_LOG2E 1.44269504088896338700e+00
#define PRM_SQRT_HALF 7.07106781186547572737e-01
#define PRM_LN10 2.30258509299404590109e+00
#define PRM_INV_LN10 4.34294481903251816668e-01

static inline double prm_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -745.1332191019412) return 0.0;
  double k = rint(x * PRM_LOG2E);
  double r = fma(-k, PRM_LN2_HI, x);
  r = fma(-k, PRM_LN2_LO, r);
  double p = 1.0 / 6227020800.0;               /* 1/13! */
  p = fma(p, r, 1.0 / 479001600.0);            /* 1/12! */
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}

static inline double prm_log(double x) {
  if (!(x > 0.0)) return (x == 0.0) ? -HUGE_VAL : (x - x) / (x - x);
  if (x == HUGE_VAL) return x;
  int e;
  double m = frexp(x, &e);
  if (m < PRM_SQRT_HALF) {
    m = m * 2.0;
    e = e - 1;
  }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, z, 1.0 / 19.0);
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  double t = z * p;                    /* s^2/3 + s^4/5 + ... */
  double s2 = s + s;
  double logm = fma(s2, t, s2);        /* 2 s (1 + t) */
  double de = (double)e;
  return fma(de, PRM_LN2_HI, fma(de, PRM_LN2_LO, logm));
}

static inline double prm_log10(double x) { return prm_log(x) * PRM_INV_LN10; }

static inline double prm_pow(double x, double y) { return prm_exp(y * prm_log(x)); }

#endif
