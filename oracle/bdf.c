/* CPU oracle: variable-order BDF with CVODE semantics, and a dense LU with partial pivoting.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Restates the algorithm PelePhysics' ReactorCvode
 * drives through SUNDIALS CVODE (CV_BDF, Newton, dense direct linear solver:
 * cvode.solve_type = denseAJ_direct, Docs/sphinx/manual/LMeXControls.rst:300-310). SUNDIALS is
 * an empty submodule in the reference (CMakeLists.txt:76-82), so this follows the published
 * description: Nordsieck history in fixed-leading-coefficient form, WRMS local error test,
 * Newton with convergence-rate test (maxcor 3, crdown 0.3, rdiv 2), Jacobian reuse (msbj 51,
 * re-evaluated on a bad-Jacobian convergence failure when gamma moved < 0.2), LU refresh rules
 * (msbp 20, |gamma/gammap - 1| > 0.3, after failures), step / order selection (bias factors 6,
 * 6, 10; threshold 1.5; etamx 1e4 / 10; etacf 0.25; etamin 0.1; etamxf 0.2 after two error
 * failures, order cut after three) and the initial-step estimate from a second-derivative
 * difference quotient. Constants are the published CVODE defaults.
 */
#include <math.h>
#include <string.h>
#include <float.h>
#include "oracle.h"
#include "prmath.h"

#define QMAXC 5
#define L_MAX (QMAXC + 1)
#define ETAMX1 10000.0
#define ETAMX2 10.0
#define ETAMX3 10.0
#define ETAMXF 0.2
#define ETAMIN 0.1
#define ETACF 0.25
#define ADDON 1e-6
#define BIAS1 6.0
#define BIAS2 6.0
#define BIAS3 10.0
#define THRESH 1.5
#define ONEPSM 1.000001
#define SMALL_NST 10
#define MXNCF 10
#define MXNEF 7
#define MXNEF1 3
#define SMALL_NEF 2
#define LONG_WAIT 10
#define CRDOWN 0.3
#define DGMAX 0.3
#define RDIV 2.0
#define MSBP 20
#define MSBJ 51
#define DGMAX_LS 0.2
#define NLSCOEF 0.1
#define MAXCOR 3

enum { FIRST_CALL, PREV_CONV_FAIL, PREV_ERR_FAIL };
enum { NO_FAILURES, FAIL_BAD_J, FAIL_OTHER };

int orc_lu_factor(int n, double* A, int* piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[i * n + k]) > fabs(A[p * n + k])) p = i;
    piv[k] = p;
    if (!(fabs(A[p * n + k]) > 0.0)) return k + 1;
    if (p != k)
      for (int j = 0; j < n; ++j) {
        double t = A[k * n + j];
        A[k * n + j] = A[p * n + j];
        A[p * n + j] = t;
      }
    double mult = 1.0 / A[k * n + k];
    for (int i = k + 1; i < n; ++i) A[i * n + k] *= mult;
    for (int i = k + 1; i < n; ++i) {
      double lik = A[i * n + k];
      for (int j = k + 1; j < n; ++j) A[i * n + j] = fma(-lik, A[k * n + j], A[i * n + j]);
    }
  }
  return 0;
}

/* Forward / back substitution; U_kk is applied as a multiplication by its reciprocal. */
void orc_lu_solve(int n, const double* LU, const int* piv, double* b) {
  for (int k = 0; k < n; ++k) {
    int p = piv[k];
    if (p != k) {
      double t = b[k];
      b[k] = b[p];
      b[p] = t;
    }
  }
  for (int k = 0; k < n - 1; ++k)
    for (int i = k + 1; i < n; ++i) b[i] = fma(-LU[i * n + k], b[k], b[i]);
  for (int k = n - 1; k >= 0; --k) {
    b[k] = b[k] * (1.0 / LU[k * n + k]);
    for (int i = 0; i < k; ++i) b[i] = fma(-LU[i * n + k], b[k], b[i]);
  }
}

typedef struct {
  int n, qmax;
  orc_rhs_fn f;
  orc_jac_fn jac;
  void* ud;
  const orc_bdf_opts* o;
  orc_bdf_stats* st;
  double zn[L_MAX + 1][ORC_NMAX];
  double ewt[ORC_NMAX], acor[ORC_NMAX], y[ORC_NMAX], ftemp[ORC_NMAX], tempv[ORC_NMAX];
  double savedJ[ORC_NMAX * ORC_NMAX], M[ORC_NMAX * ORC_NMAX];
  int piv[ORC_NMAX];
  double tn, h, hscale, hprime, eta, etamax, etaq, etaqm1, etaqp1;
  double tau[L_MAX + 1], tq[6], l[L_MAX + 1];
  double rl1, gamma, gammap, gamrat, crate, delp, acnrm, saved_tq5;
  int q, qprime, L, qwait, nscon;
  long nst, nstlp, nstlj;
  int jcur, convfail;
} bdf_t;

static double wrms(const bdf_t* b, const double* v) {
  double sq[ORC_NMAX];
  for (int i = 0; i < b->n; ++i) {
    double t = v[i] * b->ewt[i];
    sq[i] = t * t;
  }
  return sqrt(orc_psum(sq, b->n) / b->n);
}

static void set_ewt(bdf_t* b, const double* y) {
  for (int i = 0; i < b->n; ++i) {
    double at = b->o->atol_vec ? b->o->atol_vec[i] : b->o->atol;
    b->ewt[i] = 1.0 / (b->o->rtol * fabs(y[i]) + at);
  }
}

static void rescale(bdf_t* b) {
  double factor = b->eta;
  for (int j = 1; j <= b->q; ++j) {
    for (int i = 0; i < b->n; ++i) b->zn[j][i] *= factor;
    factor *= b->eta;
  }
  b->h = b->hscale * b->eta;
  b->hscale = b->h;
  b->nscon = 0;
}

static void predict(bdf_t* b) {
  b->tn += b->h;
  for (int k = 1; k <= b->q; ++k)
    for (int j = b->q; j >= k; --j)
      for (int i = 0; i < b->n; ++i) b->zn[j - 1][i] += b->zn[j][i];
}

static void restore(bdf_t* b, double saved_t) {
  b->tn = saved_t;
  for (int k = 1; k <= b->q; ++k)
    for (int j = b->q; j >= k; --j)
      for (int i = 0; i < b->n; ++i) b->zn[j - 1][i] -= b->zn[j][i];
}

static void set_bdf(bdf_t* b) {
  int q = b->q;
  double alpha0, alpha0_hat, xi_inv, xistar_inv, hsum;
  b->l[0] = b->l[1] = xi_inv = xistar_inv = 1.0;
  for (int i = 2; i <= q; ++i) b->l[i] = 0.0;
  alpha0 = alpha0_hat = -1.0;
  hsum = b->h;
  if (q > 1) {
    for (int j = 2; j < q; ++j) {
      hsum += b->tau[j - 1];
      xi_inv = b->h / hsum;
      alpha0 -= 1.0 / j;
      for (int i = j; i >= 1; --i) b->l[i] += b->l[i - 1] * xi_inv;
    }
    alpha0 -= 1.0 / q;
    xistar_inv = -b->l[1] - alpha0;
    hsum += b->tau[q - 1];
    xi_inv = b->h / hsum;
    alpha0_hat = -b->l[1] - xi_inv;
    for (int i = q; i >= 1; --i) b->l[i] += b->l[i - 1] * xistar_inv;
  }
  /* error-test constants */
  double A1 = 1.0 - alpha0_hat + alpha0;
  double A2 = 1.0 + q * A1;
  b->tq[2] = fabs(A1 / (alpha0 * A2));
  b->tq[5] = fabs(A2 * xistar_inv / (b->l[q] * xi_inv));
  if (b->qwait == 1) {
    if (q > 1) {
      double C = xistar_inv / b->l[q];
      double A3 = alpha0 + 1.0 / q;
      double A4 = alpha0_hat + xi_inv;
      double Cpinv = (1.0 - A4 + A3) / A3;
      b->tq[1] = fabs(C * Cpinv);
    } else {
      b->tq[1] = 1.0;
    }
    hsum += b->tau[q];
    xi_inv = b->h / hsum;
    double A5 = alpha0 - 1.0 / (q + 1);
    double A6 = alpha0_hat - xi_inv;
    double Cppinv = (1.0 - A6 + A5) / A2;
    b->tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
  }
  b->tq[4] = NLSCOEF / b->tq[2];
  b->rl1 = 1.0 / b->l[1];
  b->gamma = b->h * b->rl1;
  if (b->nst == 0) b->gammap = b->gamma;
  b->gamrat = (b->nst > 0) ? b->gamma / b->gammap : 1.0;
}

static void increase_bdf(bdf_t* b) {
  double alpha0, alpha1, prod, xi, xiold, hsum, A1;
  for (int i = 0; i <= b->qmax; ++i) b->l[i] = 0.0;
  b->l[2] = alpha1 = prod = xiold = 1.0;
  alpha0 = -1.0;
  hsum = b->hscale;
  if (b->q > 1) {
    for (int j = 1; j < b->q; ++j) {
      hsum += b->tau[j + 1];
      xi = hsum / b->hscale;
      prod *= xi;
      alpha0 -= 1.0 / (j + 1);
      alpha1 += 1.0 / xi;
      for (int i = j + 2; i >= 2; --i) b->l[i] = b->l[i] * xiold + b->l[i - 1];
      xiold = xi;
    }
  }
  A1 = (-alpha0 - alpha1) / prod;
  for (int i = 0; i < b->n; ++i) b->zn[b->L][i] = A1 * b->zn[b->qmax][i];
  for (int j = 2; j <= b->q; ++j)
    for (int i = 0; i < b->n; ++i) b->zn[j][i] += b->l[j] * b->zn[b->L][i];
}

static void decrease_bdf(bdf_t* b) {
  double hsum, xi;
  for (int i = 0; i <= b->qmax; ++i) b->l[i] = 0.0;
  b->l[2] = 1.0;
  hsum = 0.0;
  for (int j = 1; j <= b->q - 2; ++j) {
    hsum += b->tau[j];
    xi = hsum / b->hscale;
    for (int i = j + 2; i >= 2; --i) b->l[i] = b->l[i] * xi + b->l[i - 1];
  }
  for (int j = 2; j < b->q; ++j)
    for (int i = 0; i < b->n; ++i) b->zn[j][i] -= b->l[j] * b->zn[b->q][i];
}

static void adjust_params(bdf_t* b) {
  if (b->qprime != b->q) {
    int dq = b->qprime - b->q;
    if (!(b->q == 2 && dq != 1)) {
      if (dq == 1) increase_bdf(b);
      else if (dq == -1) decrease_bdf(b);
    }
    b->q = b->qprime;
    b->L = b->q + 1;
    b->qwait = b->L;
  }
  rescale(b);
}

/* Linear-system setup: Jacobian (fresh or saved), M = I - gamma J, LU. */
static int lsetup(bdf_t* b, int convfail) {
  int n = b->n;
  double dgamma = fabs(b->gamma / b->gammap - 1.0);
  int jbad = (b->nst == 0) || (b->nst >= b->nstlj + MSBJ) ||
             (convfail == FAIL_BAD_J && dgamma < DGMAX_LS) || (convfail == FAIL_OTHER);
  if (!jbad) {
    b->jcur = 0;
  } else {
    b->st->nje++;
    b->nstlj = b->nst;
    b->jcur = 1;
    b->jac(b->tn, b->y, b->ftemp, b->savedJ, b->ud);
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) b->M[i * n + j] = (i == j ? 1.0 : 0.0) - b->gamma * b->savedJ[i * n + j];
  b->st->nsetups++;
  int ier = orc_lu_factor(n, b->M, b->piv);
  b->gamrat = 1.0;
  b->gammap = b->gamma;
  b->crate = 1.0;
  b->nstlp = b->nst;
  return ier;
}

/* Newton iteration for the corrector; returns 0 on success, >0 recoverable failure. */
static int nls(bdf_t* b, int nflag) {
  int n = b->n;
  b->convfail = (nflag == FIRST_CALL || nflag == PREV_ERR_FAIL) ? NO_FAILURES : FAIL_OTHER;
  int callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (b->nst == 0) ||
                  (b->nst >= b->nstlp + MSBP) || (fabs(b->gamrat - 1.0) > DGMAX);
  for (int i = 0; i < n; ++i) b->acor[i] = 0.0;
  int jbad = 0;
  double res[ORC_NMAX];
  for (;;) {
    /* residual at the current correction */
    for (int i = 0; i < n; ++i) b->y[i] = b->zn[0][i] + b->acor[i];
    b->f(b->tn, b->y, b->ftemp, b->ud);
    b->st->nfe++;
    for (int i = 0; i < n; ++i) res[i] = b->rl1 * b->zn[1][i] + b->acor[i] - b->gamma * b->ftemp[i];
    int m = 0, singular = 0;
    if (callSetup) {
      int cf = jbad ? FAIL_BAD_J : b->convfail;
      singular = lsetup(b, cf) != 0; /* a zero pivot is a recoverable failure */
      callSetup = 0;
    }
    while (!singular) {
      b->st->nni++;
      for (int i = 0; i < n; ++i) res[i] = -res[i];
      orc_lu_solve(n, b->M, b->piv, res);
      if (b->gamrat != 1.0) {
        double s = 2.0 / (1.0 + b->gamrat);
        for (int i = 0; i < n; ++i) res[i] *= s;
      }
      for (int i = 0; i < n; ++i) b->acor[i] += res[i];
      double del = wrms(b, res);
      if (m > 0) b->crate = fmax(CRDOWN * b->crate, del / b->delp);
      double dcon = del * fmin(1.0, b->crate) / b->tq[4];
      if (dcon <= 1.0) {
        b->acnrm = (m == 0) ? del : wrms(b, b->acor);
        b->jcur = 0;
        for (int i = 0; i < n; ++i) b->y[i] = b->zn[0][i] + b->acor[i];
        return 0;
      }
      if (m >= 1 && del > RDIV * b->delp) break;
      b->delp = del;
      m++;
      if (m >= MAXCOR) break;
      for (int i = 0; i < n; ++i) b->y[i] = b->zn[0][i] + b->acor[i];
      b->f(b->tn, b->y, b->ftemp, b->ud);
      b->st->nfe++;
      for (int i = 0; i < n; ++i) res[i] = b->rl1 * b->zn[1][i] + b->acor[i] - b->gamma * b->ftemp[i];
    }
    if (!b->jcur) {
      callSetup = 1;
      jbad = 1;
      for (int i = 0; i < n; ++i) b->acor[i] = 0.0;
      continue;
    }
    return 1;
  }
}

static double etaq_of(double dsm, int L) { return 1.0 / (prm_pow(BIAS2 * dsm, 1.0 / L) + ADDON); }

static int step(bdf_t* b) {
  double saved_t = b->tn;
  int ncf = 0, nef = 0, nflag = FIRST_CALL;
  double dsm = 0.0;
  if (b->nst > 0 && b->hprime != b->h) adjust_params(b);
  for (;;) {
    predict(b);
    set_bdf(b);
    int nf = nls(b, nflag);
    if (nf != 0) {
      /* convergence failure */
      b->st->ncfn++;
      restore(b, saved_t);
      ncf++;
      b->etamax = 1.0;
      if (ncf == MXNCF) return ORC_CONV_FAILURE;
      b->eta = ETACF;
      nflag = PREV_CONV_FAIL;
      rescale(b);
      continue;
    }
    dsm = b->acnrm * b->tq[2];
    if (dsm <= 1.0) break;
    /* error test failure */
    nef++;
    b->st->netf++;
    nflag = PREV_ERR_FAIL;
    restore(b, saved_t);
    if (nef == MXNEF) return ORC_ERR_FAILURE;
    b->etamax = 1.0;
    if (nef <= MXNEF1) {
      b->eta = etaq_of(dsm, b->L);
      b->eta = fmax(ETAMIN, b->eta);
      if (nef >= SMALL_NEF) b->eta = fmin(b->eta, ETAMXF);
      rescale(b);
      continue;
    }
    if (b->q > 1) {
      b->eta = ETAMIN;
      decrease_bdf(b);
      b->L = b->q;
      b->q--;
      b->qwait = b->L;
      rescale(b);
      continue;
    }
    b->eta = ETAMIN;
    b->h *= b->eta;
    b->hscale = b->h;
    b->qwait = LONG_WAIT;
    b->nscon = 0;
    b->f(b->tn, b->zn[0], b->tempv, b->ud);
    b->st->nfe++;
    for (int i = 0; i < b->n; ++i) b->zn[1][i] = b->h * b->tempv[i];
  }
  /* complete the step */
  b->nst++;
  b->nscon++;
  for (int i = b->q; i >= 2; --i) b->tau[i] = b->tau[i - 1];
  if (b->q == 1 && b->nst > 1) b->tau[2] = b->tau[1];
  b->tau[1] = b->h;
  for (int j = 0; j <= b->q; ++j)
    for (int i = 0; i < b->n; ++i) b->zn[j][i] += b->l[j] * b->acor[i];
  b->qwait--;
  if (b->qwait == 1 && b->q != b->qmax) {
    for (int i = 0; i < b->n; ++i) b->zn[b->qmax][i] = b->acor[i];
    b->saved_tq5 = b->tq[5];
  }
  /* prepare the next step: order and step-size selection */
  if (b->etamax == 1.0) {
    if (b->qwait < 2) b->qwait = 2;
    b->qprime = b->q;
    b->hprime = b->h;
    b->eta = 1.0;
  } else {
    b->etaq = etaq_of(dsm, b->L);
    if (b->qwait != 0) {
      b->eta = b->etaq;
      b->qprime = b->q;
    } else {
      b->qwait = 2;
      b->etaqm1 = 0.0;
      if (b->q > 1) {
        double ddn = wrms(b, b->zn[b->q]) * b->tq[1];
        b->etaqm1 = 1.0 / (prm_pow(BIAS1 * ddn, 1.0 / b->q) + ADDON);
      }
      b->etaqp1 = 0.0;
      if (b->q != b->qmax && b->saved_tq5 != 0.0) {
        double cquot = (b->tq[5] / b->saved_tq5) * prm_pow(b->h / b->tau[2], (double)b->L);
        for (int i = 0; i < b->n; ++i) b->tempv[i] = b->acor[i] - cquot * b->zn[b->qmax][i];
        double dup = wrms(b, b->tempv) * b->tq[3];
        b->etaqp1 = 1.0 / (prm_pow(BIAS3 * dup, 1.0 / (b->L + 1)) + ADDON);
      }
      double etam = fmax(b->etaqm1, fmax(b->etaq, b->etaqp1));
      if (etam < THRESH) {
        b->eta = 1.0;
        b->qprime = b->q;
      } else if (etam == b->etaqm1) {
        b->eta = b->etaqm1;
        b->qprime = b->q - 1;
      } else if (etam == b->etaq) {
        b->eta = b->etaq;
        b->qprime = b->q;
      } else {
        b->eta = b->etaqp1;
        b->qprime = b->q + 1;
        for (int i = 0; i < b->n; ++i) b->zn[b->qmax][i] = b->acor[i];
      }
    }
    if (b->eta < THRESH) {
      b->eta = 1.0;
      b->hprime = b->h;
    } else {
      b->eta = fmin(b->eta, b->etamax);
      b->hprime = b->h * b->eta;
      if (b->qprime < b->q) b->nscon = 0;
    }
  }
  b->etamax = (b->nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
  return ORC_OK;
}

/* Initial step from |y''| estimated by a difference quotient, bounded by 0.1 |tout - t0| and
   by 0.1 (|y0| + atol)/|y0'|. */
static int initial_h(bdf_t* b, double tout) {
  int n = b->n;
  double tdiff = tout - b->tn;
  if (tdiff == 0.0) return ORC_TOO_CLOSE;
  double sign = tdiff > 0 ? 1.0 : -1.0;
  double tdist = fabs(tdiff);
  double tround = DBL_EPSILON * fmax(fabs(b->tn), fabs(tout));
  if (tdist < 2.0 * tround) return ORC_TOO_CLOSE;
  double hlb = 100.0 * tround;
  double hub_inv = 0.0;
  for (int i = 0; i < n; ++i) {
    double at = b->o->atol_vec ? b->o->atol_vec[i] : b->o->atol;
    double t1 = 0.1 * fabs(b->zn[0][i]) + at;
    double r = fabs(b->zn[1][i]) / t1;
    if (r > hub_inv) hub_inv = r;
  }
  double hub = 0.1 * tdist;
  if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
  double hg = sqrt(hlb * hub);
  if (hub < hlb) {
    b->h = sign * hg;
    return ORC_OK;
  }
  double hnew = hg;
  for (int count1 = 1; count1 <= 4; ++count1) {
    double hgs = hg * sign;
    for (int i = 0; i < n; ++i) b->y[i] = b->zn[0][i] + hgs * b->zn[1][i];
    b->f(b->tn + hgs, b->y, b->tempv, b->ud);
    b->st->nfe++;
    for (int i = 0; i < n; ++i) b->tempv[i] = (b->tempv[i] - b->zn[1][i]) / hgs;
    double yddnrm = wrms(b, b->tempv);
    hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
    if (count1 == 4) break;
    double hrat = hnew / hg;
    if (hrat > 0.5 && hrat < 2.0) break;
    if (count1 > 1 && hrat > 2.0) {
      hnew = hg;
      break;
    }
    hg = hnew;
  }
  double h0 = 0.5 * hnew;
  if (h0 < hlb) h0 = hlb;
  if (h0 > hub) h0 = hub;
  b->h = sign * h0;
  return ORC_OK;
}

int orc_bdf_solve(int n, double t0, double tout, double* y, orc_rhs_fn f, orc_jac_fn jac,
                  void* ud, const orc_bdf_opts* opts, orc_bdf_stats* st) {
  bdf_t B;
  bdf_t* b = &B;
  memset(st, 0, sizeof(*st));
  memset(b->tau, 0, sizeof(b->tau));
  memset(b->tq, 0, sizeof(b->tq));
  b->n = n;
  b->qmax = opts->max_order < 1 ? 1 : (opts->max_order > QMAXC ? QMAXC : opts->max_order);
  b->f = f;
  b->jac = jac;
  b->ud = ud;
  b->o = opts;
  b->st = st;
  b->q = 1;
  b->L = 2;
  b->qwait = b->L;
  b->etamax = ETAMX1;
  b->nst = b->nstlp = b->nstlj = 0;
  b->nscon = 0;
  b->saved_tq5 = 0.0;
  b->crate = 1.0;
  b->delp = 0.0;
  b->gamrat = 1.0;
  b->gammap = 1.0;
  b->tn = t0;
  b->jcur = 0;
  for (int i = 0; i < n; ++i) b->zn[0][i] = y[i];
  set_ewt(b, y);
  f(t0, y, b->zn[1], ud);
  st->nfe++;
  int ier = initial_h(b, tout);
  if (ier != ORC_OK) return ier;
  for (int i = 0; i < n; ++i) b->zn[1][i] *= b->h;
  b->hscale = b->h;
  b->hprime = b->h;
  long nstloc = 0;
  for (;;) {
    if (b->nst > 0) set_ewt(b, b->zn[0]);
    if (nstloc >= opts->max_steps) {
      ier = ORC_TOO_MUCH_WORK;
      break;
    }
    ier = step(b);
    if (ier != ORC_OK) break;
    nstloc++;
    if ((b->tn - tout) * b->h >= 0.0) {
      double s = (tout - b->tn) / b->h;
      for (int i = 0; i < n; ++i) {
        double v = b->zn[b->q][i];
        for (int j = b->q - 1; j >= 0; --j) v = b->zn[j][i] + s * v;
        y[i] = v;
      }
      break;
    }
  }
  if (ier != ORC_OK) /* return the last accepted state */
    for (int i = 0; i < n; ++i) y[i] = b->zn[0][i];
  st->nst = b->nst;
  st->q_last = b->q;
  st->h_last = b->h;
  return ier;
}
