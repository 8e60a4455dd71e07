/* CPU oracle: thermodynamics, reaction rates, reactor right-hand side and its Jacobian, and
 * the box-level react driver.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates (bodies absent from the reference tree; semantics from the call sites and docs):
 *  - the reactor right-hand side PelePhysics evaluates for reactor_type = 2 (constant pressure,
 *    rho h energy; Source/PeleLMeX_Setup.cpp:146): rho = sum rhoY, h = (rhoh0 + t Q_h)/rho,
 *    T from (h, Y), wdot from the mechanism, d(rhoY_k)/dt = W_k wdot_k + Q_k,
 *    dT/dt = (Q_h - sum_k h_k d(rhoY_k)/dt) / (rho cp)   (Docs/sphinx/manual/Model.rst:300-311)
 *  - productionRate / RTY2WDOT as used by reactionRateRhoY (Source/PeleLMeX_K.H:848-889):
 *    Arrhenius k_f = A T^b exp(-Ea/RT); third-body [M] = sum eps_k C_k; Lindemann and Troe
 *    falloff; k_r = k_f / K_c with K_c from NASA-7 Gibbs energies.
 *  - the analytic dense Jacobian (denseAJ_direct; PELE_COMPILE_AJACOBIAN, Exec/Make.PeleLMeX:100-104):
 *    here the exact Jacobian of the right-hand side with T taken as a state variable.
 *  - react over a box with mask == -1 cells skipped and FC = number of RHS evaluations
 *    (Source/PeleLMeX_Reactions.cpp:101-108, Source/PeleLMeX_EB.cpp:157-161, PeleLMeX.H:209).
 *
 * Arithmetic is written as explicit IEEE-754 expression trees (no implicit contraction, fma
 * where stated) with sums over components in one canonical pairwise order (orc_psum) and
 * production rates / Jacobian rows
 * gathered species by species over each species' reactions in ascending order. The HIP
 * kernels evaluate the same trees, so GPU and oracle agree bit for bit; the oracle itself is
 * pinned by tests against finite differences, conservation laws and scipy integrators.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "prmath.h"
#include "../pelelmex_amd/csrc/mech/lidryer.h"
#include "../pelelmex_amd/csrc/mech/drm19.h"
#include "../pelelmex_amd/csrc/mech/syn53.h"
#include "../pelelmex_amd/csrc/mech/dodecane_lu.h"

static const pr_mech_t* const MECHS[] = {&pr_lidryer, &pr_drm19, &pr_syn53, &pr_dodecane_lu};
#define NMECH ((int)(sizeof(MECHS) / sizeof(MECHS[0])))

int orc_mech_count(void) { return NMECH; }
const pr_mech_t* orc_mech_get(int id) { return (id >= 0 && id < NMECH) ? MECHS[id] : NULL; }
int orc_mech_find(const char* name) {
  for (int i = 0; i < NMECH; ++i)
    if (strcmp(MECHS[i]->name, name) == 0) return i;
  return -1;
}
int orc_mech_nsp(int id) { return MECHS[id]->nsp; }
int orc_mech_nr(int id) { return MECHS[id]->nr; }
void orc_mech_weights(int id, double* W) {
  for (int k = 0; k < MECHS[id]->nsp; ++k) W[k] = MECHS[id]->W[k];
}

/* Sum of x[0..n-1] in the canonical pairwise order: zero-padded to 64 slots, then pairs
   (2j, 2j+1), pairs of pairs, ... (the order the GPU's lane butterflies produce). */
double orc_psum(const double* x, int n) {
  double b[64];
  for (int i = 0; i < 64; ++i) b[i] = (i < n) ? x[i] : 0.0;
  for (int w = 1; w < 64; w *= 2)
    for (int j = 0; j < 64; j += 2 * w) b[j] = b[j] + b[j + w];
  return b[0];
}

/* Derived per-species tables: R/W, 1/W and the NASA-7 polynomials rewritten for Horner
   evaluation of cp/R, h/RT (without the a6/T term), s/R (without a1 ln T and a7) and d(cp/R)/dT.
   Range 0 = low (T < T_mid), 1 = high. */
typedef struct {
  double RW, invW, W, tmid;
  double cp[2][5], h[2][6], s[2][6], dcp[2][4];
} sp_t;

static sp_t SP[NMECH][ORC_NMAX];
static int SP_READY = 0;

static void derive_species(void) {
  for (int id = 0; id < NMECH; ++id) {
    const pr_mech_t* m = MECHS[id];
    for (int k = 0; k < m->nsp; ++k) {
      sp_t* o = &SP[id][k];
      o->W = m->W[k];
      o->RW = PR_RU / m->W[k];
      o->invW = 1.0 / m->W[k];
      o->tmid = m->nasa[15 * k];
      for (int r = 0; r < 2; ++r) {
        const double* a = m->nasa + 15 * k + (r == 0 ? 8 : 1);
        for (int c = 0; c < 5; ++c) o->cp[r][c] = a[c];
        o->h[r][0] = a[0];
        o->h[r][1] = a[1] / 2.0;
        o->h[r][2] = a[2] / 3.0;
        o->h[r][3] = a[3] / 4.0;
        o->h[r][4] = a[4] / 5.0;
        o->h[r][5] = a[5];
        o->s[r][0] = a[0];
        o->s[r][1] = a[1];
        o->s[r][2] = a[2] / 2.0;
        o->s[r][3] = a[3] / 3.0;
        o->s[r][4] = a[4] / 4.0;
        o->s[r][5] = a[6];
        o->dcp[r][0] = a[1];
        o->dcp[r][1] = 2.0 * a[2];
        o->dcp[r][2] = 3.0 * a[3];
        o->dcp[r][3] = 4.0 * a[4];
      }
    }
  }
  SP_READY = 1;
}

__attribute__((constructor)) static void orc_init(void) { derive_species(); }

/* cp/R and h/RT at T */
static void nasa_hcp(const sp_t* p, double T, double invT, double* cpR, double* hRT) {
  int r = (T < p->tmid) ? 0 : 1;
  const double* c = p->cp[r];
  const double* h = p->h[r];
  *cpR = fma(fma(fma(fma(c[4], T, c[3]), T, c[2]), T, c[1]), T, c[0]);
  double hp = fma(fma(fma(fma(h[4], T, h[3]), T, h[2]), T, h[1]), T, h[0]);
  *hRT = fma(h[5], invT, hp);
}

/* cp/R, h/RT, g/RT = h/RT - s/R and d(cp/R)/dT at T */
static void nasa_full(const sp_t* p, double T, double invT, double lnT, double* cpR, double* hRT,
                      double* gRT, double* dcpR) {
  int r = (T < p->tmid) ? 0 : 1;
  nasa_hcp(p, T, invT, cpR, hRT);
  const double* s = p->s[r];
  const double* d = p->dcp[r];
  double sp = fma(fma(fma(s[4], T, s[3]), T, s[2]), T, s[1]) * T;
  double sR = fma(s[0], lnT, sp + s[5]);
  *gRT = *hRT - sR;
  *dcpR = fma(fma(fma(d[3], T, d[2]), T, d[1]), T, d[0]);
}

void orc_thermo(int id, double T, double* cpR, double* hRT, double* sR) {
  const pr_mech_t* m = MECHS[id];
  double invT = 1.0 / T, lnT = prm_log(T);
  for (int k = 0; k < m->nsp; ++k) {
    double c, h, g, d;
    nasa_full(&SP[id][k], T, invT, lnT, &c, &h, &g, &d);
    if (cpR) cpR[k] = c;
    if (hRT) hRT[k] = h;
    if (sR) sR[k] = h - g;
  }
}

/* Per-reaction quantities needed by both the rates and the Jacobian. */
typedef struct {
  double kf, kr;     /* effective coefficients (third-body / falloff factor included) */
  double fwd, rev;   /* kf * prod C_reac, kr * prod C_prod */
  double dqdM;       /* d q / d [M] (0 for elementary reactions) */
  double dqdT;       /* d q / d T at constant concentrations */
} rx_info;

static void reaction(const pr_mech_t* m, int r, const double* C, double Ctot, double T, double lnT,
                     double invT, double lnRTP, const double* gRT, const double* hRT, int jac,
                     rx_info* o) {
  const double* a = m->arr + 3 * r;
  const int type = m->rtype[r], nre = m->nreac[r], npr = m->nprod[r];
  const int* rs = m->reac + 3 * r;
  const int* ps = m->prod + 3 * r;
  double kinf = prm_exp(fma(-a[2], invT, fma(a[1], lnT, a[0])));
  double dlnkinf = fma(a[2], invT, a[1]) * invT;
  double M = 0.0;
  if (type != PR_RX_ELEM) {
    M = m->eff_dflt[r] * Ctot;
    for (int e = m->eff_start[r]; e < m->eff_start[r + 1]; ++e) M = fma(m->eff_val[e], C[m->eff_sp[e]], M);
  }
  double pr = C[rs[0]];
  if (nre > 1) pr *= C[rs[1]];
  if (nre > 2) pr *= C[rs[2]];
  double pp = C[ps[0]];
  if (npr > 1) pp *= C[ps[1]];
  if (npr > 2) pp *= C[ps[2]];
  double invKc = 0.0, dH = 0.0;
  const int sumnu = npr - nre;
  if (m->rev[r]) {
    double dG = gRT[ps[0]];
    if (npr > 1) dG += gRT[ps[1]];
    if (npr > 2) dG += gRT[ps[2]];
    dG -= gRT[rs[0]];
    if (nre > 1) dG -= gRT[rs[1]];
    if (nre > 2) dG -= gRT[rs[2]];
    invKc = prm_exp(fma((double)sumnu, lnRTP, dG));
    if (jac) {
      dH = hRT[ps[0]];
      if (npr > 1) dH += hRT[ps[1]];
      if (npr > 2) dH += hRT[ps[2]];
      dH -= hRT[rs[0]];
      if (nre > 1) dH -= hRT[rs[1]];
      if (nre > 2) dH -= hRT[rs[2]];
    }
  }
  const double dlnKc = (dH - (double)sumnu) * invT;
  if (type == PR_RX_ELEM || type == PR_RX_THIRDBODY) {
    double kr0 = kinf * invKc;
    double fwd0 = kinf * pr, rev0 = kr0 * pp;
    double dq0 = 0.0;
    if (jac) dq0 = fwd0 * dlnkinf - rev0 * (dlnkinf - dlnKc);
    if (type == PR_RX_ELEM) {
      o->kf = kinf; o->kr = kr0; o->fwd = fwd0; o->rev = rev0; o->dqdM = 0.0; o->dqdT = dq0;
    } else {
      o->kf = M * kinf; o->kr = M * kr0; o->fwd = M * fwd0; o->rev = M * rev0;
      o->dqdM = fwd0 - rev0; o->dqdT = M * dq0;
    }
    return;
  }
  /* falloff */
  const double* lo = m->low + 3 * r;
  double k0 = prm_exp(fma(-lo[2], invT, fma(lo[1], lnT, lo[0])));
  double Pr = (k0 * M) / kinf;
  if (Pr < 1e-300) Pr = 1e-300;
  double F = 1.0, dlogF_dx = 0.0, dlogF_dT = 0.0, dlnPr = 0.0;
  if (jac) dlnPr = fma(lo[2], invT, lo[1]) * invT - dlnkinf;
  if (type == PR_RX_TROE) {
    const double* tr = m->troe + 4 * r;
    double e3 = prm_exp(-T / tr[1]), e1 = prm_exp(-T / tr[2]);
    double Fc = (1.0 - tr[0]) * e3 + tr[0] * e1;
    double dFc = 0.0;
    if (jac) dFc = -((1.0 - tr[0]) / tr[1]) * e3 - (tr[0] / tr[2]) * e1;
    if (m->ntroe[r] == 4) {
      double e2 = prm_exp(-tr[3] * invT);
      Fc += e2;
      if (jac) dFc += ((tr[3] * invT) * invT) * e2;
    }
    if (Fc < 1e-300) Fc = 1e-300;
    double logFc = prm_log10(Fc);
    double c = -0.4 - 0.67 * logFc, nn = 0.75 - 1.27 * logFc;
    double x = prm_log10(Pr);
    double xc = x + c, D = nn - 0.14 * xc, f1 = xc / D;
    double den = 1.0 + f1 * f1;
    F = prm_exp((logFc / den) * PRM_LN10);
    if (jac) {
      dlogF_dx = -logFc * 2.0 * f1 * (nn / (D * D)) / (den * den);
      double dlogFc = dFc / (Fc * PRM_LN10);
      double dx = dlnPr / PRM_LN10, dc = -0.67 * dlogFc, dnn = -1.27 * dlogFc;
      double dxc = dx + dc, dD = dnn - 0.14 * dxc;
      double df1 = (dxc * D - xc * dD) / (D * D);
      dlogF_dT = (dlogFc * den - logFc * 2.0 * f1 * df1) / (den * den);
    }
  }
  double kf = ((kinf * Pr) / (1.0 + Pr)) * F;
  double kr = kf * invKc;
  o->kf = kf;
  o->kr = kr;
  o->fwd = kf * pr;
  o->rev = kr * pp;
  o->dqdM = 0.0;
  o->dqdT = 0.0;
  if (jac) {
    double dlnkf = dlnkinf + dlnPr / (1.0 + Pr) + PRM_LN10 * dlogF_dT;
    double onep = 1.0 + Pr;
    o->dqdM = ((k0 * F) * (1.0 / (onep * onep) + dlogF_dx / onep)) * (pr - pp * invKc);
    o->dqdT = o->fwd * dlnkf - o->rev * (dlnkf - dlnKc);
  }
}

/* Thermo of every species at T into per-species arrays; concentrations from rhoY. */
typedef struct {
  double T, lnT, invT, lnRTP;
  double C[ORC_NMAX], gRT[ORC_NMAX], hRT[ORC_NMAX], hk[ORC_NMAX], cpk[ORC_NMAX], dcpk[ORC_NMAX];
  double Ctot;
} state_t;

static void state_at(int id, const double* y, double T, state_t* s) {
  const pr_mech_t* m = MECHS[id];
  s->T = T;
  s->lnT = prm_log(T);
  s->invT = 1.0 / T;
  s->lnRTP = prm_log((PR_RU * T) / PR_PATM);
  s->Ctot = 0.0;
  for (int k = 0; k < m->nsp; ++k) {
    const sp_t* p = &SP[id][k];
    double cpR, hRT, gRT, dcpR;
    nasa_full(p, T, s->invT, s->lnT, &cpR, &hRT, &gRT, &dcpR);
    s->C[k] = y[k] * p->invW;
    s->gRT[k] = gRT;
    s->hRT[k] = hRT;
    s->hk[k] = (hRT * p->RW) * T;
    s->cpk[k] = cpR * p->RW;
    s->dcpk[k] = dcpR * p->RW;
  }
  s->Ctot = orc_psum(s->C, m->nsp);
}

/* production rates by the species-wise gather over each species' reactions, in blocks of
   PR_GATHER_BLOCK entries: block sums from 0.0 in list order, added in list order */
static void gather_wdot(const pr_mech_t* m, const double* q, double* wdot) {
  for (int k = 0; k < m->nsp; ++k) {
    const int e0 = m->inc_start[k], e1 = m->inc_start[k + 1];
    double w = 0.0;
    for (int b = e0; b < e1; b += PR_GATHER_BLOCK) {
      const int be = (b + PR_GATHER_BLOCK < e1) ? b + PR_GATHER_BLOCK : e1;
      double p = 0.0;
      for (int e = b; e < be; ++e) p += (double)m->inc_nu[e] * q[m->inc_r[e]];
      w = (b == e0) ? p : w + p;
    }
    wdot[k] = w;
  }
}

void orc_rates(int id, const double* C, double T, double* q, double* wdot) {
  const pr_mech_t* m = MECHS[id];
  double y[ORC_NMAX] = {0}, qq[1024];
  for (int k = 0; k < m->nsp; ++k) y[k] = C[k] * m->W[k];
  state_t s;
  state_at(id, y, T, &s);
  /* use the given concentrations exactly */
  for (int k = 0; k < m->nsp; ++k) s.C[k] = C[k];
  s.Ctot = orc_psum(s.C, m->nsp);
  for (int r = 0; r < m->nr; ++r) {
    rx_info o;
    reaction(m, r, s.C, s.Ctot, T, s.lnT, s.invT, s.lnRTP, s.gRT, s.hRT, 0, &o);
    qq[r] = o.fwd - o.rev;
    if (q) q[r] = qq[r];
  }
  if (wdot) gather_wdot(m, qq, wdot);
}

void orc_rates_fr(int id, const double* C, double T, double* fwd, double* rev) {
  const pr_mech_t* m = MECHS[id];
  double y[ORC_NMAX] = {0};
  for (int k = 0; k < m->nsp; ++k) y[k] = C[k] * m->W[k];
  state_t s;
  state_at(id, y, T, &s);
  for (int k = 0; k < m->nsp; ++k) s.C[k] = C[k];
  s.Ctot = orc_psum(s.C, m->nsp);
  for (int r = 0; r < m->nr; ++r) {
    rx_info o;
    reaction(m, r, s.C, s.Ctot, T, s.lnT, s.invT, s.lnRTP, s.gRT, s.hRT, 0, &o);
    fwd[r] = o.fwd;
    rev[r] = o.rev;
  }
}

double orc_exp(double x) { return prm_exp(x); }
double orc_log(double x) { return prm_log(x); }

#define HY2T_MAXIT 20
#define HY2T_RTOL 1e-10

/* Newton on T for h(T, Y) = h; Y given as rhoY and rho (Y_k = rhoY_k / rho). */
static int hy2t_rhoY(int id, double h, const double* rY, double rho, double* T) {
  const pr_mech_t* m = MECHS[id];
  double t = *T;
  int it;
  for (it = 0; it < HY2T_MAXIT; ++it) {
    double invT = 1.0 / t, ph[ORC_NMAX], pc[ORC_NMAX];
    for (int k = 0; k < m->nsp; ++k) {
      const sp_t* p = &SP[id][k];
      double cpR, hRT;
      nasa_hcp(p, t, invT, &cpR, &hRT);
      double Y = rY[k] / rho;
      ph[k] = ((Y * hRT) * p->RW) * t;
      pc[k] = (Y * cpR) * p->RW;
    }
    double hm = orc_psum(ph, m->nsp), cpm = orc_psum(pc, m->nsp);
    double dT = (h - hm) / cpm;
    t = t + dT;
    if (fabs(dT) <= HY2T_RTOL * t) {
      ++it;
      break;
    }
  }
  *T = t;
  return it;
}

int orc_hy2t(int id, double h, const double* Y, double* T) { return hy2t_rhoY(id, h, Y, 1.0, T); }

/* reactionRateRhoY (Source/PeleLMeX_K.H:848-889), MKS in and out, for ncells cells stored
   component-major (rY[k * ncells + c]). Steps as the reference: rho = sum rhoY, Y = rhoY / rho,
   h = rhoH / rho * 1e4, T by HY2T from the stored T, then rhoYdot_k = W_k wdot_k(rho * 1e-3, T, Y)
   * 1000. Tout (nullable) receives the temperatures. */
void orc_rate_rhoY(int id, long ncells, const double* rY, const double* rH, const double* T0, double* out,
                   double* Tout) {
  const pr_mech_t* m = MECHS[id];
  const int N = m->nsp;
  for (long c = 0; c < ncells; ++c) {
    double ry[ORC_NMAX], Y[ORC_NMAX], yc[ORC_NMAX], q[1024], wdot[ORC_NMAX];
    for (int k = 0; k < N; ++k) ry[k] = rY[(long)k * ncells + c];
    const double rho = orc_psum(ry, N);
    const double rhoinv = 1.0 / rho;
    for (int k = 0; k < N; ++k) Y[k] = ry[k] * rhoinv;
    double T = T0[c];
    hy2t_rhoY(id, (rH[c] * rhoinv) * 1.0e4, Y, 1.0, &T);
    const double rho_cgs = rho * 0.001;
    for (int k = 0; k < N; ++k) yc[k] = rho_cgs * Y[k];
    state_t s;
    state_at(id, yc, T, &s);
    for (int r = 0; r < m->nr; ++r) {
      rx_info o;
      reaction(m, r, s.C, s.Ctot, T, s.lnT, s.invT, s.lnRTP, s.gRT, s.hRT, 0, &o);
      q[r] = o.fwd - o.rev;
    }
    gather_wdot(m, q, wdot);
    for (int k = 0; k < N; ++k) out[(long)k * ncells + c] = (wdot[k] * m->W[k]) * 1000.0;
    if (Tout) Tout[c] = T;
  }
}

/* setTemperature (Source/PeleLMeX_Eos.cpp:218-240) through getTfromHY (Source/PeleLMeX_K.H:700-723),
   MKS, component-major rY: rhoinv = 1 / rho (the DENSITY component), Y = rhoY * rhoinv,
   h = rhoH * rhoinv * 1e4 (erg/g), T = HY2T(h, Y) started at the stored T, written back. */
void orc_set_temperature(int id, long ncells, const double* rho, const double* rY, const double* rH, double* T) {
  const int N = MECHS[id]->nsp;
  for (long c = 0; c < ncells; ++c) {
    double Y[ORC_NMAX];
    const double rhoinv = 1.0 / rho[c];
    for (int k = 0; k < N; ++k) Y[k] = rY[(long)k * ncells + c] * rhoinv;
    double t = T[c];
    hy2t_rhoY(id, (rH[c] * rhoinv) * 1.0e4, Y, 1.0, &t);
    T[c] = t;
  }
}

/* floorSpecies (Source/PeleLMeX_Utils.cpp:649-690): fabMinMax (Source/PeleLMeX_K.H:894-907)
   rhoY_k = max(0, min(DBL_MAX, rhoY_k)) with std::min / std::max semantics ((b < a) ? b : a and
   (a < b) ? b : a, so NaN becomes DBL_MAX and -0 becomes +0); rho = 0 + rhoY_0 + rhoY_1 + ... in
   species order; Y = rhoY * (1 / rho); h = TY2H(T, Y) = sum over species in order of
   ((Y_k h_k/RT) R/W_k) T; rhoH = h * 1e-4 * rho. Updates rY, writes rho and rH. */
void orc_floor_species(int id, long ncells, double* rY, double* rho, double* rH, const double* T) {
  const int N = MECHS[id]->nsp;
  for (long c = 0; c < ncells; ++c) {
    double r = 0.0;
    for (int k = 0; k < N; ++k) {
      double x = rY[(long)k * ncells + c];
      x = (x < DBL_MAX) ? x : DBL_MAX;
      x = (0.0 < x) ? x : 0.0;
      rY[(long)k * ncells + c] = x;
      r = r + x;
    }
    const double rhoinv = 1.0 / r;
    const double t = T[c], invT = 1.0 / t;
    double h = 0.0;
    for (int k = 0; k < N; ++k) {
      const sp_t* p = &SP[id][k];
      double cpR, hRT;
      nasa_hcp(p, t, invT, &cpR, &hRT);
      h = h + (((rY[(long)k * ncells + c] * rhoinv) * hRT) * p->RW) * t;
    }
    rho[c] = r;
    rH[c] = (h * 1.0e-4) * r;
  }
}

/* getHeatRelease (Source/PeleLMeX_Reactions.cpp:477-506) with getHGivenT
   (Source/PeleLMeX_K.H:598-615): Hi_k = h_k(T) * 1e-4 (J/kg; h_k = ((h_k/RT) R/W_k) T erg/g),
   HR = 0 - Hi_0 I_R,0 - Hi_1 I_R,1 - ... in species order. Component-major I_R. */
void orc_heat_release(int id, long ncells, const double* T, const double* IR, double* HR) {
  const int N = MECHS[id]->nsp;
  for (long c = 0; c < ncells; ++c) {
    const double t = T[c], invT = 1.0 / t;
    double hr = 0.0;
    for (int k = 0; k < N; ++k) {
      const sp_t* p = &SP[id][k];
      double cpR, hRT;
      nasa_hcp(p, t, invT, &cpR, &hRT);
      hr = hr - (((hRT * p->RW) * t) * 0.0001) * IR[(long)k * ncells + c];
    }
    HR[c] = hr;
  }
}

/* Shared tail of both right-hand sides, evaluated at temperature T. */
static void rhs_at_T(int id, const orc_cell* c, const double* y, double T, double* f) {
  const pr_mech_t* m = MECHS[id];
  const int N = m->nsp;
  state_t s;
  double q[1024], wdot[ORC_NMAX];
  state_at(id, y, T, &s);
  for (int r = 0; r < m->nr; ++r) {
    rx_info o;
    reaction(m, r, s.C, s.Ctot, T, s.lnT, s.invT, s.lnRTP, s.gRT, s.hRT, 0, &o);
    q[r] = o.fwd - o.rev;
  }
  gather_wdot(m, q, wdot);
  for (int k = 0; k < N; ++k) f[k] = wdot[k] * m->W[k] + c->rYsrc[k];
  double a[ORC_NMAX], b[ORC_NMAX];
  for (int k = 0; k < N; ++k) {
    a[k] = s.hk[k] * f[k];
    b[k] = y[k] * s.cpk[k];
  }
  double pn = orc_psum(a, N), rcp = orc_psum(b, N);
  f[N] = (c->rhohsrc - pn) / rcp;
}

int orc_reactor_rhs(double t, const double* y, double* f, void* cell) {
  const orc_cell* c = (const orc_cell*)cell;
  const int N = MECHS[c->mech]->nsp;
  double rho = orc_psum(y, N);
  double h = (c->rhoh0 + t * c->rhohsrc) / rho;
  double T = y[N];
  hy2t_rhoY(c->mech, h, y, rho, &T);
  rhs_at_T(c->mech, c, y, T, f);
  return 0;
}

int orc_reactor_rhs_frozen(double t, const double* y, double* f, void* cell) {
  (void)t;
  const orc_cell* c = (const orc_cell*)cell;
  rhs_at_T(c->mech, c, y, y[MECHS[c->mech]->nsp], f);
  return 0;
}

int orc_reactor_jac(double t, const double* y, const double* fy, double* J, void* cell) {
  (void)t;
  (void)fy;
  const orc_cell* c = (const orc_cell*)cell;
  const int id = c->mech;
  const pr_mech_t* m = MECHS[id];
  const int N = m->nsp, n = N + 1;
  const double T = y[N];
  state_t s;
  state_at(id, y, T, &s);
  static __thread rx_info rx[1024];
  for (int r = 0; r < m->nr; ++r)
    reaction(m, r, s.C, s.Ctot, T, s.lnT, s.invT, s.lnRTP, s.gRT, s.hRT, 1, &rx[r]);
  double fk[ORC_NMAX];
  /* species rows: gather over each species' reactions */
  for (int k = 0; k < N; ++k) {
    double* row = J + k * n;
    for (int j = 0; j < n; ++j) row[j] = 0.0;
    double Sd = 0.0, JT = 0.0, w = 0.0;
    for (int e = m->inc_start[k]; e < m->inc_start[k + 1]; ++e) {
      const int r = m->inc_r[e];
      const double nu = (double)m->inc_nu[e];
      const rx_info* o = &rx[r];
      const int nre = m->nreac[r], npr = m->nprod[r];
      const int* rs = m->reac + 3 * r;
      const int* ps = m->prod + 3 * r;
      w += nu * (o->fwd - o->rev);
      double c0 = s.C[rs[0]], c1 = nre > 1 ? s.C[rs[1]] : 1.0, c2 = nre > 2 ? s.C[rs[2]] : 1.0;
      double kf = nu * o->kf;
      row[rs[0]] += (kf * c1) * c2;
      if (nre > 1) row[rs[1]] += (kf * c0) * c2;
      if (nre > 2) row[rs[2]] += (kf * c0) * c1;
      double d0 = s.C[ps[0]], d1 = npr > 1 ? s.C[ps[1]] : 1.0, d2 = npr > 2 ? s.C[ps[2]] : 1.0;
      double kr = nu * o->kr;
      row[ps[0]] -= (kr * d1) * d2;
      if (npr > 1) row[ps[1]] -= (kr * d0) * d2;
      if (npr > 2) row[ps[2]] -= (kr * d0) * d1;
      if (m->rtype[r] != PR_RX_ELEM) {
        double dm = nu * o->dqdM;
        Sd += dm * m->eff_dflt[r];
        for (int q = m->eff_start[r]; q < m->eff_start[r + 1]; ++q) row[m->eff_sp[q]] += dm * m->eff_val[q];
      }
      JT += nu * o->dqdT;
    }
    const double Wk = m->W[k];
    for (int j = 0; j < N; ++j) row[j] = ((row[j] + Sd) * Wk) * SP[id][j].invW;
    row[N] = JT * Wk;
    fk[k] = w * Wk + c->rYsrc[k];
  }
  double a1[ORC_NMAX], a2[ORC_NMAX], a3[ORC_NMAX], a4[ORC_NMAX];
  for (int k = 0; k < N; ++k) {
    a1[k] = s.hk[k] * fk[k];
    a2[k] = y[k] * s.cpk[k];
    a3[k] = y[k] * s.dcpk[k];
    a4[k] = s.cpk[k] * fk[k];
  }
  double pn = orc_psum(a1, N), rcp = orc_psum(a2, N), sdcp = orc_psum(a3, N), scf = orc_psum(a4, N);
  const double fT = (c->rhohsrc - pn) / rcp;
  for (int j = 0; j < n; ++j) {
    double cs = 0.0;
    for (int k = 0; k < N; ++k) cs += s.hk[k] * J[k * n + j];
    J[N * n + j] = (j < N) ? (-cs - fT * s.cpk[j]) / rcp : ((-cs - scf) - fT * sdcp) / rcp;
  }
  return 0;
}

int orc_react(int mech, long ncells, double* rY, const double* rYsrc, double* T, double* rH,
              const double* rHsrc, double* FC, const int* mask, double dt,
              const orc_bdf_opts* opts, int nthreads, int* status, long* nst_out) {
  const pr_mech_t* m = MECHS[mech];
  int N = m->nsp, n = N + 1;
  int nfail = 0;
  long nst_tot = 0;
#pragma omp parallel for schedule(dynamic, 8) num_threads(nthreads > 0 ? nthreads : 1) \
    reduction(+ : nfail, nst_tot)
  for (long c = 0; c < ncells; ++c) {
    if (mask && mask[c] == -1) {
      FC[c] = 0.0;
      if (status) status[c] = 0;
      continue;
    }
    double y[ORC_NMAX], src[ORC_NMAX], neg[ORC_NMAX];
    for (int k = 0; k < N; ++k) {
      y[k] = rY[k * ncells + c];
      src[k] = rYsrc[k * ncells + c];
      /* ode.clean_init_massfrac: integrate from the non-negative part, add the rest back */
      neg[k] = (opts->clean_init_massfrac && y[k] < 0.0) ? y[k] : 0.0;
      if (neg[k] < 0.0) y[k] = 0.0;
    }
    y[N] = T[c];
    orc_cell cell = {mech, rH[c], rHsrc[c], src};
    orc_bdf_stats st;
    int ier = orc_bdf_solve(n, 0.0, dt, y, orc_reactor_rhs, orc_reactor_jac, &cell, opts, &st);
    for (int k = 0; k < N; ++k) rY[k * ncells + c] = neg[k] < 0.0 ? y[k] + neg[k] : y[k];
    T[c] = y[N];
    rH[c] = rH[c] + dt * rHsrc[c];
    FC[c] = (double)st.nfe;
    if (status) status[c] = ier;
    if (ier != ORC_OK) nfail++;
    nst_tot += st.nst;
  }
  if (nst_out) *nst_out = nst_tot;
  return nfail;
}
