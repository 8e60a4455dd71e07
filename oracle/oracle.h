/* CPU oracle: a plain-C restatement of the reactor path for testing and the CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load this library, and only as the checker or the timed CPU baseline; the product
 * path (pelelmex_amd) never calls it.
 *
 * What it restates (reference call sites; the bodies live in the absent PelePhysics/SUNDIALS
 * submodules, so their behaviour is re-derived from the published algorithms):
 *  - ReactorBase::react(box, rY, rYsrc, T, rH, rHsrc, FC, mask, dt, time)
 *      call: Source/PeleLMeX_Reactions.cpp:101-108, factory Source/PeleLMeX_Setup.cpp:146-160
 *  - the constant-pressure reactor ODE of Docs/sphinx/manual/Model.rst:300-318:
 *      d(rhoY_k)/dt = W_k wdot_k + Q_k, rho h(t) = rho h_0 + t Q_h, T from (h, Y)
 *  - CVODE's variable-order BDF with Newton and a dense direct linear solver (Hindmarsh et al.,
 *    ACM TOMS 31(3) 2005; Jackson & Sacks-Davis 1980), keys of LMeXControls.rst:300-308.
 *
 * Parity pins are in tests/ (closed forms, invariants, finite differences, scipy integrators).
 */
#ifndef PR_ORACLE_H
#define PR_ORACLE_H
#include "../pelelmex_amd/csrc/mech_common.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- generic BDF (CVODE semantics) ---------------- */
typedef int (*orc_rhs_fn)(double t, const double* y, double* f, void* ud);
/* J is row-major n x n: J[i*n + j] = d f_i / d y_j */
typedef int (*orc_jac_fn)(double t, const double* y, const double* fy, double* J, void* ud);

typedef struct {
  double rtol;
  double atol;          /* scalar absolute tolerance (used when atol_vec == NULL) */
  const double* atol_vec;
  int max_order;        /* cvode.max_order, 1..5 */
  int max_steps;        /* cvode.max_substeps */
  int clean_init_massfrac; /* ode.clean_init_massfrac (box react): integrate from max(rhoY, 0) and
                              add the clipped part back to the result */
} orc_bdf_opts;

typedef struct {
  long nst, nfe, nje, nsetups, nni, netf, ncfn;
  int q_last;
  double h_last;
} orc_bdf_stats;

enum {
  ORC_OK = 0,
  ORC_TOO_MUCH_WORK = -1,
  ORC_ERR_FAILURE = -3,
  ORC_CONV_FAILURE = -4,
  ORC_TOO_CLOSE = -27
};

#define ORC_NMAX 64

/* Integrate y' = f(t,y) from t0 to tout (CV_NORMAL: step past tout, interpolate back). */
int orc_bdf_solve(int n, double t0, double tout, double* y, orc_rhs_fn f, orc_jac_fn jac,
                  void* ud, const orc_bdf_opts* opts, orc_bdf_stats* st);

/* Dense LU with partial pivoting (first maximal |pivot|), row-major, in place. */
int orc_lu_factor(int n, double* A, int* piv);
void orc_lu_solve(int n, const double* LU, const int* piv, double* b);

/* Sum in the canonical pairwise order (zero-padded to 64 slots, pairs (2j, 2j+1) first). */
double orc_psum(const double* x, int n);

/* ---------------- mechanism kernels ---------------- */
int orc_mech_count(void);
const pr_mech_t* orc_mech_get(int id);
int orc_mech_find(const char* name);
int orc_mech_nsp(int id);
int orc_mech_nr(int id);
void orc_mech_weights(int id, double* W);

/* per species: cp/R, h/RT, s/R at T */
void orc_thermo(int id, double T, double* cpR, double* hRT, double* sR);
/* rates of progress q[nr] (mol/cm^3/s) and molar production rates wdot[nsp]; C in mol/cm^3 */
void orc_rates(int id, const double* C, double T, double* q, double* wdot);
/* forward and reverse rates of progress per reaction (mol/cm^3/s) */
void orc_rates_fr(int id, const double* C, double T, double* fwd, double* rev);
/* the oracle's exp / log (the same routines the HIP kernels use) */
double orc_exp(double x);
double orc_log(double x);
/* T from mass-specific enthalpy h (erg/g) and Y by Newton from Tguess; returns iterations */
int orc_hy2t(int id, double h, const double* Y, double* T);
/* reactionRateRhoY on ncells cells, MKS, component-major arrays; Tout nullable */
void orc_rate_rhoY(int id, long ncells, const double* rY, const double* rH, const double* T0, double* out,
                   double* Tout);
/* setTemperature / floorSpecies / getHeatRelease per cell, MKS, component-major species arrays */
void orc_set_temperature(int id, long ncells, const double* rho, const double* rY, const double* rH, double* T);
void orc_floor_species(int id, long ncells, double* rY, double* rho, double* rH, const double* T);
void orc_heat_release(int id, long ncells, const double* T, const double* IR, double* HR);

/* Reactor right-hand side on y = (rhoY_1..N, T) (CGS) with rho h(t) = rhoh0 + t*rhohsrc. */
typedef struct {
  int mech;
  double rhoh0, rhohsrc;
  const double* rYsrc; /* [nsp] */
} orc_cell;
int orc_reactor_rhs(double t, const double* y, double* f, void* cell /* orc_cell* */);
/* Same right-hand side with T taken from y[N] (no enthalpy inversion); its exact Jacobian is
   what orc_reactor_jac returns. */
int orc_reactor_rhs_frozen(double t, const double* y, double* f, void* cell);
int orc_reactor_jac(double t, const double* y, const double* fy, double* J, void* cell);

/* ReactorBase::react over a flattened box: component-major arrays with stride ncells
   (Array4 with nstride = ncells). Cells with mask == -1 are skipped and get FC = 0.
   Returns 0 when every cell succeeded; status[c] holds the per-cell code when non-NULL. */
int orc_react(int mech, long ncells, double* rY, const double* rYsrc, double* T, double* rH,
              const double* rHsrc, double* FC, const int* mask, double dt,
              const orc_bdf_opts* opts, int nthreads, int* status, long* nst_out);

#ifdef __cplusplus
}
#endif
#endif
