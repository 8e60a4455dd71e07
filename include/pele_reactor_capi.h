/* C-ABI of the MI355X-native stiff chemistry reactor (libpelelmex_amd.so).
 *
 * Drop-in boundary for PelePhysics' reactor interface as PeleLMeX uses it. Every entry point
 * takes plain pointers and sizes; device arrays are HIP device pointers in AMReX Array4 layout
 * (i fastest, then j, k, component: p[(i-b.x) + (j-b.y)*jstride + (k-b.z)*kstride + n*nstride]).
 * Units inside are CGS, as in the reference (conversion at Source/PeleLMeX_Reactions.cpp:68-79).
 *
 * Reference interface each entry point replaces:
 *   pr_create        ReactorBase::create(name) + init(reactor_type, ncells)
 *                    (Source/PeleLMeX_Setup.cpp:146-155); the ode.* / cvode.* keys PelePhysics
 *                    parses from ParmParse (Docs/sphinx/manual/LMeXControls.rst:300-308) are
 *                    passed as "key = value" text.
 *   pr_set_typ_vals  ReactorBase::set_typ_vals_ode (Source/PeleLMeX_Utils.cpp:1472)
 *   pr_react_box     ReactorBase::react(box, rY, rYsrc, T, rE, rEsrc, FC, mask, dt, time, stream)
 *                    (Source/PeleLMeX_Reactions.cpp:101-108 and :267-274)
 *   pr_destroy       ReactorBase::close() + destructor (Source/PeleLMeX.cpp:26)
 *   pr_get_stats     no reference counterpart: integration counters of the last react call
 *                    (FC carries RHS counts per cell, as PeleLMeX's functC does, PeleLMeX.H:209)
 *   pr_last_error    no reference counterpart: message for the last non-zero return code
 *   pr_rhs_batch     the reactor right-hand side the integrator evaluates (PelePhysics' RHS kernel
 *                    inside ReactorCvode, fed by RTY2WDOT as in Source/PeleLMeX_K.H:848-889)
 *                    for a batch of states, one evaluation per cell
 *   pr_jac_batch     the analytic Jacobian the integrator factors (PelePhysics' aJacobian path,
 *                    PELE_COMPILE_AJACOBIAN in Exec/Make.PeleLMeX:100-104), one per cell
 *   pr_reaction_rate_box  reactionRateRhoY over a box (Source/PeleLMeX_K.H:848-889), as
 *                    PeleLM::computeInstantaneousReactionRate runs it per box
 *                    (Source/PeleLMeX_Reactions.cpp:399-435)
 */
#ifndef PELE_REACTOR_CAPI_H
#define PELE_REACTOR_CAPI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pr_handle pr_handle;

/* Array4 view: p points at element (begin[0], begin[1], begin[2], 0). */
typedef struct {
  double* p;
  int64_t jstride, kstride, nstride;
  int begin[3];
  int ncomp;
} pr_array4;

typedef struct {
  const int* p;
  int64_t jstride, kstride;
  int begin[3];
} pr_iarray4;

typedef struct {
  int64_t cells;      /* cells integrated (mask != -1) */
  int64_t failed;     /* cells whose integration failed */
  int64_t nst, nfe, nje, nsetups, nni, netf, ncfn; /* sums over cells */
  double kernel_ms;   /* device time of the integration kernel */
  int lanes_per_cell, blocks;
} pr_stats;

/* Error codes (negative); pr_react_box returns >= 0 (number of failed cells) on success. */
#define PR_ERR_ARG (-1)
#define PR_ERR_UNSUPPORTED (-2)
#define PR_ERR_HIP (-3)
#define PR_ERR_KEY (-4)

/* integrator: "ReactorCvode" (drop-in name) or "ReactorHipBDF"; mechanism: "LiDryer",
   "drm19"; reactor_type: 2 (constant pressure, rho h) is the one PeleLMeX uses. keyvals may
   be NULL. The current HIP device is the one the reactor runs on. */
int pr_create(const char* integrator, const char* mechanism, int reactor_type, int ncells,
              const char* keyvals, pr_handle** out);
int pr_set_typ_vals(pr_handle* h, const double* typ, int n);
/* Integrates every cell of the box [lo, hi] with mask != -1 over *dt_react (mask may be NULL:
   all cells active). rY, T, rH, FC are updated in place; rH <- rH + dt * rHsrc, FC <- RHS
   evaluations (0 on masked cells). Enqueues on `stream` and waits for it, as the reference
   synchronizes after each box. status_out (device, one int per box cell, may be NULL) gets
   0 or a negative integrator code per cell. *time is advanced by *dt_react. */
int pr_react_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY,
                 const pr_array4* rYsrc, const pr_array4* T, const pr_array4* rH,
                 const pr_array4* rHsrc, const pr_array4* FC, const pr_iarray4* mask,
                 double* dt_react, double* time, void* stream, int* status_out);
/* Batched evaluations on device arrays stored component-major with stride ncells:
   y[i * ncells + c] (i < NUM_SPECIES + 1: rhoY then T), rYsrc[k * ncells + c], rhoh0[c],
   rhohsrc[c]. f gets the right-hand side at time t (T recomputed from rho h(t) and Y, starting
   from y_T); J[(i * n + j) * ncells + c] = d f_i / d y_j with T = y_T. Synchronous. */
int pr_rhs_batch(pr_handle* h, int64_t ncells, const double* y, const double* rYsrc, const double* rhoh0,
                 const double* rhohsrc, double t, double* f, void* stream);
int pr_jac_batch(pr_handle* h, int64_t ncells, const double* y, const double* rYsrc, const double* rhohsrc,
                 double* J, void* stream);
/* Instantaneous production rates rhoYdot = W_k wdot_k of every cell of [lo, hi], MKS in and out
   as in the reference (rY kg m^-3, rH J m^-3, T K; rYdot kg m^-3 s^-1): T is recomputed from
   (rH / rho, Y) starting at the stored T (T itself is not written). Cells with mask == -1 get
   zero rates (mask may be NULL). Enqueues on `stream` and waits for it. */
int pr_reaction_rate_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY,
                         const pr_array4* rH, const pr_array4* T, const pr_array4* rYdot,
                         const pr_iarray4* mask, void* stream);
int pr_get_stats(const pr_handle* h, pr_stats* out);
int pr_num_species(const pr_handle* h);
void pr_destroy(pr_handle* h);
const char* pr_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
