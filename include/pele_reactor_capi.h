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
 *   pr_set_eos_parm  ReactorBase::set_eos_parm (Source/PeleLMeX_Setup.cpp:155)
 *   pr_react_box     ReactorBase::react(box, rY, rYsrc, T, rE, rEsrc, FC, mask, dt, time, stream)
 *                    (Source/PeleLMeX_Reactions.cpp:101-108 and :267-274)
 *   pr_react_batch   the per-box loop around react in advanceChemistry
 *                    (Source/PeleLMeX_Reactions.cpp:48-108, chem-BoxArray variant :243-274):
 *                    every box of a level, each with its own dt_react, in one launch
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
 *   pr_box_costs     PeleLM::computeCosts for the chemistry BoxArray (Source/PeleLMeX_Regrid.cpp:558-606)
 *   pr_chem_boxes    the chemistry BoxArray and its flags built in PeleLM::resetCoveredMask
 *                    (Source/PeleLMeX_Utils.cpp:728-788)
 *   pr_set_temperature_box  PeleLM::setTemperature per box (Source/PeleLMeX_Eos.cpp:218-240,
 *                    getTfromHY Source/PeleLMeX_K.H:700-723)
 *   pr_floor_species_box    PeleLM::floorSpecies per box (Source/PeleLMeX_Utils.cpp:649-690)
 *   pr_advance_chemistry_box  the per-box body of PeleLM::advanceChemistry(lev, dt, Forcing)
 *                    (Source/PeleLMeX_Reactions.cpp:54-160: conversions, react, I_R)
 *   pr_advance_chemistry_chem_box  the per-box body of PeleLM::advanceChemistryBAChem on the
 *                    chemistry BoxArray (Source/PeleLMeX_Reactions.cpp:220-305)
 *   pr_reaction_source_box  I_R formation in advanceChemistry (Source/PeleLMeX_Reactions.cpp:141-160,
 *                    342-347)
 *   pr_heat_release_box     PeleLM::getHeatRelease per box (Source/PeleLMeX_Reactions.cpp:477-506,
 *                    getHGivenT Source/PeleLMeX_K.H:598-615)
 *   pr_surrogate_load, pr_surrogate_features, pr_surrogate_forward, pr_mlp_*   no reference
 *                    counterpart: the learned
 *                    dt_react map that pr_react_box tries before the BDF (cells failing its FP64
 *                    trust check are integrated by the BDF), and the kernels that train it
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
  int64_t cells;      /* cells integrated by the BDF (mask != -1; with a surrogate: its rejects) */
  int64_t failed;     /* cells whose integration failed */
  int64_t nst, nfe, nje, nsetups, nni, netf, ncfn; /* sums over BDF cells */
  double kernel_ms;   /* device time of the BDF integration kernel */
  int lanes_per_cell, blocks;
  /* surrogate path (zero without one) */
  int64_t sur_accepted, sur_rejected, sur_audited, audit_failed;
  double total_ms;      /* device time of the whole react call */
  double surrogate_ms;  /* features + network + trust check */
  double audit_ms;      /* BDF re-solve of the audit sample */
  double audit_max_dT_rel, audit_max_dY;  /* worst |dT| / T and |dY_k| of audited cells vs BDF */
  int block_threads, waves_per_simd;      /* BDF kernel launch shape */
} pr_stats;

/* Error codes (negative); pr_react_box returns >= 0 (number of failed cells) on success. */
#define PR_ERR_ARG (-1)
#define PR_ERR_UNSUPPORTED (-2)
#define PR_ERR_HIP (-3)
#define PR_ERR_KEY (-4)

/* integrator: "ReactorCvode" (drop-in name) or "ReactorHipBDF"; mechanism: "LiDryer",
   "drm19", "dodecane_lu", "syn53"; reactor_type: 2 (constant pressure, rho h) is the one PeleLMeX uses.
   keyvals may be NULL. The current HIP device is the one the reactor runs on.
   Threads: calls on one handle may come from several host threads (PeleLMeX's OpenMP loop over
   boxes, Source/PeleLMeX_Reactions.cpp:48-55); they are serialized per handle, since its
   buffers and statistics are shared. Separate handles run independently. pr_last_error is
   per thread. */
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
/* No counterpart in the reference. pr_react_box for callers that cannot allocate device memory
   (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): same
   arguments, kernels and results (bit for bit), but it calls no hipMalloc / hipFree. Its
   counters, box table, cost order and BDF group slots live in one static device array of
   PR_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends. Limits: a box of at most
   4194304 cells when hipbdf.cost_order = 1 (larger boxes return PR_ERR_ARG); the persistent
   grid is capped to the group slots the array holds (on a 256-CU MI355X that is every resident
   group of every shipped kernel variant); no learned map
   (surrogate.enable = 1 returns PR_ERR_UNSUPPORTED) and no hipbdf.autotune before a first
   pr_react_box call has tuned the handle. Every call on one GPU shares that GPU's array, which
   keeps the last call's contents: calls that overlap on one GPU (two streams, two threads on
   two handles) race and must be serialized by the caller. */
#define PR_NOALLOC_SCRATCH_BYTES (192ull << 20)
int pr_react_box_noalloc(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY,
                         const pr_array4* rYsrc, const pr_array4* T, const pr_array4* rH,
                         const pr_array4* rHsrc, const pr_array4* FC, const pr_iarray4* mask,
                         double* dt_react, double* time, void* stream, int* status_out);
/* One box of pr_react_batch: the arguments of pr_react_box for that box. */
typedef struct {
  int lo[3], hi[3];
  pr_array4 rY, rYsrc, T, rH, rHsrc, FC;
  pr_iarray4 mask;   /* mask.p may be NULL: every cell active */
  double dt_react;
  int* status;       /* device, one int per box cell, may be NULL */
} pr_box;

/* Integrates every box of the list over its own dt_react with the semantics of pr_react_box,
   in one persistent launch whose work queue holds the cells of all boxes (so boxes of mixed
   sizes and stiffness balance across the GPU). With a learned map loaded (and
   surrogate.enable = 1) every box's cells go through it first - features per box, one network
   pass over all boxes, the FP64 trust check per box with its dt - and the persistent launch
   integrates the rejected cells of all boxes; results are those of pr_react_box with the same
   map, box by box. Returns the number of failed cells (>= 0) or a negative error code. nbox = 0
   (boxes may be NULL: a level with no box on this rank) integrates nothing and returns 0. */
int pr_react_batch(pr_handle* h, int nbox, const pr_box* boxes, void* stream);
/* pr_react_batch in two halves, for callers that capture the work in a hipGraph: _enqueue only
   enqueues on `stream` (a copy of the box table from the handle's pinned staging memory, the
   counter reset kernels, the learned map's kernels when one is loaded, the cost-ordering
   kernels (hipbdf.cost_order), the integration and audit kernels: no synchronization, no
   memset node, and no allocation once a first call of the same or larger size has run);
   pr_collect waits for the
   stream and returns the failed-cell count of the last integration, filling pr_get_stats. A
   captured batch replays the box table staged at capture; other react calls on the same handle
   restage it. */
int pr_react_batch_enqueue(pr_handle* h, int nbox, const pr_box* boxes, void* stream);
int pr_collect(pr_handle* h, void* stream);

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
/* Chemistry cost per box from the reactor's work counters (FC = right-hand-side evaluations per
   cell), as PeleLM::computeCosts does for the chemistry BoxArray before its knapsack / SFC
   distribution (Source/PeleLMeX_Regrid.cpp:558-606): method 0 Ncell (cells of the box), 1
   ChemFunctCallAvg (sum of FC / cells), 2 ChemFunctCallMax, 3 ChemFunctCallSum. Box b spans
   lo[3b..3b+2] .. hi[3b..3b+2]; FC[b] is its (device) FC view (unused for method 0); costs is a
   host array of nbox doubles. Waits for `stream`. */
int pr_box_costs(pr_handle* h, int nbox, const int* lo, const int* hi, const pr_array4* FC, int method,
                 double* costs, void* stream);
/* The chemistry BoxArray of a level and its flags, as PeleLM::resetCoveredMask builds them
   (Source/PeleLMeX_Utils.cpp:728-788; host only, no handle). Below the finest level: the cells
   of the level's grids that the finer level does not cover (domain minus the finer grids, given
   at the finer resolution and coarsened by ref_ratio, intersected with the grids), chopped so no
   box exceeds max_grid_size_chem in any direction when all three entries are > 0, flag 1
   (reacted); then the coarsened finer grids, flag 0 (advanceChemistryBAChem gives them FC = 0
   without reacting them). Finest level (finest != 0): the grids, chopped the same way, all flag 1.
   Boxes are cell-centred, lo / hi inclusive, 3 ints per box. Returns the number of chemistry
   boxes n and writes the first min(n, cap) (call again with cap >= n), or a negative error. */
int pr_chem_boxes(const int domain_lo[3], const int domain_hi[3], int ngrid, const int* grid_lo, const int* grid_hi,
                  int nfine, const int* fine_lo, const int* fine_hi, int ref_ratio, const int max_grid_size_chem[3],
                  int finest, int cap, int* box_lo, int* box_hi, int* flags);
/* Equation-of-state box kernels around the chemistry, MKS Array4 views as PeleLMeX's state
   (rho kg m^-3, rY kg m^-3, rH J m^-3, T K), one enqueue on `stream` then a wait for it.
   pr_set_temperature_box (setTemperature, Source/PeleLMeX_Eos.cpp:218-240 with getTfromHY,
   Source/PeleLMeX_K.H:700-723): T = HY2T(rH / rho * 1e4, rY / rho) from the stored T (the
   reactor's own temperature iteration); reads rho, rY, rH, T, writes T.
   pr_floor_species_box (floorSpecies, Source/PeleLMeX_Utils.cpp:649-690): rY_k = max(0,
   min(DBL_MAX, rY_k)), rho = sum_k rY_k, rH = TY2H(T, rY / rho) * 1e-4 * rho; reads rY, T,
   writes rY, rho, rH.
   pr_heat_release_box (getHeatRelease, Source/PeleLMeX_Reactions.cpp:477-506): HR = -sum_k
   h_k(T) I_R,k with h_k in J kg^-1 (getHGivenT, Source/PeleLMeX_K.H:598-615) and I_R in
   kg m^-3 s^-1; reads T, IR, writes HR (W m^-3). */
int pr_set_temperature_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rho,
                           const pr_array4* rY, const pr_array4* rH, const pr_array4* T, void* stream);
int pr_floor_species_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY,
                         const pr_array4* rho, const pr_array4* rH, const pr_array4* T, void* stream);
int pr_heat_release_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* T,
                        const pr_array4* IR, const pr_array4* HR, void* stream);
/* I_R formation after the advance: I_R_k = -(rhoY_old_k - rhoY_new_k) * (1 / dt) - F_rhoY_k over
   the box (MKS Array4 views, NUM_SPECIES components each; the reciprocal is formed once, as
   PeleLMeX forms dt_inv). Replaces the ParallelFor at Source/PeleLMeX_Reactions.cpp:141-160
   (and the I_R lines of :342-347 on the chemistry BoxArray). Synchronous. */
int pr_reaction_source_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY_old,
                           const pr_array4* rY_new, const pr_array4* extF_rY, const pr_array4* I_R, double dt,
                           void* stream);

/* One box of PeleLM::advanceChemistry(lev, dt, Forcing) in one call, on MKS views of the level
   state: rhoY_new = rhoY_old * 1e-3, T_new = T_old, rhoH_new = rhoH_old * 10 and the forcing
   scaled to CGS in place (Source/PeleLMeX_Reactions.cpp:68-79); pr_react_box over dt on the new
   state (:101-108; the learned map too when loaded); the new state and the forcing scaled back
   (:110-120); and, when I_R->p is not NULL, I_R = -(rhoY_old - rhoY_new) * (1 / dt) - F_rhoY from
   the restored forcing (:141-160). Every product is the reference's, so the caller's forcing
   makes the same rounding round trip. Returns the failed-cell count (>= 0) or an error code;
   after an error inside react the new state and the forcing are still scaled back to MKS. */
int pr_advance_chemistry_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY_old,
                             const pr_array4* rH_old, const pr_array4* T_old, const pr_array4* rY_new,
                             const pr_array4* rH_new, const pr_array4* T_new, const pr_array4* extF_rY,
                             const pr_array4* extF_rH, const pr_array4* FC, const pr_iarray4* mask,
                             const pr_array4* I_R, double dt, void* stream, int* status_out);
/* One box of PeleLM::advanceChemistryBAChem(lev, dt, Forcing) on the chemistry BoxArray's copy
   of the state (Source/PeleLMeX_Reactions.cpp:220-305), in place on MKS views: rY (NUM_SPECIES
   components), rH and T of the chem copy and its forcing copy extF_rY, extF_rH. rhoY and its
   forcing times 1e-3, rho h and its forcing times 10 (:228-238); when do_reaction != 0
   (m_baChemFlag: the box is not entirely EB-covered) pr_react_box over dt (:260-274), otherwise
   FC = 0 on every cell and no integration (:276-280); then the state only back to MKS (rhoY times
   1e3, rho h times 0.1, :284-291; the forcing copy stays in CGS, as the reference leaves its
   scratch copy). The ParallelCopy into and out of the chem copy and the I_R of :342-347
   (pr_reaction_source_box) stay with the caller. Returns the failed-cell count (>= 0) or an
   error code; after an error inside react the state is still scaled back. */
int pr_advance_chemistry_chem_box(pr_handle* h, const int lo[3], const int hi[3], const pr_array4* rY,
                                  const pr_array4* rH, const pr_array4* T, const pr_array4* extF_rY,
                                  const pr_array4* extF_rH, const pr_array4* FC, const pr_iarray4* mask,
                                  int do_reaction, double dt, void* stream, int* status_out);
/* ReactorBase::set_eos_parm (Source/PeleLMeX_Setup.cpp:155, called right after init with
   eos_parms.device_parm()): eos_parm is the caller's EOS parameter block (device pointer, may be
   NULL), eos_model the build's Eos_Model ("Fuego", the ideal-gas mixture EOS of the reference
   cases, or "GammaLaw"; NULL means "Fuego"). The reactor's thermodynamics is the ideal-gas mixture
   of the mechanism's NASA-7 fits, for which that block carries no parameter it reads: the pointer
   is recorded and not dereferenced. A real-gas model (e.g. "SRK") returns PR_ERR_UNSUPPORTED. */
int pr_set_eos_parm(pr_handle* h, const void* eos_parm, const char* eos_model);
/* Learned dt_react map. Network: nlayers dense layers, dims[0] = NUM_SPECIES + 2 inputs
   (log10(Y_k + 1e-20), T / 2000, log10(dt)), dims[nlayers] = NUM_SPECIES + 1 outputs (dY_k, dT),
   GELU between layers. params (host, fp32): for each layer W [dims[l] x dims[l+1]] row-major then
   b [dims[l+1]]. xnorm: mean then standard deviation of the inputs (2 x dims[0]); ynorm: mean,
   standard deviation and rate-scale floor eps of the outputs (3 x dims[nlayers]): with
   u_i = out_i sd_i + mean_i the increment is d_i = u_i (dt |f_i(y*)| / rho* + eps_i) for species
   and u_T (dt |f_T(y*)| + eps_T) for T, f the source-free reactor right-hand side at the
   source-advanced state y* = y0 + dt src (or d_i = u_i when eps_i = 0). Once loaded (and surrogate.enable != 0) pr_react_box
   evaluates it in BF16 on the matrix cores, accepts the cells that pass the FP64 trust check
   (element balance, backward-Euler residual with the analytic rates at surrogate.trust_rtol /
   surrogate.trust_atol) and integrates the others with the BDF; surrogate.audit_frac of the
   accepted cells are re-solved by the BDF and their worst error reported in pr_stats. With
   surrogate.zero_increment = 1 the same path runs with every increment zero (the exact source
   step alone, no network pass): the baseline a trained network is measured against. */
int pr_surrogate_load(pr_handle* h, int nlayers, const int* dims, const float* params, const float* xnorm,
                      const float* ynorm);
/* The loaded network on given inputs (no reference counterpart): X device bf16 [ncells x dims[0]]
   normalised features (as pr_surrogate_features returns them), Y device fp32 [ncells x dims[L]]
   the raw outputs, the same forward the learned path of pr_react_box runs. Waits for `stream`. */
int pr_surrogate_forward(pr_handle* h, int64_t ncells, const void* X, float* Y, void* stream);
/* Network inputs for a batch of states (device, component-major rY[k * ncells + c], T[c],
   dt[c]): raw (fp32 [ncells x (N + 2)], may be NULL) and normalised bf16 X (may be NULL; needs
   a loaded network). */
int pr_surrogate_features(pr_handle* h, int64_t ncells, const double* rY, const double* T, const double* dt,
                          float* raw, void* X, void* stream);
/* Training kernels (device pointers; bf16 arrays as void*; all enqueue on stream):
   C = op(A) op(B) (+ bias, then GELU when gelu != 0) with A(m, k) = A[m sam + k sak] and
   B(k, n) = B[k sbk + n sbn] in bf16, fp32 accumulation; outputs C (fp32, ldc), Cb (bf16,
   ldcb), Z (pre-activation fp32, ldz), each nullable; accumulate != 0 adds the raw product
   into C instead. */
int pr_mlp_gemm(int M, int N, int K, const void* A, int64_t sam, int64_t sak, const void* B, int64_t sbk,
                int64_t sbn, const float* bias, int gelu, float* C, int64_t ldc, void* Cb, int64_t ldcb, float* Z,
                int64_t ldz, int accumulate, void* stream);
/* dZ = dA * gelu'(Z) into bf16 dZb (and fp32 dZ when non-NULL). */
int pr_mlp_gelu_backward(int64_t rows, int cols, const float* dA, int64_t lda, const float* Z, int64_t ldz,
                         void* dZb, int64_t ldb, float* dZ, int64_t ldd, void* stream);
/* out[c] (+)= sum over rows of d[r * ld + c]. */
int pr_mlp_column_sum(int64_t rows, int cols, const float* d, int64_t ld, float* out, int accumulate, void* stream);
/* Mean-squared error: dY = 2 (Y - T) / (rows cols), *loss += mean of (Y - T)^2. */
int pr_mlp_mse(int64_t rows, int cols, const float* Y, int64_t ldy, const float* T, int64_t ldt, float* dY,
               float* loss, void* stream);
/* Adam step `step` (1-based) on fp32 parameters p with gradient g and moments m, v; pb (bf16
   copy, nullable) is refreshed. */
int pr_mlp_adam(int64_t n, float* p, const float* g, float* m, float* v, void* pb, float lr, float beta1, float beta2,
                float eps, int step, void* stream);
/* b = bf16(a). */
int pr_mlp_to_bf16(int64_t n, const float* a, void* b, void* stream);
int pr_get_stats(const pr_handle* h, pr_stats* out);
int pr_num_species(const pr_handle* h);
void pr_destroy(pr_handle* h);
const char* pr_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
