/* Mechanism table descriptor shared by the CPU oracle (C) and the HIP library (C++/HIP).
 *
 * Content corresponds to what PelePhysics' generated mechanism.H provides to the reactor
 * (species count and weights, NASA-7 thermo, Arrhenius / third-body / falloff data; reference
 * build file list CMake/BuildPelePhysicsLib.cmake:8,126-131), stored as data tables.
 * Units: CGS (mol, cm, s, K); activation energies are stored as temperatures Ea/R in K.
 */
#ifndef PR_MECH_COMMON_H
#define PR_MECH_COMMON_H

#ifdef __cplusplus
extern "C" {
#endif

/* Reaction kinds. */
enum { PR_RX_ELEM = 0, PR_RX_THIRDBODY = 1, PR_RX_LINDEMANN = 2, PR_RX_TROE = 3 };

typedef struct {
  const char* name;
  int nsp, nr, nelem, neff, ninc;
  const char* const* species;
  const char* const* elements;
  const double* W;         /* [nsp] g/mol */
  const double* nasa;      /* [nsp][15]: T_mid, high a1..a7, low a1..a7 */
  const int* elem;         /* [nsp][nelem] atom counts */
  const int* rtype;        /* [nr] PR_RX_* */
  const int* rev;          /* [nr] 1 = reversible */
  const int* nreac;        /* [nr] molecule count on the left (third body excluded) */
  const int* reac;         /* [nr][3] species index per molecule slot */
  const int* nprod;        /* [nr] */
  const int* prod;         /* [nr][3] */
  const double* arr;       /* [nr][3] ln A, beta, Ea/R */
  const double* low;       /* [nr][3] ln A0, beta0, Ea0/R (falloff only) */
  const double* troe;      /* [nr][4] a, T***, T*, T** */
  const int* ntroe;        /* [nr] 3 or 4 Troe parameters, 0 otherwise */
  const double* eff_dflt;  /* [nr] default collision efficiency (1 for +M, 0 for none) */
  const int* eff_start;    /* [nr+1] CSR into eff_sp / eff_val */
  const int* eff_sp;       /* [neff] */
  const double* eff_val;   /* [neff] efficiency minus default */
  const int* inc_start;    /* [nsp+1] CSR into inc_r / inc_nu */
  const int* inc_r;        /* [ninc] reactions touching the species */
  const int* inc_nu;       /* [ninc] net stoichiometric coefficient */
} pr_mech_t;

/* Net production rate of species k: wdot_k = sum over its incidence list (reaction order) of
 * nu_e q_e. The list is cut into consecutive blocks of PR_GATHER_BLOCK entries; each block is
 * summed from 0.0 in list order, and the block sums are added in list order to the first one. A
 * list of at most PR_GATHER_BLOCK entries is thus one sequential sum. The blocks of all species
 * are independent, which spreads a long list (H, OH, O) over several GPU lanes. */
#define PR_GATHER_BLOCK 8

/* Physical constants (CGS). */
#define PR_RU 8.31446261815324e7   /* erg / (mol K) */
#define PR_PATM 1013250.0          /* dyn / cm^2 */

#ifdef __cplusplus
}
#endif
#endif
