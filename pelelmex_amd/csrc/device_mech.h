// Packed mechanism tables as the HIP kernels read them (host packs, device stages into LDS).
//
// One reaction record is 12 doubles + 4 ints; one species record is 46 doubles. The host
// builds the blob once per reactor from the generated pr_mech_t tables (pr_create) and the
// kernels copy it into LDS at launch, so every Arrhenius / third-body / falloff parameter the
// rate and Jacobian code touches is an LDS read.
#ifndef PR_DEVICE_MECH_H
#define PR_DEVICE_MECH_H
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "mech_common.h"

#define PR_AL16(x) (((x) + 15) & ~15)

namespace prk {

struct RxRec {
  double lnA, beta, Ta;       // k_inf (or k) = exp(lnA + beta ln T - Ta / T)
  double lnA0, beta0, Ta0;    // low-pressure limit (falloff)
  double troe[4];             // a, T***, T*, T**
  double effd;                // default collision efficiency
  double pad;
  int32_t info;               // type | rev << 4 | nreac << 8 | nprod << 12 | ntroe << 16
  int32_t reac;               // three 8-bit species slots
  int32_t prod;               // three 8-bit species slots
  int32_t eff;                // eff_begin | eff_end << 16
};

// NASA-7 fits rewritten for Horner evaluation (range 0 = low, T < T_mid; 1 = high):
//   cp/R = cp[0..4] in T;  h/RT = h[0..4] in T + h[5] / T;  s/R = s[0] ln T + T (s[1..4] in T) + s[5];
//   d(cp/R)/dT = dcp[0..3] in T.  RW = R_u / W, invW = 1 / W (correctly rounded on the host).
struct SpRec {
  double cp[2][5], h[2][6], s[2][6], dcp[2][4];
  double W, invW, RW, tmid;
};

// Blob layout (byte offsets, 16-byte aligned sections):
//   SpRec sp[nsp] | RxRec rx[nr] | int eff_sp[neff] | double eff_val[neff]
//   | int inc_start[nsp+1] | int inc_rnu[ninc]  (reaction | nu << 16, nu as signed 16-bit)
//   | int inc_qoff[ninc] (byte offset of the reaction's rate slot) | double inc_nud[ninc] (nu)
//   | int blk_of_sp[nsp+1] (first gather block of each species; [nsp] = block count)
//   | int blk_start[max_blocks+1] (first incidence entry of each block)
// The gather blocks cut every species' incidence list into PR_GATHER_BLOCK-entry pieces, so
// there are at most ninc / PR_GATHER_BLOCK + nsp of them.
struct BlobLayout {
  int nsp, nr, neff, ninc;
  int off_sp, off_rx, off_effsp, off_effval, off_incs, off_inc, off_incq, off_incnu, off_kblk, off_blk, bytes;
};

__host__ __device__ inline int gather_blocks_max(int nsp, int ninc) { return ninc / PR_GATHER_BLOCK + nsp + 1; }

__host__ __device__ inline BlobLayout blob_layout(int nsp, int nr, int neff, int ninc) {
  BlobLayout L;
  L.nsp = nsp; L.nr = nr; L.neff = neff; L.ninc = ninc;
  int o = 0;
  L.off_sp = o;     o = PR_AL16(o + nsp * (int)sizeof(SpRec));
  L.off_rx = o;     o = PR_AL16(o + nr * (int)sizeof(RxRec));
  L.off_effsp = o;  o = PR_AL16(o + (neff > 0 ? neff : 1) * 4);
  L.off_effval = o; o = PR_AL16(o + (neff > 0 ? neff : 1) * 8);
  L.off_incs = o;   o = PR_AL16(o + (nsp + 1) * 4);
  L.off_inc = o;    o = PR_AL16(o + (ninc > 0 ? ninc : 1) * 4);
  L.off_incq = o;   o = PR_AL16(o + (ninc > 0 ? ninc : 1) * 4);
  L.off_incnu = o;  o = PR_AL16(o + (ninc > 0 ? ninc : 1) * 8);
  L.off_kblk = o;   o = PR_AL16(o + (nsp + 1) * 4);
  L.off_blk = o;    o = PR_AL16(o + (gather_blocks_max(nsp, ninc) + 1) * 4);
  L.bytes = o;
  return L;
}

}  // namespace prk
#endif
