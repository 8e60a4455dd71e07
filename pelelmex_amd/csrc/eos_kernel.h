// Equation-of-state box kernels around the chemistry advance, MKS Array4 views, one thread per
// cell (cells fastest, so every species row of the box is read and written coalesced; the
// species records are staged in LDS and read by all lanes at once):
//  * setTemperature  (Source/PeleLMeX_Eos.cpp:218-240 -> getTfromHY, Source/PeleLMeX_K.H:700-723):
//    rhoinv = 1 / rho, Y = rhoY * rhoinv, h = rhoH * rhoinv * 1e4, T = HY2T(h, Y) from the stored
//    T - the reactor's own iteration and sums (hy2t), so the temperature is the one its
//    right-hand side computes;
//  * floorSpecies    (Source/PeleLMeX_Utils.cpp:649-690, fabMinMax Source/PeleLMeX_K.H:894-907):
//    rhoY_k = max(0, min(DBL_MAX, rhoY_k)) with std::min / std::max semantics, rho = sum_k rhoY_k
//    and h = TY2H(T, Y) = sum_k Y_k h_k(T) in species order, rhoH = h * 1e-4 * rho;
//  * getHeatRelease  (Source/PeleLMeX_Reactions.cpp:477-506, getHGivenT Source/PeleLMeX_K.H:598-615):
//    HR = 0 - sum_k (h_k(T) * 1e-4) * I_R,k in species order.
// h_k(T) = (h_k / RT) (R / W_k) T in erg / g, the reactor's species enthalpy. HBM-bound: a cell
// moves (N + 4), (2 N + 3) and (N + 2) doubles.
#pragma once

#include <cfloat>

#include "react_kernel.h"

namespace prk {

enum { EOS_SET_T = 0, EOS_FLOOR = 1, EOS_HEAT_RELEASE = 2 };

struct EosParams {
  int lo0, lo1, lo2, nx, ny;
  long ncells;
  Arr rY, rho, rH, T, IR, HR;
  const unsigned char* blob;
  int blob_bytes, neff, ninc;
};

// hy2t of the reactor (bdf_kernel.h) by one thread: the same Newton iteration, with the mixture
// enthalpy and heat capacity summed over the canonical pairwise tree of the reactor's component
// sums (64 zero-padded slots: pairs, pairs of pairs, ...). The tree is evaluated as the slots are
// formed, keeping one pending left subtree per level, so only six partial sums per quantity are
// live; Y_k = rhoY_k / rho is re-read from the cell's row at each iteration.
template <int N>
__device__ __forceinline__ double hy2t_cell(const SpRec* sp, const double* rY, long ns, double rhoinv, double h,
                                            double T) {
  static_assert(N <= 64, "at most 64 components");
  for (int it = 0; it < 20; ++it) {
    const double invT = 1.0 / T;
    double sh[6], sc[6], hm = 0.0, cpm = 0.0;
#pragma unroll 1
    for (int s = 0; s < 64; ++s) {
      double a = 0.0, b = 0.0;
      if (s < N) {
        const double Y = rY[(long)s * ns] * rhoinv;
        double cpR, hRT;
        nasa_hcp(sp[s], T, invT, cpR, hRT);
        a = ((Y * hRT) * sp[s].RW) * T;
        b = (Y * cpR) * sp[s].RW;
      }
      bool placed = false;
#pragma unroll
      for (int l = 0; l < 6; ++l) {
        if (!placed) {
          if ((s >> l) & 1) {  // the right subtree at level l: join its pending left sibling
            a = sh[l] + a;
            b = sc[l] + b;
          } else {  // a left subtree: pending until its sibling is complete
            sh[l] = a;
            sc[l] = b;
            placed = true;
          }
        }
      }
      if (s == 63) {  // every level has joined: the totals
        hm = a;
        cpm = b;
      }
    }
    const double dT = (h - hm) / cpm;
    T = T + dT;
    if (fabs(dT) <= 1e-10 * T) break;
  }
  return T;
}

template <int N, int R, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) eos_cell_kernel(EosParams P) {
  extern __shared__ __align__(16) unsigned char smem[];
  {
    // the species records only: the blob's first section
    const BlobLayout BL = blob_layout(N, R, P.neff, P.ninc);
    const uint4* srcb = reinterpret_cast<const uint4*>(P.blob + BL.off_sp);
    uint4* dst = reinterpret_cast<uint4*>(smem);
    for (int b = threadIdx.x; b < (int)(N * sizeof(SpRec)) / 16; b += blockDim.x) dst[b] = srcb[b];
  }
  __syncthreads();
  const SpRec* sp = reinterpret_cast<const SpRec*>(smem);
  const long stride = (long)gridDim.x * blockDim.x;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < P.ncells; c += stride) {
    const unsigned u = (unsigned)c;
    const int i = P.lo0 + (int)(u % (unsigned)P.nx);
    const int j = P.lo1 + (int)((u / (unsigned)P.nx) % (unsigned)P.ny);
    const int k = P.lo2 + (int)(u / ((unsigned)P.nx * (unsigned)P.ny));
    if constexpr (MODE == EOS_SET_T) {
      const double rhoinv = 1.0 / P.rho.p[aoff(P.rho, i, j, k)];
      const long offY = aoff(P.rY, i, j, k), oT = aoff(P.T, i, j, k);
      P.T.p[oT] = hy2t_cell<N>(sp, P.rY.p + offY, P.rY.ns, rhoinv, (P.rH.p[aoff(P.rH, i, j, k)] * rhoinv) * 1.0e4,
                               P.T.p[oT]);
    } else if constexpr (MODE == EOS_FLOOR) {
      const long offY = aoff(P.rY, i, j, k);
      double rho = 0.0;
#pragma unroll 4
      for (int s = 0; s < N; ++s) {
        double x = P.rY.p[offY + (long)s * P.rY.ns];
        x = (x < DBL_MAX) ? x : DBL_MAX;  // min(DBL_MAX, x): NaN becomes DBL_MAX
        x = (0.0 < x) ? x : 0.0;          // max(0, x): negatives and -0 become +0
        P.rY.p[offY + (long)s * P.rY.ns] = x;
        rho += x;
      }
      const double rhoinv = 1.0 / rho;
      const double T = P.T.p[aoff(P.T, i, j, k)];
      const double invT = 1.0 / T;
      double h = 0.0;
#pragma unroll 4
      for (int s = 0; s < N; ++s) {  // the floored values just written, re-read from the row
        double cpR, hRT;
        nasa_hcp(sp[s], T, invT, cpR, hRT);
        h += ((P.rY.p[offY + (long)s * P.rY.ns] * rhoinv * hRT) * sp[s].RW) * T;
      }
      P.rho.p[aoff(P.rho, i, j, k)] = rho;
      P.rH.p[aoff(P.rH, i, j, k)] = (h * 1.0e-4) * rho;
    } else {
      const double T = P.T.p[aoff(P.T, i, j, k)];
      const double invT = 1.0 / T;
      const long offR = aoff(P.IR, i, j, k);
      double hr = 0.0;
#pragma unroll 4
      for (int s = 0; s < N; ++s) {
        double cpR, hRT;
        nasa_hcp(sp[s], T, invT, cpR, hRT);
        hr -= (((hRT * sp[s].RW) * T) * 0.0001) * P.IR.p[offR + (long)s * P.IR.ns];
      }
      P.HR.p[aoff(P.HR, i, j, k)] = hr;
    }
  }
}

// I_R formation after the advance (Source/PeleLMeX_Reactions.cpp:141-160, and :342-347 on the
// chemistry BoxArray): I_R_k = -(rhoY_old_k - rhoY_new_k) * dt_inv - F_rhoY_k with
// dt_inv = 1 / dt formed on the host, as the reference forms it before its ParallelFor. One
// thread per (cell, species), cells fastest so that the Array4 reads coalesce; HBM-bound (three
// reads and one write of 8 bytes per entry).
struct IrParams {
  int lo0, lo1, lo2, nx, ny;
  long ncells;
  int N;
  Arr rYo, rYn, F, IR;
  double dt_inv;
};

__global__ void __launch_bounds__(256) ir_kernel(IrParams P) {
  const long total = P.ncells * (long)P.N;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const long c = e % P.ncells;
    const long n = e / P.ncells;
    const unsigned u = (unsigned)c;
    const int i = P.lo0 + (int)(u % (unsigned)P.nx);
    const int j = P.lo1 + (int)((u / (unsigned)P.nx) % (unsigned)P.ny);
    const int k = P.lo2 + (int)(u / ((unsigned)P.nx * (unsigned)P.ny));
    const double o = P.rYo.p[aoff(P.rYo, i, j, k) + n * P.rYo.ns];
    const double w = P.rYn.p[aoff(P.rYn, i, j, k) + n * P.rYn.ns];
    const double f = P.F.p[aoff(P.F, i, j, k) + n * P.F.ns];
    P.IR.p[aoff(P.IR, i, j, k) + n * P.IR.ns] = -(o - w) * P.dt_inv - f;
  }
}

// The unit conversions around react in advanceChemistry (Source/PeleLMeX_Reactions.cpp:68-79
// before, :110-120 after), one thread per cell, MKS views:
//  before: rhoY_new = rhoY_old * 1e-3, F_rhoY *= 1e-3, T_new = T_old, rhoH_new = rhoH_old * 10,
//          F_rhoH *= 10 (the forcing is scaled in place, as the reference scales it);
//  after:  rhoY_new *= 1e3, F_rhoY *= 1e3, rhoH_new *= 0.1, F_rhoH *= 0.1, then
//          I_R = -(rhoY_old - rhoY_new) * dt_inv - F_rhoY (:141-160) from the restored forcing.
struct ChemBoxParams {
  int lo0, lo1, lo2, nx, ny;
  long ncells;
  int N;
  Arr rYo, rHo, To, rYn, rHn, Tn, FY, FH, IR;
  double dt_inv;
};

template <bool AFTER>
__global__ void __launch_bounds__(256) chem_units_kernel(ChemBoxParams P) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < P.ncells; c += stride) {
    const unsigned u = (unsigned)c;
    const int i = P.lo0 + (int)(u % (unsigned)P.nx);
    const int j = P.lo1 + (int)((u / (unsigned)P.nx) % (unsigned)P.ny);
    const int k = P.lo2 + (int)(u / ((unsigned)P.nx * (unsigned)P.ny));
    const long oYo = aoff(P.rYo, i, j, k), oYn = aoff(P.rYn, i, j, k), oF = aoff(P.FY, i, j, k);
    if (!AFTER) {
      for (int n = 0; n < P.N; ++n) {
        P.rYn.p[oYn + n * P.rYn.ns] = P.rYo.p[oYo + n * P.rYo.ns] * 1.0e-3;
        P.FY.p[oF + n * P.FY.ns] *= 1.0e-3;
      }
      P.Tn.p[aoff(P.Tn, i, j, k)] = P.To.p[aoff(P.To, i, j, k)];
      P.rHn.p[aoff(P.rHn, i, j, k)] = P.rHo.p[aoff(P.rHo, i, j, k)] * 10.0;
      P.FH.p[aoff(P.FH, i, j, k)] *= 10.0;
    } else {
      const long oI = aoff(P.IR, i, j, k);
      for (int n = 0; n < P.N; ++n) {
        const double w = P.rYn.p[oYn + n * P.rYn.ns] * 1.0e3;
        const double f = P.FY.p[oF + n * P.FY.ns] * 1.0e3;
        P.rYn.p[oYn + n * P.rYn.ns] = w;
        P.FY.p[oF + n * P.FY.ns] = f;
        if (P.IR.p) P.IR.p[oI + n * P.IR.ns] = -(P.rYo.p[oYo + n * P.rYo.ns] - w) * P.dt_inv - f;
      }
      P.rHn.p[aoff(P.rHn, i, j, k)] *= 0.1;
      P.FH.p[aoff(P.FH, i, j, k)] *= 0.1;
    }
  }
}

// advanceChemistryBAChem's unit conversions on the chemistry BoxArray's copy, in place
// (Source/PeleLMeX_Reactions.cpp:228-238 and :284-291): before the react, rhoY and its forcing
// times 1e-3, rho h and its forcing times 10 (P.rYn, P.rHn, P.FY, P.FH); after it, the state only
// (rhoY times 1e3, rho h times 0.1: the forcing copy stays in CGS). ZERO_FC additionally writes
// FC = 0 (a covered box, which skips the react).
template <int MODE>   // 0 before, 1 after, 2 after with FC = 0
__global__ void __launch_bounds__(256) chem_units_inplace_kernel(ChemBoxParams P, Arr FC) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < P.ncells; c += stride) {
    const unsigned u = (unsigned)c;
    const int i = P.lo0 + (int)(u % (unsigned)P.nx);
    const int j = P.lo1 + (int)((u / (unsigned)P.nx) % (unsigned)P.ny);
    const int k = P.lo2 + (int)(u / ((unsigned)P.nx * (unsigned)P.ny));
    const long oY = aoff(P.rYn, i, j, k);
    if (MODE == 0) {
      const long oF = aoff(P.FY, i, j, k);
      for (int n = 0; n < P.N; ++n) {
        P.rYn.p[oY + n * P.rYn.ns] *= 1.0e-3;
        P.FY.p[oF + n * P.FY.ns] *= 1.0e-3;
      }
      P.rHn.p[aoff(P.rHn, i, j, k)] *= 10.0;
      P.FH.p[aoff(P.FH, i, j, k)] *= 10.0;
    } else {
      for (int n = 0; n < P.N; ++n) P.rYn.p[oY + n * P.rYn.ns] *= 1.0e3;
      P.rHn.p[aoff(P.rHn, i, j, k)] *= 0.1;
      if (MODE == 2) FC.p[aoff(FC, i, j, k)] = 0.0;
    }
  }
}

}  // namespace prk
