// Equation-of-state box kernels around the chemistry advance, MKS Array4 views, one group of G
// lanes per cell (species rows as in the reactor):
//  * setTemperature  (Source/PeleLMeX_Eos.cpp:218-240 -> getTfromHY, Source/PeleLMeX_K.H:700-723):
//    rhoinv = 1 / rho, Y = rhoY * rhoinv, h = rhoH * rhoinv * 1e4, T = HY2T(h, Y) from the stored
//    T - the reactor's own hy2t, so the temperature is the one its right-hand side computes;
//  * floorSpecies    (Source/PeleLMeX_Utils.cpp:649-690, fabMinMax Source/PeleLMeX_K.H:894-907):
//    rhoY_k = max(0, min(DBL_MAX, rhoY_k)) with std::min / std::max semantics, rho = sum_k rhoY_k
//    and h = TY2H(T, Y) = sum_k Y_k h_k(T) in species order, rhoH = h * 1e-4 * rho;
//  * getHeatRelease  (Source/PeleLMeX_Reactions.cpp:477-506, getHGivenT Source/PeleLMeX_K.H:598-615):
//    HR = 0 - sum_k (h_k(T) * 1e-4) * I_R,k in species order.
// h_k(T) = (h_k / RT) (R / W_k) T in erg / g, the reactor's species enthalpy.
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

// Sum of the per-row values over species 0..N-1 in species order, starting from 0.0 (SUB:
// 0.0 - v_0 - v_1 - ...), through the group's rate area in LDS; returned to every lane.
template <int N, int R, int G, bool SUB>
__device__ __forceinline__ double seq_sum(const Smem<N, R>& S, int lane, const double (&v)[Cell<N, R, G>::RPL]) {
  group_sync();  // earlier readers of the area are done
#pragma unroll
  for (int s = 0; s < Cell<N, R, G>::RPL; ++s) {
    const int i = lane + G * s;
    if (i < N) S.q[i] = v[s];
  }
  group_sync();
  double a = 0.0;
  for (int k = 0; k < N; ++k) a = SUB ? a - S.q[k] : a + S.q[k];
  return a;
}

template <int N, int R, int G, int MODE>
__global__ void __launch_bounds__(BLOCK) eos_kernel(EosParams P) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  extern __shared__ __align__(16) unsigned char smem[];
  {
    const uint4* srcb = reinterpret_cast<const uint4*>(P.blob);
    uint4* dst = reinterpret_cast<uint4*>(smem);
    for (int b = threadIdx.x; b < P.blob_bytes / 16; b += blockDim.x) dst[b] = srcb[b];
  }
  __syncthreads();
  const BlobLayout BL = blob_layout(N, R, P.neff, P.ninc);
  Smem<N, R> S;
  const int lane = threadIdx.x % G;
  const int grp = threadIdx.x / G;
  setup_views<N, R, G>(S, smem, BL, grp, nullptr);
  const long stride = (long)gridDim.x * (blockDim.x / G);
  for (long c = (long)blockIdx.x * (blockDim.x / G) + grp; c < P.ncells; c += stride) {  // whole groups
    const unsigned u = (unsigned)c;
    const int i = P.lo0 + (int)(u % (unsigned)P.nx);
    const int j = P.lo1 + (int)((u / (unsigned)P.nx) % (unsigned)P.ny);
    const int k = P.lo2 + (int)(u / ((unsigned)P.nx * (unsigned)P.ny));
    if constexpr (MODE == EOS_SET_T) {
      const double rhoinv = 1.0 / P.rho.p[aoff(P.rho, i, j, k)];
      const long offY = aoff(P.rY, i, j, k), oT = aoff(P.T, i, j, k);
      double Y[RPL];
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        const int ii = lane + G * s;
        Y[s] = (ii < N) ? P.rY.p[offY + (long)ii * P.rY.ns] * rhoinv : 0.0;
      }
      const double T = hy2t<N, R, G>(S, lane, Y, (P.rH.p[aoff(P.rH, i, j, k)] * rhoinv) * 1.0e4, P.T.p[oT]);
      if (lane == 0) P.T.p[oT] = T;
    } else if constexpr (MODE == EOS_FLOOR) {
      const long offY = aoff(P.rY, i, j, k);
      double v[RPL];
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        const int ii = lane + G * s;
        v[s] = 0.0;
        if (ii < N) {
          double x = P.rY.p[offY + (long)ii * P.rY.ns];
          x = (x < DBL_MAX) ? x : DBL_MAX;  // min(DBL_MAX, x): NaN becomes DBL_MAX
          x = (0.0 < x) ? x : 0.0;          // max(0, x): negatives and -0 become +0
          v[s] = x;
          P.rY.p[offY + (long)ii * P.rY.ns] = x;
        }
      }
      const double rho = seq_sum<N, R, G, false>(S, lane, v);
      const double rhoinv = 1.0 / rho;
      const double T = P.T.p[aoff(P.T, i, j, k)];
      const double invT = 1.0 / T;
      double t[RPL];
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        const int ii = lane + G * s;
        t[s] = 0.0;
        if (ii < N) {
          double cpR, hRT;
          nasa_hcp(S.sp[ii], T, invT, cpR, hRT);
          t[s] = ((v[s] * rhoinv * hRT) * S.sp[ii].RW) * T;
        }
      }
      const double h = seq_sum<N, R, G, false>(S, lane, t);
      if (lane == 0) {
        P.rho.p[aoff(P.rho, i, j, k)] = rho;
        P.rH.p[aoff(P.rH, i, j, k)] = (h * 1.0e-4) * rho;
      }
    } else {
      const double T = P.T.p[aoff(P.T, i, j, k)];
      const double invT = 1.0 / T;
      const long offR = aoff(P.IR, i, j, k);
      double p[RPL];
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        const int ii = lane + G * s;
        p[s] = 0.0;
        if (ii < N) {
          double cpR, hRT;
          nasa_hcp(S.sp[ii], T, invT, cpR, hRT);
          p[s] = (((hRT * S.sp[ii].RW) * T) * 0.0001) * P.IR.p[offR + (long)ii * P.IR.ns];
        }
      }
      const double hr = seq_sum<N, R, G, true>(S, lane, p);
      if (lane == 0) P.HR.p[aoff(P.HR, i, j, k)] = hr;
    }
    group_sync();  // the next cell's stores into the rate area wait for this cell's readers
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
