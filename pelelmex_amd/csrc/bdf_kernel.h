// Persistent batched BDF reactor kernel for gfx950 (FP64 throughout).
//
// Replaces the integration inside PelePhysics ReactorBase::react (call site
// Source/PeleLMeX_Reactions.cpp:101-108): every unmasked cell of a box is advanced over
// dt_react by its own variable-order BDF with CVODE step / order / Newton / Jacobian-reuse
// semantics (the CPU path's per-cell integration, ncells = 1 in Source/PeleLMeX_Setup.cpp:147),
// on the constant-pressure (rhoY, T) system with rho h(t) = rho h0 + t Q_h.
//
// Mapping onto CDNA4:
//  * a cell is owned by a group of G lanes (G = 8..64, so 64/G cells per wavefront); lane l
//    owns state components / matrix rows l, l+G, ... All vectors of the BDF (Nordsieck
//    history, error weights, corrections) and the rows of the LU factors stay in VGPRs.
//  * the right-hand side and the analytic Jacobian are evaluated cooperatively: species
//    thermo lane-parallel over species, Arrhenius / third-body / falloff rates lane-parallel
//    over reactions with the mechanism tables staged in LDS, production rates gathered
//    species-parallel through per-species incidence lists (deterministic, no atomics), the
//    Jacobian assembled row-by-row in a per-cell LDS matrix that is also the saved Jacobian.
//  * the dense LU of M = I - gamma J uses partial pivoting without moving data: the pivot row
//    of each step is broadcast across the group and each lane eliminates its own rows.
//  * the integrator is one state machine per group with a single RHS call site per iteration,
//    so the cells of a wavefront evaluate their right-hand sides together even when they are
//    in different phases (initial step, Newton, restart); a finished group pulls the next cell
//    from an atomic work queue, so stiff cells do not hold back their neighbours.
#ifndef PR_BDF_KERNEL_H
#define PR_BDF_KERNEL_H
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "device_mech.h"

namespace prk {

constexpr int QMAXC = 5;
constexpr int BLOCK = 256;

struct Arr {
  double* p;
  long js, ks, ns;
  int bx, by, bz;
};
struct IArr {
  const int* p;
  long js, ks;
  int bx, by, bz;
};

struct KParams {
  int lo0, lo1, lo2, nx, ny;
  long ncells;
  Arr rY, rYsrc, T, rH, rHsrc, FC;
  IArr mask;
  int has_mask;
  double dt;
  double rtol, atol;
  const double* atolv;  // n entries or null
  int max_order, max_steps;
  const unsigned char* blob;
  int blob_bytes, neff, ninc;
  unsigned int* queue;           // work counter, zero before launch
  unsigned long long* stats;     // cells, nst, nfe, nje, nsetups, nni, netf, ncfn
  int* status;                   // per cell (linear box index), nullable
  unsigned int* nfail;
};

enum { ST_OK = 0, ST_TOO_MUCH_WORK = -1, ST_ERR_FAILURE = -3, ST_CONV_FAILURE = -4, ST_GUARD = -99 };

__device__ __forceinline__ void group_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

template <int G>
__device__ __forceinline__ double gsum(double v) {
#pragma unroll
  for (int m = G / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, G);
  return v;
}

template <int G>
__device__ __forceinline__ double gbcast(double v, int src) {
  return __shfl(v, src, G);
}

template <int G>
__device__ __forceinline__ void gargmax(double& v, int& idx) {
#pragma unroll
  for (int m = G / 2; m >= 1; m >>= 1) {
    double ov = __shfl_xor(v, m, G);
    int oi = __shfl_xor(idx, m, G);
    bool take = (ov > v) || (ov == v && oi < idx);
    v = take ? ov : v;
    idx = take ? oi : idx;
  }
}

// Select element `s` (runtime, < RPL) of a register array without dynamic indexing.
template <int RPL>
__device__ __forceinline__ double rsel(const double (&a)[RPL], int s) {
  double v = a[0];
#pragma unroll
  for (int t = 1; t < RPL; ++t) v = (s == t) ? a[t] : v;
  return v;
}

__device__ __forceinline__ void nasa_eval(const SpRec& sp, double T, double lnT, double& cpR,
                                          double& hRT, double& sR, double& dcpR) {
  const double* a = (T < sp.nasa[0]) ? sp.nasa + 8 : sp.nasa + 1;
  double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  cpR = a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4;
  hRT = a[0] + a[1] * T / 2 + a[2] * T2 / 3 + a[3] * T3 / 4 + a[4] * T4 / 5 + a[5] / T;
  sR = a[0] * lnT + a[1] * T + a[2] * T2 / 2 + a[3] * T3 / 3 + a[4] * T4 / 4 + a[6];
  dcpR = a[1] + 2 * a[2] * T + 3 * a[3] * T2 + 4 * a[4] * T3;
}

__device__ __forceinline__ void nasa_hcp(const SpRec& sp, double T, double& cpR, double& hRT) {
  const double* a = (T < sp.nasa[0]) ? sp.nasa + 8 : sp.nasa + 1;
  double T2 = T * T, T3 = T2 * T, T4 = T3 * T;
  cpR = a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4;
  hRT = a[0] + a[1] * T / 2 + a[2] * T2 / 3 + a[3] * T3 / 4 + a[4] * T4 / 5 + a[5] / T;
}

struct RxOut {
  double kf, kr, fwd, rev, dqdM, dqdT;
};

// One reaction at (C, T): effective rate coefficients, forward/reverse rates and, when JAC,
// d q / d[M] and d q / dT at constant concentrations.
template <bool JAC>
__device__ __forceinline__ void rx_eval(const RxRec& rx, const double* sC, const double* sG,
                                        const double* sH, double Ctot, double T, double lnT,
                                        double invT, double lnRTP, const int* effsp,
                                        const double* effval, RxOut& o) {
  const int info = rx.info;
  const int type = info & 15, rev = (info >> 4) & 15, nre = (info >> 8) & 15, npr = (info >> 12) & 15;
  const int ntroe = (info >> 16) & 15;
  double kinf = exp(rx.lnA + rx.beta * lnT - rx.Ta * invT);
  double dlnkinf = (rx.beta + rx.Ta * invT) * invT;
  double M = 0.0;
  if (type != PR_RX_ELEM) {
    M = rx.effd * Ctot;
    int eb = rx.eff & 0xFFFF, ee = (rx.eff >> 16) & 0xFFFF;
    for (int e = eb; e < ee; ++e) M += effval[e] * sC[effsp[e]];
  }
  int r0 = rx.reac & 0xFF, r1 = (rx.reac >> 8) & 0xFF, r2 = (rx.reac >> 16) & 0xFF;
  int p0 = rx.prod & 0xFF, p1 = (rx.prod >> 8) & 0xFF, p2 = (rx.prod >> 16) & 0xFF;
  double pr = sC[r0];
  if (nre > 1) pr *= sC[r1];
  if (nre > 2) pr *= sC[r2];
  double pp = sC[p0];
  if (npr > 1) pp *= sC[p1];
  if (npr > 2) pp *= sC[p2];
  double invKc = 0.0, dH = 0.0;
  const int sumnu = npr - nre;
  if (rev) {
    double dG = sG[p0] - sG[r0];
    if (npr > 1) dG += sG[p1];
    if (npr > 2) dG += sG[p2];
    if (nre > 1) dG -= sG[r1];
    if (nre > 2) dG -= sG[r2];
    invKc = exp(dG + sumnu * lnRTP);
    if (JAC) {
      dH = sH[p0] - sH[r0];
      if (npr > 1) dH += sH[p1];
      if (npr > 2) dH += sH[p2];
      if (nre > 1) dH -= sH[r1];
      if (nre > 2) dH -= sH[r2];
    }
  }
  const double dlnKc = (dH - sumnu) * invT;
  if (type <= PR_RX_THIRDBODY) {
    double kr0 = kinf * invKc;
    double fwd0 = kinf * pr, rev0 = kr0 * pp;
    if (type == PR_RX_ELEM) {
      o.kf = kinf; o.kr = kr0; o.fwd = fwd0; o.rev = rev0;
      if (JAC) { o.dqdM = 0.0; o.dqdT = fwd0 * dlnkinf - rev0 * (dlnkinf - dlnKc); }
    } else {
      o.kf = M * kinf; o.kr = M * kr0; o.fwd = M * fwd0; o.rev = M * rev0;
      if (JAC) { o.dqdM = fwd0 - rev0; o.dqdT = M * (fwd0 * dlnkinf - rev0 * (dlnkinf - dlnKc)); }
    }
    return;
  }
  double k0 = exp(rx.lnA0 + rx.beta0 * lnT - rx.Ta0 * invT);
  double Pr = k0 * M / kinf;
  Pr = Pr < 1e-300 ? 1e-300 : Pr;
  double F = 1.0, dlogF_dx = 0.0, dlogF_dT = 0.0;
  const double LN10 = 2.302585092994045684;
  double dlnPr = 0.0;
  if (JAC) dlnPr = (rx.beta0 + rx.Ta0 * invT) * invT - dlnkinf;
  if (type == PR_RX_TROE) {
    double e3 = exp(-T / rx.troe[1]), e1 = exp(-T / rx.troe[2]);
    double Fc = (1.0 - rx.troe[0]) * e3 + rx.troe[0] * e1;
    double dFc = 0.0;
    if (JAC) dFc = -(1.0 - rx.troe[0]) / rx.troe[1] * e3 - rx.troe[0] / rx.troe[2] * e1;
    if (ntroe == 4) {
      double e2 = exp(-rx.troe[3] * invT);
      Fc += e2;
      if (JAC) dFc += rx.troe[3] * invT * invT * e2;
    }
    Fc = Fc < 1e-300 ? 1e-300 : Fc;
    double logFc = log10(Fc);
    double c = -0.4 - 0.67 * logFc, nn = 0.75 - 1.27 * logFc;
    double x = log10(Pr);
    double xc = x + c, D = nn - 0.14 * xc, f1 = xc / D;
    double den = 1.0 + f1 * f1;
    F = exp(logFc / den * LN10);
    if (JAC) {
      dlogF_dx = -logFc * 2.0 * f1 * (nn / (D * D)) / (den * den);
      double dlogFc = dFc / (Fc * LN10);
      double dx = dlnPr / LN10, dc = -0.67 * dlogFc, dnn = -1.27 * dlogFc;
      double dxc = dx + dc, dD = dnn - 0.14 * dxc;
      double df1 = (dxc * D - xc * dD) / (D * D);
      dlogF_dT = (dlogFc * den - logFc * 2.0 * f1 * df1) / (den * den);
    }
  }
  double kf = kinf * Pr / (1.0 + Pr) * F;
  double kr = kf * invKc;
  o.kf = kf;
  o.kr = kr;
  o.fwd = kf * pr;
  o.rev = kr * pp;
  if (JAC) {
    double dlnkf = dlnkinf + dlnPr / (1.0 + Pr) + LN10 * dlogF_dT;
    o.dqdM = k0 * F * (1.0 / ((1.0 + Pr) * (1.0 + Pr)) + dlogF_dx / (1.0 + Pr)) * (pr - pp * invKc);
    o.dqdT = o.fwd * dlnkf - o.rev * (dlnkf - dlnKc);
  }
}

// Views of the LDS-staged mechanism and one group's scratch.
template <int N, int R>
struct Smem {
  const SpRec* sp;
  const RxRec* rx;
  const int* effsp;
  const double* effval;
  const int* incs;
  const int* inc;
  double* Js;    // (N+1)^2, row-major: saved Jacobian
  double* sC;    // N concentrations
  double* sG;    // N g/RT
  double* sH;    // N h/RT
  double* sHm;   // N mass-specific h
  double* sCpm;  // N mass-specific cp
  double* rec;   // 5 R: kf, kr, dqdM, dqdT, q
  int* perm;     // N+1 pivot rows
  static constexpr int n = N + 1;
  static constexpr int scratch_doubles = n * n + 5 * N + 5 * R + (n + 1) / 2 + 1;
};

template <int N, int R, int G>
struct Cell {
  static constexpr int n = N + 1;
  static constexpr int RPL = (n + G - 1) / G;
  static constexpr int RXL = (R + G - 1) / G;
};

// Reactor right-hand side f(t, y) on the group; T from (h(t), Y) by Newton, started at y_T.
template <int N, int R, int G>
__device__ __forceinline__ void rhs_eval(const Smem<N, R>& S, int lane, double t, const double (&y)[Cell<N, R, G>::RPL],
                         const double (&src)[Cell<N, R, G>::RPL], double rhoh0, double rhohsrc,
                         double (&f)[Cell<N, R, G>::RPL]) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  constexpr int n = N + 1;
  double part = 0.0;
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    int i = lane + G * s;
    if (i < N) part += y[s];
  }
  const double rho = gsum<G>(part);
  const double hmix = (rhoh0 + t * rhohsrc) / rho;
  double T = gbcast<G>(y[N / G], N % G);
  const double RU = PR_RU;
  for (int it = 0; it < 20; ++it) {
    double ph = 0.0, pc = 0.0;
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      int i = lane + G * s;
      if (i < N) {
        double cpR, hRT;
        nasa_hcp(S.sp[i], T, cpR, hRT);
        double Y = y[s] / rho, RW = RU * S.sp[i].invW;
        ph += Y * hRT * RW * T;
        pc += Y * cpR * RW;
      }
    }
    double hm = gsum<G>(ph), cpm = gsum<G>(pc);
    double dT = (hmix - hm) / cpm;
    T += dT;
    if (fabs(dT) <= 1e-10 * T) break;
  }
  const double lnT = log(T), invT = 1.0 / T, lnRTP = log(RU * T / PR_PATM);
  double hk[RPL], cpk[RPL], ctot = 0.0;
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    int i = lane + G * s;
    hk[s] = 0.0;
    cpk[s] = 0.0;
    if (i < N) {
      double cpR, hRT, sR, dcp;
      nasa_eval(S.sp[i], T, lnT, cpR, hRT, sR, dcp);
      double RW = RU * S.sp[i].invW;
      double C = y[s] * S.sp[i].invW;
      S.sC[i] = C;
      S.sG[i] = hRT - sR;
      ctot += C;
      hk[s] = hRT * RW * T;
      cpk[s] = cpR * RW;
    }
  }
  const double Ctot = gsum<G>(ctot);
  group_sync();
#pragma unroll 1
  for (int m = 0; m < Cell<N, R, G>::RXL; ++m) {
    int r = lane + G * m;
    if (r < R) {
      RxOut o;
      rx_eval<false>(S.rx[r], S.sC, S.sG, S.sH, Ctot, T, lnT, invT, lnRTP, S.effsp, S.effval, o);
      S.rec[4 * R + r] = o.fwd - o.rev;
    }
  }
  group_sync();
  double pn = 0.0, pcp = 0.0;
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    int i = lane + G * s;
    f[s] = 0.0;
    if (i < N) {
      double w = 0.0;
      for (int e = S.incs[i]; e < S.incs[i + 1]; ++e) {
        int rn = S.inc[e];
        w += (double)(rn >> 16) * S.rec[4 * R + (rn & 0xFFFF)];
      }
      f[s] = w * S.sp[i].W + src[s];
      pn += hk[s] * f[s];
      pcp += y[s] * cpk[s];
    }
  }
  const double num = rhohsrc - gsum<G>(pn);
  const double rcp = gsum<G>(pcp);
  const double fT = num / rcp;
  if (lane == N % G) f[N / G] = fT;
  (void)n;
}

// Analytic Jacobian of the right-hand side with T = y_T, written into S.Js (row-major).
template <int N, int R, int G>
__device__ __forceinline__ void jac_eval(const Smem<N, R>& S, int lane, const double (&y)[Cell<N, R, G>::RPL],
                         const double (&src)[Cell<N, R, G>::RPL], double rhohsrc) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  constexpr int n = N + 1;
  const double RU = PR_RU;
  const double T = gbcast<G>(y[N / G], N % G);
  const double lnT = log(T), invT = 1.0 / T, lnRTP = log(RU * T / PR_PATM);
  double ctot = 0.0, cpk[RPL], dcpk[RPL];
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    int i = lane + G * s;
    cpk[s] = 0.0;
    dcpk[s] = 0.0;
    if (i < N) {
      double cpR, hRT, sR, dcp;
      nasa_eval(S.sp[i], T, lnT, cpR, hRT, sR, dcp);
      double RW = RU * S.sp[i].invW;
      double C = y[s] * S.sp[i].invW;
      S.sC[i] = C;
      S.sG[i] = hRT - sR;
      S.sH[i] = hRT;
      S.sHm[i] = hRT * RW * T;
      S.sCpm[i] = cpR * RW;
      cpk[s] = cpR * RW;
      dcpk[s] = dcp * RW;
      ctot += C;
    }
  }
  const double Ctot = gsum<G>(ctot);
  group_sync();
#pragma unroll 1
  for (int m = 0; m < Cell<N, R, G>::RXL; ++m) {
    int r = lane + G * m;
    if (r < R) {
      RxOut o;
      rx_eval<true>(S.rx[r], S.sC, S.sG, S.sH, Ctot, T, lnT, invT, lnRTP, S.effsp, S.effval, o);
      double* rc = S.rec + 4 * r;
      rc[0] = o.kf;
      rc[1] = o.kr;
      rc[2] = o.dqdM;
      rc[3] = o.dqdT;
      S.rec[4 * R + r] = o.fwd - o.rev;
    }
  }
  group_sync();
  double pn = 0.0, prcp = 0.0, psd = 0.0, pcf = 0.0;
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    int k = lane + G * s;
    if (k < N) {
      double* row = S.Js + k * n;
      for (int j = 0; j < n; ++j) row[j] = 0.0;
      double Sd = 0.0, JT = 0.0, w = 0.0;
      for (int e = S.incs[k]; e < S.incs[k + 1]; ++e) {
        const int rn = S.inc[e];
        const int r = rn & 0xFFFF;
        const double nu = (double)(rn >> 16);
        const RxRec& rx = S.rx[r];
        const double* rc = S.rec + 4 * r;
        const int info = rx.info, type = info & 15, nre = (info >> 8) & 15, npr = (info >> 12) & 15;
        w += nu * S.rec[4 * R + r];
        int a0 = rx.reac & 0xFF, a1 = (rx.reac >> 8) & 0xFF, a2 = (rx.reac >> 16) & 0xFF;
        double c0 = S.sC[a0], c1 = nre > 1 ? S.sC[a1] : 1.0, c2 = nre > 2 ? S.sC[a2] : 1.0;
        double kf = nu * rc[0];
        row[a0] += kf * c1 * c2;
        if (nre > 1) row[a1] += kf * c0 * c2;
        if (nre > 2) row[a2] += kf * c0 * c1;
        int b0 = rx.prod & 0xFF, b1 = (rx.prod >> 8) & 0xFF, b2 = (rx.prod >> 16) & 0xFF;
        double d0 = S.sC[b0], d1 = npr > 1 ? S.sC[b1] : 1.0, d2 = npr > 2 ? S.sC[b2] : 1.0;
        double kr = nu * rc[1];
        row[b0] -= kr * d1 * d2;
        if (npr > 1) row[b1] -= kr * d0 * d2;
        if (npr > 2) row[b2] -= kr * d0 * d1;
        if (type != PR_RX_ELEM) {
          double dm = nu * rc[2];
          Sd += dm * rx.effd;
          int eb = rx.eff & 0xFFFF, ee = (rx.eff >> 16) & 0xFFFF;
          for (int q = eb; q < ee; ++q) row[S.effsp[q]] += dm * S.effval[q];
        }
        JT += nu * rc[3];
      }
      const double Wk = S.sp[k].W;
      for (int j = 0; j < N; ++j) row[j] = (row[j] + Sd) * Wk * S.sp[j].invW;
      row[N] = JT * Wk;
      double fk = w * Wk + src[s];
      pn += S.sHm[k] * fk;
      prcp += y[s] * cpk[s];
      psd += y[s] * dcpk[s];
      pcf += cpk[s] * fk;
    }
  }
  const double num = rhohsrc - gsum<G>(pn);
  const double rcp = gsum<G>(prcp);
  const double sdcp = gsum<G>(psd);
  const double scf = gsum<G>(pcf);
  const double fT = num / rcp;
  group_sync();
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    int j = lane + G * s;
    if (j < n) {
      double cs = 0.0;
      for (int k = 0; k < N; ++k) cs += S.sHm[k] * S.Js[k * n + j];
      S.Js[N * n + j] = (j < N) ? (-cs - fT * S.sCpm[j]) / rcp : (-cs - scf - fT * sdcp) / rcp;
    }
  }
  group_sync();
}

// LU of M = I - gamma * Js with partial pivoting (first maximal |pivot| in row order).
// Returns true when a zero pivot is met. Rows stay in the owning lanes.
template <int N, int R, int G>
__device__ __forceinline__ bool lu_factor(const Smem<N, R>& S, int lane, double gamma,
                          double (&lu)[Cell<N, R, G>::RPL][N + 1], int (&pst)[Cell<N, R, G>::RPL],
                          double (&rinv)[Cell<N, R, G>::RPL]) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  constexpr int n = N + 1;
  constexpr int UNPIV = 1 << 20;
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    int i = lane + G * s;
    pst[s] = (i < n) ? UNPIV : -1;
    rinv[s] = 0.0;
#pragma unroll
    for (int j = 0; j < n; ++j) lu[s][j] = (i < n) ? ((i == j ? 1.0 : 0.0) - gamma * S.Js[i * n + j]) : 0.0;
  }
  bool singular = false;
#pragma unroll
  for (int k = 0; k < n; ++k) {
    double best = -1.0;
    int bi = 1 << 29;
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      if (pst[s] == UNPIV) {
        double v = fabs(lu[s][k]);
        v = (v == v) ? v : -2.0;
        if (v > best) {
          best = v;
          bi = lane + G * s;
        }
      }
    }
    gargmax<G>(best, bi);
    if (!(best > 0.0)) {
      singular = true;
      break;
    }
    const int owner = bi % G, ps = bi / G;
    if (lane == owner) {
#pragma unroll
      for (int s = 0; s < RPL; ++s)
        if (s == ps) pst[s] = k;
      S.perm[k] = bi;
    }
    double pk = 0.0;
#pragma unroll
    for (int s = 0; s < RPL; ++s) pk = (s == ps) ? lu[s][k] : pk;
    const double piv = gbcast<G>(pk, owner);
    const double inv = 1.0 / piv;
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      if (lane == owner && s == ps) rinv[s] = inv;
      if (pst[s] > k) lu[s][k] *= inv;
    }
#pragma unroll
    for (int j = k + 1; j < n; ++j) {
      double pj = 0.0;
#pragma unroll
      for (int s = 0; s < RPL; ++s) pj = (s == ps) ? lu[s][j] : pj;
      pj = gbcast<G>(pj, owner);
#pragma unroll
      for (int s = 0; s < RPL; ++s)
        if (pst[s] > k) lu[s][j] -= lu[s][k] * pj;
    }
  }
  group_sync();
  return singular;
}

// Solve M x = b with the factors from lu_factor; b is distributed by row, x by component.
template <int N, int R, int G>
__device__ __forceinline__ void lu_solve(const Smem<N, R>& S, int lane, const double (&lu)[Cell<N, R, G>::RPL][N + 1],
                         const int (&pst)[Cell<N, R, G>::RPL], const double (&rinv)[Cell<N, R, G>::RPL],
                         double (&b)[Cell<N, R, G>::RPL], double (&x)[Cell<N, R, G>::RPL]) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  constexpr int n = N + 1;
#pragma unroll
  for (int k = 0; k < n; ++k) {
    const int p = S.perm[k];
    const int owner = p % G, ps = p / G;
    double v = 0.0;
#pragma unroll
    for (int s = 0; s < RPL; ++s) v = (s == ps) ? b[s] : v;
    const double yk = gbcast<G>(v, owner);
#pragma unroll
    for (int s = 0; s < RPL; ++s)
      if (pst[s] > k) b[s] -= lu[s][k] * yk;
  }
#pragma unroll
  for (int s = 0; s < RPL; ++s) x[s] = 0.0;
#pragma unroll
  for (int k = n - 1; k >= 0; --k) {
    const int p = S.perm[k];
    const int owner = p % G, ps = p / G;
    double v = 0.0;
#pragma unroll
    for (int s = 0; s < RPL; ++s) v = (s == ps) ? b[s] * rinv[s] : v;
    const double xk = gbcast<G>(v, owner);
#pragma unroll
    for (int s = 0; s < RPL; ++s)
      if (pst[s] >= 0 && pst[s] < k) b[s] -= lu[s][k] * xk;
    if (lane == k % G) x[k / G] = xk;
  }
}

template <int N, int R, int G>
__device__ __forceinline__ double wrms(int lane, const double (&v)[Cell<N, R, G>::RPL],
                                       const double (&w)[Cell<N, R, G>::RPL]) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  double p = 0.0;
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    int i = lane + G * s;
    if (i < N + 1) {
      double t = v[s] * w[s];
      p += t * t;
    }
  }
  return sqrt(gsum<G>(p) / (N + 1));
}

}  // namespace prk
#endif
