// Persistent batched BDF reactor kernel for gfx950 (FP64 throughout): device building blocks.
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
//    Jacobian assembled row-by-row (each lane owns its rows) in the group's global scratch
//    slot, which also keeps it as the saved Jacobian between refreshes (L2-resident; it is
//    touched only on Jacobian / LU refreshes, not in the Newton iterations).
//  * the dense LU of M = I - gamma J uses partial pivoting without moving data: the pivot row
//    of each step is broadcast across the group and each lane eliminates its own rows.
//
// Arithmetic: every expression is an explicit IEEE-754 tree (the library is built without
// floating-point contraction; fma only where written), exp / log are this file's own
// routines, and every sum over components follows one canonical pairwise tree (gpsum).
// Results therefore do not depend on G or on the launch shape, and they are bit-identical to
// the CPU oracle, which evaluates the same trees.
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

// One box of a react call: its index range, Array4 views, mask, per-cell status and dt_react.
struct BoxDesc {
  int lo0, lo1, lo2, nx, ny;
  long ncells;
  Arr rY, rYsrc, T, rH, rHsrc, FC;
  IArr mask;
  int has_mask;
  int* status;  // per cell (linear box index), nullable
  double dt;
  const double* dtv;  // per-cell dt_react (linear box index), nullable: dt for every cell
};

struct KParams {
  BoxDesc one;           // host side: the box of a single-box call (staged into boxes[0])
  const BoxDesc* boxes;  // device table of the launch's boxes: queue entry q is cell
  const long* box_start; // q - box_start[b] of box b, box_start[b] <= q < box_start[b + 1]
  int nbox;
  long ncells;           // queue entries (cells of all boxes)
  double rtol, atol;
  const double* atolv;  // n entries or null
  int max_order, max_steps;
  const unsigned char* blob;
  int blob_bytes, neff, ninc;
  unsigned int* queue;           // work counter, zero before launch
  unsigned long long* stats;     // cells, nst, nfe, nje, nsetups, nni, netf, ncfn
  unsigned int* nfail;
  double* gscr;                  // per-group global scratch slots (saved Jacobian, Jacobian records)
  int clean;                     // ode.clean_init_massfrac: integrate from max(rhoY, 0), add the rest back
  const int* cell_list;          // when set: queue entry q integrates box cell cell_list[q]
  const unsigned int* ncells_dev;  // when set: min(*ncells_dev, ncells) entries are integrated
};

enum { ST_OK = 0, ST_TOO_MUCH_WORK = -1, ST_ERR_FAILURE = -3, ST_CONV_FAILURE = -4, ST_GUARD = -99 };

// ---------------------------------------------------------------- math (own exp / log)
// exp: x = k ln2 + r (Cody-Waite, ln2 in two parts), degree-13 Taylor polynomial of e^r, 2^k.
// log: x = 2^e m, m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1), log m = 2 s (1 + s^2/3 + ... + s^20/21).
#define PRK_LN2_HI 6.93147180559945286227e-01
#define PRK_LN2_LO 2.31904681384629955842e-17
#define PRK_LOG2E 1.44269504088896338700e+00
#define PRK_SQRT_HALF 7.07106781186547572737e-01
#define PRK_LN10 2.30258509299404590109e+00
#define PRK_INV_LN10 4.34294481903251816668e-01

// A constant materialized in scalar registers where it is used (an empty asm that may change it):
// polynomial steps then issue as one v_fma_f64 with a scalar addend, instead of a vector copy of
// a constant hoisted into vector registers followed by v_fmac_f64.
__device__ __forceinline__ double kc(double c) {
  asm volatile("" : "+s"(c));
  return c;
}

__device__ __forceinline__ double pr_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return __builtin_huge_val();
  if (x < -745.1332191019412) return 0.0;
  const double k = rint(x * PRK_LOG2E);
  double r = fma(-k, PRK_LN2_HI, x);
  r = fma(-k, PRK_LN2_LO, r);
  double p = 1.0 / 6227020800.0;
  p = fma(p, r, kc(1.0 / 479001600.0));
  p = fma(p, r, kc(1.0 / 39916800.0));
  p = fma(p, r, kc(1.0 / 3628800.0));
  p = fma(p, r, kc(1.0 / 362880.0));
  p = fma(p, r, kc(1.0 / 40320.0));
  p = fma(p, r, kc(1.0 / 5040.0));
  p = fma(p, r, kc(1.0 / 720.0));
  p = fma(p, r, kc(1.0 / 120.0));
  p = fma(p, r, kc(1.0 / 24.0));
  p = fma(p, r, kc(1.0 / 6.0));
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}

__device__ __forceinline__ double pr_log(double x) {
  if (!(x > 0.0)) return (x == 0.0) ? -__builtin_huge_val() : __builtin_nan("");
  if (x == __builtin_huge_val()) return x;
  int e;
  double m = frexp(x, &e);
  if (m < PRK_SQRT_HALF) {
    m = m * 2.0;
    e = e - 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double p = 1.0 / 21.0;
  p = fma(p, z, kc(1.0 / 19.0));
  p = fma(p, z, kc(1.0 / 17.0));
  p = fma(p, z, kc(1.0 / 15.0));
  p = fma(p, z, kc(1.0 / 13.0));
  p = fma(p, z, kc(1.0 / 11.0));
  p = fma(p, z, kc(1.0 / 9.0));
  p = fma(p, z, kc(1.0 / 7.0));
  p = fma(p, z, kc(1.0 / 5.0));
  p = fma(p, z, kc(1.0 / 3.0));
  const double t = z * p;
  const double s2 = s + s;
  const double logm = fma(s2, t, s2);
  const double de = (double)e;
  return fma(de, PRK_LN2_HI, fma(de, PRK_LN2_LO, logm));
}

__device__ __forceinline__ double pr_log10(double x) { return pr_log(x) * PRK_INV_LN10; }
__device__ __forceinline__ double pr_pow(double x, double y) { return pr_exp(y * pr_log(x)); }

// v through an empty volatile asm: expressions of it are computed where they are used, not
// hoisted out of the integrator loop into registers held for the whole kernel.
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

// ---------------------------------------------------------------- group primitives
// A group of lanes lives inside one wavefront, whose LDS instructions complete in program order:
// keeping the compiler from moving memory operations across this point is all the ordering the
// group's LDS traffic needs (no wait for outstanding LDS writes).
__device__ __forceinline__ void group_sync() {
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

// One cell per wavefront (G == 64): cross-lane moves stay on the VALU. DPP permutes within rows
// of 16 lanes (quad_perm [1,0,3,2] = lane ^ 1, quad_perm [2,3,0,1] = lane ^ 2, row_half_mirror,
// row_mirror); readlane moves a lane's value into scalar registers, so reductions and
// broadcasts return wave-uniform values. Every caller runs with all 64 lanes active.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// value of lane l (l wave-uniform)
__device__ __forceinline__ double rlane_d(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// value of lane ^ 16 (v_permlane16_swap exchanges rows 0 <-> 1 and 2 <-> 3)
__device__ __forceinline__ double xor16_d(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  const bool up = (threadIdx.x & 16) != 0;
  return __hiloint2double(up ? b[0] : b[1], up ? a[0] : a[1]);
}

// value of the lane whose byte address (4 x lane) is addr, through the LDS crossbar
__device__ __forceinline__ double bperm_d(double v, int addr) {
  const int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(v));
  const int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(v));
  return __hiloint2double(hi, lo);
}

// Two cells per wavefront (G == 32): lane l of the group is wave lane l (lanes 0..31) or
// 32 + l (lanes 32..63). hsel picks this lane's half of a pair of wave-uniform values.
__device__ __forceinline__ bool upper_half() { return (threadIdx.x & 32) != 0; }
__device__ __forceinline__ double hsel(double lo, double hi) { return upper_half() ? hi : lo; }

// max of two values that are never NaN, as one v_max_f64: fmax would first canonicalize an
// operand the compiler cannot prove quiet (one more instruction per step of every reduction)
__device__ __forceinline__ double max_nn(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// butterfly over the 32 lanes of each half: DPP within rows, permlane16 swap across rows
// (MAX: the values must not be NaN)
template <bool MAX>
__device__ __forceinline__ double half_reduce(double v) {
  v = MAX ? max_nn(v, dpp_d<0xB1>(v)) : v + dpp_d<0xB1>(v);
  v = MAX ? max_nn(v, dpp_d<0x4E>(v)) : v + dpp_d<0x4E>(v);
  v = MAX ? max_nn(v, dpp_d<0x141>(v)) : v + dpp_d<0x141>(v);
  v = MAX ? max_nn(v, dpp_d<0x140>(v)) : v + dpp_d<0x140>(v);
  return MAX ? max_nn(v, xor16_d(v)) : v + xor16_d(v);
}

// wave-uniform copy of a value that is identical in all lanes
template <int G>
__device__ __forceinline__ int uni(int x) {
  if constexpr (G == 64) return __builtin_amdgcn_readfirstlane(x);
  else return x;
}

template <int G>
__device__ __forceinline__ double gbcast(double v, int src) {
  if constexpr (G == 64) {
    return rlane_d(v, __builtin_amdgcn_readfirstlane(src));
  } else if constexpr (G == 32) {
    const int s0 = __builtin_amdgcn_readlane(src, 0), s1 = __builtin_amdgcn_readlane(src, 32) + 32;
    return hsel(rlane_d(v, s0), rlane_d(v, s1));
  } else {
    return __shfl(v, src, G);
  }
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

// maximum over the group of values that are never NaN
template <int G>
__device__ __forceinline__ double gmax(double v) {
  if constexpr (G == 64) {
    v = max_nn(v, dpp_d<0xB1>(v));
    v = max_nn(v, dpp_d<0x4E>(v));
    v = max_nn(v, dpp_d<0x141>(v));  // quad-uniform values: the mirrored quad is lane ^ 4
    v = max_nn(v, dpp_d<0x140>(v));  // 8-uniform values: the mirrored half row is lane ^ 8
    return fmax(fmax(rlane_d(v, 0), rlane_d(v, 16)), fmax(rlane_d(v, 32), rlane_d(v, 48)));
  } else if constexpr (G == 32) {
    return half_reduce<true>(v);
  } else {
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, G));
    return v;
  }
}

// Sum over components 0..CNT-1 of per-row values, returned to every lane of the group, in the
// canonical pairwise order: components zero-padded to 64 slots, then summed in pairs (2j, 2j+1),
// pairs of pairs, ... . An ascending-xor butterfly over the G lanes forms exactly that tree
// within each block of G components (lane l of block s holds component l + G s); the block sums
// are then combined pairwise up to 64 slots. The result is independent of G and equals the
// oracle's pairwise sum bit for bit.
template <int G, int RPL, int CNT>
__device__ __forceinline__ double gpsum(int lane, const double (&a)[RPL]) {
  constexpr int NB = 64 / G;
  static_assert(RPL <= NB, "at most 64 components");
  if constexpr (G == 64) {
    // the same tree: pairs, quads, octets and rows of 16 by DPP, then (row0 + row1) + (row2 + row3)
    double v = (lane < CNT) ? a[0] : 0.0;
    v = v + dpp_d<0xB1>(v);
    v = v + dpp_d<0x4E>(v);
    v = v + dpp_d<0x141>(v);
    v = v + dpp_d<0x140>(v);
    return (rlane_d(v, 0) + rlane_d(v, 16)) + (rlane_d(v, 32) + rlane_d(v, 48));
  }
  double t[NB];
#pragma unroll
  for (int s = 0; s < NB; ++s) {
    if (s >= RPL || G * s >= CNT) {  // an all-zero block sums to +0.0
      t[s] = 0.0;
      continue;
    }
    double v = (lane + G * s < CNT) ? a[s] : 0.0;
    if constexpr (G == 32) {
      v = half_reduce<false>(v);  // the same pairwise tree over the 32 lanes of the half
    } else {
#pragma unroll
      for (int m = 1; m < G; m <<= 1) v += __shfl_xor(v, m, G);
    }
    t[s] = v;
  }
#pragma unroll
  for (int w = 1; w < NB; w <<= 1)
#pragma unroll
    for (int j = 0; j < NB; j += 2 * w) t[j] = t[j] + t[j + w];
  return t[0];
}

template <int G, int RPL, int CNT>
__device__ __forceinline__ void gpsum2(int lane, const double (&a)[RPL], const double (&b)[RPL], double& sa,
                                       double& sb) {
  sa = gpsum<G, RPL, CNT>(lane, a);
  sb = gpsum<G, RPL, CNT>(lane, b);
}

// ---------------------------------------------------------------- thermo and rates
__device__ __forceinline__ void nasa_hcp(const SpRec& sp, double T, double invT, double& cpR, double& hRT) {
  const int r = (T < sp.tmid) ? 0 : 1;
  const double* c = sp.cp[r];
  const double* h = sp.h[r];
  cpR = fma(fma(fma(fma(c[4], T, c[3]), T, c[2]), T, c[1]), T, c[0]);
  const double hp = fma(fma(fma(fma(h[4], T, h[3]), T, h[2]), T, h[1]), T, h[0]);
  hRT = fma(h[5], invT, hp);
}

__device__ __forceinline__ void nasa_full(const SpRec& sp, double T, double invT, double lnT, double& cpR,
                                          double& hRT, double& gRT, double& dcpR) {
  const int r = (T < sp.tmid) ? 0 : 1;
  nasa_hcp(sp, T, invT, cpR, hRT);
  const double* s = sp.s[r];
  const double* d = sp.dcp[r];
  const double spv = fma(fma(fma(s[4], T, s[3]), T, s[2]), T, s[1]) * T;
  const double sR = fma(s[0], lnT, spv + s[5]);
  gRT = hRT - sR;
  dcpR = fma(fma(fma(d[3], T, d[2]), T, d[1]), T, d[0]);
}

struct RxOut {
  double kf, kr, fwd, rev, dqdM, dqdT;
};

// One reaction at (C, T): effective rate coefficients, forward / reverse rates and, when JAC,
// d q / d[M] and d q / dT at constant concentrations.
template <bool JAC>
__device__ __forceinline__ void rx_eval(const RxRec& rx, const double* sC, const double* sG,
                                        const double* sH, double Ctot, double T, double lnT,
                                        double invT, double lnRTP, const int* effsp,
                                        const double* effval, RxOut& o) {
  const int info = rx.info;
  const int type = info & 15, rev = (info >> 4) & 15, nre = (info >> 8) & 15, npr = (info >> 12) & 15;
  const int ntroe = (info >> 16) & 15;
  const double kinf = pr_exp(fma(-rx.Ta, invT, fma(rx.beta, lnT, rx.lnA)));
  const double dlnkinf = fma(rx.Ta, invT, rx.beta) * invT;
  double M = 0.0;
  if (type != PR_RX_ELEM) {
    M = rx.effd * Ctot;
    const int eb = rx.eff & 0xFFFF, ee = (rx.eff >> 16) & 0xFFFF;
    // efficiency list in its order; the table reads of four entries are issued together
    for (int e = eb; e < ee; e += 4) {
      const int e1 = (e + 1 < ee) ? e + 1 : e, e2 = (e + 2 < ee) ? e + 2 : e, e3 = (e + 3 < ee) ? e + 3 : e;
      const double c0 = sC[effsp[e]], c1 = sC[effsp[e1]], c2 = sC[effsp[e2]], c3 = sC[effsp[e3]];
      const double v0 = effval[e], v1 = effval[e1], v2 = effval[e2], v3 = effval[e3];
      M = fma(v0, c0, M);
      if (e + 1 < ee) M = fma(v1, c1, M);
      if (e + 2 < ee) M = fma(v2, c2, M);
      if (e + 3 < ee) M = fma(v3, c3, M);
    }
  }
  const int r0 = rx.reac & 0xFF, r1 = (rx.reac >> 8) & 0xFF, r2 = (rx.reac >> 16) & 0xFF;
  const int p0 = rx.prod & 0xFF, p1 = (rx.prod >> 8) & 0xFF, p2 = (rx.prod >> 16) & 0xFF;
  // The six species slots are read unconditionally and together (an absent slot holds species 0,
  // whose value is read and not used): one LDS round trip instead of one per present slot.
  const double cr0 = sC[r0], cr1 = sC[r1], cr2 = sC[r2], cp0 = sC[p0], cp1 = sC[p1], cp2 = sC[p2];
  const double gr0 = sG[r0], gr1 = sG[r1], gr2 = sG[r2], gp0 = sG[p0], gp1 = sG[p1], gp2 = sG[p2];
  double pr = cr0;
  if (nre > 1) pr *= cr1;
  if (nre > 2) pr *= cr2;
  double pp = cp0;
  if (npr > 1) pp *= cp1;
  if (npr > 2) pp *= cp2;
  double invKc = 0.0, dH = 0.0;
  const int sumnu = npr - nre;
  if (rev) {
    double dG = gp0;
    if (npr > 1) dG += gp1;
    if (npr > 2) dG += gp2;
    dG -= gr0;
    if (nre > 1) dG -= gr1;
    if (nre > 2) dG -= gr2;
    invKc = pr_exp(fma((double)sumnu, lnRTP, dG));
    if (JAC) {
      const double hr0 = sH[r0], hr1 = sH[r1], hr2 = sH[r2], hp0 = sH[p0], hp1 = sH[p1], hp2 = sH[p2];
      dH = hp0;
      if (npr > 1) dH += hp1;
      if (npr > 2) dH += hp2;
      dH -= hr0;
      if (nre > 1) dH -= hr1;
      if (nre > 2) dH -= hr2;
    }
  }
  const double dlnKc = (dH - (double)sumnu) * invT;
  if (type <= PR_RX_THIRDBODY) {
    const double kr0 = kinf * invKc;
    const double fwd0 = kinf * pr, rev0 = kr0 * pp;
    double dq0 = 0.0;
    if (JAC) dq0 = fwd0 * dlnkinf - rev0 * (dlnkinf - dlnKc);
    if (type == PR_RX_ELEM) {
      o.kf = kinf; o.kr = kr0; o.fwd = fwd0; o.rev = rev0; o.dqdM = 0.0; o.dqdT = dq0;
    } else {
      o.kf = M * kinf; o.kr = M * kr0; o.fwd = M * fwd0; o.rev = M * rev0;
      o.dqdM = fwd0 - rev0; o.dqdT = M * dq0;
    }
    return;
  }
  const double k0 = pr_exp(fma(-rx.Ta0, invT, fma(rx.beta0, lnT, rx.lnA0)));
  double Pr = (k0 * M) / kinf;
  if (Pr < 1e-300) Pr = 1e-300;
  double F = 1.0, dlogF_dx = 0.0, dlogF_dT = 0.0, dlnPr = 0.0;
  if (JAC) dlnPr = fma(rx.Ta0, invT, rx.beta0) * invT - dlnkinf;
  if (type == PR_RX_TROE) {
    const double e3 = pr_exp(-T / rx.troe[1]), e1 = pr_exp(-T / rx.troe[2]);
    double Fc = (1.0 - rx.troe[0]) * e3 + rx.troe[0] * e1;
    double dFc = 0.0;
    if (JAC) dFc = -((1.0 - rx.troe[0]) / rx.troe[1]) * e3 - (rx.troe[0] / rx.troe[2]) * e1;
    if (ntroe == 4) {
      const double e2 = pr_exp(-rx.troe[3] * invT);
      Fc += e2;
      if (JAC) dFc += ((rx.troe[3] * invT) * invT) * e2;
    }
    if (Fc < 1e-300) Fc = 1e-300;
    const double logFc = pr_log10(Fc);
    const double c = -0.4 - 0.67 * logFc, nn = 0.75 - 1.27 * logFc;
    const double x = pr_log10(Pr);
    const double xc = x + c, D = nn - 0.14 * xc, f1 = xc / D;
    const double den = 1.0 + f1 * f1;
    F = pr_exp((logFc / den) * PRK_LN10);
    if (JAC) {
      dlogF_dx = -logFc * 2.0 * f1 * (nn / (D * D)) / (den * den);
      const double dlogFc = dFc / (Fc * PRK_LN10);
      const double dx = dlnPr / PRK_LN10, dc = -0.67 * dlogFc, dnn = -1.27 * dlogFc;
      const double dxc = dx + dc, dD = dnn - 0.14 * dxc;
      const double df1 = (dxc * D - xc * dD) / (D * D);
      dlogF_dT = (dlogFc * den - logFc * 2.0 * f1 * df1) / (den * den);
    }
  }
  const double kf = ((kinf * Pr) / (1.0 + Pr)) * F;
  const double kr = kf * invKc;
  o.kf = kf;
  o.kr = kr;
  o.fwd = kf * pr;
  o.rev = kr * pp;
  o.dqdM = 0.0;
  o.dqdT = 0.0;
  if (JAC) {
    const double dlnkf = dlnkinf + dlnPr / (1.0 + Pr) + PRK_LN10 * dlogF_dT;
    const double onep = 1.0 + Pr;
    o.dqdM = ((k0 * F) * (1.0 / (onep * onep) + dlogF_dx / onep)) * (pr - pp * invKc);
    o.dqdT = o.fwd * dlnkf - o.rev * (dlnkf - dlnKc);
  }
}

// ---------------------------------------------------------------- LDS views
template <int N, int R>
struct Smem {
  const SpRec* sp;
  const RxRec* rx;
  const int* effsp;
  const double* effval;
  const int* incs;
  const int* inc;
  const int* incq;       // byte offset of each incidence entry's rate slot in q
  const double* incnu;   // its stoichiometric coefficient
  const int* kblk;       // first gather block of each species; kblk[N] = number of blocks
  const int* blk;        // first incidence entry of each gather block (blk[nblocks] = ninc)
  double* Js;    // global scratch slot: (N+1)^2 row-major saved Jacobian
  double* jrec;  // global scratch slot: 4 R per-reaction Jacobian records (kf, kr, dqdM, dqdT)
  double* sC;    // N concentrations
  double* sG;    // N g/RT
  double* sH;    // N h/RT
  double* sHm;   // N mass-specific h
  double* sCpm;  // N mass-specific cp
  double* q;     // R rates of progress
  int* perm;     // N+1 pivot rows
  double* LU;    // (N+1)^2 LU factors, row-major; rows live in registers only during factor / solve
  double* Lp;    // N+1 doubles of LDS: the pivot-row broadcast of a 32-lane factorization whose
                 // factors are kept in the global slot
  static constexpr int n = N + 1;
  // per-group LDS (doubles) and per-group global scratch slot (doubles)
  static constexpr int scratch_doubles = 5 * N + R + (n + 1) / 2 + 1;
  static constexpr int slot_doubles = n * n + 4 * R;
};

template <int N, int R, int G>
struct Cell {
  static constexpr int n = N + 1;
  static constexpr int RPL = (n + G - 1) / G;
  static constexpr int RXL = (R + G - 1) / G;
  // One row per lane at 32 lanes, and above 32 components at 64 lanes: the group's LDS holds no
  // n x n area. The LU factors live in the group's global slot, stored entry by entry with the n
  // rows side by side (entry j of slot s at j * n + s, so every access coalesces; lanes past n
  // read the last slot's address, which adds no memory traffic), and the Jacobian is assembled
  // in the slot itself. A 32-lane factorization broadcasts its pivot rows through n doubles of
  // LDS.
  static constexpr bool lu_global = (G == 64 && n > 32) || (G == 32 && RPL == 1);
  // Otherwise, with one cell per wavefront the saved Jacobian and its records sit in LDS after
  // the group's scratch (few groups per block); narrower groups keep them in a global slot.
  static constexpr bool js_lds = (G == 64 && !lu_global);
  static constexpr int lds_doubles = Smem<N, R>::scratch_doubles + (lu_global ? (G == 32 ? n : 0) : n * n) +
                                     (js_lds ? Smem<N, R>::slot_doubles : 0);
  static constexpr int gslot_doubles = js_lds ? 0 : Smem<N, R>::slot_doubles + (lu_global ? n * n : 0);
};

// ---------------------------------------------------------------- right-hand side
// T at fixed mass-specific enthalpy h and mass fractions Y (rows of the group), by Newton from
// the initial value T: dT = (h - sum Y_k h_k(T)) / sum Y_k cp_k(T) until |dT| <= 1e-10 T
// (at most 20 iterations). Eos HY2T as used at Source/PeleLMeX_K.H:879 and inside the reactor RHS.
// Newton converges quadratically from the state's T, so the T after the last update is exact to
// far below that criterion.
template <int N, int R, int G>
__device__ __forceinline__ double hy2t(const Smem<N, R>& S, int lane, const double (&Y)[Cell<N, R, G>::RPL],
                                       double h, double T) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  for (int it = 0; it < 20; ++it) {
    const double invT = 1.0 / T;
    double ph[RPL], pc[RPL];
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int i = lane + G * s;
      ph[s] = 0.0;
      pc[s] = 0.0;
      if (i < N) {
        const SpRec& sp = S.sp[i];
        double cpR, hRT;
        nasa_hcp(sp, T, invT, cpR, hRT);
        ph[s] = ((Y[s] * hRT) * sp.RW) * T;
        pc[s] = (Y[s] * cpR) * sp.RW;
      }
    }
    double hm, cpm;
    gpsum2<G, RPL, N>(lane, ph, pc, hm, cpm);
    const double dT = (h - hm) / cpm;
    T = T + dT;
    if (fabs(dT) <= 1e-10 * T) break;
  }
  return T;
}

// Net molar production rates w_k (mol cm^-3 s^-1) at temperature T and concentrations cc
// (mol cm^-3, rows of the group) - productionRate of the mechanism - plus the mass-specific
// enthalpies hk (erg g^-1) and heat capacities cpk (erg g^-1 K^-1) at T.
template <int N, int R, int G>
__device__ __forceinline__ void net_rates(const Smem<N, R>& S, int lane, double T,
                                          const double (&cc)[Cell<N, R, G>::RPL], double (&w)[Cell<N, R, G>::RPL],
                                          double (&hk)[Cell<N, R, G>::RPL], double (&cpk)[Cell<N, R, G>::RPL]) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  const double lnT = pr_log(T), invT = 1.0 / T, lnRTP = pr_log((PR_RU * T) / PR_PATM);
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    const int i = lane + G * s;
    hk[s] = 0.0;
    cpk[s] = 0.0;
    if (i < N) {
      const SpRec& sp = S.sp[i];
      double cpR, hRT, gRT, dcp;
      nasa_full(sp, T, invT, lnT, cpR, hRT, gRT, dcp);
      S.sC[i] = cc[s];
      S.sG[i] = gRT;
      hk[s] = (hRT * sp.RW) * T;
      cpk[s] = cpR * sp.RW;
    }
  }
  const double Ctot = gpsum<G, RPL, N>(lane, cc);
  group_sync();
#pragma unroll 1
  for (int m = 0; m < Cell<N, R, G>::RXL; ++m) {
    const int r = lane + G * m;
    if (r < R) {
      RxOut o;
      rx_eval<false>(S.rx[r], S.sC, S.sG, S.sH, Ctot, T, lnT, invT, lnRTP, S.effsp, S.effval, o);
      S.q[r] = o.fwd - o.rev;
    }
  }
  group_sync();
  // Gather in PR_GATHER_BLOCK-entry blocks (mech_common.h): lane l sums blocks l, l + G, ...,
  // each from 0.0 in list order (up to PR_GATHER_BLOCK entries, the absent ones masked off),
  // into the species arrays, which are dead once the rates are formed; then each species lane
  // adds its blocks in order. The longest lists (H, OH, O) no longer set the loop length alone.
  double* const part = S.sC;
  const int nblk = S.kblk[N];
  const unsigned char* const qb = reinterpret_cast<const unsigned char*>(S.q);
  for (int b = lane; b - lane < nblk; b += G) {
    if (b < nblk) {
      const int e0 = S.blk[b], ne = S.blk[b + 1] - e0;
      // every table read of the block first (entries past its end re-read its last one), then
      // the sum in entry order: two LDS round trips per block instead of two per entry
      int off[PR_GATHER_BLOCK];
      double nu[PR_GATHER_BLOCK], qv[PR_GATHER_BLOCK];
#pragma unroll
      for (int j = 0; j < PR_GATHER_BLOCK; ++j) {
        const int e = e0 + (j < ne ? j : ne - 1);
        off[j] = S.incq[e];
        nu[j] = S.incnu[e];
      }
#pragma unroll
      for (int j = 0; j < PR_GATHER_BLOCK; ++j) qv[j] = *reinterpret_cast<const double*>(qb + off[j]);
      double p = 0.0;
#pragma unroll
      for (int j = 0; j < PR_GATHER_BLOCK; ++j)
        if (j < ne) p += nu[j] * qv[j];
      part[b] = p;
    }
  }
  group_sync();
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    const int i = lane + G * s;
    w[s] = 0.0;
    if (i < N) {
      const int b0 = S.kblk[i], b1 = S.kblk[i + 1];
      double a = (b1 > b0) ? part[b0] : 0.0;
      for (int b = b0 + 1; b < b1; ++b) a += part[b];
      w[s] = a;
    }
  }
  group_sync();
}

// Reactor right-hand side f(t, y) on the group; T from (h(t), Y) by Newton, started at y_T.
// y = (rho Y_1 .. rho Y_N, T) in CGS; f_k = W_k w_k + src_k, f_T = (src_h - sum h_k f_k) / (sum rho Y_k cp_k).
template <int N, int R, int G>
__device__ __forceinline__ void rhs_eval(const Smem<N, R>& S, int lane, double t,
                                         const double (&y)[Cell<N, R, G>::RPL],
                                         const double (&src)[Cell<N, R, G>::RPL], double rhoh0,
                                         double rhohsrc, double (&f)[Cell<N, R, G>::RPL]) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  const double rho = gpsum<G, RPL, N>(lane, y);
  const double hmix = (rhoh0 + t * rhohsrc) / rho;
  double Y[RPL], cc[RPL];
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    const int i = lane + G * s;
    Y[s] = (i < N) ? y[s] / rho : 0.0;
    cc[s] = (i < N) ? y[s] * S.sp[i].invW : 0.0;
  }
  const double T = hy2t<N, R, G>(S, lane, Y, hmix, gbcast<G>(y[N / G], N % G));
  double w[RPL], hk[RPL], cpk[RPL];
  net_rates<N, R, G>(S, lane, T, cc, w, hk, cpk);
  double pn[RPL], pcp[RPL];
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    const int i = lane + G * s;
    f[s] = 0.0;
    pn[s] = 0.0;
    pcp[s] = 0.0;
    if (i < N) {
      f[s] = w[s] * S.sp[i].W + src[s];
      pn[s] = hk[s] * f[s];
      pcp[s] = y[s] * cpk[s];
    }
  }
  double sn, rcp;
  gpsum2<G, RPL, N>(lane, pn, pcp, sn, rcp);
  const double fT = (rhohsrc - sn) / rcp;
  if (lane == N % G) f[N / G] = fT;
}

// reactionRateRhoY (Source/PeleLMeX_K.H:848-889) on one cell, MKS in and out: rho = sum rhoY,
// Y = rhoY / rho, h = rhoH / rho * 1e4 (erg / g), T by hy2t from the stored T, then
// rhoYdot_k = W_k w_k(rho * 1e-3, T, Y) * 1000.
template <int N, int R, int G>
__device__ __forceinline__ void rate_rhoY_eval(const Smem<N, R>& S, int lane, const double (&rY)[Cell<N, R, G>::RPL],
                                               double rH, double T0, double (&out)[Cell<N, R, G>::RPL]) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  const double rho = gpsum<G, RPL, N>(lane, rY);
  const double rhoinv = 1.0 / rho;
  double Y[RPL], cc[RPL];
#pragma unroll
  for (int s = 0; s < RPL; ++s) Y[s] = (lane + G * s < N) ? rY[s] * rhoinv : 0.0;
  const double T = hy2t<N, R, G>(S, lane, Y, (rH * rhoinv) * 1.0e4, T0);
  const double rho_cgs = rho * 0.001;
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    const int i = lane + G * s;
    cc[s] = (i < N) ? (rho_cgs * Y[s]) * S.sp[i].invW : 0.0;
  }
  double w[RPL], hk[RPL], cpk[RPL];
  net_rates<N, R, G>(S, lane, T, cc, w, hk, cpk);
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    const int i = lane + G * s;
    out[s] = (i < N) ? (w[s] * S.sp[i].W) * 1000.0 : 0.0;
  }
}

// ---------------------------------------------------------------- analytic Jacobian
// Exact Jacobian of the right-hand side with T = y_T, written into S.Js (row-major). Rows are
// assembled in LDS (the saved Jacobian itself at G = 64, else the LU area, which is dead here:
// a factorization always follows) and then copied to the global slot; above 32 components there
// is no LU area in LDS and they are assembled in the global slot itself (each lane updates its
// own row; a wavefront's global accesses to one address complete in order).
template <int N, int R, int G>
__device__ __forceinline__ void jac_eval(const Smem<N, R>& S, int lane, const double (&y)[Cell<N, R, G>::RPL],
                                         const double (&src)[Cell<N, R, G>::RPL], double rhohsrc) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  constexpr int n = N + 1;
  double* const A = (Cell<N, R, G>::js_lds || Cell<N, R, G>::lu_global) ? S.Js : S.LU;  // assembly rows
  const double T = gbcast<G>(y[N / G], N % G);
  const double lnT = pr_log(T), invT = 1.0 / T, lnRTP = pr_log((PR_RU * T) / PR_PATM);
  double cc[RPL], cpk[RPL], dcpk[RPL];
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    const int i = lane + G * s;
    cc[s] = 0.0;
    cpk[s] = 0.0;
    dcpk[s] = 0.0;
    if (i < N) {
      const SpRec& sp = S.sp[i];
      double cpR, hRT, gRT, dcp;
      nasa_full(sp, T, invT, lnT, cpR, hRT, gRT, dcp);
      const double C = y[s] * sp.invW;
      S.sC[i] = C;
      S.sG[i] = gRT;
      S.sH[i] = hRT;
      S.sHm[i] = (hRT * sp.RW) * T;
      S.sCpm[i] = cpR * sp.RW;
      cc[s] = C;
      cpk[s] = cpR * sp.RW;
      dcpk[s] = dcp * sp.RW;
    }
  }
  const double Ctot = gpsum<G, RPL, N>(lane, cc);
#pragma unroll 1
  for (int m = 0; m < Cell<N, R, G>::RXL; ++m) {
    const int r = lane + G * m;
    if (r < R) {
      RxOut o;
      rx_eval<true>(S.rx[r], S.sC, S.sG, S.sH, Ctot, T, lnT, invT, lnRTP, S.effsp, S.effval, o);
      double* rc = S.jrec + 4 * r;
      rc[0] = o.kf;
      rc[1] = o.kr;
      rc[2] = o.dqdM;
      rc[3] = o.dqdT;
      S.q[r] = o.fwd - o.rev;
    }
  }
  group_sync();
  double pn[RPL], prcp[RPL], psd[RPL], pcf[RPL];
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    const int k = lane + G * s;
    pn[s] = 0.0;
    prcp[s] = 0.0;
    psd[s] = 0.0;
    pcf[s] = 0.0;
    if (k < N) {
      double* row = A + k * n;
      for (int j = 0; j < n; ++j) row[j] = 0.0;
      double Sd = 0.0, JT = 0.0, w = 0.0;
      for (int e = S.incs[k]; e < S.incs[k + 1]; ++e) {
        const int rn = S.inc[e];
        const int r = rn & 0xFFFF;
        const double nu = (double)(rn >> 16);
        const RxRec& rx = S.rx[r];
        const double* rc = S.jrec + 4 * r;
        const int info = rx.info, type = info & 15, nre = (info >> 8) & 15, npr = (info >> 12) & 15;
        w += nu * S.q[r];
        const int a0 = rx.reac & 0xFF, a1 = (rx.reac >> 8) & 0xFF, a2 = (rx.reac >> 16) & 0xFF;
        const double c0 = S.sC[a0], c1 = nre > 1 ? S.sC[a1] : 1.0, c2 = nre > 2 ? S.sC[a2] : 1.0;
        const double kf = nu * rc[0];
        row[a0] += (kf * c1) * c2;
        if (nre > 1) row[a1] += (kf * c0) * c2;
        if (nre > 2) row[a2] += (kf * c0) * c1;
        const int b0 = rx.prod & 0xFF, b1 = (rx.prod >> 8) & 0xFF, b2 = (rx.prod >> 16) & 0xFF;
        const double d0 = S.sC[b0], d1 = npr > 1 ? S.sC[b1] : 1.0, d2 = npr > 2 ? S.sC[b2] : 1.0;
        const double kr = nu * rc[1];
        row[b0] -= (kr * d1) * d2;
        if (npr > 1) row[b1] -= (kr * d0) * d2;
        if (npr > 2) row[b2] -= (kr * d0) * d1;
        if (type != PR_RX_ELEM) {
          const double dm = nu * rc[2];
          Sd += dm * rx.effd;
          const int eb = rx.eff & 0xFFFF, ee = (rx.eff >> 16) & 0xFFFF;
          for (int q = eb; q < ee; ++q) row[S.effsp[q]] += dm * S.effval[q];
        }
        JT += nu * rc[3];
      }
      const double Wk = S.sp[k].W;
      for (int j = 0; j < N; ++j) row[j] = ((row[j] + Sd) * Wk) * S.sp[j].invW;
      row[N] = JT * Wk;
      const double fk = w * Wk + src[s];
      pn[s] = S.sHm[k] * fk;
      prcp[s] = y[s] * cpk[s];
      psd[s] = y[s] * dcpk[s];
      pcf[s] = cpk[s] * fk;
    }
  }
  double sn, rcp, sdcp, scf;
  gpsum2<G, RPL, N>(lane, pn, prcp, sn, rcp);
  gpsum2<G, RPL, N>(lane, psd, pcf, sdcp, scf);
  const double fT = (rhohsrc - sn) / rcp;
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    const int j = lane + G * s;
    if (j < n) {
      double cs = 0.0;
      for (int k = 0; k < N; ++k) cs += S.sHm[k] * A[k * n + j];
      A[N * n + j] = (j < N) ? (-cs - fT * S.sCpm[j]) / rcp : ((-cs - scf) - fT * sdcp) / rcp;
    }
  }
  group_sync();
  if constexpr (!Cell<N, R, G>::js_lds && !Cell<N, R, G>::lu_global) {
    for (int e = lane; e < n * n; e += G) S.Js[e] = A[e];
    group_sync();
  }
}

// ---------------------------------------------------------------- dense LU and solve
// LU of M = I - gamma * Js with partial pivoting (first maximal |pivot| in row order).
// Returns true when a zero pivot is met. Rows stay in the owning lanes; pst[s] is the step at
// which row lane + G s became pivot row, rinv the reciprocal of its pivot, and the row pivoted
// at step k is kept in S.perm[k] (G < 64) or in lane k of permv (G == 64).
template <int N, int R, int G>
__device__ __forceinline__ bool lu_factor(const Smem<N, R>& S, int lane, double gamma,
                                          double (&lu)[Cell<N, R, G>::RPL][N + 1],
                                          int (&pst)[Cell<N, R, G>::RPL], double (&rinv)[Cell<N, R, G>::RPL],
                                          int& permv) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  constexpr int n = N + 1;
  constexpr int UNPIV = 1 << 20;
  if constexpr (G == 64) {
    // The pivot owner is wave-uniform: the pivot row is read with readlane, no LDS round trips.
    pst[0] = (lane < n) ? UNPIV : -1;
    rinv[0] = 0.0;
    const int ol = opaque(lane);
#pragma unroll
    for (int j = 0; j < n; ++j)  // lanes >= n hold a copy of row n - 1, never read (pst = -1)
      lu[0][j] = (ol == j ? 1.0 : 0.0) - gamma * S.Js[(lane < n ? lane : n - 1) * n + j];
    bool singular = false;
#pragma unroll
    for (int k = 0; k < n; ++k) {
      double cand = -1.0;
      if (pst[0] == UNPIV) {
        const double v = fabs(lu[0][k]);
        cand = (v == v) ? v : -2.0;
      }
      const double best = gmax<64>(cand);
      if (!(best > 0.0)) {
        singular = true;
        break;
      }
      const int owner = (int)__builtin_ctzll(__ballot(cand == best));  // first maximal row
      if (lane == owner) pst[0] = k;
      if (lane == k) permv = owner;
      const double inv = 1.0 / rlane_d(lu[0][k], owner);
      if (lane == owner) rinv[0] = inv;
      if (pst[0] > k) lu[0][k] *= inv;
#pragma unroll
      for (int j = k + 1; j < n; ++j) {
        const double pj = rlane_d(lu[0][j], owner);
        if (pst[0] > k) lu[0][j] = fma(-lu[0][k], pj, lu[0][j]);
      }
    }
    return singular;
  }
  if constexpr (G == 32 && RPL == 1) {
    // as above, one pivot owner per half; the pivot row is read with readlane from both halves
    pst[0] = (lane < n) ? UNPIV : -1;
    rinv[0] = 0.0;
    const int ol = opaque(lane);
#pragma unroll
    for (int j = 0; j < n; ++j)  // lanes >= n hold a copy of row n - 1, never read (pst = -1)
      lu[0][j] = (ol == j ? 1.0 : 0.0) - gamma * S.Js[(lane < n ? lane : n - 1) * n + j];
    bool singular = false;
#pragma unroll
    for (int k = 0; k < n; ++k) {
      double cand = -1.0;
      if (pst[0] == UNPIV) {
        const double v = fabs(lu[0][k]);
        cand = (v == v) ? v : -2.0;
      }
      const double best = gmax<32>(cand);
      if (!(best > 0.0)) {
        singular = true;
        break;
      }
      const unsigned long long m = __ballot(cand == best);
      const unsigned int m0 = (unsigned int)m, m1 = (unsigned int)(m >> 32);
      const int o0 = m0 ? __builtin_ctz(m0) : 0, o1 = (m1 ? __builtin_ctz(m1) : 0) + 32;  // first maximal rows
      const int owner = upper_half() ? o1 - 32 : o0;
      if (lane == owner) pst[0] = k;
      if (ol == k) permv = owner;
      // the pivot row (final row k of U) goes through LDS: the group's LU area, which holds
      // nothing during the factorization (the factors are stored after it), or the pivot-row
      // buffer when the factors are kept in the global slot (a wavefront's LDS operations complete
      // in order, so step k + 1's write follows step k's reads)
      double* const prow = Cell<N, R, G>::lu_global ? S.Lp : S.LU + k * n;
      if (lane == owner)
#pragma unroll
        for (int j = k; j < n; ++j) prow[j] = lu[0][j];
      group_sync();
      const double inv = 1.0 / prow[k];
      if (lane == owner) rinv[0] = inv;
      if (pst[0] > k) {  // rows still to be pivoted: one masked region, no per-element selects
        lu[0][k] *= inv;
        const double m = lu[0][k];
#pragma unroll
        for (int j = k + 1; j < n; ++j) lu[0][j] = fma(-m, prow[j], lu[0][j]);
      }
    }
    group_sync();
    return singular;
  }
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    const int i = lane + G * s;
    pst[s] = (i < n) ? UNPIV : -1;
    rinv[s] = 0.0;
    const int oi = opaque(i);
#pragma unroll
    for (int j = 0; j < n; ++j)  // rows >= n hold a copy of row n - 1, never read (pst = -1)
      lu[s][j] = (oi == j ? 1.0 : 0.0) - gamma * S.Js[(i < n ? i : n - 1) * n + j];
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
        if (pst[s] > k) lu[s][j] = fma(-lu[s][k], pj, lu[s][j]);
    }
  }
  group_sync();
  return singular;
}

// The LU rows persist between factorizations in the group's LDS area (or, above 32 components,
// in its global slot); registers hold them only while a factorization or a solve runs.
// At 32 lanes (one row per lane), and in the global slot above 32 components, the rows are
// stored in pivot order (slot k = row pivoted at step k), so that the solve finds the pivot row of
// step k in lane k.
template <int N, int R, int G>
constexpr bool lu_pivot_order() {
  return (G == 32 && Cell<N, R, G>::RPL == 1) || Cell<N, R, G>::lu_global;
}

template <int N, int R, int G>
__device__ __forceinline__ void store_lu(const Smem<N, R>& S, int lane, const double (&lu)[Cell<N, R, G>::RPL][N + 1],
                                         const int (&pst)[Cell<N, R, G>::RPL]) {
  constexpr int n = N + 1;
  if constexpr (Cell<N, R, G>::lu_global) {  // slot pst of entry j at j * n + pst
    if (pst[0] >= 0 && pst[0] < n)
#pragma unroll
      for (int j = 0; j < n; ++j) S.LU[j * n + pst[0]] = lu[0][j];
    return;
  }
#pragma unroll
  for (int s = 0; s < Cell<N, R, G>::RPL; ++s) {
    const int i = lane + G * s;
    const int slot = lu_pivot_order<N, R, G>() ? pst[s] : i;
    if (i < n && slot >= 0 && slot < n)
#pragma unroll
      for (int j = 0; j < n; ++j) S.LU[slot * n + j] = lu[s][j];
  }
}

template <int N, int R, int G>
__device__ __forceinline__ void load_lu(const Smem<N, R>& S, int lane, double (&lu)[Cell<N, R, G>::RPL][N + 1]) {
  constexpr int n = N + 1;
  if constexpr (Cell<N, R, G>::lu_global) {  // lane k gets slot k; lanes >= n (a copy of slot n - 1) are never read
    const int sl = lane < n ? lane : n - 1;
#pragma unroll
    for (int j = 0; j < n; ++j) lu[0][j] = S.LU[j * n + sl];
    return;
  }
#pragma unroll
  for (int s = 0; s < Cell<N, R, G>::RPL; ++s) {
    const int i = lane + G * s;
#pragma unroll
    for (int j = 0; j < n; ++j) lu[s][j] = S.LU[(i < n ? i : n - 1) * n + j];  // rows >= n: never read
  }
}

// Solve M x = b with the factors from lu_factor; b is distributed by row, x by component.
template <int N, int R, int G>
__device__ __forceinline__ void lu_solve(const Smem<N, R>& S, int lane, const double (&lu)[Cell<N, R, G>::RPL][N + 1],
                                         const int (&pst)[Cell<N, R, G>::RPL], const double (&rinv)[Cell<N, R, G>::RPL],
                                         int permv, double (&b)[Cell<N, R, G>::RPL], double (&x)[Cell<N, R, G>::RPL]) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  constexpr int n = N + 1;
  if constexpr (G == 64 && !lu_pivot_order<N, R, G>()) {
#pragma unroll
    for (int k = 0; k < n; ++k) {
      const int p = __builtin_amdgcn_readlane(permv, k);
      const double yk = rlane_d(b[0], p);
      if (pst[0] > k) b[0] = fma(-lu[0][k], yk, b[0]);
    }
    x[0] = 0.0;
#pragma unroll
    for (int k = n - 1; k >= 0; --k) {
      const int p = __builtin_amdgcn_readlane(permv, k);
      const double xk = rlane_d(b[0], p) * rlane_d(rinv[0], p);
      if (pst[0] >= 0 && pst[0] < k) b[0] = fma(-lu[0][k], xk, b[0]);
      if (lane == k) x[0] = xk;
    }
    return;
  }
  if constexpr (lu_pivot_order<N, R, G>()) {
    // lane k holds the row pivoted at step k (pst = k); its right-hand side entry is b[p_k].
    // Step k's value comes from lane k of the group through the LDS crossbar (ds_bpermute, the
    // lane in the immediate offset): no vector instructions for the broadcast.
    const int base = (G == 32) ? (int)(threadIdx.x & 32) * 4 : 0;
    const int ol = opaque(lane);  // the step masks are formed here, not hoisted and spilled
    double v = __shfl(b[0], permv, G);
    // 1 / pivot from the stored (unscaled) pivot entry: the factorization's reciprocal
    const double rdiag =
        (lane < n) ? 1.0 / S.LU[lane * n + lane] : 0.0;  // the diagonal entry in either layout
#pragma unroll
    for (int k = 0; k < n; ++k) {
      const double yk = bperm_d(v, base + 4 * k);
      if (ol > k && ol < n) v = fma(-lu[0][k], yk, v);
    }
#pragma unroll
    for (int k = n - 1; k >= 1; --k) {
      const double xk = bperm_d(v * rdiag, base + 4 * k);
      if (ol < k) v = fma(-lu[0][k], xk, v);
    }
    x[0] = v * rdiag;  // lane k's v is final from step k on: x_k = v_k / pivot_k
    (void)pst;
    (void)rinv;
    return;
  }
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
      if (pst[s] > k) b[s] = fma(-lu[s][k], yk, b[s]);
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
      if (pst[s] >= 0 && pst[s] < k) b[s] = fma(-lu[s][k], xk, b[s]);
    if (lane == k % G) x[k / G] = xk;
  }
}

// Weighted RMS norm over the n components, summed in component order.
template <int N, int R, int G>
__device__ __forceinline__ double wrms(const Smem<N, R>& S, int lane, const double (&v)[Cell<N, R, G>::RPL],
                                       const double (&w)[Cell<N, R, G>::RPL]) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  double sq[RPL];
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    const double t = v[s] * w[s];
    sq[s] = t * t;
  }
  const double ss = gpsum<G, RPL, N + 1>(lane, sq);
  return sqrt(ss / (N + 1));
}

}  // namespace prk
#endif
