// The persistent react kernel: one BDF state machine per group of G lanes, one cell at a time,
// cells pulled from an atomic work queue. Semantics follow CVODE's CV_BDF / Newton / dense
// direct path as PelePhysics' ReactorCvode uses it (called from
// Source/PeleLMeX_Reactions.cpp:101-108), integrated to tout = dt_react and interpolated back
// from the first step that passes it.
//
// Register budget: the vectors a step touches every iteration (Nordsieck history, weights,
// corrections, LU rows) live in VGPRs; the group's scalar control state (step size, order,
// counters, method coefficients) lives in a per-group LDS control block, re-read after each
// right-hand side, so it does not occupy registers across the expensive phases.
#ifndef PR_REACT_KERNEL_H
#define PR_REACT_KERNEL_H
#include <float.h>
#include "bdf_kernel.h"

namespace prk {

enum { PH_FETCH = 0, PH_F0, PH_HIN, PH_NEWTON, PH_RESTART };
enum { NF_FIRST = 0, NF_CONV = 1, NF_ERR = 2 };
enum { CF_NONE = 0, CF_BADJ = 1, CF_OTHER = 2 };

// Per-group control block (LDS). All lanes of a group hold and write identical values.
struct Ctl {
  double tn, h, hscale, hprime, eta, etamax, rl1, gamma, gammap, gamrat, crate, delp, acnrm;
  double saved_tq5, saved_t, t_eval, hlb, hub, hg, rhoh0, rhohsrc;
  double tau[QMAXC + 1], tq[6], l[QMAXC + 1];
  double rq, rq1, rq2;  // 1 / q, 1 / (q + 1), 1 / (q + 2), formed when q changes (L = q + 1 throughout)
  int q, qprime, L, qwait, nscon, ncf, nef, nflag, convfail, mnewt, jcur, jbad, callSetup, count1;
  int phase, cell, box, nst, nstlp, nstlj, nstloc, nfe, nje, nset, nni, netf, ncfn, guard;
  double tout;
  unsigned int s_cells, s_nst, s_nfe, s_nje, s_nset, s_nni, s_netf, s_ncfn, s_fail;
};
constexpr int CTL_DOUBLES = (int)((sizeof(Ctl) + 7) / 8);

__device__ __forceinline__ long aoff(const Arr& a, int i, int j, int k) {
  return (long)(i - a.bx) + (long)(j - a.by) * a.js + (long)(k - a.bz) * a.ks;
}

__device__ __forceinline__ void clobber() { asm volatile("" ::: "memory"); }

template <int N, int R, int G>
constexpr int group_doubles() {
  return Cell<N, R, G>::lds_doubles + CTL_DOUBLES;
}

// Views of the mechanism tables and of the group's scratch (LDS control block first, then the
// per-group arrays, then - for one cell per wavefront - the saved Jacobian and its records).
template <int N, int R, int G>
__device__ __forceinline__ double* setup_views(Smem<N, R>& S, unsigned char* smem, const BlobLayout& BL, int grp,
                                               double* gscr) {
  constexpr int n = N + 1;
  S.sp = reinterpret_cast<const SpRec*>(smem + BL.off_sp);
  S.rx = reinterpret_cast<const RxRec*>(smem + BL.off_rx);
  S.effsp = reinterpret_cast<const int*>(smem + BL.off_effsp);
  S.effval = reinterpret_cast<const double*>(smem + BL.off_effval);
  S.incs = reinterpret_cast<const int*>(smem + BL.off_incs);
  S.inc = reinterpret_cast<const int*>(smem + BL.off_inc);
  S.incq = reinterpret_cast<const int*>(smem + BL.off_incq);
  S.incnu = reinterpret_cast<const double*>(smem + BL.off_incnu);
  S.kblk = reinterpret_cast<const int*>(smem + BL.off_kblk);
  S.blk = reinterpret_cast<const int*>(smem + BL.off_blk);
  double* scr = reinterpret_cast<double*>(smem + BL.bytes) + (long)grp * group_doubles<N, R, G>();
  S.sC = scr + CTL_DOUBLES;
  S.sG = S.sC + N;
  S.sH = S.sG + N;
  S.sHm = S.sH + N;
  S.sCpm = S.sHm + N;
  S.q = S.sCpm + N;
  S.perm = reinterpret_cast<int*>(S.q + R);
  S.LU = Cell<N, R, G>::lu_global ? nullptr : S.q + R + (n + 1) / 2 + 1;
  S.Lp = (Cell<N, R, G>::lu_global && G == 32) ? S.q + R + (n + 1) / 2 + 1 : nullptr;
  if constexpr (Cell<N, R, G>::js_lds) {
    S.Js = S.LU + n * n;
    S.jrec = S.Js + n * n;
  } else if (gscr) {
    const long slot = (long)blockIdx.x * (blockDim.x / G) + grp;
    S.Js = gscr + slot * Cell<N, R, G>::gslot_doubles;
    S.jrec = S.Js + n * n;
    if constexpr (Cell<N, R, G>::lu_global) S.LU = S.jrec + 4 * R;  // n n doubles after the records
  } else {
    S.Js = nullptr;  // kernels without a Jacobian
    S.jrec = nullptr;
  }
  return scr;
}

// W is the number of waves per SIMD the instance is compiled for: its VGPR budget is 512 / W
// per lane and its blocks hold up to 256 W threads (W waves per SIMD from one block per CU).
template <int N, int R, int G, int W>
__global__ void __launch_bounds__(256 * W) __attribute__((amdgpu_waves_per_eu(W, W))) react_kernel(KParams P) {
  constexpr int n = N + 1;
  constexpr int RPL = Cell<N, R, G>::RPL;
  constexpr int Q = QMAXC;
  extern __shared__ __align__(16) unsigned char smem[];
  {
    const uint4* srcb = reinterpret_cast<const uint4*>(P.blob);
    uint4* dst = reinterpret_cast<uint4*>(smem);
    for (int b = threadIdx.x; b < P.blob_bytes / 16; b += blockDim.x) dst[b] = srcb[b];
  }
  __syncthreads();
  const BlobLayout BL = blob_layout(N, R, P.neff, P.ninc);
  Smem<N, R> S;
  int lane = threadIdx.x % G;  // re-derived opaquely each iteration (see the loop head)
  const int grp = threadIdx.x / G;
  Ctl* ct = reinterpret_cast<Ctl*>(setup_views<N, R, G>(S, smem, BL, grp, P.gscr));

  long nwork = P.ncells;
  if (P.ncells_dev) {
    const long d = (long)*P.ncells_dev;
    nwork = d < nwork ? d : nwork;
  }
  const int qmax = P.max_order < 1 ? 1 : (P.max_order > Q ? Q : P.max_order);

  ct->s_cells = ct->s_nst = ct->s_nfe = ct->s_nje = ct->s_nset = ct->s_nni = ct->s_netf = ct->s_ncfn = ct->s_fail = 0u;
  ct->phase = PH_FETCH;

  // vectors kept in registers (rows lane, lane + G, ...)
  double zn[Q + 1][RPL], ewt[RPL], acor[RPL], yv[RPL], f[RPL], src[RPL], atol_r[RPL], neg[RPL];
  int pst[RPL];
  double rinv[RPL];
  int permv = 0;
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    int i = lane + G * s;
    atol_r[s] = (P.atolv && i < n) ? P.atolv[i] : P.atol;
  }

  auto box_of = [&](int b) __attribute__((always_inline)) -> const BoxDesc& { return P.boxes[b]; };

  auto cell_ijk = [&](const BoxDesc& B, int cl, int& i, int& j, int& k) __attribute__((always_inline)) {
    const unsigned u = (unsigned)cl;
    i = B.lo0 + (int)(u % (unsigned)B.nx);
    j = B.lo1 + (int)((u / (unsigned)B.nx) % (unsigned)B.ny);
    k = B.lo2 + (int)(u / ((unsigned)B.nx * (unsigned)B.ny));
  };

  auto finish = [&](int code, const double (&out)[RPL]) __attribute__((always_inline)) {
    const BoxDesc& B = box_of(ct->box);
    int i, j, k;
    cell_ijk(B, ct->cell, i, j, k);
    const long offY = aoff(B.rY, i, j, k), offT = aoff(B.T, i, j, k);
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      int ii = lane + G * s;
      if (ii < N) B.rY.p[offY + (long)ii * B.rY.ns] = neg[s] < 0.0 ? out[s] + neg[s] : out[s];
      else if (ii == N) B.T.p[offT] = out[s];
    }
    if (lane == 0) {
      B.rH.p[aoff(B.rH, i, j, k)] = ct->rhoh0 + ct->tout * ct->rhohsrc;
      B.FC.p[aoff(B.FC, i, j, k)] = (double)ct->nfe;
      if (B.status) B.status[ct->cell] = code;
    }
    ct->s_cells += 1u;
    ct->s_nst += ct->nst; ct->s_nfe += ct->nfe; ct->s_nje += ct->nje; ct->s_nset += ct->nset; ct->s_nni += ct->nni;
    ct->s_netf += ct->netf; ct->s_ncfn += ct->ncfn;
    if (code != ST_OK) ct->s_fail += 1u;
    ct->phase = PH_FETCH;
  };

  auto set_ewt = [&](const double (&y)[RPL]) __attribute__((always_inline)) {
#pragma unroll
    for (int s = 0; s < RPL; ++s) ewt[s] = 1.0 / (P.rtol * fabs(y[s]) + atol_r[s]);
  };

  // reciprocals of the order the step control divides by, formed once per order change (the same
  // correctly rounded values the divisions at each use gave)
  auto set_recips = [&]() __attribute__((always_inline)) {
    const double q = (double)ct->q;
    ct->rq = 1.0 / q;
    ct->rq1 = 1.0 / (q + 1.0);
    ct->rq2 = 1.0 / (q + 2.0);
  };

  auto rescale = [&]() __attribute__((always_inline)) {
    const double eta = ct->eta;
    const int q = ct->q;
    double factor = eta;
#pragma unroll
    for (int j = 1; j <= Q; ++j) {
      if (j <= q) {
#pragma unroll
        for (int s = 0; s < RPL; ++s) zn[j][s] *= factor;
        factor *= eta;
      }
    }
    ct->h = ct->hscale * eta;
    ct->hscale = ct->h;
    ct->nscon = 0;
  };

  auto predict = [&]() __attribute__((always_inline)) {
    ct->tn += ct->h;
    const int q = ct->q;
#pragma unroll
    for (int k = 1; k <= Q; ++k)
#pragma unroll
      for (int j = Q; j >= k; --j)
        if (j <= q)
#pragma unroll
          for (int s = 0; s < RPL; ++s) zn[j - 1][s] += zn[j][s];
  };

  auto restore = [&]() __attribute__((always_inline)) {
    ct->tn = ct->saved_t;
    const int q = ct->q;
#pragma unroll
    for (int k = 1; k <= Q; ++k)
#pragma unroll
      for (int j = Q; j >= k; --j)
        if (j <= q)
#pragma unroll
          for (int s = 0; s < RPL; ++s) zn[j - 1][s] -= zn[j][s];
  };

  // l (method coefficients), tq (error constants), gamma for the current q, h, tau history.
  // Computed in registers (loops unrolled to the maximum order, predicated on q) and written
  // back once; the operations and their order are those of the loop form.
  auto set_bdf = [&]() __attribute__((always_inline)) {
    const int q = ct->q;
    const double h = ct->h;
    double tau[Q + 1], l[Q + 1];
#pragma unroll
    for (int i = 0; i <= Q; ++i) {
      tau[i] = ct->tau[i];
      l[i] = 0.0;
    }
    double alpha0 = -1.0, alpha0_hat = -1.0, xi_inv = 1.0, xistar_inv = 1.0, hsum = h;
    l[0] = 1.0;
    l[1] = 1.0;
    if (q > 1) {
#pragma unroll
      for (int j = 2; j < Q; ++j) {
        if (j < q) {
          hsum += tau[j - 1];
          xi_inv = h / hsum;
          alpha0 -= 1.0 / j;
#pragma unroll
          for (int i = j; i >= 1; --i) l[i] += l[i - 1] * xi_inv;
        }
      }
      alpha0 -= ct->rq;
      xistar_inv = -l[1] - alpha0;
      double tq1 = 0.0;
#pragma unroll
      for (int j = 1; j <= Q; ++j)
        if (j == q) tq1 = tau[j - 1];
      hsum += tq1;
      xi_inv = h / hsum;
      alpha0_hat = -l[1] - xi_inv;
#pragma unroll
      for (int i = Q; i >= 1; --i)
        if (i <= q) l[i] += l[i - 1] * xistar_inv;
    }
    double lq = l[1];
#pragma unroll
    for (int i = 2; i <= Q; ++i)
      if (i == q) lq = l[i];
    const double A1 = 1.0 - alpha0_hat + alpha0;
    const double A2 = 1.0 + q * A1;
    ct->tq[2] = fabs(A1 / (alpha0 * A2));
    ct->tq[5] = fabs(A2 * xistar_inv / (lq * xi_inv));
    if (ct->qwait == 1) {
      if (q > 1) {
        double C = xistar_inv / lq;
        double A3 = alpha0 + ct->rq;
        double A4 = alpha0_hat + xi_inv;
        double Cpinv = (1.0 - A4 + A3) / A3;
        ct->tq[1] = fabs(C * Cpinv);
      } else {
        ct->tq[1] = 1.0;
      }
      double tqq = 0.0;
#pragma unroll
      for (int j = 1; j <= Q; ++j)
        if (j == q) tqq = tau[j];
      hsum += tqq;
      xi_inv = h / hsum;
      double A5 = alpha0 - ct->rq1;
      double A6 = alpha0_hat - xi_inv;
      double Cppinv = (1.0 - A6 + A5) / A2;
      ct->tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    ct->tq[4] = 0.1 / ct->tq[2];
#pragma unroll
    for (int i = 0; i <= Q; ++i) ct->l[i] = l[i];
    ct->rl1 = 1.0 / l[1];
    ct->gamma = h * ct->rl1;
    if (ct->nst == 0) ct->gammap = ct->gamma;
    ct->gamrat = (ct->nst > 0) ? ct->gamma / ct->gammap : 1.0;
  };

  auto decrease_bdf = [&]() __attribute__((always_inline)) {
    const int q = ct->q;
    double* l = ct->l;
    for (int i = 0; i <= Q; ++i) l[i] = 0.0;
    l[2] = 1.0;
    double hsum = 0.0;
    for (int j = 1; j <= q - 2; ++j) {
      hsum += ct->tau[j];
      double xi = hsum / ct->hscale;
      for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xi + l[i - 1];
    }
    double zq[RPL];
#pragma unroll
    for (int s = 0; s < RPL; ++s) zq[s] = 0.0;
#pragma unroll
    for (int j = 0; j <= Q; ++j)
      if (j == q)
#pragma unroll
        for (int s = 0; s < RPL; ++s) zq[s] = zn[j][s];
#pragma unroll
    for (int j = 2; j < Q; ++j)
      if (j < q) {
        const double lj = l[j];
#pragma unroll
        for (int s = 0; s < RPL; ++s) zn[j][s] -= lj * zq[s];
      }
  };

  auto increase_bdf = [&]() __attribute__((always_inline)) {
    const int q = ct->q, L = ct->L;
    double* l = ct->l;
    for (int i = 0; i <= Q; ++i) l[i] = 0.0;
    double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0, hsum = ct->hscale;
    l[2] = 1.0;
    if (q > 1) {
      for (int j = 1; j < q; ++j) {
        hsum += ct->tau[j + 1];
        double xi = hsum / ct->hscale;
        prod *= xi;
        alpha0 -= 1.0 / (j + 1);
        alpha1 += 1.0 / xi;
        for (int i = j + 2; i >= 2; --i) l[i] = l[i] * xiold + l[i - 1];
        xiold = xi;
      }
    }
    const double A1 = (-alpha0 - alpha1) / prod;
    double zL[RPL], zmax[RPL];
#pragma unroll
    for (int s = 0; s < RPL; ++s) zmax[s] = 0.0;
#pragma unroll
    for (int j = 0; j <= Q; ++j)
      if (j == qmax)
#pragma unroll
        for (int s = 0; s < RPL; ++s) zmax[s] = zn[j][s];
#pragma unroll
    for (int s = 0; s < RPL; ++s) zL[s] = A1 * zmax[s];
#pragma unroll
    for (int j = 0; j <= Q; ++j)
      if (j == L)
#pragma unroll
        for (int s = 0; s < RPL; ++s) zn[j][s] = zL[s];
#pragma unroll
    for (int j = 2; j <= Q; ++j)
      if (j <= q) {
        const double lj = l[j];
#pragma unroll
        for (int s = 0; s < RPL; ++s) zn[j][s] += lj * zL[s];
      }
  };

  auto adjust_params = [&]() __attribute__((always_inline)) {
    if (ct->qprime != ct->q) {
      const int dq = ct->qprime - ct->q;
      if (!(ct->q == 2 && dq != 1)) {
        if (dq == 1) increase_bdf();
        else if (dq == -1) decrease_bdf();
      }
      ct->q = ct->qprime;
      ct->L = ct->q + 1;
      ct->qwait = ct->L;
      set_recips();
    }
    rescale();
  };

  auto attempt = [&]() __attribute__((always_inline)) {
    predict();
    set_bdf();
    const int nflag = ct->nflag;
    ct->convfail = (nflag == NF_FIRST || nflag == NF_ERR) ? CF_NONE : CF_OTHER;
    ct->callSetup = (nflag == NF_CONV) || (nflag == NF_ERR) || (ct->nst == 0) || (ct->nst >= ct->nstlp + 20) ||
                  (fabs(ct->gamrat - 1.0) > 0.3);
    ct->jbad = 0;
    ct->mnewt = 0;
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      acor[s] = 0.0;
      yv[s] = zn[0][s];
    }
    ct->t_eval = ct->tn;
    ct->phase = PH_NEWTON;
  };

  auto begin_step = [&]() __attribute__((always_inline)) {
    if (ct->nst > 0) set_ewt(zn[0]);
    if (ct->nstloc >= P.max_steps) {
      finish(ST_TOO_MUCH_WORK, zn[0]);
      return;
    }
    if (ct->nst > 0 && ct->hprime != ct->h) adjust_params();
    ct->saved_t = ct->tn;
    ct->ncf = 0;
    ct->nef = 0;
    ct->nflag = NF_FIRST;
    attempt();
  };

  auto start_integration = [&]() __attribute__((always_inline)) {
    const double h = ct->h;
#pragma unroll
    for (int s = 0; s < RPL; ++s) zn[1][s] *= h;
    ct->hscale = h;
    ct->hprime = h;
    ct->nstloc = 0;
    begin_step();
  };

  const int guard_max = 20 * P.max_steps + 1000;

  for (;;) {
    clobber();
    // lane-derived masks and addresses are formed inside each iteration: hoisted out of the loop
    // they would be held (and spilled) for the whole kernel
    lane = opaque(lane);
    // At 64 lanes the group's LDS views and slot addresses are re-derived the same way, from an
    // opaque group index: as loop invariants the many addresses into the control block and the
    // species arrays were held (and spilled) for the whole kernel (-1.3 % on n-dodecane; at 32
    // lanes it costs more registers than it saves, +2 % on drm19)
    if constexpr (G == 64) ct = reinterpret_cast<Ctl*>(setup_views<N, R, G>(S, smem, BL, opaque(grp), P.gscr));
    if (uni<G>(ct->phase) == PH_FETCH) {
      unsigned int cl = 0;
      if (lane == 0) cl = atomicAdd(P.queue, 1u);
      cl = __shfl(cl, 0, G);
      if ((long)cl >= nwork) break;
      int b = 0;
      if (P.cell_list) cl = (unsigned int)P.cell_list[cl];  // a listed entry: box-local (one box) or global
      if (P.nbox > 1) {  // the box holding entry cl
        int lo = 0, hi = P.nbox;
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (P.box_start[mid] <= (long)cl) lo = mid;
          else hi = mid;
        }
        b = lo;
        cl = (unsigned int)((long)cl - P.box_start[b]);
      }
      const BoxDesc& B = box_of(b);
      int i, j, k;
      cell_ijk(B, (int)cl, i, j, k);
      ct->cell = (int)cl;
      ct->box = b;
      ct->tout = B.dtv ? B.dtv[cl] : B.dt;
      if (B.has_mask) {
        long mo = (long)(i - B.mask.bx) + (long)(j - B.mask.by) * B.mask.js + (long)(k - B.mask.bz) * B.mask.ks;
        if (B.mask.p[mo] == -1) {
          if (lane == 0) {
            B.FC.p[aoff(B.FC, i, j, k)] = 0.0;
            if (B.status) B.status[cl] = ST_OK;
          }
          continue;
        }
      }
      const long offY = aoff(B.rY, i, j, k), offS = aoff(B.rYsrc, i, j, k);
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        int ii = lane + G * s;
        double y0 = 0.0;
        src[s] = 0.0;
        neg[s] = 0.0;
        if (ii < N) {
          y0 = B.rY.p[offY + (long)ii * B.rY.ns];
          src[s] = B.rYsrc.p[offS + (long)ii * B.rYsrc.ns];
          if (P.clean && y0 < 0.0) {  // integrate from the non-negative part
            neg[s] = y0;
            y0 = 0.0;
          }
        } else if (ii == N) {
          y0 = B.T.p[aoff(B.T, i, j, k)];
        }
        zn[0][s] = y0;
        yv[s] = y0;
#pragma unroll
        for (int j2 = 1; j2 <= Q; ++j2) zn[j2][s] = 0.0;
      }
      ct->rhoh0 = B.rH.p[aoff(B.rH, i, j, k)];
      ct->rhohsrc = B.rHsrc.p[aoff(B.rHsrc, i, j, k)];
      ct->nfe = ct->nje = ct->nset = ct->nni = ct->netf = ct->ncfn = 0;
      ct->nst = ct->nstlp = ct->nstlj = 0;
      ct->q = 1; ct->L = 2; ct->qwait = 2; ct->qprime = 1;
      ct->rq = 1.0; ct->rq1 = 0.5; ct->rq2 = 1.0 / 3.0;
      ct->etamax = 10000.0; ct->nscon = 0; ct->saved_tq5 = 0.0; ct->crate = 1.0; ct->delp = 0.0;
      ct->gamrat = 1.0; ct->gammap = 1.0; ct->tn = 0.0; ct->jcur = 0; ct->eta = 1.0;
      for (int j2 = 0; j2 <= Q; ++j2) ct->tau[j2] = 0.0;
      for (int j2 = 0; j2 < 6; ++j2) ct->tq[j2] = 0.0;
      set_ewt(zn[0]);
      ct->t_eval = 0.0;
      ct->guard = 0;
      ct->phase = PH_F0;
    }
    if (++ct->guard > guard_max) {
      finish(ST_GUARD, zn[0]);
      continue;
    }

    rhs_eval<N, R, G>(S, lane, ct->t_eval, yv, src, ct->rhoh0, ct->rhohsrc, f);
    clobber();
    ct->nfe += 1;
    const int phase = uni<G>(ct->phase);

    if (phase == PH_F0) {
      double hinv = 0.0;
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        zn[1][s] = f[s];
        int i = lane + G * s;
        if (i < n) {
          double r = fabs(f[s]) / (0.1 * fabs(zn[0][s]) + atol_r[s]);
          hinv = r > hinv ? r : hinv;
        }
      }
      hinv = gmax<G>(hinv);
      const double tout = ct->tout;
      const double tround = DBL_EPSILON * fabs(tout);
      const double hlb = 100.0 * tround;
      double hub = 0.1 * fabs(tout);
      if (hub * hinv > 1.0) hub = 1.0 / hinv;
      const double hg = sqrt(hlb * hub);
      ct->hlb = hlb;
      ct->hub = hub;
      ct->hg = hg;
      if (hub < hlb) {
        ct->h = hg;
        start_integration();
      } else {
        ct->count1 = 1;
#pragma unroll
        for (int s = 0; s < RPL; ++s) yv[s] = zn[0][s] + hg * zn[1][s];
        ct->t_eval = hg;
        ct->phase = PH_HIN;
      }
    } else if (phase == PH_HIN) {
      const double hg = ct->hg, hub = ct->hub;
      double tmp[RPL];
#pragma unroll
      for (int s = 0; s < RPL; ++s) tmp[s] = (f[s] - zn[1][s]) / hg;
      const double yddnrm = wrms<N, R, G>(S, lane, tmp, ewt);
      double hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
      bool done = false;
      if (ct->count1 == 4) {
        done = true;
      } else {
        double hrat = hnew / hg;
        if (hrat > 0.5 && hrat < 2.0) done = true;
        else if (ct->count1 > 1 && hrat > 2.0) {
          hnew = hg;
          done = true;
        }
      }
      if (done) {
        double h0 = 0.5 * hnew;
        if (h0 < ct->hlb) h0 = ct->hlb;
        if (h0 > hub) h0 = hub;
        ct->h = h0;
        start_integration();
      } else {
        ct->hg = hnew;
        ct->count1 += 1;
#pragma unroll
        for (int s = 0; s < RPL; ++s) yv[s] = zn[0][s] + hnew * zn[1][s];
        ct->t_eval = hnew;
      }
    } else if (phase == PH_RESTART) {
      const double h = ct->h;
#pragma unroll
      for (int s = 0; s < RPL; ++s) zn[1][s] = h * f[s];
      attempt();
    } else {  // PH_NEWTON: f = f(tn, zn0 + acor)
      double res[RPL];
      {
        const double rl1 = ct->rl1, gamma = ct->gamma;
#pragma unroll
        for (int s = 0; s < RPL; ++s) res[s] = rl1 * zn[1][s] + acor[s] - gamma * f[s];
      }
      bool singular = false;
      if (uni<G>(ct->callSetup)) {
        const int cf = ct->jbad ? CF_BADJ : ct->convfail;
        const double dgamma = fabs(ct->gamma / ct->gammap - 1.0);
        const bool jnew = (ct->nst == 0) || (ct->nst >= ct->nstlj + 51) || (cf == CF_BADJ && dgamma < 0.2) || (cf == CF_OTHER);
        if (jnew) {
          ct->nje += 1;
          ct->nstlj = ct->nst;
          ct->jcur = 1;
          jac_eval<N, R, G>(S, lane, yv, src, ct->rhohsrc);
          clobber();
        } else {
          ct->jcur = 0;
        }
        {
          double lu[RPL][n];
          singular = lu_factor<N, R, G>(S, lane, ct->gamma, lu, pst, rinv, permv);
          store_lu<N, R, G>(S, lane, lu, pst);
        }
        clobber();
        ct->nset += 1;
        ct->gamrat = 1.0;
        ct->gammap = ct->gamma;
        ct->crate = 1.0;
        ct->nstlp = ct->nst;
        ct->callSetup = 0;
      }
      int outcome = 0;  // 0 continue, 1 converged, 2 failed
      if (!singular) {
        ct->nni += 1;
        double b[RPL], dlt[RPL];
#pragma unroll
        for (int s = 0; s < RPL; ++s) b[s] = -res[s];
        {
          double lu[RPL][n];
          load_lu<N, R, G>(S, lane, lu);
          lu_solve<N, R, G>(S, lane, lu, pst, rinv, permv, b, dlt);
        }
        clobber();
        const double gamrat = ct->gamrat;
        if (gamrat != 1.0) {
          const double sc = 2.0 / (1.0 + gamrat);
#pragma unroll
          for (int s = 0; s < RPL; ++s) dlt[s] *= sc;
        }
#pragma unroll
        for (int s = 0; s < RPL; ++s) acor[s] += dlt[s];
        const double del = wrms<N, R, G>(S, lane, dlt, ewt);
        const int mnewt = ct->mnewt;
        if (mnewt > 0) ct->crate = fmax(0.3 * ct->crate, del / ct->delp);
        const double dcon = del * fmin(1.0, ct->crate) / ct->tq[4];
        if (dcon <= 1.0) {
          ct->acnrm = (mnewt == 0) ? del : wrms<N, R, G>(S, lane, acor, ewt);
          ct->jcur = 0;
          outcome = 1;
        } else if (mnewt >= 1 && del > 2.0 * ct->delp) {
          outcome = 2;
        } else {
          ct->delp = del;
          ct->mnewt = mnewt + 1;
          if (mnewt + 1 >= 3) outcome = 2;
        }
      } else {
        outcome = 2;
      }
      if (outcome == 0) {
#pragma unroll
        for (int s = 0; s < RPL; ++s) yv[s] = zn[0][s] + acor[s];
      } else if (outcome == 2) {
        if (!ct->jcur) {
          ct->callSetup = 1;
          ct->jbad = 1;
          ct->mnewt = 0;
#pragma unroll
          for (int s = 0; s < RPL; ++s) {
            acor[s] = 0.0;
            yv[s] = zn[0][s];
          }
        } else {
          ct->ncfn += 1;
          restore();
          ct->ncf += 1;
          ct->etamax = 1.0;
          if (ct->ncf == 10) {
            finish(ST_CONV_FAILURE, zn[0]);
          } else {
            ct->eta = 0.25;
            ct->nflag = NF_CONV;
            rescale();
            attempt();
          }
        }
      } else {
        const double dsm = ct->acnrm * ct->tq[2];
        if (dsm <= 1.0) {
          // complete the step
          ct->nst += 1;
          ct->nscon += 1;
          const int q = ct->q;
          for (int i = q; i >= 2; --i) ct->tau[i] = ct->tau[i - 1];
          if (q == 1 && ct->nst > 1) ct->tau[2] = ct->tau[1];
          ct->tau[1] = ct->h;
#pragma unroll
          for (int j = 0; j <= Q; ++j)
            if (j <= q) {
              const double lj = ct->l[j];
#pragma unroll
              for (int s = 0; s < RPL; ++s) zn[j][s] += lj * acor[s];
            }
          ct->qwait -= 1;
          if (ct->qwait == 1 && q != qmax) {
#pragma unroll
            for (int j = 0; j <= Q; ++j)
              if (j == qmax)
#pragma unroll
                for (int s = 0; s < RPL; ++s) zn[j][s] = acor[s];
            ct->saved_tq5 = ct->tq[5];
          }
          // choose the next step size and order
          if (ct->etamax == 1.0) {
            if (ct->qwait < 2) ct->qwait = 2;
            ct->qprime = q;
            ct->hprime = ct->h;
            ct->eta = 1.0;
          } else {
            const int L = ct->L;
            const double etaq = 1.0 / (pr_pow(6.0 * dsm, ct->rq1) + 1e-6);
            if (ct->qwait != 0) {
              ct->eta = etaq;
              ct->qprime = q;
            } else {
              ct->qwait = 2;
              double etaqm1 = 0.0, etaqp1 = 0.0;
              if (q > 1) {
                double zq[RPL];
#pragma unroll
                for (int s = 0; s < RPL; ++s) zq[s] = 0.0;
#pragma unroll
                for (int j = 0; j <= Q; ++j)
                  if (j == q)
#pragma unroll
                    for (int s = 0; s < RPL; ++s) zq[s] = zn[j][s];
                const double ddn = wrms<N, R, G>(S, lane, zq, ewt) * ct->tq[1];
                etaqm1 = 1.0 / (pr_pow(6.0 * ddn, ct->rq) + 1e-6);
              }
              if (q != qmax && ct->saved_tq5 != 0.0) {
                const double cquot = (ct->tq[5] / ct->saved_tq5) * pr_pow(ct->h / ct->tau[2], (double)L);
                double tmp[RPL];
#pragma unroll
                for (int s = 0; s < RPL; ++s) tmp[s] = acor[s];
#pragma unroll
                for (int j = 0; j <= Q; ++j)
                  if (j == qmax)
#pragma unroll
                    for (int s = 0; s < RPL; ++s) tmp[s] = acor[s] - cquot * zn[j][s];
                const double dup = wrms<N, R, G>(S, lane, tmp, ewt) * ct->tq[3];
                etaqp1 = 1.0 / (pr_pow(10.0 * dup, ct->rq2) + 1e-6);
              }
              const double etam = fmax(etaqm1, fmax(etaq, etaqp1));
              if (etam < 1.5) {
                ct->eta = 1.0;
                ct->qprime = q;
              } else if (etam == etaqm1) {
                ct->eta = etaqm1;
                ct->qprime = q - 1;
              } else if (etam == etaq) {
                ct->eta = etaq;
                ct->qprime = q;
              } else {
                ct->eta = etaqp1;
                ct->qprime = q + 1;
#pragma unroll
                for (int j = 0; j <= Q; ++j)
                  if (j == qmax)
#pragma unroll
                    for (int s = 0; s < RPL; ++s) zn[j][s] = acor[s];
              }
            }
            if (ct->eta < 1.5) {
              ct->eta = 1.0;
              ct->hprime = ct->h;
            } else {
              ct->eta = fmin(ct->eta, ct->etamax);
              ct->hprime = ct->h * ct->eta;
              if (ct->qprime < q) ct->nscon = 0;
            }
          }
          ct->etamax = 10.0;
          ct->nstloc += 1;
          const double tn = ct->tn, h = ct->h, tout = ct->tout;
          if ((tn - tout) * h >= 0.0) {
            const double sfac = (tout - tn) / h;
            double out[RPL];
#pragma unroll
            for (int s = 0; s < RPL; ++s) out[s] = 0.0;
#pragma unroll
            for (int j = Q; j >= 0; --j) {
              if (j == q) {
#pragma unroll
                for (int s = 0; s < RPL; ++s) out[s] = zn[j][s];
              } else if (j < q) {
#pragma unroll
                for (int s = 0; s < RPL; ++s) out[s] = zn[j][s] + sfac * out[s];
              }
            }
            finish(ST_OK, out);
          } else {
            begin_step();
          }
        } else {
          // error test failure
          ct->nef += 1;
          ct->netf += 1;
          ct->nflag = NF_ERR;
          restore();
          if (ct->nef == 7) {
            finish(ST_ERR_FAILURE, zn[0]);
          } else {
            ct->etamax = 1.0;
            if (ct->nef <= 3) {
              double eta = 1.0 / (pr_pow(6.0 * dsm, ct->rq1) + 1e-6);
              eta = fmax(0.1, eta);
              if (ct->nef >= 2) eta = fmin(eta, 0.2);
              ct->eta = eta;
              rescale();
              attempt();
            } else if (ct->q > 1) {
              ct->eta = 0.1;
              decrease_bdf();
              ct->L = ct->q;
              ct->q -= 1;
              ct->qwait = ct->L;
              set_recips();
              rescale();
              attempt();
            } else {
              ct->eta = 0.1;
              ct->h *= 0.1;
              ct->hscale = ct->h;
              ct->qwait = 10;
              ct->nscon = 0;
#pragma unroll
              for (int s = 0; s < RPL; ++s) yv[s] = zn[0][s];
              ct->t_eval = ct->tn;
              ct->phase = PH_RESTART;
            }
          }
        }
      }
    }
  }

  if (lane == 0) {
    atomicAdd(&P.stats[0], (unsigned long long)ct->s_cells);
    atomicAdd(&P.stats[1], (unsigned long long)ct->s_nst);
    atomicAdd(&P.stats[2], (unsigned long long)ct->s_nfe);
    atomicAdd(&P.stats[3], (unsigned long long)ct->s_nje);
    atomicAdd(&P.stats[4], (unsigned long long)ct->s_nset);
    atomicAdd(&P.stats[5], (unsigned long long)ct->s_nni);
    atomicAdd(&P.stats[6], (unsigned long long)ct->s_netf);
    atomicAdd(&P.stats[7], (unsigned long long)ct->s_ncfn);
    if (ct->s_fail) atomicAdd(P.nfail, ct->s_fail);
  }
}


// One group per cell: evaluates the reactor right-hand side (MODE 0) or the analytic Jacobian
// (MODE 1) for a batch of states stored component-major (y[i * ncells + c]).
template <int N, int R, int G, int MODE>
__global__ void __launch_bounds__(BLOCK) eval_kernel(const unsigned char* blob, int blob_bytes, int neff, int ninc,
                                                     long ncells, const double* y, const double* rYsrc,
                                                     const double* rhoh0, const double* rhohsrc, double t,
                                                     double* out, double* gscr) {
  constexpr int n = N + 1;
  constexpr int RPL = Cell<N, R, G>::RPL;
  extern __shared__ __align__(16) unsigned char smem[];
  {
    const uint4* srcb = reinterpret_cast<const uint4*>(blob);
    uint4* dst = reinterpret_cast<uint4*>(smem);
    for (int b = threadIdx.x; b < blob_bytes / 16; b += blockDim.x) dst[b] = srcb[b];
  }
  __syncthreads();
  const BlobLayout BL = blob_layout(N, R, neff, ninc);
  Smem<N, R> S;
  const int lane = threadIdx.x % G;
  const int grp = threadIdx.x / G;
  setup_views<N, R, G>(S, smem, BL, grp, gscr);
  const long slot = (long)blockIdx.x * (blockDim.x / G) + grp;
  const long stride = (long)gridDim.x * (blockDim.x / G);
  for (long c = slot; c < ncells; c += stride) {  // whole groups iterate together
  double yv[RPL], src[RPL], f[RPL];
#pragma unroll
  for (int s = 0; s < RPL; ++s) {
    const int i = lane + G * s;
    yv[s] = (i < n) ? y[(long)i * ncells + c] : 0.0;
    src[s] = (i < N) ? rYsrc[(long)i * ncells + c] : 0.0;
  }
  if (MODE == 0) {
    rhs_eval<N, R, G>(S, lane, t, yv, src, rhoh0[c], rhohsrc[c], f);
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int i = lane + G * s;
      if (i < n) out[(long)i * ncells + c] = f[s];
    }
  } else {
    jac_eval<N, R, G>(S, lane, yv, src, rhohsrc[c]);
    for (int e = lane; e < n * n; e += G) out[(long)e * ncells + c] = S.Js[e];
    group_sync();
  }
  }
}

// reactionRateRhoY over a box (Source/PeleLMeX_Reactions.cpp:399-435 calling
// Source/PeleLMeX_K.H:848-889): one group per cell, MKS arrays in Array4 layout; cells with
// mask == -1 (covered) get zero rates, as the reference's covered-cell branch does.
struct RateParams {
  int lo0, lo1, lo2, nx, ny;
  long ncells;
  Arr rY, rH, T, out;
  IArr mask;
  int has_mask;
  const unsigned char* blob;
  int blob_bytes, neff, ninc;
};

template <int N, int R, int G>
__global__ void __launch_bounds__(BLOCK) rate_kernel(RateParams P) {
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
  for (long c = (long)blockIdx.x * (blockDim.x / G) + grp; c < P.ncells; c += stride) {
    const unsigned u = (unsigned)c;
    const int i = P.lo0 + (int)(u % (unsigned)P.nx);
    const int j = P.lo1 + (int)((u / (unsigned)P.nx) % (unsigned)P.ny);
    const int k = P.lo2 + (int)(u / ((unsigned)P.nx * (unsigned)P.ny));
    const long offO = aoff(P.out, i, j, k);
    if (P.has_mask) {
      const long mo = (long)(i - P.mask.bx) + (long)(j - P.mask.by) * P.mask.js + (long)(k - P.mask.bz) * P.mask.ks;
      if (P.mask.p[mo] == -1) {
#pragma unroll
        for (int s = 0; s < RPL; ++s) {
          const int ii = lane + G * s;
          if (ii < N) P.out.p[offO + (long)ii * P.out.ns] = 0.0;
        }
        continue;
      }
    }
    const long offY = aoff(P.rY, i, j, k);
    double ry[RPL], out[RPL];
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int ii = lane + G * s;
      ry[s] = (ii < N) ? P.rY.p[offY + (long)ii * P.rY.ns] : 0.0;
    }
    rate_rhoY_eval<N, R, G>(S, lane, ry, P.rH.p[aoff(P.rH, i, j, k)], P.T.p[aoff(P.T, i, j, k)], out);
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int ii = lane + G * s;
      if (ii < N) P.out.p[offO + (long)ii * P.out.ns] = out[s];
    }
  }
}

// S evenly spaced cells of a box, component-major with stride S: rhoY (N), sources (N), T, rho h,
// its source, FC (zero).
__global__ void gather_sample_kernel(BoxDesc B, int N, long S, double* out) {
  const long a = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= S) return;
  const long cl = a * (B.ncells / S);
  const unsigned u = (unsigned)cl;
  const int i = B.lo0 + (int)(u % (unsigned)B.nx);
  const int j = B.lo1 + (int)((u / (unsigned)B.nx) % (unsigned)B.ny);
  const int k = B.lo2 + (int)(u / ((unsigned)B.nx * (unsigned)B.ny));
  const long offY = aoff(B.rY, i, j, k), offS = aoff(B.rYsrc, i, j, k);
  for (int n = 0; n < N; ++n) {
    out[(long)n * S + a] = B.rY.p[offY + (long)n * B.rY.ns];
    out[(long)(N + n) * S + a] = B.rYsrc.p[offS + (long)n * B.rYsrc.ns];
  }
  out[(2L * N) * S + a] = B.T.p[aoff(B.T, i, j, k)];
  out[(2L * N + 1) * S + a] = B.rH.p[aoff(B.rH, i, j, k)];
  out[(2L * N + 2) * S + a] = B.rHsrc.p[aoff(B.rHsrc, i, j, k)];
  out[(2L * N + 3) * S + a] = 0.0;
}

// Cost order of a box's cells (or of a list of them) from the input FC (right-hand-side
// evaluations of the previous advance, PeleLMeX's chemistry cost measure, Source/PeleLMeX.H:209):
// 128 buckets, 4 per octave of FC + 1, the most expensive first, so the persistent kernel starts
// the stiff cells early and its tail is short. order[] lists the cell indices bucket by bucket
// (arbitrary order within a bucket). hist holds 2 x COST_BUCKETS counters, zero on entry.
constexpr int COST_BUCKETS = 128;

__device__ __forceinline__ int cost_bucket(double fc) {
  const double f = (fc > 0.0) ? (fc < 2.0e9 ? fc : 2.0e9) : 0.0;
  const unsigned u = (unsigned)f + 1u;
  const int msb = 31 - __builtin_clz(u);
  const int frac = (msb >= 2) ? (int)((u >> (msb - 2)) & 3u) : (int)((u << (2 - msb)) & 3u);
  const int key = msb * 4 + frac;  // < 128
  return COST_BUCKETS - 1 - key;
}

// The cells to order: list[0 .. min(*nlist, nmax)) when list is set, else queue entries 0 .. nmax.
__device__ __forceinline__ long cost_count(const int* list, const unsigned int* nlist, long nmax) {
  if (!list) return nmax;
  const long n = (long)*nlist;
  return n < nmax ? n : nmax;
}

// FC of queue entry c of a launch's box table (box-local index when there is one box).
__device__ __forceinline__ double cost_fc(const BoxDesc* boxes, const long* box_start, int nbox, long c) {
  int b = 0;
  if (nbox > 1) {
    int lo = 0, hi = nbox;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (box_start[mid] <= c) lo = mid;
      else hi = mid;
    }
    b = lo;
    c -= box_start[b];
  }
  const BoxDesc& B = boxes[b];
  const unsigned u = (unsigned)c;
  const int i = B.lo0 + (int)(u % (unsigned)B.nx);
  const int j = B.lo1 + (int)((u / (unsigned)B.nx) % (unsigned)B.ny);
  const int k = B.lo2 + (int)(u / ((unsigned)B.nx * (unsigned)B.ny));
  return B.FC.p[aoff(B.FC, i, j, k)];
}

__global__ void __launch_bounds__(256) cost_hist_kernel(const BoxDesc* boxes, const long* box_start, int nbox,
                                                        const int* list, const unsigned int* nlist, long nmax,
                                                        unsigned int* hist) {
  __shared__ unsigned int h[COST_BUCKETS];
  for (int b = threadIdx.x; b < COST_BUCKETS; b += blockDim.x) h[b] = 0u;
  __syncthreads();
  const long n = cost_count(list, nlist, nmax);
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x) {
    const long c = list ? (long)list[x] : x;
    atomicAdd(&h[cost_bucket(cost_fc(boxes, box_start, nbox, c))], 1u);  // block-private counts in LDS
  }
  __syncthreads();
  for (int b = threadIdx.x; b < COST_BUCKETS; b += blockDim.x)
    if (h[b]) atomicAdd(&hist[b], h[b]);
}

__global__ void cost_scan_kernel(unsigned int* hist) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    unsigned int a = 0;
    for (int b = 0; b < COST_BUCKETS; ++b) {
      hist[COST_BUCKETS + b] = a;
      a += hist[b];
    }
  }
}

// Each block counts its cells per bucket in LDS, reserves one range per bucket from the global
// cursors, and places its cells inside its ranges (one grid-stride round per launch: the grid
// covers every cell, so each thread handles at most one).
__global__ void __launch_bounds__(256) cost_scatter_kernel(const BoxDesc* boxes, const long* box_start, int nbox,
                                                           const int* list, const unsigned int* nlist, long nmax,
                                                           unsigned int* hist, int* order) {
  __shared__ unsigned int cnt[COST_BUCKETS], base[COST_BUCKETS];
  for (int b = threadIdx.x; b < COST_BUCKETS; b += blockDim.x) cnt[b] = 0u;
  __syncthreads();
  const long n = cost_count(list, nlist, nmax);
  const long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long c = -1;
  int b = 0;
  unsigned int r = 0;
  if (x < n) {
    c = list ? (long)list[x] : x;
    b = cost_bucket(cost_fc(boxes, box_start, nbox, c));
    r = atomicAdd(&cnt[b], 1u);
  }
  __syncthreads();
  for (int q = threadIdx.x; q < COST_BUCKETS; q += blockDim.x)
    base[q] = (cnt[q] && hist[q] != (unsigned)n) ? atomicAdd(&hist[COST_BUCKETS + q], cnt[q]) : 0u;
  __syncthreads();
  if (x < n) {
    // one bucket holding every cell (e.g. FC all zero on a first advance): the given order
    const unsigned pos = (hist[b] == (unsigned)n) ? (unsigned)x : base[b] + r;
    order[pos] = (int)c;
  }
}

// Per-box chemistry cost from the FC arrays (PeleLM::computeCosts, Source/PeleLMeX_Regrid.cpp:558-606):
// one block per box, the sum and the maximum of FC over the box's cells (a tree sum: exact for the
// whole-number FC the reactor writes), then the method's figure.
enum { COST_NCELL = 0, COST_FC_AVG = 1, COST_FC_MAX = 2, COST_FC_SUM = 3 };

struct CostBox {
  int lo0, lo1, lo2, nx, ny;
  long ncells;
  Arr FC;
};

__global__ void __launch_bounds__(256) box_cost_kernel(const CostBox* boxes, int method, double* out) {
  __shared__ double ssum[256], smax[256];
  const CostBox B = boxes[blockIdx.x];
  double sum = 0.0, mx = -DBL_MAX;
  for (long c = threadIdx.x; c < B.ncells; c += blockDim.x) {
    const unsigned u = (unsigned)c;
    const int i = B.lo0 + (int)(u % (unsigned)B.nx);
    const int j = B.lo1 + (int)((u / (unsigned)B.nx) % (unsigned)B.ny);
    const int k = B.lo2 + (int)(u / ((unsigned)B.nx * (unsigned)B.ny));
    const double v = B.FC.p[aoff(B.FC, i, j, k)];
    sum += v;
    mx = v > mx ? v : mx;
  }
  ssum[threadIdx.x] = sum;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      ssum[threadIdx.x] += ssum[threadIdx.x + w];
      smax[threadIdx.x] = smax[threadIdx.x + w] > smax[threadIdx.x] ? smax[threadIdx.x + w] : smax[threadIdx.x];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double n = (double)B.ncells;
    out[blockIdx.x] = (method == COST_NCELL) ? n
                      : (method == COST_FC_AVG) ? ssum[0] / n
                      : (method == COST_FC_MAX) ? smax[0]
                                                : ssum[0];
  }
}

__global__ void zero_u32(unsigned int* p, int n) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) p[t] = 0u;
}

// Up to kStoreWords 8-byte words carried in the kernel arguments, stored to dst[0..n): host data
// staged this way is part of the launch itself, so a captured graph replays the values of its
// capture whatever the host memory holds later.
constexpr int kStoreWords = 256;
struct StoreWords {
  unsigned long long w[kStoreWords];
};
__global__ void store_words(unsigned long long* dst, int n, StoreWords src) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) dst[t] = src.w[t];
}

__global__ void zero_counters(unsigned int* queue, unsigned long long* stats, unsigned int* nfail) {
  int t = threadIdx.x;
  if (t < 8) stats[t] = 0ull;
  if (t == 0) {
    *queue = 0u;
    *nfail = 0u;
  }
}

}  // namespace prk
#endif
