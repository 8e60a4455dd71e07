// Learned dt_react map in front of the BDF: features, FP64 trust check, fallback list, audit.
// No counterpart in the reference (its react always integrates, Source/PeleLMeX_Reactions.cpp
// :101-108); this is an addition whose accepted cells must stay within the accuracy the caller
// asks for, while every cell it cannot vouch for is integrated by the BDF kernel.
//
// Per cell, the network maps x = [log10(Y_k + 1e-20) (N), T / 2000, log10(dt)] (z-normalised)
// to the chemical increments [dY_k (N), dT] over dt relative to the explicit estimate from the
// analytic rates at the source-advanced state, u_i = d_i / (dt |f_i(y*)| (/ rho*) + eps_i), then
// z-normalised: frozen cells give d = 0 and near-equilibrium cells O(1) targets. The trust kernel then
//  * removes the increments' element imbalance by the least-squares projection
//    dY <- dY - M^T (M M^T)^-1 M dY, M_ek = a_ek / W_k (elements per gram), and rejects a
//    cell whose raw imbalance exceeds elem_tol of its gross change plus 1e-10 mol / g,
//  * applies the sources first and exactly (prep kernel): y* = rhoY0 + dt src, rhoh1 =
//    rhoh0 + dt src_h, T* = HY2T(rhoh1 / rho*, Y*); the network sees (Y*, T*, dt), and the cell
//    becomes rhoY1 = rho* (Y* + dY), T1 = HY2T(rhoh1 / rho1, Y1) started at T* + dT,
//  * evaluates the backward-Euler residual r = y1 - y0 - dt f(dt, y1) with the analytic rates
//    (f is the reactor right-hand side the BDF integrates) and accepts the cell when
//    ||r||_wrms <= 1 with weights 1 / (trust_rtol |y1_i| + trust_atol), rhoY1 >= -trust_atol
//    and T1 within 200..6000 K.
// Accepted cells are written in place (FC = 1); rejected cells are appended to a list the BDF
// kernel integrates from the untouched inputs; a deterministic pseudo-random subset of the
// accepted cells (audit_frac) is copied, inputs and surrogate result, to the audit buffers
// where the BDF re-solves them and the host reports the error.
#ifndef PR_SURROGATE_KERNEL_H
#define PR_SURROGATE_KERNEL_H
#include "mlp_kernels.h"
#include "react_kernel.h"

namespace prk {

constexpr int SURR_MAXE = 8;  // elements of a mechanism

struct FeatParams {
  int lo0, lo1, lo2, nx, ny;
  long ncells;
  Arr rY, T;         // box arrays (CGS), or component-major batch when batch != 0
  int batch;         // 1: rY[k * ncells + c], T[c]
  const double* dtv; // per-cell dt (batch) or null
  double dt;
  const float* xmu;  // normalisation of the N + 2 features
  const float* xsd;
  float* raw;        // [ncells x (N + 2)] raw features, nullable
  __bf16* X;         // [ncells x (N + 2)] normalised bf16 features, nullable
};

// feature f of a cell: log10(max(Y_f, 0) + 1e-20) for f < N, T / 2000, log10(dt)
__device__ __forceinline__ double feature_value(int f, int N, double Y, double T, double dt) {
  if (f < N) return log10(fmax(Y, 0.0) + 1e-20);
  if (f == N) return T / 2000.0;
  return log10(dt);
}

__device__ __forceinline__ __bf16 feature_norm(double x, float mu, float sd) { return (__bf16)(((float)x - mu) / sd); }

template <int N>
__global__ void featurize_kernel(FeatParams P) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= P.ncells) return;
  long offY, offT, ns;
  if (P.batch) {
    offY = c;
    offT = c;
    ns = P.ncells;
  } else {
    const unsigned u = (unsigned)c;
    const int i = P.lo0 + (int)(u % (unsigned)P.nx);
    const int j = P.lo1 + (int)((u / (unsigned)P.nx) % (unsigned)P.ny);
    const int k = P.lo2 + (int)(u / ((unsigned)P.nx * (unsigned)P.ny));
    offY = aoff(P.rY, i, j, k);
    offT = aoff(P.T, i, j, k);
    ns = P.rY.ns;
  }
  double ry[N];
  double rho = 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    ry[k] = P.rY.p[offY + (long)k * ns];
    rho += ry[k];
  }
  const double dt = P.dtv ? P.dtv[c] : P.dt;
  const int F = N + 2;
#pragma unroll
  for (int f = 0; f < F; ++f) {
    const double x = feature_value(f, N, f < N ? ry[f] / rho : 0.0, P.T.p[offT], dt);
    if (P.raw) P.raw[c * F + f] = (float)x;
    if (P.X) P.X[c * F + f] = feature_norm(x, P.xmu[f], P.xsd[f]);
  }
}

// Box path, first pass: sources are applied first and exactly, y* = rhoY0 + dt src,
// rho h* = rho h0 + dt src_h, T* = HY2T(h*, Y*) from T0; the network then sees (Y*, T*, dt) and
// predicts the chemistry over dt from y* (so fast modes end re-equilibrated, as in the full
// problem). Writes the normalised features and T*.
struct PrepParams {
  int lo0, lo1, lo2, nx, ny;
  long ncells;
  Arr rY, rYsrc, T, rH, rHsrc;
  IArr mask;
  int has_mask;
  int clean;  // ode.clean_init_massfrac
  double dt;
  const float* xmu;
  const float* xsd;
  __bf16* X;
  double* Tstar;
  double* scale;  // [ncells x (N + 1)]: dt |f_k(y*)| / rho* (species), dt |f_T(y*)| (T)
  const unsigned char* blob;
  int blob_bytes, neff, ninc;
};

template <int N, int R, int G>
__global__ void __launch_bounds__(BLOCK) prep_kernel(PrepParams P) {
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
  const double dt = P.dt;
  const int F = N + 2;
  const long stride = (long)gridDim.x * (blockDim.x / G);
  for (long c = (long)blockIdx.x * (blockDim.x / G) + grp; c < P.ncells; c += stride) {
    const unsigned u = (unsigned)c;
    const int i = P.lo0 + (int)(u % (unsigned)P.nx);
    const int j = P.lo1 + (int)((u / (unsigned)P.nx) % (unsigned)P.ny);
    const int k = P.lo2 + (int)(u / ((unsigned)P.nx * (unsigned)P.ny));
    if (P.has_mask) {
      const long mo = (long)(i - P.mask.bx) + (long)(j - P.mask.by) * P.mask.js + (long)(k - P.mask.bz) * P.mask.ks;
      if (P.mask.p[mo] == -1) {
        for (int f = lane; f < F; f += G) P.X[c * F + f] = (__bf16)0.0f;
        continue;
      }
    }
    const long offY = aoff(P.rY, i, j, k), offS = aoff(P.rYsrc, i, j, k);
    double ys[RPL], Ys[RPL];
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int ii = lane + G * s;
      double y0 = (ii < N) ? P.rY.p[offY + (long)ii * P.rY.ns] : 0.0;
      if (P.clean && y0 < 0.0) y0 = 0.0;
      ys[s] = (ii < N) ? y0 + dt * P.rYsrc.p[offS + (long)ii * P.rYsrc.ns] : 0.0;
    }
    const double rhos = gpsum<G, RPL, N>(lane, ys);
#pragma unroll
    for (int s = 0; s < RPL; ++s) Ys[s] = (lane + G * s < N) ? ys[s] / rhos : 0.0;
    const double hs = (P.rH.p[aoff(P.rH, i, j, k)] + dt * P.rHsrc.p[aoff(P.rHsrc, i, j, k)]) / rhos;
    const double Ts = hy2t<N, R, G>(S, lane, Ys, hs, P.T.p[aoff(P.T, i, j, k)]);
    {
      // chemistry-only rates at y* (no sources): the output scales
      double yv[RPL], z[RPL], f[RPL];
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        yv[s] = ys[s];
        z[s] = 0.0;
      }
      if (lane == N % G) yv[N / G] = Ts;
      rhs_eval<N, R, G>(S, lane, 0.0, yv, z, hs * rhos, 0.0, f);
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        const int ii = lane + G * s;
        if (ii < N) P.scale[c * (long)(N + 1) + ii] = dt * fabs(f[s]) / rhos;
        else if (ii == N) P.scale[c * (long)(N + 1) + N] = dt * fabs(f[s]);
      }
    }
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int ii = lane + G * s;
      if (ii < N) P.X[c * F + ii] = feature_norm(feature_value(ii, N, Ys[s], Ts, dt), P.xmu[ii], P.xsd[ii]);
    }
    if (lane == 0) {
      P.X[c * F + N] = feature_norm(feature_value(N, N, 0.0, Ts, dt), P.xmu[N], P.xsd[N]);
      P.X[c * F + N + 1] = feature_norm(feature_value(N + 1, N, 0.0, Ts, dt), P.xmu[N + 1], P.xsd[N + 1]);
      P.Tstar[c] = Ts;
    }
  }
}

struct TrustParams {
  int lo0, lo1, lo2, nx, ny;
  long ncells;
  Arr rY, rYsrc, T, rH, rHsrc, FC;
  IArr mask;
  int has_mask;
  int clean;  // ode.clean_init_massfrac: the cell starts from max(rhoY, 0); the rest is added back
  int* status;
  double dt;
  const float* Yout;  // network output [ncells x (N + 1)], normalised
  const float* ymu;
  const float* ysd;
  const float* yeps;    // rate-scale floors (0: the output is the increment itself)
  const double* Tstar;  // T after the source step (prep kernel)
  const double* scale;  // [ncells x (N + 1)] rate scales dt |f_i(y*)| (/ rho* for species)
  int zero;             // 1: every increment is zero (the source step alone; Yout is not read)
  const double* Mel;  // [E x N] elements per gram of species k
  const double* Minv; // [E x E] (M M^T)^-1
  int E;
  double trust_rtol, trust_atol, elem_tol;
  const unsigned char* blob;
  int blob_bytes, neff, ninc;
  int* rej_list;           // rejected cells as cell_base + box-local index
  int cell_base;
  unsigned int* counters;  // [0] rejected, [1] audited, [2] accepted
  unsigned int audit_cap, audit_thresh, audit_seed;
  double* aud_in;   // [cap x (2 N + 3)] rhoY0, src, T0, rhoh0, src_h
  double* aud_out;  // [cap x (N + 1)] surrogate rhoY1, T1
  int* aud_cell;
  double* aud_dt;   // [cap] the audited cell's dt_react
};

// network output -> increment: u = o sd + mu; d = u (scale + eps) with the cell's rate scale
// (eps > 0), or d = u (eps = 0)
__device__ __forceinline__ double decode(float o, float mu, float sd, float eps, double scale) {
  const double u = (double)o * (double)sd + (double)mu;
  return eps > 0.f ? u * (scale + (double)eps) : u;
}

__device__ __forceinline__ unsigned int mix32(unsigned int x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

template <int N, int R, int G>
__global__ void __launch_bounds__(BLOCK) trust_kernel(TrustParams P) {
  constexpr int RPL = Cell<N, R, G>::RPL;
  constexpr int n = N + 1;
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
  const double dt = P.dt;
  const long stride = (long)gridDim.x * (blockDim.x / G);
  for (long c = (long)blockIdx.x * (blockDim.x / G) + grp; c < P.ncells; c += stride) {
    const unsigned u = (unsigned)c;
    const int i = P.lo0 + (int)(u % (unsigned)P.nx);
    const int j = P.lo1 + (int)((u / (unsigned)P.nx) % (unsigned)P.ny);
    const int k = P.lo2 + (int)(u / ((unsigned)P.nx * (unsigned)P.ny));
    if (P.has_mask) {
      const long mo = (long)(i - P.mask.bx) + (long)(j - P.mask.by) * P.mask.js + (long)(k - P.mask.bz) * P.mask.ks;
      if (P.mask.p[mo] == -1) {
        if (lane == 0) {
          P.FC.p[aoff(P.FC, i, j, k)] = 0.0;
          if (P.status) P.status[c] = ST_OK;
        }
        continue;
      }
    }
    const long offY = aoff(P.rY, i, j, k), offS = aoff(P.rYsrc, i, j, k);
    double y0[RPL], src[RPL], dY[RPL], neg[RPL];
    const float* o = P.Yout + c * (long)n;
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int ii = lane + G * s;
      y0[s] = 0.0;
      src[s] = 0.0;
      dY[s] = 0.0;
      neg[s] = 0.0;
      if (ii < N) {
        y0[s] = P.rY.p[offY + (long)ii * P.rY.ns];
        if (P.clean && y0[s] < 0.0) {
          neg[s] = y0[s];
          y0[s] = 0.0;
        }
        src[s] = P.rYsrc.p[offS + (long)ii * P.rYsrc.ns];
        dY[s] = P.zero ? 0.0 : decode(o[ii], P.ymu[ii], P.ysd[ii], P.yeps[ii], P.scale[c * (long)n + ii]);
      }
    }
    const double T0 = P.T.p[aoff(P.T, i, j, k)];
    const double rhoh0 = P.rH.p[aoff(P.rH, i, j, k)];
    const double srch = P.rHsrc.p[aoff(P.rHsrc, i, j, k)];
    const double dT = P.zero ? 0.0 : decode(o[N], P.ymu[N], P.ysd[N], P.yeps[N], P.scale[c * (long)n + N]);
    // element balance of the increments and its least-squares removal
    double a[SURR_MAXE], gross[SURR_MAXE];
#pragma unroll
    for (int e = 0; e < SURR_MAXE; ++e) {
      a[e] = 0.0;
      gross[e] = 0.0;
      if (e >= P.E) continue;
      double pa[RPL], pg[RPL];
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        const int ii = lane + G * s;
        const double me = (ii < N) ? P.Mel[e * N + ii] : 0.0;
        pa[s] = me * dY[s];
        pg[s] = fabs(me * dY[s]);
      }
      gpsum2<G, RPL, N>(lane, pa, pg, a[e], gross[e]);
    }
    bool ok = true;
#pragma unroll
    for (int e = 0; e < SURR_MAXE; ++e) ok = ok && (e >= P.E || fabs(a[e]) <= P.elem_tol * gross[e] + 1e-10);
#pragma unroll
    for (int e = 0; e < SURR_MAXE; ++e) {
      if (e >= P.E) continue;
      double lam = 0.0;
#pragma unroll
      for (int f = 0; f < SURR_MAXE; ++f)
        if (f < P.E) lam += P.Minv[e * P.E + f] * a[f];
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        const int ii = lane + G * s;
        if (ii < N) dY[s] -= P.Mel[e * N + ii] * lam;
      }
    }
    // y* = y0 + dt src (sources first), then the chemistry increment at constant rho*
    double ys[RPL];
#pragma unroll
    for (int s = 0; s < RPL; ++s) ys[s] = (lane + G * s < N) ? y0[s] + dt * src[s] : 0.0;
    const double rhos = gpsum<G, RPL, N>(lane, ys);
    double y1[RPL], Y1[RPL];
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int ii = lane + G * s;
      y1[s] = (ii < N) ? rhos * (ys[s] / rhos + dY[s]) : 0.0;
      ok = ok && (ii >= N || y1[s] >= -P.trust_atol);
    }
    const double rho1 = gpsum<G, RPL, N>(lane, y1);
    const double rhoh1 = rhoh0 + dt * srch;
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int ii = lane + G * s;
      Y1[s] = (ii < N) ? y1[s] / rho1 : 0.0;
    }
    const double Ts = P.Tstar[c];
    double Tg = Ts + dT;
    if (!(Tg > 200.0 && Tg < 6000.0)) Tg = Ts;
    const double T1 = hy2t<N, R, G>(S, lane, Y1, rhoh1 / rho1, Tg);
    ok = ok && (T1 > 200.0 && T1 < 6000.0);
    if (lane == N % G) y1[N / G] = T1;
    double f1[RPL];
    rhs_eval<N, R, G>(S, lane, dt, y1, src, rhoh0, srch, f1);
    double r[RPL], w[RPL];
#pragma unroll
    for (int s = 0; s < RPL; ++s) {
      const int ii = lane + G * s;
      double base = 0.0;
      if (ii < N) base = y0[s];
      else if (ii == N) base = T0;
      r[s] = (ii < n) ? (y1[s] - base) - dt * f1[s] : 0.0;
      w[s] = (ii < n) ? 1.0 / (P.trust_rtol * fabs(y1[s]) + P.trust_atol) : 0.0;
    }
    const double nrm = wrms<N, R, G>(S, lane, r, w);
    ok = ok && (nrm <= 1.0);
    const bool accept = gmax<G>(ok ? 0.0 : 1.0) == 0.0;  // every lane's rows passed
    unsigned int slot = 0;
    if (accept) {
      const long offT = aoff(P.T, i, j, k);
      const bool audit = P.audit_thresh == 0xFFFFFFFFu || mix32((unsigned)c ^ P.audit_seed) < P.audit_thresh;
      if (audit) {
        if (lane == 0) slot = atomicAdd(&P.counters[1], 1u);
        slot = __shfl(slot, 0, G);
        if (slot < P.audit_cap) {
          double* ai = P.aud_in + (long)slot * (2 * N + 3);
          double* ao = P.aud_out + (long)slot * n;
#pragma unroll
          for (int s = 0; s < RPL; ++s) {
            const int ii = lane + G * s;
            if (ii < N) {
              ai[ii] = y0[s] + neg[s];  // the caller's input
              ai[N + ii] = src[s];
            }
            if (ii < N) ao[ii] = neg[s] < 0.0 ? y1[s] + neg[s] : y1[s];
            else if (ii == N) ao[ii] = y1[s];
          }
          if (lane == 0) {
            ai[2 * N] = T0;
            ai[2 * N + 1] = rhoh0;
            ai[2 * N + 2] = srch;
            P.aud_cell[slot] = (int)c;
            P.aud_dt[slot] = dt;
          }
        }
      }
#pragma unroll
      for (int s = 0; s < RPL; ++s) {
        const int ii = lane + G * s;
        if (ii < N) P.rY.p[offY + (long)ii * P.rY.ns] = neg[s] < 0.0 ? y1[s] + neg[s] : y1[s];
        else if (ii == N) P.T.p[offT] = y1[s];
      }
      if (lane == 0) {
        P.rH.p[aoff(P.rH, i, j, k)] = rhoh1;
        P.FC.p[aoff(P.FC, i, j, k)] = 1.0;
        if (P.status) P.status[c] = ST_OK;
        atomicAdd(&P.counters[2], 1u);
      }
    } else if (lane == 0) {
      slot = atomicAdd(&P.counters[0], 1u);
      P.rej_list[slot] = P.cell_base + (int)c;
    }
  }
}

// Audit comparison: err[2 a] = |T_bdf - T_sur| / T_bdf, err[2 a + 1] = max_k |Y_bdf - Y_sur|
// (mass fractions) for audit slot a < min(count, cap); bdf arrays component-major with stride cap.
__global__ void audit_err_kernel(int N, const unsigned int* count, unsigned int cap, const double* bdf_rY,
                                 const double* bdf_T, const double* sur, double* err) {
  const unsigned a = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned cnt = *count < cap ? *count : cap;
  if (a >= cnt) return;
  const double* s = sur + (long)a * (N + 1);
  double rb = 0.0, rs = 0.0;
  for (int k = 0; k < N; ++k) {
    rb += bdf_rY[(long)k * cap + a];
    rs += s[k];
  }
  double dy = 0.0;
  for (int k = 0; k < N; ++k) dy = fmax(dy, fabs(bdf_rY[(long)k * cap + a] / rb - s[k] / rs));
  err[2 * a] = fabs(bdf_T[a] - s[N]) / bdf_T[a];
  err[2 * a + 1] = dy;
}

// Unpacks the audit inputs [cap x (2 N + 3)] into the component-major 1D batch (stride cap) the
// BDF kernel integrates; slots past the count are left as they are (the BDF never visits them).
__global__ void audit_unpack_kernel(int N, const unsigned int* count, unsigned int cap, const double* in,
                                    double* rY, double* src, double* T, double* rH, double* rHsrc, double* FC) {
  const unsigned a = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned cnt = *count < cap ? *count : cap;
  if (a >= cnt) return;
  const double* p = in + (long)a * (2 * N + 3);
  for (int k = 0; k < N; ++k) {
    rY[(long)k * cap + a] = p[k];
    src[(long)k * cap + a] = p[N + k];
  }
  T[a] = p[2 * N];
  rH[a] = p[2 * N + 1];
  rHsrc[a] = p[2 * N + 2];
  FC[a] = 0.0;
}

}  // namespace prk
#endif
