// Surrogate MLP kernels for gfx950: BF16 operands on the matrix cores (v_mfma_f32_16x16x32_bf16),
// FP32 accumulation, fused bias / GELU epilogues, and the elementwise kernels of training
// (GELU backward, bias gradients, MSE loss, Adam). No counterpart in the reference: the learned
// dt_react map is an addition in front of the BDF (cells it cannot vouch for go to the BDF).
//
// Layouts: activations are row-major [rows x features] (features fastest), weights row-major
// [in x out]. The GEMM takes element strides for both operands, so the forward product
// (X W), the data gradient (dZ W^T) and the weight gradient (X^T dZ) all use the one kernel.
#ifndef PR_MLP_KERNELS_H
#define PR_MLP_KERNELS_H
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace prk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

enum { EPI_NONE = 0, EPI_GELU = 1 };

struct GemmParams {
  int M, N, K;
  const __bf16* A;  // element (m, k) at A[m * sam + k * sak]
  long sam, sak;
  const __bf16* B;  // element (k, n) at B[k * sbk + n * sbn]
  long sbk, sbn;
  const float* bias;  // [N] or null
  int epi;            // EPI_NONE / EPI_GELU
  float* C;           // fp32 result after the epilogue [M x ldc] or null
  long ldc;
  __bf16* Cb;         // bf16 copy of the result or null
  long ldcb;
  float* Z;           // pre-activation (bias added, before GELU) or null
  long ldz;
  int accumulate;     // C += result (fp32 only, no epilogue), for gradient accumulation
};

__device__ __forceinline__ float gelu_f(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

// d gelu / dz = Phi(z) + z phi(z)
__device__ __forceinline__ float gelu_grad_f(float z) {
  const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

// Block tile 64 x 64, K step 32, 256 threads = 4 waves in a 2 x 2 arrangement of 32 x 32 wave
// tiles, each wave tile = 2 x 2 MFMA 16 x 16 x 32 tiles. Operand tiles are staged in LDS with k
// fastest (row pitch 40 bf16 = 80 bytes, so the 16-byte operand reads of neighbouring rows fall
// in different banks). MFMA lane layout: lane l supplies A row l % 16 and B column l % 16 with
// k = 8 (l / 16) .. 8 (l / 16) + 7, and receives D rows 4 (l / 16) .. + 3 of column l % 16.
constexpr int GT_M = 64, GT_N = 64, GT_K = 32, G_PITCH = 40;

__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmParams P) {
  __shared__ __align__(16) __bf16 As[GT_M * G_PITCH];
  __shared__ __align__(16) __bf16 Bs[GT_N * G_PITCH];
  const int t = threadIdx.x;
  const int wave = t / 64, l = t % 64;
  const int m0 = blockIdx.y * GT_M, n0 = blockIdx.x * GT_N;
  const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bool a_kfast = (P.sak == 1), b_kfast = (P.sbk == 1);
  for (int k0 = 0; k0 < P.K; k0 += GT_K) {
    // stage: 2048 elements per operand tile, 8 per thread, consecutive threads on the
    // operand's contiguous dimension
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int e = t + 256 * j;
      int m, k;
      if (a_kfast) { m = e / GT_K; k = e % GT_K; } else { k = e / GT_M; m = e % GT_M; }
      const int gm = m0 + m, gk = k0 + k;
      As[m * G_PITCH + k] = (gm < P.M && gk < P.K) ? P.A[(long)gm * P.sam + (long)gk * P.sak] : (__bf16)0.0f;
      int n;
      if (b_kfast) { n = e / GT_K; k = e % GT_K; } else { k = e / GT_N; n = e % GT_N; }
      const int gn = n0 + n;
      const int gk2 = k0 + k;
      Bs[n * G_PITCH + k] = (gn < P.N && gk2 < P.K) ? P.B[(long)gk2 * P.sbk + (long)gn * P.sbn] : (__bf16)0.0f;
    }
    __syncthreads();
    bf16x8 a[2], b[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&As[(wm + 16 * i + l % 16) * G_PITCH + 8 * (l / 16)]);
#pragma unroll
    for (int j = 0; j < 2; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[(wn + 16 * j + l % 16) * G_PITCH + 8 * (l / 16)]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm + 16 * i + 4 * (l / 16) + r;
        const int gn = n0 + wn + 16 * j + l % 16;
        if (gm >= P.M || gn >= P.N) continue;
        float v = acc[i][j][r];
        if (P.accumulate) {
          P.C[(long)gm * P.ldc + gn] += v;
          continue;
        }
        if (P.bias) v += P.bias[gn];
        if (P.Z) P.Z[(long)gm * P.ldz + gn] = v;
        if (P.epi == EPI_GELU) v = gelu_f(v);
        if (P.C) P.C[(long)gm * P.ldc + gn] = v;
        if (P.Cb) P.Cb[(long)gm * P.ldcb + gn] = (__bf16)v;
      }
}

// Forward product of a hidden layer on the inference path: A [M x K] row-major (k fastest, lda =
// K), Bt = W^T [N x K] row-major (k fastest), K % 32 == 0, N % 128 == 0. Block tile 128 x 128,
// 256 threads = 4 waves of 64 x 64 (4 x 4 MFMA 16 x 16 x 32 tiles); operand tiles move as 16-byte
// vectors into double-buffered LDS (pitch 40 bf16), the next K step's loads in flight while the
// current one multiplies. Per output element the K steps and the MFMA are those of
// gemm_bf16_kernel, so the two give the same bits. Epilogue: bias, GELU (epi), bf16 store.
constexpr int FT_M = 128, FT_N = 128;

__global__ void __launch_bounds__(256) gemm_bf16_fwd_kernel(int M, int N, int K, const __bf16* __restrict__ A,
                                                            const __bf16* __restrict__ Bt,
                                                            const float* __restrict__ bias, int epi,
                                                            __bf16* __restrict__ Cb, long ldcb) {
  __shared__ __align__(16) __bf16 As[2][FT_M * G_PITCH];
  __shared__ __align__(16) __bf16 Bs[2][FT_N * G_PITCH];
  const int t = threadIdx.x;
  const int wave = t / 64, l = t % 64;
  const int m0 = blockIdx.y * FT_M, n0 = blockIdx.x * FT_N;
  const int wm = (wave / 2) * 64, wn = (wave % 2) * 64;
  // staging: 512 16-byte chunks per operand tile (128 rows x 4 chunks of 8 k); thread t moves
  // chunks t and t + 256 of each operand: row (c / 4), k offset 8 (c % 4)
  int arow[2], brow[2], kc[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int c = t + 256 * u;
    const int r = c / 4;
    kc[u] = 8 * (c % 4);
    const int gm = m0 + r;
    arow[u] = gm < M ? gm : M - 1;  // rows past M read a valid row; their results are not stored
    brow[u] = n0 + r;
  }
  bf16x8 ra[2], rb[2];
  auto gload = [&](int k0) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      ra[u] = *reinterpret_cast<const bf16x8*>(A + (long)arow[u] * K + k0 + kc[u]);
      rb[u] = *reinterpret_cast<const bf16x8*>(Bt + (long)brow[u] * K + k0 + kc[u]);
    }
  };
  auto sstore = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int r = (t + 256 * u) / 4;
      *reinterpret_cast<bf16x8*>(&As[buf][r * G_PITCH + kc[u]]) = ra[u];
      *reinterpret_cast<bf16x8*>(&Bs[buf][r * G_PITCH + kc[u]]) = rb[u];
    }
  };
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  gload(0);
  sstore(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += GT_K) {
    const bool more = k0 + GT_K < K;
    if (more) gload(k0 + GT_K);
    bf16x8 a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const bf16x8*>(&As[buf][(wm + 16 * i + l % 16) * G_PITCH + 8 * (l / 16)]);
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const bf16x8*>(&Bs[buf][(wn + 16 * j + l % 16) * G_PITCH + 8 * (l / 16)]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    if (more) sstore(buf ^ 1);  // the other buffer: its last readers passed the previous barrier
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int gn = n0 + wn + 16 * j + l % 16;
    const float bj = bias ? bias[gn] : 0.f;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm + 16 * i + 4 * (l / 16) + r;
        if (gm >= M) continue;
        float v = acc[i][j][r];
        if (bias) v += bj;
        if (epi == EPI_GELU) v = gelu_f(v);
        Cb[(long)gm * ldcb + gn] = (__bf16)v;
      }
  }
}

// dZ = dA * gelu'(Z) (elementwise over rows x cols, row pitches given), as bf16 for the next
// products, optionally also fp32.
__global__ void gelu_bwd_kernel(long rows, int cols, const float* dA, long lda, const float* Z, long ldz,
                                __bf16* dZb, long ldb, float* dZ, long ldd) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= rows * cols) return;
  const long r = e / cols;
  const int c = (int)(e % cols);
  const float g = dA[r * lda + c] * gelu_grad_f(Z[r * ldz + c]);
  dZb[r * ldb + c] = (__bf16)g;
  if (dZ) dZ[r * ldd + c] = g;
}

// db[c] = sum over rows of d[r, c] (one block per 64 columns, rows strided over the block).
__global__ void __launch_bounds__(256) col_sum_kernel(long rows, int cols, const float* d, long ld, float* out,
                                                      int accumulate) {
  __shared__ float part[4][64];
  const int c = blockIdx.x * 64 + threadIdx.x % 64;
  const int rg = threadIdx.x / 64;
  float s = 0.f;
  if (c < cols)
    for (long r = rg; r < rows; r += 4) s += d[r * ld + c];
  part[rg][threadIdx.x % 64] = s;
  __syncthreads();
  if (rg == 0 && c < cols) {
    const float v = (part[0][threadIdx.x] + part[1][threadIdx.x]) + (part[2][threadIdx.x] + part[3][threadIdx.x]);
    out[c] = accumulate ? out[c] + v : v;
  }
}

// Mean-squared error over the first `cols` outputs: dY = 2 (Y - T) / (rows cols) (fp32, row
// pitch ldy) and the loss summed into *loss (one atomic per block).
__global__ void __launch_bounds__(256) mse_kernel(long rows, int cols, const float* Y, long ldy, const float* T,
                                                  long ldt, float* dY, float* loss) {
  __shared__ float red[256];
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (e < rows * cols) {
    const long r = e / cols;
    const int c = (int)(e % cols);
    const float d = Y[r * ldy + c] - T[r * ldt + c];
    dY[r * ldy + c] = 2.0f * d / (float)(rows * cols);
    l = d * d / (float)(rows * cols);
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// Adam (Kingma & Ba 2015) on fp32 master parameters, with the bf16 copy the products read.
__global__ void adam_kernel(long n, float* p, const float* g, float* m, float* v, __bf16* pb, float lr, float b1,
                            float b2, float eps, float bc1, float bc2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  p[i] = pi;
  if (pb) pb[i] = (__bf16)pi;
}

__global__ void f32_to_bf16_kernel(long n, const float* a, __bf16* b) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (__bf16)a[i];
}

}  // namespace prk
#endif
