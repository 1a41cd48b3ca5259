// bmx_mlp.hip -- gfx950 kernels of the dynamic-aperture surrogate: a BF16
// matrix-core GEMM with FP32 accumulation and fused epilogues (bias, SiLU,
// SiLU backward), a bias-gradient column reduction and an Adam step.  The
// surrogate's forward and backward passes are compositions of these calls
// (bmad-ecosystem_amd/bmx/surrogate.py).  No counterpart in the reference.
//
// GEMM tile: 256 threads = 4 waves per 64 x 64 output tile; K steps of 32
// through LDS; each wave owns a 32 x 32 quadrant as 2 x 2 blocks of
// v_mfma_f32_16x16x32_bf16 (A: 16 x 32, B: 32 x 16, C: 16 x 16 FP32).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>

#include "../../include/bmx.h"

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4_t;

namespace {

constexpr int BM = 64, BN = 64, BK = 32;
constexpr int LDA_S = BK + 8;  // LDS row pitch (bf16) for A[m][k]
constexpr int LDB_S = BK + 8;  // LDS row pitch (bf16) for B^T[n][k]

__device__ __forceinline__ float silu(float x) { return x / (1.0f + __expf(-x)); }
__device__ __forceinline__ float silu_grad(float x) {
  float s = 1.0f / (1.0f + __expf(-x));
  return s * (1.0f + x * (1.0f - s));
}

// C[M x N] = op(A)[M x K] op(B)[K x N] with op = transpose when TA / TB.
// A element (m, k) is A[k * lda + m] when TA else A[m * lda + k]; same for B.
// Epilogue on acc (+ beta_acc * C_old when C_old != nullptr):
//   0: out = acc
//   1: out = acc + bias[n]
//   2: pre = acc + bias[n] -> C (fp32, if given); C_bf16 = silu(pre)
//   3: out = acc * silu'(aux[m, n])   (aux = pre-activation, fp32, pitch ldc)
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, const __bf16* __restrict__ A, int lda,
                                                       const __bf16* __restrict__ B, int ldb, float* __restrict__ C,
                                                       __bf16* __restrict__ Cb, int ldc, const float* __restrict__ bias,
                                                       int epi, const float* __restrict__ aux) {
  __shared__ __bf16 As[BM * LDA_S];
  __shared__ __bf16 Bs[BN * LDB_S];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;  // wave quadrant
  f32x4_t acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};

  for (int k0 = 0; k0 < K; k0 += BK) {
    // stage the A tile (BM x BK) as As[m][k] and the B tile (BK x BN) as
    // Bs[n][k]; consecutive threads walk the operand's contiguous index.
    for (int e = tid; e < BM * BK; e += 256) {
      int m = TA ? e % BM : e / BK, k = TA ? e / BM : e % BK;
      int gm = m0 + m, gk = k0 + k;
      __bf16 v = (__bf16)0.0f;
      if (gm < M && gk < K) v = TA ? A[(int64_t)gk * lda + gm] : A[(int64_t)gm * lda + gk];
      As[m * LDA_S + k] = v;
    }
    for (int e = tid; e < BN * BK; e += 256) {
      int n = TB ? e / BK : e % BN, k = TB ? e % BK : e / BN;
      int gn = n0 + n, gk = k0 + k;
      __bf16 v = (__bf16)0.0f;
      if (gn < N && gk < K) v = TB ? B[(int64_t)gn * ldb + gk] : B[(int64_t)gk * ldb + gn];
      Bs[n * LDB_S + k] = v;
    }
    __syncthreads();
    // lane -> (row / column r = lane % 16, k group q = lane / 16 holding k = 8q .. 8q+7)
    const int r = lane & 15, q = lane >> 4;
    bf16x8_t af[2], bfr[2];
    for (int i = 0; i < 2; i++) {
      const __bf16* pa = &As[(wm + 16 * i + r) * LDA_S + 8 * q];
      const __bf16* pb = &Bs[(wn + 16 * i + r) * LDB_S + 8 * q];
      for (int t = 0; t < 8; t++) {
        af[i][t] = pa[t];
        bfr[i][t] = pb[t];
      }
    }
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // C fragment: lane holds C[4 (lane / 16) + t][lane % 16], t = 0..3
  const int cn = lane & 15, cq = lane >> 4;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int t = 0; t < 4; t++) {
        int m = m0 + wm + 16 * i + 4 * cq + t;
        int n = n0 + wn + 16 * j + cn;
        if (m >= M || n >= N) continue;
        float v = acc[i][j][t];
        int64_t o = (int64_t)m * ldc + n;
        if (epi == 1 || epi == 2) v += bias[n];
        if (epi == 3) v *= silu_grad(aux[o]);
        if (epi == 2) {
          if (C) C[o] = v;
          if (Cb) Cb[o] = (__bf16)silu(v);
        } else {
          if (C) C[o] = v;
          if (Cb) Cb[o] = (__bf16)v;
        }
      }
}

// column sums of a row-major fp32 matrix: out[n] = sum_m X[m * ld + n]
__global__ __launch_bounds__(256) void colsum_kernel(int M, int N, const float* __restrict__ X, int ld,
                                                    float* __restrict__ out) {
  int n = blockIdx.x;
  float s = 0.f;
  for (int m = threadIdx.x; m < M; m += 256) s += X[(int64_t)m * ld + n];
  __shared__ float red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[n] = red[0];
}

// Adam (Kingma & Ba 2015) on fp32 master weights, refreshing the bf16 copy.
__global__ __launch_bounds__(256) void adam_kernel(int64_t n, float* __restrict__ w, __bf16* __restrict__ wb,
                                                  const float* __restrict__ g, float* __restrict__ m,
                                                  float* __restrict__ v, float lr, float b1, float b2, float eps,
                                                  float c1, float c2) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = (__bf16)wi;
}

__global__ __launch_bounds__(256) void f32_to_bf16_kernel(int64_t n, const float* __restrict__ x,
                                                         __bf16* __restrict__ y) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = (__bf16)x[i];
}

// Mean-squared-error loss and its gradient, one workgroup:
// loss = sum_i (p_i - t_i)^2 / n, d_i = 2 (p_i - t_i) / n (fp32 and bf16).
__global__ __launch_bounds__(1024) void mse_kernel(int n, const float* __restrict__ p, const float* __restrict__ t,
                                                   float* __restrict__ d, __bf16* __restrict__ db,
                                                   float* __restrict__ loss) {
  float s = 0.f, inv = 1.0f / (float)n;
  for (int i = threadIdx.x; i < n; i += 1024) {
    float e = p[i] - t[i];
    s += e * e;
    float g = 2.0f * e * inv;
    d[i] = g;
    db[i] = (__bf16)g;
  }
  __shared__ float red[1024];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *loss = red[0] * inv;
}

}  // namespace

// defined in bmx_track.hip (hidden visibility)
extern "C" hipStream_t bmx_ctx_stream(const bmx_ctx* ctx);
extern "C" int32_t bmx_ctx_set_error(bmx_ctx* ctx, const char* msg);

extern "C" {

int32_t bmx_gemm_bf16(bmx_ctx* ctx, int32_t trans_a, int32_t trans_b, int32_t M, int32_t N, int32_t K, const void* A,
                      int32_t lda, const void* B, int32_t ldb, float* C, void* C_bf16, int32_t ldc, const float* bias,
                      int32_t epilogue, const float* aux) {
  if (!ctx || !A || !B || M <= 0 || N <= 0 || K <= 0 || (!C && !C_bf16))
    return bmx_ctx_set_error(ctx, "bmx_gemm_bf16: invalid arguments");
  if ((epilogue == 1 || epilogue == 2) && !bias) return bmx_ctx_set_error(ctx, "bmx_gemm_bf16: bias required");
  if (epilogue == 3 && !aux) return bmx_ctx_set_error(ctx, "bmx_gemm_bf16: aux required");
  if (epilogue < 0 || epilogue > 3) return bmx_ctx_set_error(ctx, "bmx_gemm_bf16: bad epilogue");
  dim3 grid((unsigned)((N + BN - 1) / BN), (unsigned)((M + BM - 1) / BM));
  hipStream_t s = bmx_ctx_stream(ctx);
  const __bf16* a = (const __bf16*)A;
  const __bf16* b = (const __bf16*)B;
  __bf16* cb = (__bf16*)C_bf16;
  if (!trans_a && !trans_b)
    hipLaunchKernelGGL((gemm_bf16_kernel<false, false>), grid, dim3(256), 0, s, M, N, K, a, lda, b, ldb, C, cb, ldc,
                       bias, epilogue, aux);
  else if (!trans_a && trans_b)
    hipLaunchKernelGGL((gemm_bf16_kernel<false, true>), grid, dim3(256), 0, s, M, N, K, a, lda, b, ldb, C, cb, ldc,
                       bias, epilogue, aux);
  else if (trans_a && !trans_b)
    hipLaunchKernelGGL((gemm_bf16_kernel<true, false>), grid, dim3(256), 0, s, M, N, K, a, lda, b, ldb, C, cb, ldc,
                       bias, epilogue, aux);
  else
    hipLaunchKernelGGL((gemm_bf16_kernel<true, true>), grid, dim3(256), 0, s, M, N, K, a, lda, b, ldb, C, cb, ldc,
                       bias, epilogue, aux);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bmx_ctx_set_error(ctx, hipGetErrorString(e));
  return 0;
}

int32_t bmx_colsum_f32(bmx_ctx* ctx, int32_t M, int32_t N, const float* X, int32_t ld, float* out) {
  if (!ctx || !X || !out || M <= 0 || N <= 0) return bmx_ctx_set_error(ctx, "bmx_colsum_f32: invalid arguments");
  hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)N), dim3(256), 0, bmx_ctx_stream(ctx), M, N, X, ld, out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : bmx_ctx_set_error(ctx, hipGetErrorString(e));
}

int32_t bmx_adam_step(bmx_ctx* ctx, int64_t n, float* w, void* w_bf16, const float* g, float* m, float* v, float lr,
                      float beta1, float beta2, float eps, int32_t step) {
  if (!ctx || !w || !g || !m || !v || n <= 0 || step < 1) return bmx_ctx_set_error(ctx, "bmx_adam_step: invalid arguments");
  float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, bmx_ctx_stream(ctx), n, w,
                     (__bf16*)w_bf16, g, m, v, lr, beta1, beta2, eps, c1, c2);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : bmx_ctx_set_error(ctx, hipGetErrorString(e));
}

int32_t bmx_f32_to_bf16(bmx_ctx* ctx, int64_t n, const float* x, void* y) {
  if (!ctx || !x || !y || n <= 0) return bmx_ctx_set_error(ctx, "bmx_f32_to_bf16: invalid arguments");
  hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, bmx_ctx_stream(ctx), n, x,
                     (__bf16*)y);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : bmx_ctx_set_error(ctx, hipGetErrorString(e));
}

int32_t bmx_mse_grad(bmx_ctx* ctx, int32_t n, const float* pred, const float* target, float* d, void* d_bf16,
                     float* loss) {
  if (!ctx || !pred || !target || !d || !d_bf16 || !loss || n <= 0)
    return bmx_ctx_set_error(ctx, "bmx_mse_grad: invalid arguments");
  hipLaunchKernelGGL(mse_kernel, dim3(1), dim3(1024), 0, bmx_ctx_stream(ctx), n, pred, target, d, (__bf16*)d_bf16,
                     loss);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : bmx_ctx_set_error(ctx, hipGetErrorString(e));
}

int32_t bmx_ctx_sync(bmx_ctx* ctx) {
  if (!ctx) return -1;
  hipError_t e = hipStreamSynchronize(bmx_ctx_stream(ctx));
  return e == hipSuccess ? 0 : bmx_ctx_set_error(ctx, hipGetErrorString(e));
}

}  // extern "C"
