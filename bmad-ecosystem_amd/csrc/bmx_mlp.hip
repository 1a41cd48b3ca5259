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

#include <array>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <utility>

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

// Epilogue on the FP32 accumulator v of C(m, n):
//   0: out = v
//   1: out = v + bias[n]
//   2: pre = v + bias[n] -> C (fp32, if given); C_bf16 = silu(pre)
//   3: out = v * silu'(aux[m, n])   (aux = pre-activation, fp32, pitch ldc)
__device__ __forceinline__ void epilogue_store(float v, int64_t o, int n, int epi, float* __restrict__ C,
                                               __bf16* __restrict__ Cb, const float* __restrict__ bias,
                                               const float* __restrict__ aux) {
  if (epi == 1 || epi == 2) v += bias[n];
  if (epi == 3) v *= silu_grad(aux[o]);
  if (C) C[o] = v;
  if (Cb) Cb[o] = (__bf16)((epi == 2) ? silu(v) : v);
}

typedef __attribute__((__vector_size__(8 * sizeof(unsigned short)))) unsigned short u16x8_t;

// Stage a 64 x 32 operand tile into LDS as T[r][k] (r = row of the output
// tile side, k = reduction index), from a row-major matrix whose element
// (r, k) lives at P[k * ld + r] when KMAJOR (r contiguous) or P[r * ld + k]
// (k contiguous).  VEC: 16-byte loads (the caller checked alignment and that
// the contiguous extent is a multiple of 8).
template <bool KMAJOR, bool VEC>
__device__ __forceinline__ void stage_tile(__bf16* __restrict__ T, const __bf16* __restrict__ P, int ld, int r0,
                                           int R, int k0, int K, int tid) {
  if (VEC) {
    if (!KMAJOR) {  // k contiguous: thread -> (r = tid / 4, 8 k's)
      int r = tid >> 2, kk = (tid & 3) * 8;
      int gr = r0 + r, gk = k0 + kk;
      u16x8_t v = {0, 0, 0, 0, 0, 0, 0, 0};
      if (gr < R && gk < K) v = *reinterpret_cast<const u16x8_t*>(P + (int64_t)gr * ld + gk);
      *reinterpret_cast<u16x8_t*>(T + r * LDA_S + kk) = v;
    } else {  // r contiguous: thread -> (k = tid / 8, 8 r's)
      int kk = tid >> 3, r = (tid & 7) * 8;
      int gr = r0 + r, gk = k0 + kk;
      u16x8_t v = {0, 0, 0, 0, 0, 0, 0, 0};
      if (gr < R && gk < K) v = *reinterpret_cast<const u16x8_t*>(P + (int64_t)gk * ld + gr);
      unsigned short* t = reinterpret_cast<unsigned short*>(T);
      for (int j = 0; j < 8; j++) t[(r + j) * LDA_S + kk] = v[j];
    }
  } else {
    for (int e = tid; e < BM * BK; e += 256) {
      int r = KMAJOR ? e % BM : e / BK, kk = KMAJOR ? e / BM : e % BK;
      int gr = r0 + r, gk = k0 + kk;
      __bf16 v = (__bf16)0.0f;
      if (gr < R && gk < K) v = KMAJOR ? P[(int64_t)gk * ld + gr] : P[(int64_t)gr * ld + gk];
      T[r * LDA_S + kk] = v;
    }
  }
}

// C[M x N] = op(A)[M x K] op(B)[K x N] with op = transpose when TA / TB.
// A element (m, k) is A[k * lda + m] when TA else A[m * lda + k]; B element
// (k, n) is B[n * ldb + k] when TB else B[k * ldb + n].  blockIdx.z = K
// split: with ksplit > 1 each split writes its raw partial tile to
// W[z][m][n] and splitk_reduce_kernel applies the epilogue.
template <bool TA, bool TB, bool VA, bool VB>
__global__ __launch_bounds__(256) void gemm_bf16_kernel(int M, int N, int K, int k_chunk, const __bf16* __restrict__ A,
                                                       int lda, const __bf16* __restrict__ B, int ldb,
                                                       float* __restrict__ C, __bf16* __restrict__ Cb, int ldc,
                                                       const float* __restrict__ bias, int epi,
                                                       const float* __restrict__ aux, float* __restrict__ W) {
  __shared__ __bf16 As[BM * LDA_S];
  __shared__ __bf16 Bs[BN * LDB_S];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;  // wave quadrant
  const int kb = blockIdx.z * k_chunk, ke = min(K, kb + k_chunk);
  f32x4_t acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};

  for (int k0 = kb; k0 < ke; k0 += BK) {
    stage_tile<TA, VA>(As, A, lda, m0, M, k0, ke, tid);
    stage_tile<!TB, VB>(Bs, B, ldb, n0, N, k0, ke, tid);
    __syncthreads();
    // lane -> (row / column r = lane % 16, k group q = lane / 16 holding k = 8q .. 8q+7)
    const int r = lane & 15, q = lane >> 4;
    bf16x8_t af[2], bfr[2];
    for (int i = 0; i < 2; i++) {
      af[i] = *reinterpret_cast<const bf16x8_t*>(&As[(wm + 16 * i + r) * LDA_S + 8 * q]);
      bfr[i] = *reinterpret_cast<const bf16x8_t*>(&Bs[(wn + 16 * i + r) * LDB_S + 8 * q]);
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
        if (W) W[((int64_t)blockIdx.z * M + m) * N + n] = v;
        else epilogue_store(v, (int64_t)m * ldc + n, n, epi, C, Cb, bias, aux);
      }
}

// Sum of the K-split partial tiles (in split order, so the result does not
// depend on scheduling) followed by the epilogue.
__global__ __launch_bounds__(256) void splitk_reduce_kernel(int M, int N, int ksplit, const float* __restrict__ W,
                                                           float* __restrict__ C, __bf16* __restrict__ Cb, int ldc,
                                                           const float* __restrict__ bias, int epi,
                                                           const float* __restrict__ aux) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)M * N) return;
  int m = (int)(i / N), n = (int)(i % N);
  float v = 0.f;
  for (int z = 0; z < ksplit; z++) v += W[(int64_t)z * M * N + i];
  epilogue_store(v, (int64_t)m * ldc + n, n, epi, C, Cb, bias, aux);
}

// Column sums of a row-major fp32 matrix, out[n] = sum_m X[m * ld + n] (the
// bias gradients), in two passes with a fixed summation order.  Pass 1 gives
// one partial per (row block, column): for N >= 32 a workgroup sums
// COLSUM_ROWS rows of 256 consecutive columns, lane n reading column n (each
// row a coalesced load); for narrow matrices a workgroup sums COLSUM_ROWS1
// rows of one column.  Pass 2 adds each column's partials.
constexpr int COLSUM_ROWS = 16;
constexpr int COLSUM_ROWS1 = 1024;

__global__ __launch_bounds__(256) void colsum_partial_kernel(int M, int N, const float* __restrict__ X, int ld,
                                                            float* __restrict__ P) {
  int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  int m0 = blockIdx.y * COLSUM_ROWS, m1 = min(M, m0 + COLSUM_ROWS);
  float s = 0.f;
  for (int m = m0; m < m1; m++) s += X[(int64_t)m * ld + n];
  P[(int64_t)blockIdx.y * N + n] = s;
}

// narrow pass 1: workgroup (b, n) sums rows [b COLSUM_ROWS1, (b+1) COLSUM_ROWS1) of column n
__global__ __launch_bounds__(256) void colsum_partial1_kernel(int M, int N, const float* __restrict__ X, int ld,
                                                             float* __restrict__ P) {
  int n = blockIdx.y;
  int m0 = blockIdx.x * COLSUM_ROWS1;
  float s = 0.f;
  for (int k = 0; k < COLSUM_ROWS1 / 256; k++) {
    int m = m0 + k * 256 + (int)threadIdx.x;
    if (m < M) s += X[(int64_t)m * ld + n];
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  __shared__ float red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) P[(int64_t)blockIdx.x * N + n] = (red[0] + red[1]) + (red[2] + red[3]);
}

// one wave per column: lane j adds the partials j, j + 64, ..., then a fixed
// shuffle tree over the wave (the column index is uniform over a wave, so a
// wave leaves or stays whole)
__global__ __launch_bounds__(256) void colsum_final_kernel(int nb, int N, const float* __restrict__ P,
                                                          float* __restrict__ out) {
  int n = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  int j = (int)(threadIdx.x & 63);
  if (n >= N) return;
  float s = 0.f;
  for (int b = j; b < nb; b += 64) s += P[(int64_t)b * N + n];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (j == 0) out[n] = s;
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

// Surrogate inputs of a DA scan, straight to BF16: row l * n_ic + i holds the
// knob vector of lattice l scaled by 1 / knob_range, then the n_geo grid
// coordinates of IC i.
__global__ __launch_bounds__(256) void da_features_kernel(int64_t n_rows, int n_ic, int n_knob, int n_geo,
                                                         const float* __restrict__ knobs, float inv_range,
                                                         const float* __restrict__ geo, __bf16* __restrict__ X) {
  int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (row >= n_rows) return;
  int64_t l = row / n_ic;
  int i = (int)(row % n_ic);
  int w = n_knob + n_geo;
  __bf16* x = X + row * w;
  for (int k = 0; k < n_knob; k++) x[k] = (__bf16)(knobs[l * n_knob + k] * inv_range);
  for (int g = 0; g < n_geo; g++) x[n_knob + g] = (__bf16)geo[(int64_t)i * n_geo + g];
}

// Least-squares step fit of each ray's predicted survival profile: the k in
// 0..K minimising sum_{j<k} (p_j - alive)^2 + sum_{j>=k} (p_j - mean_{j>=k} p)^2
// (a step from the surviving level to a free lower level), ties to the
// larger k.  One lane per ray, sums in FP64; a workgroup stages its RAYS
// profiles through LDS (coalesced loads, rows padded by one word against
// bank conflicts).
template <int RAYS>
__global__ __launch_bounds__(RAYS) void step_fit_kernel(int n_rays, int K, const float* __restrict__ pred, float alive,
                                                        int32_t* __restrict__ out) {
  extern __shared__ float tile[];
  int r0 = blockIdx.x * RAYS;
  int nr = min(RAYS, n_rays - r0);
  int ld = K + 1;
  for (int e = threadIdx.x; e < nr * K; e += RAYS) tile[(e / K) * ld + e % K] = pred[(int64_t)r0 * K + e];
  __syncthreads();
  if ((int)threadIdx.x >= nr) return;
  const float* p = tile + threadIdx.x * ld;
  double S = 0.0, Q = 0.0;
  for (int j = 0; j < K; j++) {
    double v = p[j];
    S += v;
    Q += v * v;
  }
  double pre = 0.0, ps = 0.0, pq = 0.0, best = __builtin_inf();
  int kbest = 0;
  for (int k = 0; k <= K; k++) {
    int cnt = K - k;
    double sk = S - ps, qk = Q - pq;
    double cost = pre + (cnt > 0 ? qk - sk * sk / cnt : 0.0);
    if (cost <= best) {
      best = cost;
      kbest = k;
    }
    if (k < K) {
      double v = p[k], d = v - (double)alive;
      pre += d * d;
      ps += v;
      pq += v * v;
    }
  }
  out[r0 + threadIdx.x] = kbest;
}

}  // namespace

// defined in bmx_track.hip (hidden visibility)
extern "C" hipStream_t bmx_ctx_stream(const bmx_ctx* ctx);
extern "C" int32_t bmx_ctx_set_error(bmx_ctx* ctx, const char* msg);

// per-context workspaces (slot 0: split-K partial products, slot 1:
// column-sum partials), grown on demand, released by bmx_ctx_destroy
static std::mutex g_ws_mu;
static std::unordered_map<const bmx_ctx*, std::array<std::pair<float*, size_t>, 2>> g_ws;

static float* splitk_workspace(const bmx_ctx* ctx, size_t floats, int slot = 0) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  auto& e = g_ws[ctx][slot];
  if (e.second < floats) {
    if (e.first) (void)hipFree(e.first);
    e.first = nullptr;
    e.second = 0;
    if (hipMalloc(&e.first, floats * sizeof(float)) != hipSuccess) return nullptr;
    e.second = floats;
  }
  return e.first;
}

template <bool TA, bool TB>
static void launch_gemm(dim3 grid, hipStream_t s, bool va, bool vb, int M, int N, int K, int k_chunk, const __bf16* a,
                        int lda, const __bf16* b, int ldb, float* C, __bf16* cb, int ldc, const float* bias, int epi,
                        const float* aux, float* W) {
#define BMX_GEMM_LAUNCH(VA_, VB_)                                                                                \
  hipLaunchKernelGGL((gemm_bf16_kernel<TA, TB, VA_, VB_>), grid, dim3(256), 0, s, M, N, K, k_chunk, a, lda, b, ldb, \
                     C, cb, ldc, bias, epi, aux, W)
  if (va && vb) BMX_GEMM_LAUNCH(true, true);
  else if (va) BMX_GEMM_LAUNCH(true, false);
  else if (vb) BMX_GEMM_LAUNCH(false, true);
  else BMX_GEMM_LAUNCH(false, false);
#undef BMX_GEMM_LAUNCH
}

// Called by bmx_ctx_destroy (hidden: not part of the exported ABI).
extern "C" __attribute__((visibility("hidden"))) void bmx_mlp_release(const bmx_ctx* ctx) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  auto it = g_ws.find(ctx);
  if (it == g_ws.end()) return;
  for (auto& e : it->second)
    if (e.first) (void)hipFree(e.first);
  g_ws.erase(it);
}

extern "C" {

int32_t bmx_gemm_bf16(bmx_ctx* ctx, int32_t trans_a, int32_t trans_b, int32_t M, int32_t N, int32_t K, const void* A,
                      int32_t lda, const void* B, int32_t ldb, float* C, void* C_bf16, int32_t ldc, const float* bias,
                      int32_t epilogue, const float* aux) {
  if (!ctx || !A || !B || M <= 0 || N <= 0 || K <= 0 || (!C && !C_bf16))
    return bmx_ctx_set_error(ctx, "bmx_gemm_bf16: invalid arguments");
  if ((epilogue == 1 || epilogue == 2) && !bias) return bmx_ctx_set_error(ctx, "bmx_gemm_bf16: bias required");
  if (epilogue == 3 && !aux) return bmx_ctx_set_error(ctx, "bmx_gemm_bf16: aux required");
  if (epilogue < 0 || epilogue > 3) return bmx_ctx_set_error(ctx, "bmx_gemm_bf16: bad epilogue");
  hipStream_t s = bmx_ctx_stream(ctx);
  const __bf16* a = (const __bf16*)A;
  const __bf16* b = (const __bf16*)B;
  __bf16* cb = (__bf16*)C_bf16;
  // 16-byte staging needs the contiguous extent and leading dimension in
  // multiples of 8 elements and 16-byte aligned base pointers
  auto al16 = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  bool va = al16(A) && (lda % 8 == 0) && ((trans_a ? M : K) % 8 == 0);
  bool vb = al16(B) && (ldb % 8 == 0) && ((trans_b ? K : N) % 8 == 0);
  int tiles = ((N + BN - 1) / BN) * ((M + BM - 1) / BM);
  // split K when the output has too few tiles to fill the chip
  int ksplit = 1;
  if (tiles < 512 && K >= 2048) ksplit = std::min(std::min(K / 512, (1024 + tiles - 1) / tiles), 64);
  int k_chunk = ((K + ksplit - 1) / ksplit + BK - 1) / BK * BK;
  ksplit = (K + k_chunk - 1) / k_chunk;
  float* W = nullptr;
  if (ksplit > 1) {
    W = splitk_workspace(ctx, (size_t)ksplit * M * N);
    if (!W) return bmx_ctx_set_error(ctx, "bmx_gemm_bf16: split-K workspace allocation failed");
  }
  dim3 grid((unsigned)((N + BN - 1) / BN), (unsigned)((M + BM - 1) / BM), (unsigned)ksplit);
  if (!trans_a && !trans_b)
    launch_gemm<false, false>(grid, s, va, vb, M, N, K, k_chunk, a, lda, b, ldb, C, cb, ldc, bias, epilogue, aux, W);
  else if (!trans_a && trans_b)
    launch_gemm<false, true>(grid, s, va, vb, M, N, K, k_chunk, a, lda, b, ldb, C, cb, ldc, bias, epilogue, aux, W);
  else if (trans_a && !trans_b)
    launch_gemm<true, false>(grid, s, va, vb, M, N, K, k_chunk, a, lda, b, ldb, C, cb, ldc, bias, epilogue, aux, W);
  else
    launch_gemm<true, true>(grid, s, va, vb, M, N, K, k_chunk, a, lda, b, ldb, C, cb, ldc, bias, epilogue, aux, W);
  if (ksplit > 1) {
    int64_t total = (int64_t)M * N;
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, M, N, ksplit, W,
                       C, cb, ldc, bias, epilogue, aux);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return bmx_ctx_set_error(ctx, hipGetErrorString(e));
  return 0;
}

int32_t bmx_colsum_f32(bmx_ctx* ctx, int32_t M, int32_t N, const float* X, int32_t ld, float* out) {
  if (!ctx || !X || !out || M <= 0 || N <= 0 || ld < N)
    return bmx_ctx_set_error(ctx, "bmx_colsum_f32: invalid arguments");
  hipStream_t s = bmx_ctx_stream(ctx);
  bool wide = N >= 32;
  int rows = wide ? COLSUM_ROWS : COLSUM_ROWS1;
  int nb = (M + rows - 1) / rows;
  float* P = splitk_workspace(ctx, (size_t)nb * N, 1);
  if (!P) return bmx_ctx_set_error(ctx, "bmx_colsum_f32: workspace allocation failed");
  if (wide)
    hipLaunchKernelGGL(colsum_partial_kernel, dim3((unsigned)((N + 255) / 256), (unsigned)nb), dim3(256), 0, s, M, N,
                       X, ld, P);
  else
    hipLaunchKernelGGL(colsum_partial1_kernel, dim3((unsigned)nb, (unsigned)N), dim3(256), 0, s, M, N, X, ld, P);
  hipLaunchKernelGGL(colsum_final_kernel, dim3((unsigned)((N + 3) / 4)), dim3(256), 0, s, nb, N, P, out);
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

int32_t bmx_da_features(bmx_ctx* ctx, int32_t n_lat, int32_t n_ic, int32_t n_knob, int32_t n_geo,
                        const float* knobs, float knob_range, const float* geo, void* X_bf16) {
  if (!ctx || !knobs || !geo || !X_bf16 || n_lat <= 0 || n_ic <= 0 || n_knob < 0 || n_geo < 0 || knob_range <= 0)
    return bmx_ctx_set_error(ctx, "bmx_da_features: invalid arguments");
  int64_t rows = (int64_t)n_lat * n_ic;
  hipLaunchKernelGGL(da_features_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, bmx_ctx_stream(ctx), rows,
                     n_ic, n_knob, n_geo, knobs, 1.0f / knob_range, geo, (__bf16*)X_bf16);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : bmx_ctx_set_error(ctx, hipGetErrorString(e));
}

int32_t bmx_step_fit(bmx_ctx* ctx, int32_t n_rays, int32_t n_amp, const float* pred, float alive_level,
                     int32_t* k_out) {
  if (!ctx || !pred || !k_out || n_rays <= 0 || n_amp <= 0 || n_amp > 2048)
    return bmx_ctx_set_error(ctx, "bmx_step_fit: invalid arguments");
  // rays per workgroup so that the staged profiles stay within 33 KiB of LDS
  hipStream_t st = bmx_ctx_stream(ctx);
  size_t row = (size_t)(n_amp + 1) * sizeof(float);
  if (n_amp <= 128)
    hipLaunchKernelGGL(step_fit_kernel<64>, dim3((unsigned)((n_rays + 63) / 64)), dim3(64), 64 * row, st, n_rays,
                       n_amp, pred, alive_level, k_out);
  else if (n_amp <= 256)
    hipLaunchKernelGGL(step_fit_kernel<32>, dim3((unsigned)((n_rays + 31) / 32)), dim3(32), 32 * row, st, n_rays,
                       n_amp, pred, alive_level, k_out);
  else
    hipLaunchKernelGGL(step_fit_kernel<4>, dim3((unsigned)((n_rays + 3) / 4)), dim3(4), 4 * row, st, n_rays, n_amp,
                       pred, alive_level, k_out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : bmx_ctx_set_error(ctx, hipGetErrorString(e));
}

int32_t bmx_ctx_sync(bmx_ctx* ctx) {
  if (!ctx) return -1;
  hipError_t e = hipStreamSynchronize(bmx_ctx_stream(ctx));
  return e == hipSuccess ? 0 : bmx_ctx_set_error(ctx, hipGetErrorString(e));
}

}  // extern "C"
