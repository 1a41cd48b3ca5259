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

#include <algorithm>
#include <array>
#include <vector>
#include <random>
#include <cmath>
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

// SiLU x sigma(x) with sigma = 1 / (1 + e^-x) from the hardware exponential and
// reciprocal (v_exp_f32, v_rcp_f32: about 1 ulp each), the activation's cost
// in the epilogues being what bounds the fused forward pass; every kernel
// (GEMM epilogues, fused pass, backward) uses these same two functions.
__device__ __forceinline__ float sigmoid(float x) { return __builtin_amdgcn_rcpf(1.0f + __expf(-x)); }
__device__ __forceinline__ float silu(float x) { return x * sigmoid(x); }
__device__ __forceinline__ float silu_grad(float x) {
  float s = sigmoid(x);
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

// ---------------------------------------------------------------------------
// The whole forward pass in one kernel, for inference: a workgroup takes 64
// rows and keeps their activations in LDS through every layer, so that only
// the inputs are read from and one output per row written to HBM.  Hidden
// layers are FH = 256 wide; the input is at most FK0 = 64 wide (zero-padded),
// given as BF16 rows or formed in place from a DA scan's knob vectors and
// grid coordinates (as da_features_kernel forms them).  Wave w computes output
// columns 64 w .. 64 w + 63 of every layer for all 64 rows as 4 x 4 blocks of
// v_mfma_f32_16x16x32_bf16, A from LDS, B (the weights, W[n][k] row-major)
// straight from L2; the next k-step's weights are loaded before this one's
// MFMAs.  The MFMA operands, their k order and the epilogues (bias, SiLU,
// BF16 rounding) are those of gemm_bf16_kernel, so the outputs equal the
// layer-by-layer pass bit for bit.  The linear head (one output) is a 16 x 16
// block per wave whose B holds the head weights in column 0.
constexpr int FH = 256, FK0 = 64, FP = FH + 8;

struct FusedNet {
  int d0;            // input width, <= FK0
  int n_hidden;      // hidden layers, each FH wide
  const __bf16* wb;  // BF16 parameters, per layer W [out][in] then b [out]
  const float* w;    // FP32 master parameters (the biases are read here)
};

struct FusedIn {
  const __bf16* X;  // rows of d0 BF16 values; null: DA features below
  const float* knobs;
  float inv_range;
  int n_knob;
  const float* geo;
  int n_geo;
  int n_ic;
};

__device__ __forceinline__ bf16x8_t load_w8(const __bf16* __restrict__ W, int64_t base, int k, int n_in, bool vec) {
  if (vec) return *reinterpret_cast<const bf16x8_t*>(W + base + k);
  bf16x8_t v;
  for (int e = 0; e < 8; e++) v[e] = (k + e < n_in) ? W[base + k + e] : (__bf16)0.0f;
  return v;
}

// 128 VGPRs (the accumulators included) so that four waves share each SIMD:
// one wave's SiLU epilogue and LDS stores overlap the others' MFMAs (at two
// waves per SIMD the matrix cores idled through the epilogues).
template <int RB>  // row blocks of 16 per workgroup: FR = 16 RB rows
__global__ __launch_bounds__(256, 4) void mlp_fused_forward_kernel(int64_t n, FusedNet net, FusedIn in,
                                                               float* __restrict__ y) {
  constexpr int FR = 16 * RB;
  __shared__ __bf16 act[FR * FP];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t row0 = (int64_t)blockIdx.x * FR;
  // DA features: the lattice l and grid point i of row0 by one division per
  // workgroup; the rows after it step through i and wrap into the next lattice
  int64_t l0 = 0;
  int i0 = 0;
  if (!in.X) {
    l0 = row0 / in.n_ic;
    i0 = (int)(row0 - l0 * in.n_ic);
  }
  for (int e = tid; e < FR * FK0; e += 256) {
    int m = e / FK0, k = e % FK0;
    int64_t row = row0 + m;
    __bf16 v = (__bf16)0.0f;
    if (row < n && k < net.d0) {
      if (in.X) {
        v = in.X[row * net.d0 + k];
      } else {
        int64_t l = l0;
        int i = i0 + m;
        if (i >= in.n_ic) {
          l += i / in.n_ic;
          i %= in.n_ic;
        }
        v = (k < in.n_knob) ? (__bf16)(in.knobs[l * in.n_knob + k] * in.inv_range)
                            : (__bf16)in.geo[(int64_t)i * in.n_geo + (k - in.n_knob)];
      }
    }
    act[m * FP + k] = v;
  }
  __syncthreads();
  const int r = lane & 15, q = lane >> 4;
  int64_t off = 0;
  int n_in = net.d0;
  for (int l = 0; l < net.n_hidden; l++) {
    const __bf16* W = net.wb + off;
    const float* bias = net.w + off + (int64_t)FH * n_in;
    const int K = (l == 0) ? FK0 : FH;
    const bool vec = l > 0;
    f32x4_t acc[RB][4];
    for (int i = 0; i < RB; i++)
      for (int j = 0; j < 4; j++) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    bf16x8_t bfr[4], nxt[4];
    for (int j = 0; j < 4; j++) bfr[j] = load_w8(W, (int64_t)(64 * wave + 16 * j + r) * n_in, 8 * q, n_in, vec);
    for (int k0 = 0; k0 < K; k0 += 32) {
      if (k0 + 32 < K)
        for (int j = 0; j < 4; j++)
          nxt[j] = load_w8(W, (int64_t)(64 * wave + 16 * j + r) * n_in, k0 + 32 + 8 * q, n_in, vec);
      for (int i = 0; i < RB; i++) {
        bf16x8_t af = *reinterpret_cast<const bf16x8_t*>(&act[(16 * i + r) * FP + k0 + 8 * q]);
        for (int j = 0; j < 4; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr[j], acc[i][j], 0, 0, 0);
      }
      for (int j = 0; j < 4; j++) bfr[j] = nxt[j];
    }
    __syncthreads();  // every wave has read this layer's input
    for (int i = 0; i < RB; i++)
      for (int j = 0; j < 4; j++) {
        int nn = 64 * wave + 16 * j + r;
        float b = bias[nn];
        for (int t = 0; t < 4; t++) act[(16 * i + 4 * q + t) * FP + nn] = (__bf16)silu(acc[i][j][t] + b);
      }
    __syncthreads();
    off += (int64_t)FH * n_in + FH;
    n_in = FH;
  }
  const __bf16* Wh = net.wb + off;
  const float bh = net.w[off + FH];
  for (int ib = wave; ib < RB; ib += 4) {
    f32x4_t acc = f32x4_t{0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < FH; k0 += 32) {
      bf16x8_t af = *reinterpret_cast<const bf16x8_t*>(&act[(16 * ib + r) * FP + k0 + 8 * q]);
      bf16x8_t b = {};
      if (r == 0) b = *reinterpret_cast<const bf16x8_t*>(Wh + k0 + 8 * q);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, b, acc, 0, 0, 0);
    }
    if (r == 0)
      for (int t = 0; t < 4; t++) {
        int64_t row = row0 + 16 * ib + 4 * q + t;
        if (row < n) y[row] = acc[t] + bh;
      }
  }
}

// Training steps of a bmx_net as a replayable graph: the step's indices live
// in device memory (TrainState), so that the captured kernels read the batch,
// the learning rate and Adam's bias corrections for the step they run as.
struct TrainState {
  int64_t t;  // global step count before this step (Adam's step is t + 1)
  int32_t k;  // step within the epoch
};

// xb[i] = bf16(X[perm[k batch + i]]), yb[i] = Y[perm[k batch + i]]
__global__ __launch_bounds__(256) void gather_batch_kernel(const float* __restrict__ X, const float* __restrict__ Y,
                                                          const int64_t* __restrict__ perm,
                                                          const TrainState* __restrict__ st, int batch, int d0,
                                                          __bf16* __restrict__ xb, float* __restrict__ yb) {
  int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)batch * d0) return;
  int i = (int)(e / d0), c = (int)(e % d0);
  int64_t src = perm[(int64_t)st->k * batch + i];
  xb[e] = (__bf16)X[src * d0 + c];
  if (c == 0) yb[i] = Y[src];
}

// adam_kernel with the step's learning rate (lr_tab[t - t0]) and bias
// corrections taken from the step count in device memory
__global__ __launch_bounds__(256) void adam_graph_kernel(int64_t n, float* __restrict__ w, __bf16* __restrict__ wb,
                                                        const float* __restrict__ g, float* __restrict__ m,
                                                        float* __restrict__ v, const float* __restrict__ lr_tab,
                                                        int64_t t0, const TrainState* __restrict__ st, float b1,
                                                        float b2, float eps) {
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int64_t t = st->t;
  float lr = lr_tab[t - t0];
  float c1 = 1.f - powf(b1, (float)(t + 1)), c2 = 1.f - powf(b2, (float)(t + 1));
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  wb[i] = (__bf16)wi;
}

__global__ void train_advance_kernel(TrainState* st) {
  st->t += 1;
  st->k += 1;
}

__global__ void train_reset_kernel(TrainState* st, int64_t t) {
  st->t = t;
  st->k = 0;
}

}  // namespace

// defined in bmx_track.hip (hidden visibility)
extern "C" hipStream_t bmx_ctx_stream(const bmx_ctx* ctx);
extern "C" int bmx_ctx_device(const bmx_ctx* ctx);
extern "C" int32_t bmx_ctx_set_error(bmx_ctx* ctx, const char* msg);

// per-context workspaces (slot 0: split-K partial products, slot 1:
// column-sum partials, slot 2: network outputs of a screening chunk), grown
// on demand, released by bmx_ctx_destroy
static std::mutex g_ws_mu;
static std::unordered_map<const bmx_ctx*, std::array<std::pair<float*, size_t>, 3>> g_ws;

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

// dims[0] <= FK0 -> FH -> ... -> FH -> 1 (at least one hidden layer): the shapes the fused pass takes
static bool fused_shape(int32_t n_dims, const int32_t* dims, FusedNet* net) {
  if (!dims || n_dims < 3 || dims[0] < 1 || dims[0] > FK0 || dims[n_dims - 1] != 1) return false;
  for (int i = 1; i < n_dims - 1; i++)
    if (dims[i] != FH) return false;
  net->d0 = dims[0];
  net->n_hidden = n_dims - 2;
  return true;
}

#ifndef BMX_FUSED_RB
#define BMX_FUSED_RB 4
#endif
static void launch_fused(hipStream_t s, int64_t n, const FusedNet& net, const FusedIn& in, float* y) {
  constexpr int RB = BMX_FUSED_RB, FR = 16 * RB;
  hipLaunchKernelGGL(mlp_fused_forward_kernel<RB>, dim3((unsigned)((n + FR - 1) / FR)), dim3(256), 0, s, n, net, in,
                     y);
}

static void launch_step_fit(hipStream_t st, int32_t n_rays, int32_t n_amp, const float* pred, float alive_level,
                            int32_t* k_out) {
  // rays per workgroup so that the staged profiles stay within 33 KiB of LDS
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
  launch_step_fit(bmx_ctx_stream(ctx), n_rays, n_amp, pred, alive_level, k_out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : bmx_ctx_set_error(ctx, hipGetErrorString(e));
}

int32_t bmx_mlp_forward(bmx_ctx* ctx, int32_t n_dims, const int32_t* dims, const float* w, const void* w_bf16,
                        const void* X_bf16, int64_t n, float* y) {
  FusedNet net;
  if (!ctx || !w || !w_bf16 || !X_bf16 || !y || n <= 0) return bmx_ctx_set_error(ctx, "bmx_mlp_forward: invalid arguments");
  if (!fused_shape(n_dims, dims, &net))
    return bmx_ctx_set_error(ctx, "bmx_mlp_forward: the network must be dims[0] <= 64 -> 256 -> ... -> 256 -> 1");
  if (((uintptr_t)w_bf16 & 15) != 0) return bmx_ctx_set_error(ctx, "bmx_mlp_forward: w_bf16 must be 16-byte aligned");
  net.wb = (const __bf16*)w_bf16;
  net.w = w;
  FusedIn in{};
  in.X = (const __bf16*)X_bf16;
  launch_fused(bmx_ctx_stream(ctx), n, net, in, y);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : bmx_ctx_set_error(ctx, hipGetErrorString(e));
}

int32_t bmx_da_screen(bmx_ctx* ctx, int32_t n_dims, const int32_t* dims, const float* w, const void* w_bf16,
                      int32_t n_lat, int32_t n_knob, const float* knobs, float knob_range, int32_t n_ic, int32_t n_geo,
                      const float* geo, int32_t n_amp, float alive_level, int32_t* kb) {
  FusedNet net;
  if (!ctx || !w || !w_bf16 || !knobs || !geo || !kb || n_lat <= 0 || n_knob < 0 || n_geo < 0 || n_ic <= 0 ||
      n_amp <= 0 || n_amp > 2048 || n_ic % n_amp != 0 || knob_range <= 0)
    return bmx_ctx_set_error(ctx, "bmx_da_screen: invalid arguments");
  if (!fused_shape(n_dims, dims, &net) || n_knob + n_geo != dims[0])
    return bmx_ctx_set_error(ctx, "bmx_da_screen: the network must be (n_knob + n_geo <= 64) -> 256 -> ... -> 256 -> 1");
  if (((uintptr_t)w_bf16 & 15) != 0) return bmx_ctx_set_error(ctx, "bmx_da_screen: w_bf16 must be 16-byte aligned");
  net.wb = (const __bf16*)w_bf16;
  net.w = w;
  // lattices per chunk: at most 2^22 network outputs in the workspace
  int64_t chunk = std::max<int64_t>(1, ((int64_t)1 << 22) / n_ic);
  chunk = std::min<int64_t>(chunk, n_lat);
  float* yw = splitk_workspace(ctx, (size_t)(chunk * n_ic), 2);
  if (!yw) return bmx_ctx_set_error(ctx, "bmx_da_screen: workspace allocation failed");
  hipStream_t s = bmx_ctx_stream(ctx);
  int rays_lat = n_ic / n_amp;
  for (int64_t a = 0; a < n_lat; a += chunk) {
    int64_t nl = std::min<int64_t>(chunk, n_lat - a);
    FusedIn in{};
    in.knobs = knobs + a * n_knob;
    in.inv_range = 1.0f / knob_range;
    in.n_knob = n_knob;
    in.geo = geo;
    in.n_geo = n_geo;
    in.n_ic = n_ic;
    launch_fused(s, nl * n_ic, net, in, yw);
    launch_step_fit(s, (int32_t)(nl * rays_lat), n_amp, yw, alive_level, kb + a * rays_lat);
  }
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? 0 : bmx_ctx_set_error(ctx, hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// bmx_net: a surrogate network owned by the library
// ---------------------------------------------------------------------------
}  // extern "C"

struct bmx_net {
  bmx_ctx* ctx = nullptr;
  int device = 0;
  std::vector<int32_t> dims;
  struct Layer {
    int64_t wo, bo;
    int n_in, n_out;
  };
  std::vector<Layer> layout;
  int64_t n_params = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr, *loss = nullptr;
  __bf16* wb = nullptr;
  int64_t step = 0;  // Adam steps taken
  // one resident batch size
  int nb = 0;
  __bf16 *xb = nullptr, *db = nullptr, *dpre_b[2] = {nullptr, nullptr};
  std::vector<float*> pre;
  std::vector<__bf16*> h;
  float *y = nullptr, *d = nullptr, *yb = nullptr, *dpre[2] = {nullptr, nullptr};
  std::vector<void*> batch_bufs;
  ~bmx_net() {
    for (void* p : batch_bufs) (void)hipFree(p);
    for (void* p : {(void*)w, (void*)g, (void*)m, (void*)v, (void*)loss, (void*)wb})
      if (p) (void)hipFree(p);
  }
  template <class T>
  bool alloc(T** p, size_t n) {
    if (hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return false;
    batch_bufs.push_back(*p);
    return true;
  }
  bool buffers(int n) {
    if (n == nb) return true;
    for (void* p : batch_bufs) (void)hipFree(p);
    batch_bufs.clear();
    pre.clear();
    h.clear();
    nb = 0;
    int wmax = 1;
    for (size_t l = 1; l + 1 < dims.size(); l++) wmax = std::max(wmax, dims[l]);
    bool ok = alloc(&xb, (size_t)n * dims[0]) && alloc(&y, n) && alloc(&yb, n) && alloc(&d, n) && alloc(&db, n);
    for (int k = 0; k < 2 && ok; k++) ok = alloc(&dpre[k], (size_t)n * wmax) && alloc(&dpre_b[k], (size_t)n * wmax);
    for (size_t l = 1; l + 1 < dims.size() && ok; l++) {
      float* p;
      __bf16* q;
      ok = alloc(&p, (size_t)n * dims[l]) && alloc(&q, (size_t)n * dims[l]);
      if (ok) {
        pre.push_back(p);
        h.push_back(q);
      }
    }
    if (ok) nb = n;
    return ok;
  }
};

namespace {

// The layer-by-layer forward pass of the training step (as the Python MLP
// forms it) on the batch in xb: hidden layers pre = h W^T + b, h = bf16(SiLU),
// then the linear head into y.
int32_t net_forward(bmx_net* net, int n) {
  int L = (int)net->layout.size() - 1;
  const __bf16* hin = net->xb;
  for (int l = 0; l < L; l++) {
    const auto& ly = net->layout[l];
    if (bmx_gemm_bf16(net->ctx, 0, 1, n, ly.n_out, ly.n_in, hin, ly.n_in, net->wb + ly.wo, ly.n_in, net->pre[l],
                      net->h[l], ly.n_out, net->w + ly.bo, 2, nullptr) != 0)
      return -1;
    hin = net->h[l];
  }
  const auto& hd = net->layout[L];
  return bmx_gemm_bf16(net->ctx, 0, 1, n, 1, hd.n_in, hin, hd.n_in, net->wb + hd.wo, hd.n_in, net->y, nullptr, 1,
                       net->w + hd.bo, 1, nullptr);
}

// MSE loss against yb and the gradients of every parameter into g.
int32_t net_backward(bmx_net* net, int n) {
  bmx_ctx* ctx = net->ctx;
  int L = (int)net->layout.size() - 1;
  if (bmx_mse_grad(ctx, n, net->y, net->yb, net->d, net->db, net->loss) != 0) return -1;
  const auto& hd = net->layout[L];
  const __bf16* h_in = L > 0 ? net->h[L - 1] : net->xb;
  if (bmx_gemm_bf16(ctx, 1, 0, 1, hd.n_in, n, net->db, 1, h_in, hd.n_in, net->g + hd.wo, nullptr, hd.n_in, nullptr, 0,
                    nullptr) != 0 ||
      bmx_colsum_f32(ctx, n, 1, net->d, 1, net->g + hd.bo) != 0)
    return -1;
  if (L == 0) return 0;
  int cur = 0;
  if (bmx_gemm_bf16(ctx, 0, 0, n, hd.n_in, 1, net->db, 1, net->wb + hd.wo, hd.n_in, net->dpre[cur], net->dpre_b[cur],
                    hd.n_in, nullptr, 3, net->pre[L - 1]) != 0)
    return -1;
  for (int l = L - 1; l >= 0; l--) {
    const auto& ly = net->layout[l];
    const __bf16* hi = l > 0 ? net->h[l - 1] : net->xb;
    if (bmx_gemm_bf16(ctx, 1, 0, ly.n_out, ly.n_in, n, net->dpre_b[cur], ly.n_out, hi, ly.n_in, net->g + ly.wo,
                      nullptr, ly.n_in, nullptr, 0, nullptr) != 0 ||
        bmx_colsum_f32(ctx, n, ly.n_out, net->dpre[cur], ly.n_out, net->g + ly.bo) != 0)
      return -1;
    if (l > 0) {
      int nxt = 1 - cur;
      if (bmx_gemm_bf16(ctx, 0, 0, n, ly.n_in, ly.n_out, net->dpre_b[cur], ly.n_out, net->wb + ly.wo, ly.n_in,
                        net->dpre[nxt], net->dpre_b[nxt], ly.n_in, nullptr, 3, net->pre[l - 1]) != 0)
        return -1;
      cur = nxt;
    }
  }
  return 0;
}

}  // namespace

extern "C" {

int32_t bmx_surrogate_create(bmx_ctx* ctx, int32_t n_dims, const int32_t* dims, const float* params,
                             int64_t* n_params, bmx_net** out) {
  if (!ctx || !dims || !params || !out || n_dims < 2 || dims[n_dims - 1] != 1)
    return bmx_ctx_set_error(ctx, "bmx_surrogate_create: invalid arguments (the head must have one output)");
  for (int i = 0; i < n_dims; i++)
    if (dims[i] < 1 || dims[i] > 65536) return bmx_ctx_set_error(ctx, "bmx_surrogate_create: bad layer width");
  if (hipSetDevice(bmx_ctx_device(ctx)) != hipSuccess) return bmx_ctx_set_error(ctx, "bmx_surrogate_create: device");
  bmx_net* net = new bmx_net();
  net->ctx = ctx;
  net->device = bmx_ctx_device(ctx);
  net->dims.assign(dims, dims + n_dims);
  int64_t off = 0;
  for (int i = 0; i + 1 < n_dims; i++) {
    bmx_net::Layer ly{off, off + (int64_t)dims[i + 1] * dims[i], dims[i], dims[i + 1]};
    net->layout.push_back(ly);
    off = ly.bo + dims[i + 1];
  }
  net->n_params = off;
  size_t nf = (size_t)off * sizeof(float);
  bool ok = hipMalloc(&net->w, nf) == hipSuccess && hipMalloc(&net->g, nf) == hipSuccess &&
            hipMalloc(&net->m, nf) == hipSuccess && hipMalloc(&net->v, nf) == hipSuccess &&
            hipMalloc(&net->loss, sizeof(float)) == hipSuccess &&
            hipMalloc(&net->wb, (size_t)off * sizeof(__bf16)) == hipSuccess;
  if (!ok) {
    delete net;
    return bmx_ctx_set_error(ctx, "bmx_surrogate_create: device allocation failed");
  }
  *out = net;
  if (n_params) *n_params = off;
  return bmx_surrogate_set_params(net, params);
}

void bmx_surrogate_destroy(bmx_net* net) {  // hipFree waits for the device; the context is not needed
  if (!net) return;
  (void)hipSetDevice(net->device);
  delete net;
}

int32_t bmx_surrogate_set_params(bmx_net* net, const float* params) {
  if (!net || !params) return -1;
  bmx_ctx* ctx = net->ctx;
  hipStream_t s = bmx_ctx_stream(ctx);
  size_t nf = (size_t)net->n_params * sizeof(float);
  hipError_t e = hipMemcpyAsync(net->w, params, nf, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(net->m, 0, nf, s);
  if (e == hipSuccess) e = hipMemsetAsync(net->v, 0, nf, s);
  if (e == hipSuccess) e = hipMemsetAsync(net->g, 0, nf, s);
  if (e == hipSuccess)
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((unsigned)((net->n_params + 255) / 256)), dim3(256), 0, s,
                       net->n_params, net->w, net->wb);
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  net->step = 0;
  return e == hipSuccess ? 0 : bmx_ctx_set_error(ctx, hipGetErrorString(e));
}

int32_t bmx_surrogate_get_params(const bmx_net* net, float* params) {
  if (!net || !params) return -1;
  hipStream_t s = bmx_ctx_stream(net->ctx);
  hipError_t e = hipMemcpyAsync(params, net->w, (size_t)net->n_params * sizeof(float), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? 0 : bmx_ctx_set_error(net->ctx, hipGetErrorString(e));
}

int32_t bmx_surrogate_train(bmx_net* net, const float* X, const float* Y, int64_t n, int32_t epochs, int32_t batch,
                            float lr, float lr_final, uint64_t seed, int32_t flags, float* epoch_loss) {
  if (!net) return -1;
  bmx_ctx* ctx = net->ctx;
  if (!X || !Y || n <= 0 || epochs < 0 || batch <= 0 || batch > n || !(lr > 0) || !(lr_final > 0))
    return bmx_ctx_set_error(ctx, "bmx_surrogate_train: invalid arguments");
  if (hipSetDevice(bmx_ctx_device(ctx)) != hipSuccess) return bmx_ctx_set_error(ctx, "bmx_surrogate_train: device");
  if (!net->buffers(batch)) return bmx_ctx_set_error(ctx, "bmx_surrogate_train: batch buffer allocation failed");
  hipStream_t s = bmx_ctx_stream(ctx);
  const int d0 = net->dims[0];
  const int64_t spe = n / batch;  // steps per epoch (the remainder of a shuffle is left out)
  const int64_t total = spe * epochs;
  if (total == 0) return 0;
  // cosine learning-rate decay from lr to lr_final over all steps
  std::vector<float> lr_h(total);
  for (int64_t t = 0; t < total; t++) {
    double frac = total > 1 ? (double)t / (double)(total - 1) : 0.0;
    lr_h[t] = (float)(lr_final + 0.5 * ((double)lr - lr_final) * (1.0 + std::cos(3.14159265358979323846 * frac)));
  }
  float* lr_d = nullptr;
  int64_t* perm_d = nullptr;
  TrainState* st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string err;
  auto hip_ok = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && err.empty()) err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipSuccess;
  };
  std::vector<int64_t> perm(n);
  std::mt19937_64 rng(seed);
  const int64_t t0 = net->step;
  const int64_t n_params = net->n_params;
  auto step_kernels = [&]() -> bool {  // one training step, all on stream s
    hipLaunchKernelGGL(gather_batch_kernel, dim3((unsigned)(((int64_t)batch * d0 + 255) / 256)), dim3(256), 0, s, X,
                       Y, perm_d, st, batch, d0, net->xb, net->yb);
    if (net_forward(net, batch) != 0 || net_backward(net, batch) != 0) return false;
    hipLaunchKernelGGL(adam_graph_kernel, dim3((unsigned)((n_params + 255) / 256)), dim3(256), 0, s, n_params, net->w,
                       net->wb, net->g, net->m, net->v, lr_d, t0, st, 0.9f, 0.999f, 1e-8f);
    hipLaunchKernelGGL(train_advance_kernel, dim3(1), dim3(1), 0, s, st);
    return hip_ok(hipGetLastError(), "training step");
  };
  bool ok = hip_ok(hipMalloc(&lr_d, total * sizeof(float)), "allocation") &&
            hip_ok(hipMalloc(&perm_d, n * sizeof(int64_t)), "allocation") &&
            hip_ok(hipMalloc(&st, sizeof(TrainState)), "allocation") &&
            hip_ok(hipMemcpyAsync(lr_d, lr_h.data(), total * sizeof(float), hipMemcpyHostToDevice, s), "copy");
  bool use_graph = !(flags & 1);
  for (int ep = 0; ep < epochs && ok; ep++) {
    for (int64_t i = 0; i < n; i++) perm[i] = i;
    for (int64_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[(int64_t)(rng() % (uint64_t)(i + 1))]);
    ok = hip_ok(hipMemcpyAsync(perm_d, perm.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, s), "copy");
    hipLaunchKernelGGL(train_reset_kernel, dim3(1), dim3(1), 0, s, st, net->step);
    ok = ok && hip_ok(hipStreamSynchronize(s), "synchronize");
    double acc = 0.0;
    int cnt = 0;
    for (int64_t k = 0; k < spe && ok; k++) {
      bool eager = !use_graph || (ep == 0 && k == 0);  // the first step runs eagerly: it sizes the workspaces
      if (eager) {
        ok = step_kernels();
      } else {
        if (!exec) {
          ok = hip_ok(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "capture");
          bool built = ok && step_kernels();
          hipError_t ce = hipStreamEndCapture(s, &graph);
          ok = built && hip_ok(ce, "capture") && hip_ok(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph");
          if (!ok) break;
        }
        ok = hip_ok(hipGraphLaunch(exec, s), "graph launch");
      }
      net->step++;
      if (ok && (k % 16 == 15 || k == spe - 1)) {
        float l = 0.f;
        ok = hip_ok(hipMemcpyAsync(&l, net->loss, sizeof(float), hipMemcpyDeviceToHost, s), "copy") &&
             hip_ok(hipStreamSynchronize(s), "synchronize");
        acc += l;
        cnt++;
      }
    }
    if (epoch_loss) epoch_loss[ep] = cnt ? (float)(acc / cnt) : 0.f;
  }
  (void)hipStreamSynchronize(s);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (lr_d) (void)hipFree(lr_d);
  if (perm_d) (void)hipFree(perm_d);
  if (st) (void)hipFree(st);
  if (!ok) return bmx_ctx_set_error(ctx, ("bmx_surrogate_train: " + (err.empty() ? std::string(bmx_last_error(ctx)) : err)).c_str());
  return 0;
}

int32_t bmx_surrogate_infer(bmx_net* net, const void* X_bf16, int64_t n, float* y) {
  if (!net) return -1;
  bmx_ctx* ctx = net->ctx;
  if (!X_bf16 || !y || n <= 0) return bmx_ctx_set_error(ctx, "bmx_surrogate_infer: invalid arguments");
  FusedNet fn;
  if (fused_shape((int32_t)net->dims.size(), net->dims.data(), &fn)) {
    fn.wb = net->wb;
    fn.w = net->w;
    FusedIn in{};
    in.X = (const __bf16*)X_bf16;
    launch_fused(bmx_ctx_stream(ctx), n, fn, in, y);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : bmx_ctx_set_error(ctx, hipGetErrorString(e));
  }
  // other shapes: layer by layer through the batch buffers, in batches
  const int64_t B = 1 << 16;
  if (!net->buffers((int)std::min<int64_t>(B, n))) return bmx_ctx_set_error(ctx, "bmx_surrogate_infer: allocation failed");
  hipStream_t s = bmx_ctx_stream(ctx);
  for (int64_t a = 0; a < n; a += net->nb) {
    int m = (int)std::min<int64_t>(net->nb, n - a);
    if (hipMemcpyAsync(net->xb, (const __bf16*)X_bf16 + a * net->dims[0], (size_t)m * net->dims[0] * sizeof(__bf16),
                       hipMemcpyDeviceToDevice, s) != hipSuccess ||
        net_forward(net, m) != 0 ||
        hipMemcpyAsync(y + a, net->y, (size_t)m * sizeof(float), hipMemcpyDeviceToDevice, s) != hipSuccess)
      return bmx_ctx_set_error(ctx, "bmx_surrogate_infer: forward pass failed");
  }
  return 0;
}

int32_t bmx_surrogate_screen(bmx_net* net, int32_t n_lat, int32_t n_knob, const float* knobs, float knob_range,
                             int32_t n_ic, int32_t n_geo, const float* geo, int32_t n_amp, float alive_level,
                             int32_t* kb) {
  if (!net) return -1;
  return bmx_da_screen(net->ctx, (int32_t)net->dims.size(), net->dims.data(), net->w, net->wb, n_lat, n_knob, knobs,
                       knob_range, n_ic, n_geo, geo, n_amp, alive_level, kb);
}

int32_t bmx_ctx_sync(bmx_ctx* ctx) {
  if (!ctx) return -1;
  hipError_t e = hipStreamSynchronize(bmx_ctx_stream(ctx));
  return e == hipSuccess ? 0 : bmx_ctx_set_error(ctx, hipGetErrorString(e));
}

}  // extern "C"
