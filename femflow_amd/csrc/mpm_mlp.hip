// BF16 MFMA training and inference kernels of the grid-corrector MLP.
//
// gemm_nt:  C[m][n] = epi(sum_k A[m][k] B[n][k])   (activations x weights^T)
// gemm_tn:  C[m][n] += sum_k A[k][m] B[k][n]        (weight gradients, split
//           over k = the rows of the batch, FP32 atomics per output tile)
// Operands are BF16 (uint16 storage), accumulation FP32 on
// v_mfma_f32_16x16x32_bf16.  A workgroup is four waves; each wave owns 16
// output rows and NT tiles of 16 columns.
#include <hip/hip_runtime.h>

#include "mpm_mlp.h"

namespace mpm {

struct NTArgs {
  int64_t M;
  int N, K;
  const uint16_t* A;
  int64_t lda;
  const uint16_t* B;
  int64_t ldb;
  const float* bias;     // [N] or null
  const uint16_t* mask;  // keep C[m][n] only where mask[m][n] > 0 (ReLU'), or null
  int64_t ldm;
  int relu;
  uint16_t* Cb;  // BF16 output or null
  int64_t ldcb;
  float* Cf;  // FP32 output or null
  int64_t ldcf;
};

template <int NT>
__global__ __launch_bounds__(256) void gemm_nt_kernel(NTArgs g) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = lane & 15, lk = (lane >> 4) * 8;
  const int64_t row0 = (int64_t)blockIdx.x * 64 + wv * 16;
  const int n0 = blockIdx.y * 16 * NT;
  const int64_t arow = row0 + li;
  floatx4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
  for (int k0 = 0; k0 < g.K; k0 += 32) {
    const uint4 a = arow < g.M ? *reinterpret_cast<const uint4*>(g.A + arow * g.lda + k0 + lk) : zero;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int col = n0 + 16 * t + li;
      const uint4 b = col < g.N ? *reinterpret_cast<const uint4*>(g.B + (int64_t)col * g.ldb + k0 + lk) : zero;
      acc[t] = mfma16x16x32(a, b, acc[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int col = n0 + 16 * t + li;
    if (col >= g.N) continue;
    const float bb = g.bias ? g.bias[col] : 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t row = row0 + (lane >> 4) * 4 + r;
      if (row >= g.M) continue;
      float v = acc[t][r] + bb;
      if (g.mask && !(bf2f(g.mask[row * g.ldm + col]) > 0.f)) v = 0.f;
      if (g.relu) v = v > 0.f ? v : 0.f;
      if (g.Cb) g.Cb[row * g.ldcb + col] = f2bf(v);
      if (g.Cf) g.Cf[row * g.ldcf + col] = v;
    }
  }
}

struct TNArgs {
  int64_t K;
  int M, N;
  const uint16_t* A;
  int64_t lda;
  const uint16_t* B;
  int64_t ldb;
  float* C;
  int64_t ldc;
  int64_t kchunk;  // rows per split (multiple of 32)
};

template <int NT>
__global__ __launch_bounds__(256) void gemm_tn_kernel(TNArgs g) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = lane & 15, lk = (lane >> 4) * 8;
  const int m0 = blockIdx.x * 64 + wv * 16;
  const int n0 = blockIdx.y * 16 * NT;
  if (m0 >= g.M) return;
  const int64_t kb = (int64_t)blockIdx.z * g.kchunk;
  const int64_t ke = kb + g.kchunk < g.K ? kb + g.kchunk : g.K;
  const int m = m0 + li;
  floatx4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int64_t k0 = kb; k0 < ke; k0 += 32) {
    uint16_t av[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int64_t k = k0 + lk + j;
      av[j] = (k < ke && m < g.M) ? g.A[k * g.lda + m] : (uint16_t)0;
    }
    const uint4 a = make_uint4(av[0] | ((uint32_t)av[1] << 16), av[2] | ((uint32_t)av[3] << 16),
                               av[4] | ((uint32_t)av[5] << 16), av[6] | ((uint32_t)av[7] << 16));
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int n = n0 + 16 * t + li;
      uint16_t bv[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int64_t k = k0 + lk + j;
        bv[j] = (k < ke && n < g.N) ? g.B[k * g.ldb + n] : (uint16_t)0;
      }
      const uint4 b = make_uint4(bv[0] | ((uint32_t)bv[1] << 16), bv[2] | ((uint32_t)bv[3] << 16),
                                 bv[4] | ((uint32_t)bv[5] << 16), bv[6] | ((uint32_t)bv[7] << 16));
      acc[t] = mfma16x16x32(a, b, acc[t]);
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int n = n0 + 16 * t + li;
    if (n >= g.N) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int mm = m0 + (lane >> 4) * 4 + r;
      if (mm < g.M) atomicAdd(g.C + (int64_t)mm * g.ldc + n, acc[t][r]);
    }
  }
}

// out[n] += sum over rows of A[row][n] (n < N), BF16 input
__global__ __launch_bounds__(256) void colsum_kernel(const uint16_t* A, int64_t lda, int64_t M, int N,
                                                     int64_t rows_per, float* out) {
  __shared__ float part[4][64];
  const int c = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int n = blockIdx.x * 64 + c;
  const int64_t r0 = (int64_t)blockIdx.y * rows_per;
  const int64_t r1 = r0 + rows_per < M ? r0 + rows_per : M;
  float s = 0.f;
  if (n < N)
    for (int64_t r = r0 + q; r < r1; r += 4) s += bf2f(A[r * lda + n]);
  part[q][c] = s;
  __syncthreads();
  if (q == 0 && n < N) atomicAdd(out + n, part[0][c] + part[1][c] + part[2][c] + part[3][c]);
}

// Weighted mean squared error of the 3 outputs; writes dL/dY (BF16, ld 32,
// zero padding) and adds the loss to *loss.
__global__ __launch_bounds__(256) void mlp_loss_kernel(const float* Y, const float* T, const float* w,
                                                       int64_t M, float inv_norm, uint16_t* dY, double* loss) {
  __shared__ double red[4];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (i < M) {
    const float wt = w ? w[i] : 1.f;
#pragma unroll
    for (int o = 0; o < MLP_OUT; ++o) {
      const float d = Y[i * MLP_OUTP + o] - T[i * MLP_OUT + o];
      dY[i * MLP_DYP + o] = f2bf(2.f * wt * d * inv_norm);
      l += (double)wt * d * d;
    }
#pragma unroll
    for (int o = MLP_OUT; o < MLP_DYP; ++o) dY[i * MLP_DYP + o] = 0;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(loss, (red[0] + red[1] + red[2] + red[3]) * (double)inv_norm);
}

__global__ void mlp_adam_kernel(float* p, const float* g, float* m, float* v, int n, float lr, float b1,
                                float b2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// FP32 master -> BF16 operand copies (W1, W2, W3 and the transposes the
// backward pass reads), zero padding included.
__global__ void mlp_pack_kernel(const float* P, uint16_t* K) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K_COUNT) return;
  float v = 0.f;
  if (i < K_W2) {
    const int r = (i - K_W1) / MLP_INP, c = (i - K_W1) % MLP_INP;
    v = c < MLP_IN ? P[P_W1 + r * MLP_IN + c] : 0.f;
  } else if (i < K_W3) {
    v = P[P_W2 + (i - K_W2)];
  } else if (i < K_W2T) {
    const int r = (i - K_W3) / MLP_H, c = (i - K_W3) % MLP_H;
    v = r < MLP_OUT ? P[P_W3 + r * MLP_H + c] : 0.f;
  } else if (i < K_W3T) {
    const int r = (i - K_W2T) / MLP_H, c = (i - K_W2T) % MLP_H;  // W2T[r][c] = W2[c][r]
    v = P[P_W2 + c * MLP_H + r];
  } else {
    const int r = (i - K_W3T) / MLP_DYP, c = (i - K_W3T) % MLP_DYP;  // W3T[r][c] = W3[c][r]
    v = c < MLP_OUT ? P[P_W3 + c * MLP_H + r] : 0.f;
  }
  K[i] = f2bf(v);
}

static inline unsigned nblk64(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

template <int NT>
static void launch_nt(hipStream_t st, const NTArgs& g) {
  dim3 grid(nblk64(g.M, 64), (unsigned)((g.N + 16 * NT - 1) / (16 * NT)));
  if (g.M > 0) hipLaunchKernelGGL(gemm_nt_kernel<NT>, grid, dim3(256), 0, st, g);
}

static void launch_tn(hipStream_t st, int64_t K, int M, int N, const uint16_t* A, int64_t lda, const uint16_t* B,
                      int64_t ldb, float* C, int64_t ldc) {
  if (K <= 0) return;
  int64_t splits = (K + 2047) / 2048;  // ~2k rows per workgroup
  if (splits > 1024) splits = 1024;
  int64_t kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + 31) / 32 * 32;
  splits = (K + kchunk - 1) / kchunk;
  TNArgs g{K, M, N, A, lda, B, ldb, C, ldc, kchunk};
  dim3 grid((unsigned)((M + 63) / 64), (unsigned)((N + 31) / 32), (unsigned)splits);
  hipLaunchKernelGGL(gemm_tn_kernel<2>, grid, dim3(256), 0, st, g);
}

static void launch_colsum(hipStream_t st, const uint16_t* A, int64_t lda, int64_t M, int N, float* out) {
  if (M <= 0) return;
  int64_t splits = (M + 4095) / 4096;
  if (splits > 4096) splits = 4096;
  const int64_t per = (M + splits - 1) / splits;
  dim3 grid((unsigned)((N + 63) / 64), (unsigned)splits);
  hipLaunchKernelGGL(colsum_kernel, grid, dim3(256), 0, st, A, lda, M, N, per, out);
}

void mlp_pack(hipStream_t st, const float* P, uint16_t* K) {
  hipLaunchKernelGGL(mlp_pack_kernel, dim3(nblk64(K_COUNT, 256)), dim3(256), 0, st, P, K);
}

void mlp_forward(hipStream_t st, const uint16_t* K, const float* P, const uint16_t* X, int64_t M, uint16_t* H1,
                 uint16_t* H2, float* Y) {
  NTArgs g{};
  g.M = M;
  g.N = MLP_H;
  g.K = MLP_INP;
  g.A = X;
  g.lda = MLP_INP;
  g.B = K + K_W1;
  g.ldb = MLP_INP;
  g.bias = P + P_B1;
  g.relu = 1;
  g.Cb = H1;
  g.ldcb = MLP_H;
  launch_nt<8>(st, g);
  g.K = MLP_H;
  g.A = H1;
  g.lda = MLP_H;
  g.B = K + K_W2;
  g.ldb = MLP_H;
  g.bias = P + P_B2;
  g.Cb = H2;
  launch_nt<8>(st, g);
  g.N = MLP_OUT;
  g.A = H2;
  g.B = K + K_W3;
  g.bias = P + P_B3;
  g.relu = 0;
  g.Cb = nullptr;
  g.Cf = Y;
  g.ldcf = MLP_OUTP;
  launch_nt<1>(st, g);
}

void mlp_backward(hipStream_t st, const uint16_t* K, const uint16_t* X, int64_t M, const uint16_t* H1,
                  const uint16_t* H2, const uint16_t* dY, uint16_t* dH2, uint16_t* dH1, float* G) {
  // output layer
  launch_tn(st, M, MLP_OUT, MLP_H, dY, MLP_DYP, H2, MLP_H, G + P_W3, MLP_H);
  launch_colsum(st, dY, MLP_DYP, M, MLP_OUT, G + P_B3);
  NTArgs g{};
  g.M = M;
  g.N = MLP_H;
  g.K = MLP_DYP;
  g.A = dY;
  g.lda = MLP_DYP;
  g.B = K + K_W3T;
  g.ldb = MLP_DYP;
  g.mask = H2;
  g.ldm = MLP_H;
  g.Cb = dH2;
  g.ldcb = MLP_H;
  launch_nt<8>(st, g);
  // hidden layer 2
  launch_tn(st, M, MLP_H, MLP_H, dH2, MLP_H, H1, MLP_H, G + P_W2, MLP_H);
  launch_colsum(st, dH2, MLP_H, M, MLP_H, G + P_B2);
  g.K = MLP_H;
  g.A = dH2;
  g.lda = MLP_H;
  g.B = K + K_W2T;
  g.ldb = MLP_H;
  g.mask = H1;
  g.Cb = dH1;
  launch_nt<8>(st, g);
  // hidden layer 1
  launch_tn(st, M, MLP_H, MLP_IN, dH1, MLP_H, X, MLP_INP, G + P_W1, MLP_IN);
  launch_colsum(st, dH1, MLP_H, M, MLP_H, G + P_B1);
}

void mlp_loss(hipStream_t st, const float* Y, const float* T, const float* w, int64_t M, float inv_norm,
              uint16_t* dY, double* loss) {
  if (M > 0) hipLaunchKernelGGL(mlp_loss_kernel, dim3(nblk64(M, 256)), dim3(256), 0, st, Y, T, w, M, inv_norm, dY, loss);
}

void mlp_adam(hipStream_t st, float* P, const float* G, float* m, float* v, int n, float lr, float b1, float b2,
              float eps, int step) {
  const float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(mlp_adam_kernel, dim3(nblk64(n, 256)), dim3(256), 0, st, P, G, m, v, n, lr, b1, b2, eps, bc1,
                     bc2);
}

}  // namespace mpm
