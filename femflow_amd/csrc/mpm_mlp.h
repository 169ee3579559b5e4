// Grid-corrector MLP: shapes, parameter layouts and BF16 MFMA device helpers.
//
// Network: 351 node-stencil features -> 128 -> 128 -> 3 (ReLU hidden
// layers).  FP32 master parameters (the optimiser's copy) are packed into
// BF16 operand copies for the MFMA kernels; every product accumulates in
// FP32.  Matrices are stored as the linear layers' (out x in) row-major
// weights, padded so every reduction length is a multiple of 32 (the K of
// one v_mfma_f32_16x16x32_bf16).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

constexpr int MLP_IN = 351, MLP_INP = 352, MLP_H = 128, MLP_OUT = 3, MLP_OUTP = 16, MLP_DYP = 32;

// FP32 master parameters
constexpr int P_W1 = 0;                          // [128][351]
constexpr int P_B1 = P_W1 + MLP_H * MLP_IN;      // [128]
constexpr int P_W2 = P_B1 + MLP_H;               // [128][128]
constexpr int P_B2 = P_W2 + MLP_H * MLP_H;       // [128]
constexpr int P_W3 = P_B2 + MLP_H;               // [3][128]
constexpr int P_B3 = P_W3 + MLP_OUT * MLP_H;     // [3]
constexpr int P_COUNT = P_B3 + MLP_OUT;

// BF16 operand copies (uint16 elements)
constexpr int K_W1 = 0;                          // [128][352]   W1, zero column 351
constexpr int K_W2 = K_W1 + MLP_H * MLP_INP;     // [128][128]   W2
constexpr int K_W3 = K_W2 + MLP_H * MLP_H;       // [16][128]    W3, zero rows 3..15
constexpr int K_W2T = K_W3 + MLP_OUTP * MLP_H;   // [128][128]   W2^T
constexpr int K_W3T = K_W2T + MLP_H * MLP_H;     // [128][32]    W3^T, zero columns 3..31
constexpr int K_COUNT = K_W3T + MLP_H * MLP_DYP;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 u) { return __builtin_bit_cast(bf16x8, u); }

// D(16x16) += A(16x32) B(32x16).  Lane l supplies A[l % 16][8 (l / 16) .. +7]
// and B[8 (l / 16) .. +7][l % 16]; it holds D[4 (l / 16) + r][l % 16], r < 4.
__device__ __forceinline__ floatx4 mfma16x16x32(uint4 a, uint4 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(a), as_bf16x8(b), c, 0, 0, 0);
}

void mlp_pack(hipStream_t st, const float* P, uint16_t* K);
void mlp_forward(hipStream_t st, const uint16_t* K, const float* P, const uint16_t* X, int64_t M, uint16_t* H1,
                 uint16_t* H2, float* Y);
void mlp_backward(hipStream_t st, const uint16_t* K, const uint16_t* X, int64_t M, const uint16_t* H1,
                  const uint16_t* H2, const uint16_t* dY, uint16_t* dH2, uint16_t* dH1, float* G);
void mlp_loss(hipStream_t st, const float* Y, const float* T, const float* w, int64_t M, float inv_norm,
              uint16_t* dY, double* loss);
void mlp_adam(hipStream_t st, float* P, const float* G, float* m, float* v, int n, float lr, float b1, float b2,
              float eps, int step);

}  // namespace mpm
