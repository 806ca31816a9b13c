// BF16 MFMA GEMM with fused MLP epilogues (forward and backward of dense layers).
//
// C[M,N] = A[M,K] * B[K,N], A and B in bf16 with arbitrary element strides
// (A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]), FP32 accumulation on the
// CDNA4 matrix cores (v_mfma_f32_16x16x32_bf16). One kernel covers the three GEMMs of
// a dense layer: forward X*W, backward G*W^T and the weight gradient X^T*G.
//
// Tiling: a 256-thread workgroup computes a 64x64 tile of C; each wave computes 32x32
// as 2x2 MFMA tiles of 16x16. The K loop stages a 64x32 slab of A and a 32x64 slab of
// B in LDS per iteration (each thread loads 8 elements along whichever dimension is
// contiguous in memory, 16-byte vector loads when the stride is 1), stored k-contiguous
// so that every lane reads its 8-element MFMA fragment with one 16-byte LDS read.
//
// MFMA 16x16x32 fragment layout (lane l, 64 lanes): A row l%16, k = 8*(l/16)+j;
// B column l%16, same k; accumulator D rows 4*(l/16)+j, column l%16, j = 0..3.
//
// Epilogues (mode):
//   GH_EPI_STORE      C32 = acc                                   (weight gradients)
//   GH_EPI_BIAS_SILU  Z32 = acc + bias, Y16 = silu(Z32)            (hidden layers)
//   GH_EPI_BIAS       Z32 = acc + bias, Y16 = bf16(Z32)            (output layer)
//   GH_EPI_DSILU      G32 = acc * silu'(Zprev), Y16 = bf16(G32)     (backward through
//                     the activation of the previous layer)
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "../../include/gatehip.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32;
constexpr int LDK = BK + 8;  // LDS row pitch in bf16 (16-byte aligned rows, fewer bank conflicts)

__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }
__device__ __forceinline__ __bf16 f2bf(float v) { return (__bf16)v; }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// loads 8 elements of an operand slab: rows r..(fixed), along the contiguous dimension
__device__ __forceinline__ void load8(const __bf16 *base, int64_t s_row, int64_t s_col, int row, int col0,
                                      int rows, int cols, bool along_col, __bf16 out[8]) {
  // element (row, col) at base[row*s_row + col*s_col]; loads cols col0..col0+7 of one row
  // (along_col) or rows row..row+7 of one column (!along_col), zero outside [rows, cols)
  if (along_col) {
    if (row < rows && col0 + 7 < cols && s_col == 1 && (s_row & 7) == 0) {
      const bf16x8 v = *(const bf16x8 *)(base + (int64_t)row * s_row + col0);
#pragma unroll
      for (int j = 0; j < 8; ++j) out[j] = v[j];
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j)
        out[j] = (row < rows && col0 + j < cols) ? base[(int64_t)row * s_row + (int64_t)(col0 + j) * s_col]
                                                 : f2bf(0.0f);
    }
  } else {
    if (col0 < cols && row + 7 < rows && s_row == 1 && (s_col & 7) == 0) {
      const bf16x8 v = *(const bf16x8 *)(base + (int64_t)col0 * s_col + row);
#pragma unroll
      for (int j = 0; j < 8; ++j) out[j] = v[j];
    } else {
#pragma unroll
      for (int j = 0; j < 8; ++j)
        out[j] = (row + j < rows && col0 < cols) ? base[(int64_t)(row + j) * s_row + (int64_t)col0 * s_col]
                                                 : f2bf(0.0f);
    }
  }
}

__global__ void __launch_bounds__(256) gemm_bf16_kernel(const __bf16 *__restrict__ A, int64_t sam, int64_t sak,
                                                        const __bf16 *__restrict__ B, int64_t sbk, int64_t sbn,
                                                        int M, int N, int K, const float *__restrict__ bias,
                                                        const float *__restrict__ zprev, int64_t ldz,
                                                        float *__restrict__ C32, __bf16 *__restrict__ C16,
                                                        int64_t ldc, int mode) {
  __shared__ __bf16 As[BM * LDK];  // [m][k]
  __shared__ __bf16 Bs[BN * LDK];  // [n][k]
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

  // which dimension each thread loads along: the contiguous one
  const bool a_along_k = sak == 1 || sam != 1;
  const bool b_along_k = sbk == 1 || sbn != 1;

  for (int k0 = 0; k0 < K; k0 += BK) {
    // ---- stage A slab [64 x 32]
    {
      __bf16 v[8];
      if (a_along_k) {
        const int r = tid >> 2, c = (tid & 3) * 8;  // 64 rows x 4 groups of 8 k
        load8(A, sam, sak, m0 + r, k0 + c, M, K, true, v);
#pragma unroll
        for (int j = 0; j < 8; ++j) As[r * LDK + c + j] = v[j];
      } else {
        const int c = tid >> 3, r = (tid & 7) * 8;  // 32 k x 8 groups of 8 rows
        load8(A, sam, sak, m0 + r, k0 + c, M, K, false, v);
#pragma unroll
        for (int j = 0; j < 8; ++j) As[(r + j) * LDK + c] = v[j];
      }
    }
    // ---- stage B slab [32 x 64] as Bs[n][k]
    {
      __bf16 v[8];
      if (b_along_k) {
        const int n = tid >> 2, c = (tid & 3) * 8;  // B(k, n): rows = k
        // load along k: element (k, n) at B[k*sbk + n*sbn]; treat n as "row", k as "col"
        load8(B, sbn, sbk, n0 + n, k0 + c, N, K, true, v);
#pragma unroll
        for (int j = 0; j < 8; ++j) Bs[n * LDK + c + j] = v[j];
      } else {
        const int c = tid >> 3, n = (tid & 7) * 8;
        load8(B, sbn, sbk, n0 + n, k0 + c, N, K, false, v);
#pragma unroll
        for (int j = 0; j < 8; ++j) Bs[(n + j) * LDK + c] = v[j];
      }
    }
    __syncthreads();
    const int fr = lane & 15, fk = (lane >> 4) * 8;
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) af[i] = *(const bf16x8 *)(&As[(wm + i * 16 + fr) * LDK + fk]);
#pragma unroll
    for (int j = 0; j < 2; ++j) bfr[j] = *(const bf16x8 *)(&Bs[(wn + j * 16 + fr) * LDK + fk]);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }

  // ---- epilogue
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn + j * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        if (m >= M || n >= N) continue;
        float v = acc[i][j][r];
        const int64_t o = (int64_t)m * ldc + n;
        if (mode == GH_EPI_STORE) {
          C32[o] = v;
        } else if (mode == GH_EPI_BIAS_SILU) {
          v += bias[n];
          if (C32) C32[o] = v;
          C16[o] = f2bf(silu(v));
        } else if (mode == GH_EPI_BIAS) {
          v += bias[n];
          if (C32) C32[o] = v;
          if (C16) C16[o] = f2bf(v);
        } else {  // GH_EPI_DSILU
          v *= dsilu(zprev[(int64_t)m * ldz + n]);
          if (C32) C32[o] = v;
          C16[o] = f2bf(v);
        }
      }
    }
}

}  // namespace

extern "C" int gh_gemm_bf16(const void *A, int64_t sam, int64_t sak, const void *B, int64_t sbk, int64_t sbn,
                            int32_t M, int32_t N, int32_t K, const float *bias, const float *zprev, int64_t ldz,
                            float *C32, void *C16, int64_t ldc, int32_t mode, void *stream) {
  if (!A || !B || M < 0 || N < 0 || K < 0) return GH_EINVAL;
  if (mode < GH_EPI_STORE || mode > GH_EPI_DSILU) return GH_EINVAL;
  if (mode == GH_EPI_STORE && !C32) return GH_EINVAL;
  if ((mode == GH_EPI_BIAS_SILU || mode == GH_EPI_BIAS) && !bias) return GH_EINVAL;
  if ((mode == GH_EPI_BIAS_SILU || mode == GH_EPI_DSILU) && !C16) return GH_EINVAL;
  if (mode == GH_EPI_BIAS && !C16 && !C32) return GH_EINVAL;
  if (mode == GH_EPI_DSILU && !zprev) return GH_EINVAL;
  if (M == 0 || N == 0) return GH_OK;
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM);
  hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, (hipStream_t)stream, (const __bf16 *)A, sam, sak,
                     (const __bf16 *)B, sbk, sbn, M, N, K, bias, zprev, ldz, C32, (__bf16 *)C16, ldc, mode);
  return hipGetLastError() == hipSuccess ? GH_OK : GH_EHIP;
}
