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
//   GH_EPI_BIAS_RELU  Z32 = acc + bias, Y16 = relu(Z32)            (hidden layers)
//   GH_EPI_BIAS       Z32 = acc + bias, Y16 = bf16(Z32)            (output layer)
//   GH_EPI_DRELU      G32 = acc * relu'(Zprev), Y16 = bf16(G32)     (backward through
//                     the activation of the previous layer)
#include <hip/hip_runtime.h>
#include <type_traits>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <string>

#include "../../include/gatehip.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32;
constexpr int LDK = BK + 8;  // LDS row pitch in bf16 (16-byte aligned rows, fewer bank conflicts)

__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }
__device__ __forceinline__ __bf16 f2bf(float v) { return (__bf16)v; }

// Hidden-layer activation: ReLU, and its derivative (1 where z > 0) for the backward pass.
// (An epilogue with z sigma(z) instead, through the hardware exp2 and reciprocal, measured
// 682 TFLOP/s for the fused generator forward against 816 for ReLU: the two transcendental
// instructions per element issue at about a third of the rate of other vector instructions.)
__device__ __forceinline__ float relu(float z) { return fmaxf(z, 0.0f); }
__device__ __forceinline__ float drelu(float z) { return z > 0.0f ? 1.0f : 0.0f; }

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

// blockIdx.z selects a slice [z k_chunk, (z+1) k_chunk) of the reduction (split-K, store
// mode only): slice z writes its partial sums at C32 + z * part_stride.
__global__ void __launch_bounds__(256) gemm_bf16_kernel(const __bf16 *__restrict__ A, int64_t sam, int64_t sak,
                                                        const __bf16 *__restrict__ B, int64_t sbk, int64_t sbn,
                                                        int M, int N, int K, const float *__restrict__ bias,
                                                        const float *__restrict__ zprev, int64_t ldz,
                                                        float *__restrict__ C32, __bf16 *__restrict__ C16,
                                                        int64_t ldc, int mode, int k_chunk, int64_t part_stride) {
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

  const int k_begin = blockIdx.z * k_chunk;
  const int k_end = min(K, k_begin + k_chunk);
  if (part_stride) C32 += (int64_t)blockIdx.z * part_stride;
  for (int k0 = k_begin; k0 < k_end; k0 += BK) {
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
        } else if (mode == GH_EPI_BIAS_RELU) {
          v += bias[n];
          if (C32) C32[o] = v;
          C16[o] = f2bf(relu(v));
        } else if (mode == GH_EPI_BIAS) {
          v += bias[n];
          if (C32) C32[o] = v;
          if (C16) C16[o] = f2bf(v);
        } else {  // GH_EPI_DRELU
          v *= drelu(zprev[(int64_t)m * ldz + n]);
          if (C32) C32[o] = v;
          C16[o] = f2bf(v);
        }
      }
    }
}

// C32[m, n] = sum over the split-K slices in order (deterministic)
__global__ void __launch_bounds__(256) splitk_reduce_kernel(const float *__restrict__ part, int S, int64_t stride,
                                                            int M, int N, float *__restrict__ C32, int64_t ldc) {
  const int64_t total = (int64_t)M * N;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int m = (int)(i / N), n = (int)(i - (int64_t)m * N);
    float acc = 0.0f;
    for (int z = 0; z < S; ++z) acc += part[(int64_t)z * stride + i];
    C32[(int64_t)m * ldc + n] = acc;
  }
}

// split-K plan of a GEMM: S slices of kc (a multiple of BK) reduction elements each;
// S = 1 means no split. Weight gradients (few output tiles, long reduction) are split so
// that the launch fills the chip; the slices are then summed in a fixed order.
void splitk_plan(int M, int N, int K, int mode, int &S, int &kc) {
  const int tiles = ((N + BN - 1) / BN) * ((M + BM - 1) / BM);
  S = 1;
  kc = ((K + BK - 1) / BK) * BK;
  if (mode == GH_EPI_STORE && tiles < 512 && K >= 2048) {
    S = std::max(1, std::min((K + 511) / 512, std::max(1, 1024 / tiles)));
    kc = (K + S - 1) / S;
    kc = (kc + BK - 1) / BK * BK;
    S = (K + kc - 1) / kc;
  }
}

int check_gemm_args(const void *A, const void *B, int M, int N, int K, const float *bias, const float *zprev,
                    float *C32, void *C16, int mode) {
  if (!A || !B || M < 0 || N < 0 || K < 0) return GH_EINVAL;
  if (mode < GH_EPI_STORE || mode > GH_EPI_DRELU) return GH_EINVAL;
  if (mode == GH_EPI_STORE && !C32) return GH_EINVAL;
  if ((mode == GH_EPI_BIAS_RELU || mode == GH_EPI_BIAS) && !bias) return GH_EINVAL;
  if ((mode == GH_EPI_BIAS_RELU || mode == GH_EPI_DRELU) && !C16) return GH_EINVAL;
  if (mode == GH_EPI_BIAS && !C16 && !C32) return GH_EINVAL;
  if (mode == GH_EPI_DRELU && !zprev) return GH_EINVAL;
  return GH_OK;
}

// Launches the GEMM of a plan; ws holds S*M*N floats when S > 1 (unused otherwise).
int launch_gemm(const void *A, int64_t sam, int64_t sak, const void *B, int64_t sbk, int64_t sbn, int M, int N,
                int K, const float *bias, const float *zprev, int64_t ldz, float *C32, void *C16, int64_t ldc,
                int mode, int S, int kc, float *ws, hipStream_t stream) {
  if (S > 1) {
    dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, S);
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, stream, (const __bf16 *)A, sam, sak,
                       (const __bf16 *)B, sbk, sbn, M, N, K, bias, zprev, ldz, ws, (__bf16 *)nullptr,
                       (int64_t)N, mode, kc, (int64_t)M * N);
    if (hipGetLastError() != hipSuccess) return GH_EHIP;
    const int64_t total = (int64_t)M * N;
    const int rgrid = (int)std::min<int64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(rgrid), dim3(256), 0, stream, ws, S, (int64_t)M * N, M, N, C32,
                       ldc);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_EHIP;
  }
  dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM);
  hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, stream, (const __bf16 *)A, sam, sak, (const __bf16 *)B,
                     sbk, sbn, M, N, K, bias, zprev, ldz, C32, (__bf16 *)C16, ldc, mode, kc, (int64_t)0);
  return hipGetLastError() == hipSuccess ? GH_OK : GH_EHIP;
}

// Split-K workspaces of gh_gemm_bf16, one per (device, stream) and grown on demand: calls
// on one stream are ordered, so a workspace is never used by two kernels at once; before
// a workspace is replaced its stream is drained, so no queued kernel still reads it.
struct WsKey {
  int device;
  hipStream_t stream;
  bool operator<(const WsKey &o) const {
    return device != o.device ? device < o.device : (uintptr_t)stream < (uintptr_t)o.stream;
  }
};
struct WsBuf {
  float *p = nullptr;
  size_t bytes = 0;
};
std::mutex g_ws_mutex;
std::map<WsKey, WsBuf> g_ws;

// message of the last failed call of the allocation-free entry point (per host thread)
thread_local std::string g_free_err;

}  // namespace

// Scratch memory of gh_gemm_bf16_noalloc: the split-K slices live in this array, which
// every GPU the library runs on holds once, at its own address, until the process ends.
static __device__ __align__(256) float scratch[GH_GEMM_NOALLOC_WS_BYTES / sizeof(float)];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
float *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on the current device (hipGetSymbolAddress once per device,
// with that device current; the host-side name of the array is not a device address).
int scratch_address(float **out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    g_free_err = "gh_gemm_bf16_noalloc: hipGetDevice failed";
    return GH_EHIP;
  }
  if (dev < 0 || dev >= MAX_DEVICES) {
    g_free_err = "gh_gemm_bf16_noalloc: device ordinal beyond the library's limit";
    return GH_ERANGE;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      g_free_err = std::string("gh_gemm_bf16_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return GH_EHIP;
    }
    g_scratch_ptr[dev] = (float *)p;
  }
  *out = g_scratch_ptr[dev];
  return GH_OK;
}
}  // namespace

// internal to the library (not part of the C-ABI): read by gh_last_error(NULL)
extern "C" __attribute__((visibility("hidden"))) const char *gh_free_call_error() { return g_free_err.c_str(); }

extern "C" int gh_gemm_bf16(const void *A, int64_t sam, int64_t sak, const void *B, int64_t sbk, int64_t sbn,
                            int32_t M, int32_t N, int32_t K, const float *bias, const float *zprev, int64_t ldz,
                            float *C32, void *C16, int64_t ldc, int32_t mode, void *stream) {
  const int rc = check_gemm_args(A, B, M, N, K, bias, zprev, C32, C16, mode);
  if (rc != GH_OK) return rc;
  if (M == 0 || N == 0) return GH_OK;
  int S, kc;
  splitk_plan(M, N, K, mode, S, kc);
  float *ws = nullptr;
  if (S > 1) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return GH_EHIP;
    const size_t need = (size_t)S * M * N * sizeof(float);
    std::lock_guard<std::mutex> lock(g_ws_mutex);
    WsBuf &b = g_ws[WsKey{dev, (hipStream_t)stream}];
    if (b.bytes < need) {
      if (b.p) {
        if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess) return GH_EHIP;
        (void)hipFree(b.p);
      }
      b.p = nullptr;
      b.bytes = 0;
      if (hipMalloc(&b.p, need) != hipSuccess) return GH_EHIP;
      b.bytes = need;
    }
    ws = b.p;
  }
  return launch_gemm(A, sam, sak, B, sbk, sbn, M, N, K, bias, zprev, ldz, C32, C16, ldc, mode, S, kc, ws,
                     (hipStream_t)stream);
}

extern "C" int gh_gemm_bf16_noalloc(const void *A, int64_t sam, int64_t sak, const void *B, int64_t sbk,
                                    int64_t sbn, int32_t M, int32_t N, int32_t K, const float *bias,
                                    const float *zprev, int64_t ldz, float *C32, void *C16, int64_t ldc,
                                    int32_t mode, void *stream) {
  const int rc = check_gemm_args(A, B, M, N, K, bias, zprev, C32, C16, mode);
  if (rc != GH_OK) {
    g_free_err = "gh_gemm_bf16_noalloc: invalid argument";
    return rc;
  }
  if (M == 0 || N == 0) return GH_OK;
  int S, kc;
  splitk_plan(M, N, K, mode, S, kc);
  float *ws = nullptr;
  if (S > 1) {
    const size_t need = (size_t)S * M * N * sizeof(float);
    if (need > (size_t)GH_GEMM_NOALLOC_WS_BYTES) {
      g_free_err = "gh_gemm_bf16_noalloc: the split-K workspace of this shape (" + std::to_string(need) +
                   " bytes) exceeds the library's static scratch (" +
                   std::to_string((size_t)GH_GEMM_NOALLOC_WS_BYTES) + " bytes)";
      return GH_ERANGE;
    }
    const int r = scratch_address(&ws);
    if (r != GH_OK) return r;
  }
  const int r = launch_gemm(A, sam, sak, B, sbk, sbn, M, N, K, bias, zprev, ldz, C32, C16, ldc, mode, S, kc, ws,
                            (hipStream_t)stream);
  if (r != GH_OK) g_free_err = "gh_gemm_bf16_noalloc: kernel launch failed";
  return r;
}

// ---------------------------------------------------------------------------------------
// Fused MLP inference: all layers of a dense network in one kernel, activations kept in
// LDS. Layer l: Z = X W_l + b_l; hidden layers X_{l+1} = bf16(relu(Z)), the last layer is
// written as FP32. Weights are given transposed, Wt_l[n][k] (k contiguous), so that each
// lane's MFMA B fragment (column n, 8 consecutive k) is one 16-byte load from L2-resident
// memory. This 4-wave form serves hidden width 128 (width 256 has the 8-wave form below).
// A workgroup of 4 waves owns 16 RT rows; in a hidden layer wave w computes the
// columns [w H/4, (w+1) H/4) for all rows (RT row tiles x H/64 column tiles of 16x16),
// reading its A fragments (row, 8 consecutive k) from the LDS activation tile. Both
// operands of k step ks+1 are loaded while the MFMAs of step ks run, into a second set of
// registers (k steps go in pairs, so no register is copied): a rotation by register
// copies would make the compiler wait for the next step's loads before this step's MFMAs.
// One LDS activation buffer serves as input and output of a layer: the accumulators hold
// the whole output tile, so after a barrier the epilogue overwrites the input in place,
// which halves the LDS per workgroup (and lets 112-row tiles run at 2 waves per SIMD).
// The output layer (16 columns) spreads its row tiles over the 4 waves.
// The per-element arithmetic (MFMA over k slabs of 32 in order, bias, ReLU, bf16 rounding)
// is the same as the layer-by-layer GEMM path, so both give bit-identical outputs.
namespace {

constexpr int MAXL = 8;          // layers
constexpr int K0 = 32;           // first-layer input width (padded)
constexpr int NOUT = 16;         // last-layer output width (padded)
// 16-row tiles per workgroup, per hidden width (measured best on MI355X)
#ifndef GH_MLP_RT128
#define GH_MLP_RT128 8           // 128 rows
#endif
#ifndef GH_MLP_SB
#define GH_MLP_SB 1              // one activation buffer, overwritten after a barrier
#endif

struct MlpArgs {
  const __bf16 *wt[MAXL];        // [N_l][K_l]
  const float *bias[MAXL];       // [N_l]
  int n_layers;
  int k_in;                      // valid input columns (<= K0)
  int n_out;                     // valid output columns (<= NOUT)
  int x_vec;                     // 1: input rows 16-byte aligned (16-byte loads)
};

template <int H, int RT, int SB>
constexpr size_t mlp_lds_bytes() { return (SB ? 1ull : 2ull) * (16 * RT) * (H + 8) * sizeof(__bf16); }

template <int H, int RT, int SB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) mlp_forward_kernel(const __bf16 *__restrict__ X, int64_t ldx, int64_t M,
                                                          MlpArgs a, float *__restrict__ Z, int64_t ldz) {
  constexpr int P = H + 8;                 // LDS row pitch (bf16)
  constexpr int CT = H / 64;               // 16-wide column tiles per wave (hidden layers)
  constexpr int ROWS = 16 * RT;
  extern __shared__ __align__(16) unsigned char mlp_lds[];
  __bf16 *act0 = (__bf16 *)mlp_lds;
  __bf16 *act1 = SB ? act0 : act0 + ROWS * P;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t row0 = (int64_t)blockIdx.x * ROWS;
  const int fr = lane & 15, fk = (lane >> 4) * 8;

  // input tile -> act0[:, 0:K0] (zero beyond k_in and beyond M): one 16-byte load per
  // thread and 8 columns when the rows are 16-byte aligned (a.x_vec), else element-wise
  if (a.x_vec) {
    for (int i = tid; i < ROWS * (K0 / 8); i += 256) {
      const int r = i / (K0 / 8), q = (i - r * (K0 / 8)) * 8;
      const int64_t gr = row0 + r;
      bf16x8 v;
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = (__bf16)0.0f;
      if (gr < M && q < a.k_in) {
        v = *(const bf16x8 *)(X + gr * ldx + q);
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (q + j >= a.k_in) v[j] = (__bf16)0.0f;
      }
      *(bf16x8 *)(act0 + r * P + q) = v;
    }
  } else {
    for (int i = tid; i < ROWS * K0; i += 256) {
      const int r = i / K0, k = i - r * K0;
      const int64_t gr = row0 + r;
      act0[r * P + k] = (gr < M && k < a.k_in) ? X[gr * ldx + k] : (__bf16)0.0f;
    }
  }
  __syncthreads();

  __bf16 *in = act0, *out = act1;
  // weights of the next layer's first k step (hidden layers) or of the whole output layer,
  // loaded before the epilogue of the layer before it, so that their L2 latency passes
  // while the epilogue runs
  bf16x8 wpre[CT];
  bool have_pre = false;
  for (int l = 0; l < a.n_layers; ++l) {
    const bool last = l == a.n_layers - 1;
    const int K = l == 0 ? K0 : H;
    const __bf16 *__restrict__ W = a.wt[l];
    const float *__restrict__ bias = a.bias[l];
    if (!last) {
      floatx4 acc[RT][CT];
#pragma unroll
      for (int i = 0; i < RT; ++i)
#pragma unroll
        for (int j = 0; j < CT; ++j) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
      const int n0 = wave * (H / 4);
      const int nk = K / 32;
      const __bf16 *wrow[CT];
#pragma unroll
      for (int j = 0; j < CT; ++j) wrow[j] = W + (int64_t)(n0 + j * 16 + fr) * K + fk;
      const __bf16 *arow = in + fr * P + fk;
      bf16x8 bA[CT], bB[CT], aA[RT], aB[RT];
      auto load = [&](bf16x8 *b, bf16x8 *af, int ks) {
#pragma unroll
        for (int j = 0; j < CT; ++j) b[j] = *(const bf16x8 *)(wrow[j] + ks * 32);
#pragma unroll
        for (int i = 0; i < RT; ++i) af[i] = *(const bf16x8 *)(arow + i * 16 * P + ks * 32);
      };
      // operands swapped (W tile as A, activations as B): the accumulator holds the
      // transposed tile, lane l owning 4 consecutive output columns of one row, so the
      // epilogue writes 4 packed bf16 per 8-byte LDS store
      auto mma = [&](const bf16x8 *b, const bf16x8 *af) {
#pragma unroll
        for (int i = 0; i < RT; ++i)
#pragma unroll
          for (int j = 0; j < CT; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j], af[i], acc[i][j], 0, 0, 0);
      };
      // the number of k steps is a compile-time constant (first layer K0 / 32, then
      // H / 32), so the loop unrolls into straight-line code whose load waits are exact
      auto steps = [&](auto nk_c) {
        constexpr int NK = decltype(nk_c)::value;
        // scheduling barriers keep each step's loads ahead of the previous step's MFMAs
        // (left alone, the scheduler sinks them next to their first use)
        if (have_pre) {
#pragma unroll
          for (int j = 0; j < CT; ++j) bA[j] = wpre[j];
#pragma unroll
          for (int i = 0; i < RT; ++i) aA[i] = *(const bf16x8 *)(arow + i * 16 * P);
        } else {
          load(bA, aA, 0);
        }
#pragma unroll
        for (int ks = 0; ks < NK; ks += 2) {
          if (ks + 1 < NK) load(bB, aB, ks + 1);
          __builtin_amdgcn_sched_barrier(0);
          mma(bA, aA);
          __builtin_amdgcn_sched_barrier(0);
          if (ks + 1 < NK) {
            if (ks + 2 < NK) load(bA, aA, ks + 2);
            __builtin_amdgcn_sched_barrier(0);
            mma(bB, aB);
            __builtin_amdgcn_sched_barrier(0);
          }
        }
      };
      if (nk == K0 / 32) steps(std::integral_constant<int, K0 / 32>{});
      else steps(std::integral_constant<int, H / 32>{});
      // next layer's weights (its input does not exist yet, its weights do)
      have_pre = false;
      if (l + 2 < a.n_layers) {
        const __bf16 *Wn = a.wt[l + 1];
#pragma unroll
        for (int j = 0; j < CT; ++j) wpre[j] = *(const bf16x8 *)(Wn + (int64_t)(n0 + j * 16 + fr) * H + fk);
        have_pre = true;
      }
      // with one buffer, every wave must be done reading the layer input before any
      // wave overwrites it with the output (the accumulators hold the whole output tile)
      if (SB) __syncthreads();
      // epilogue: bias + ReLU -> bf16 activation tile; lane (fr, g = lane / 16) holds
      // rows i*16 + fr, columns n0 + 16 j + 4 g + r (r = 0..3)
      const int g4 = (lane >> 4) * 4;
#pragma unroll
      for (int j = 0; j < CT; ++j) {
        const int c = n0 + j * 16 + g4;
        const floatx4 b = *(const floatx4 *)(bias + c);
#pragma unroll
        for (int i = 0; i < RT; ++i) {
          bf16x4 v;
#pragma unroll
          for (int r = 0; r < 4; ++r) v[r] = f2bf(relu(acc[i][j][r] + b[r]));
          *(bf16x4 *)(out + (i * 16 + fr) * P + c) = v;
        }
      }
      __syncthreads();
      __bf16 *t = in; in = out; out = t;
    } else {
      // output layer: NOUT = 16 columns, one column tile; row tiles spread over the waves,
      // a wave's tiles interleaved (independent accumulation chains, each in k order), the
      // weights loaded at once (one L2 latency instead of one per k step)
      bf16x8 wout[H / 32];
#pragma unroll
      for (int k = 0; k < H / 32; ++k) wout[k] = *(const bf16x8 *)(W + (int64_t)fr * H + k * 32 + fk);
      constexpr int NTW = (RT + 3) / 4;
      floatx4 acc[NTW];
#pragma unroll
      for (int t = 0; t < NTW; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int k = 0; k < H / 32; ++k) {
#pragma unroll
        for (int t = 0; t < NTW; ++t) {
          const int i = wave + 4 * t;
          if (i < RT) {
            const bf16x8 af = *(const bf16x8 *)(in + (i * 16 + fr) * P + k * 32 + fk);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, wout[k], acc[t], 0, 0, 0);
          }
        }
      }
      const int c = fr;
      if (c < a.n_out) {
        const float b = bias[c];
#pragma unroll
        for (int t = 0; t < NTW; ++t) {
          const int i = wave + 4 * t;
          if (i >= RT) continue;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int64_t row = row0 + i * 16 + (lane >> 4) * 4 + r;
            if (row < M) Z[row * ldz + c] = acc[t][r] + b;
          }
        }
      }
    }
  }
}

// Fused MLP inference for hidden width 256: 8 waves (512 threads) own 256 rows, in two row
// groups of 128 (rg = wave / 4) and four column groups of 64 (cg = wave % 4), each wave
// 8 x 4 tiles of 16 x 16. The weights of each k step of 32 (all 256 columns, 16 KB) are
// staged in LDS once for the 256 rows (a quarter of the L2 traffic per row of the 4-wave
// form) by direct global-to-LDS loads (global_load_lds_dwordx4, no registers in flight),
// double-buffered. Per k step: wait for this thread's loads of the next step, a barrier
// (every thread's loads visible, every wave's reads of the stage to be refilled done), the
// loads of the step after next, then the LDS reads of the next step's fragments while the
// MFMAs of this step run on fragments read one step earlier. The activation tile
// (256 x 256 bf16, no padding: 16-byte unit u of row r is stored at unit u ^ (r mod 16), so
// that the 16 rows of a fragment read fall in different banks) is overwritten in place by
// each layer's epilogue; with the two stages it fills the 160 KiB of LDS of a CU. Same MFMA
// sequence per output element as the 4-wave form and the layer-by-layer GEMMs:
// bit-identical results.
constexpr int W8_RT = 8;                                   // row tiles per wave
constexpr int W8_ROWS = 2 * 16 * W8_RT;                    // 256 rows per workgroup
constexpr int W8_P = 256;                                  // activation row pitch (bf16)
constexpr size_t W8_ACT = (size_t)W8_ROWS * W8_P * 2;      // 131 072 bytes
constexpr size_t W8_STAGE = 256 * 32 * 2;                  // one k step of weights: 16 KB
constexpr size_t W8_LDS = W8_ACT + 2 * W8_STAGE;           // 163 840 bytes

// element (r, k) of the swizzled activation tile
__device__ __forceinline__ int w8_at(int r, int k) { return r * W8_P + ((((k >> 3) ^ (r & 15)) << 3) | (k & 7)); }

// k step ks of a 256-column weight matrix Wt[n][k] (row length K) into a stage: 16-byte
// slot s holds k group s / 256 (8 values) of column s % 256, so that the 16 lanes of an
// MFMA fragment read (16 consecutive columns, one k group) read 256 consecutive bytes.
// Each of the 512 threads moves two slots; lane i of an instruction writes its 16 bytes
// at the instruction's LDS base + 16 i.
__device__ __forceinline__ void w8_stage_load(unsigned char *stage, const __bf16 *W, int K, int ks, int tid) {
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int s = i * 512 + wave * 64 + lane;
    const int n = s & 255, kg = s >> 8;
    const __bf16 *src = W + (int64_t)n * K + ks * 32 + kg * 8;
    unsigned char *dst = stage + (size_t)(i * 512 + wave * 64) * 16;
    __builtin_amdgcn_global_load_lds((const void *)src, (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
  }
}

__global__ void __launch_bounds__(512) mlp_forward_w8(const __bf16 *__restrict__ X, int64_t ldx, int64_t M,
                                                      MlpArgs a, float *__restrict__ Z, int64_t ldz) {
  constexpr int H = 256, RT = W8_RT, CT = 4, ROWS = W8_ROWS;
  extern __shared__ __align__(16) unsigned char mlp_lds[];
  __bf16 *act = (__bf16 *)mlp_lds;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int rg = wave >> 2, cg = wave & 3;
  const int64_t row0 = (int64_t)blockIdx.x * ROWS;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  // k steps numbered over the layers (step 0: the input layer's single step, then 8 per
  // hidden layer) alternate between the two stages: the input layer's step uses stage 0,
  // step ks of a hidden layer stage (ks + 1) % 2. Every stage index and weight offset below
  // is a compile-time constant; only the layer's weight pointers are read per layer.
  auto stage = [&](int g) { return mlp_lds + W8_ACT + (size_t)(g & 1) * W8_STAGE; };
  w8_stage_load(stage(0), a.wt[0], K0, 0, tid);  // the input layer's weights load while the input tile does
  if (a.x_vec) {
    for (int i = tid; i < ROWS * (K0 / 8); i += 512) {
      const int r = i / (K0 / 8), q = (i - r * (K0 / 8)) * 8;
      const int64_t gr = row0 + r;
      bf16x8 v;
#pragma unroll
      for (int j = 0; j < 8; ++j) v[j] = (__bf16)0.0f;
      if (gr < M && q < a.k_in) {
        v = *(const bf16x8 *)(X + gr * ldx + q);
#pragma unroll
        for (int j = 0; j < 8; ++j)
          if (q + j >= a.k_in) v[j] = (__bf16)0.0f;
      }
      *(bf16x8 *)(act + w8_at(r, q)) = v;
    }
  } else {
    for (int i = tid; i < ROWS * K0; i += 512) {
      const int r = i / K0, k = i - r * K0;
      const int64_t gr = row0 + r;
      act[w8_at(r, k)] = (gr < M && k < a.k_in) ? X[gr * ldx + k] : (__bf16)0.0f;
    }
  }
  const int arow0 = rg * 16 * RT + fr;
  for (int l = 0; l + 1 < a.n_layers; ++l) {
    const float *__restrict__ bias = a.bias[l];
    floatx4 acc[RT][CT];
#pragma unroll
    for (int i = 0; i < RT; ++i)
#pragma unroll
      for (int j = 0; j < CT; ++j) acc[i][j] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    auto loadA = [&](bf16x8 *af, int ks) {
#pragma unroll
      for (int i = 0; i < RT; ++i) af[i] = *(const bf16x8 *)(act + w8_at(arow0 + i * 16, ks * 32 + fk));
    };
    auto loadB = [&](bf16x8 *b, int g) {
      const unsigned char *st = stage(g);
#pragma unroll
      for (int j = 0; j < CT; ++j) b[j] = *(const bf16x8 *)(st + (size_t)((lane >> 4) * 256 + cg * 64 + j * 16 + fr) * 16);
    };
    // this layer's weights and, when the next layer is a hidden one, the next layer's (its
    // first step is staged during this layer's second-to-last step)
    const __bf16 *Wl = a.wt[l];
    const __bf16 *Wn = (l + 2 < a.n_layers) ? a.wt[l + 1] : nullptr;
    auto steps = [&](auto nk_c) {
      constexpr int NK = decltype(nk_c)::value;  // 1: the input layer, 8: a hidden layer
      auto stg = [](int ks) { return NK == 1 ? 0 : ((ks + 1) & 1); };
      bf16x8 aA[RT], aB[RT], bA[CT], bB[CT];
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();  // the layer input is complete, and so are the weights of its first step
      if (NK == 1) {
        if (Wn) w8_stage_load(stage(1), Wn, H, 0, tid);
      } else {
        w8_stage_load(stage(stg(1)), Wl, H, 1, tid);
      }
      loadB(bA, stg(0));
      loadA(aA, 0);
      auto step = [&](int ks, const bf16x8 *aCur, bf16x8 *bCur, bf16x8 *aNext, bf16x8 *bNext) {
        if (ks + 1 < NK) {
          asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
          __syncthreads();
          // the MFMAs of this step read registers only; tie them to the barrier, or the
          // compiler hoists them into the step before, next to the reads that feed them
#pragma unroll
          for (int j = 0; j < CT; ++j) asm volatile("" : "+v"(bCur[j]) :: "memory");
          // refill the stage that step ks used (its fragments were read one step earlier)
          if (ks + 2 < NK) w8_stage_load(stage(stg(ks + 2)), Wl, H, ks + 2, tid);
          else if (Wn) w8_stage_load(stage(1), Wn, H, 0, tid);
          loadB(bNext, stg(ks + 1));
          loadA(aNext, ks + 1);
        }
#pragma unroll
        for (int i = 0; i < RT; ++i)
#pragma unroll
          for (int j = 0; j < CT; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bCur[j], aCur[i], acc[i][j], 0, 0, 0);
      };
#pragma unroll
      for (int ks = 0; ks < NK; ks += 2) {
        step(ks, aA, bA, aB, bB);
        if (ks + 1 < NK) step(ks + 1, aB, bB, aA, bA);
      }
    };
    if (l == 0) steps(std::integral_constant<int, K0 / 32>{});
    else steps(std::integral_constant<int, H / 32>{});
    // every wave has read the layer input: overwrite it with the output
    __syncthreads();
    const int g4 = (lane >> 4) * 4;
#pragma unroll
    for (int j = 0; j < CT; ++j) {
      const int c = cg * 64 + j * 16 + g4;
      const floatx4 bb = *(const floatx4 *)(bias + c);
#pragma unroll
      for (int i = 0; i < RT; ++i) {
        bf16x4 v;
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] = f2bf(relu(acc[i][j][r] + bb[r]));
        *(bf16x4 *)(act + w8_at(arow0 + i * 16, c)) = v;
      }
    }
  }
  __syncthreads();
  // output layer (NOUT = 16 columns): the 16 row tiles over the 8 waves
  const int l = a.n_layers - 1;
  const __bf16 *__restrict__ W = a.wt[l];
  const float *__restrict__ bias = a.bias[l];
  bf16x8 wout[H / 32];
#pragma unroll
  for (int k = 0; k < H / 32; ++k) wout[k] = *(const bf16x8 *)(W + (int64_t)fr * H + k * 32 + fk);
  constexpr int NT = 2 * RT, NTW = (NT + 7) / 8;
  floatx4 acc[NTW];
#pragma unroll
  for (int t = 0; t < NTW; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int k = 0; k < H / 32; ++k) {
#pragma unroll
    for (int t = 0; t < NTW; ++t) {
      const int i = wave + 8 * t;
      if (i < NT) {
        const bf16x8 af = *(const bf16x8 *)(act + w8_at(i * 16 + fr, k * 32 + fk));
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, wout[k], acc[t], 0, 0, 0);
      }
    }
  }
  const int c = fr;
  if (c < a.n_out) {
    const float bv = bias[c];
#pragma unroll
    for (int t = 0; t < NTW; ++t) {
      const int i = wave + 8 * t;
      if (i >= NT) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = row0 + i * 16 + (lane >> 4) * 4 + r;
        if (row < M) Z[row * ldz + c] = acc[t][r] + bv;
      }
    }
  }
}

template <int H>
int launch_mlp(const __bf16 *X, int64_t ldx, int64_t M, const MlpArgs &a, float *Z, int64_t ldz, hipStream_t st) {
  if constexpr (H == 256) {
    const void *fn = (const void *)&mlp_forward_w8;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)W8_LDS) != hipSuccess)
      return GH_EHIP;
    const dim3 grid((unsigned)((M + W8_ROWS - 1) / W8_ROWS));
    hipLaunchKernelGGL(mlp_forward_w8, grid, dim3(512), W8_LDS, st, X, ldx, M, a, Z, ldz);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_EHIP;
  } else {
    constexpr int RT = GH_MLP_RT128, SB = GH_MLP_SB;
    constexpr size_t lds = mlp_lds_bytes<H, RT, SB>();
    const void *fn = (const void *)&mlp_forward_kernel<H, RT, SB>;
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
      return GH_EHIP;
    const dim3 grid((unsigned)((M + 16 * RT - 1) / (16 * RT)));
    hipLaunchKernelGGL((mlp_forward_kernel<H, RT, SB>), grid, dim3(256), lds, st, X, ldx, M, a, Z, ldz);
    return hipGetLastError() == hipSuccess ? GH_OK : GH_EHIP;
  }
}

}  // namespace

extern "C" int gh_mlp_forward_bf16(const void *X, int64_t ldx, int64_t M, int32_t n_layers, int32_t k_in,
                                   int32_t hidden, int32_t n_out, const void *const *wt, const float *const *bias,
                                   float *Z, int64_t ldz, void *stream) {
  if (!X || !wt || !bias || !Z || M < 0 || n_layers < 2 || n_layers > MAXL || k_in <= 0 || k_in > K0 ||
      n_out <= 0 || n_out > NOUT || ldz < n_out || ldx < k_in)
    return GH_EINVAL;
  if (hidden != 128 && hidden != 256) return GH_EINVAL;
  MlpArgs a{};
  a.n_layers = n_layers;
  a.k_in = k_in;
  a.n_out = n_out;
  for (int l = 0; l < n_layers; ++l) {
    if (!wt[l] || !bias[l]) return GH_EINVAL;
    a.wt[l] = (const __bf16 *)wt[l];
    a.bias[l] = bias[l];
  }
  if (M == 0) return GH_OK;
  a.x_vec = ((uintptr_t)X % 16 == 0 && ldx % 8 == 0) ? 1 : 0;
  if (hidden == 256) return launch_mlp<256>((const __bf16 *)X, ldx, M, a, Z, ldz, (hipStream_t)stream);
  return launch_mlp<128>((const __bf16 *)X, ldx, M, a, Z, ldz, (hipStream_t)stream);
}

// ------------------------------------------------------------------ optimiser
// One Adam step over every parameter tensor of a network in one launch (the training loop
// otherwise issues about ten elementwise kernels per tensor). Entry e of the flattened
// launch belongs to tensor i with offset[i] <= e < offset[i+1]; inside the tensor's
// valid_rows x valid_cols corner it takes the Adam update (with L2 weight decay added to the
// gradient when decay != 0), outside it the parameter and both moments stay
// zero (the padding of the bf16 GEMM operands). The bf16 copy, when given, is the parameter
// rounded to nearest even, ready for the next forward pass.
namespace {
constexpr int ADAM_MAX = 32;
struct AdamArgs {
  gh_adam_tensor t[ADAM_MAX];
  int64_t offset[ADAM_MAX + 1];
  int n;
  float lr, b1, b2, eps, c1, c2;  // c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
};

__global__ void __launch_bounds__(256) adam_kernel(AdamArgs a) {
  const int64_t total = a.offset[a.n];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    int i = 0;
    while (e >= a.offset[i + 1]) ++i;  // at most ADAM_MAX tensors, offsets in kernel arguments
    const gh_adam_tensor &t = a.t[i];
    const int64_t j = e - a.offset[i];
    const int64_t r = j / t.cols, c = j - r * t.cols;
    float p = 0.0f, m = 0.0f, v = 0.0f;
    if (r < t.valid_rows && c < t.valid_cols) {
      p = t.param[j];
      const float g = t.decay != 0.0f ? fmaf(t.decay, p, t.grad[j]) : t.grad[j];
      m = fmaf(a.b1, t.m[j], (1.0f - a.b1) * g);
      v = fmaf(a.b2, t.v[j], (1.0f - a.b2) * (g * g));
      const float mh = m * a.c1, vh = v * a.c2;
      p -= a.lr * mh / (sqrtf(vh) + a.eps);
    }
    t.param[j] = p;
    t.m[j] = m;
    t.v[j] = v;
    if (t.param_bf16) ((__bf16 *)t.param_bf16)[j] = f2bf(p);
  }
}
}  // namespace

extern "C" int gh_adam_step(const gh_adam_tensor *tensors, int32_t n_tensors, float lr, float beta1, float beta2,
                            float eps, int32_t step, void *stream) {
  if (!tensors || n_tensors <= 0 || n_tensors > ADAM_MAX || step <= 0 || !(lr >= 0.0f) || !(beta1 >= 0.0f) ||
      !(beta1 < 1.0f) || !(beta2 >= 0.0f) || !(beta2 < 1.0f) || !(eps > 0.0f))
    return GH_EINVAL;
  AdamArgs a{};
  a.n = n_tensors;
  a.offset[0] = 0;
  for (int i = 0; i < n_tensors; ++i) {
    const gh_adam_tensor &t = tensors[i];
    if (!t.param || !t.grad || !t.m || !t.v || t.rows <= 0 || t.cols <= 0 || t.valid_rows < 0 ||
        t.valid_cols < 0 || t.valid_rows > t.rows || t.valid_cols > t.cols)
      return GH_EINVAL;
    a.t[i] = t;
    a.offset[i + 1] = a.offset[i] + (int64_t)t.rows * t.cols;
  }
  a.lr = lr;
  a.b1 = beta1;
  a.b2 = beta2;
  a.eps = eps;
  a.c1 = (float)(1.0 / (1.0 - std::pow((double)beta1, (double)step)));
  a.c2 = (float)(1.0 / (1.0 - std::pow((double)beta2, (double)step)));
  const int64_t total = a.offset[n_tensors];
  const unsigned grid = (unsigned)std::min<int64_t>((total + 255) / 256, 2048);
  hipLaunchKernelGGL(adam_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, a);
  return hipGetLastError() == hipSuccess ? GH_OK : GH_EHIP;
}
