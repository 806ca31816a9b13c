// Learned exit-plane source: generator noise and the decoder from generator outputs
// to phase-space records.
//
// gh_gaussian_bf16: N(0,1) noise for the generator input, one Philox4x32-10 block
// (key = seed, counter = (row lo, row hi, "NOIS", column/4)) per 4 columns of a row,
// Box-Muller on the two uniform pairs. A row depends only on (seed, row index), so
// chunked generation reproduces one big batch.
//
// gh_exitplane_decode: the generator emits 6 values per photon in a Gaussianised
// space: (E, x, y, u', v', w'). Each value is mapped back through its feature's
// quantile table (knots at equally spaced Gaussian values g_k in [-g_max, g_max],
// linear between knots, clamped at the ends), then
//   position  (x, y, plane_z) on the exit plane,
//   direction d = u' e1 + v' e2 + w' n (normalised), with n the unit vector from the
//             focal point to (x, y, plane_z) and (e1, e2, n) the right-handed frame
//             e1 = (nx nz, ny nz, -up)/up, e2 = (-ny, nx, 0)/up, up = |(nx, ny)|
//             (up -> 0: e1 = (nz, 0, 0), e2 = (0, 1, 0)).
// Photons that travel along the ray from the focal point therefore have
// (u', v', w') = (0, 0, 1): the generator learns deviations from the point-source
// geometry instead of the steep position-direction correlation itself.
// Records are written as 8 float arrays (x, y, z, u, v, w, E, weight) shifted by an
// offset into the frame of the next simulation.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include <algorithm>

#include "../../include/gatehip.h"
#include "gh_device.h"

namespace {

using namespace gh;

constexpr uint32_t NOISE_TAG = 0x4E4F4953u;  // "NOIS"

__global__ void __launch_bounds__(256) gaussian_bf16_kernel(uint32_t k0, uint32_t k1, uint64_t first, int64_t n,
                                                            int cols, int ld, __bf16 *__restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
    const uint64_t id = first + (uint64_t)r;
    for (int c0 = 0; c0 < ld; c0 += 4) {
      uint32_t x0 = (uint32_t)id, x1 = (uint32_t)(id >> 32), x2 = NOISE_TAG, x3 = (uint32_t)(c0 >> 2);
      philox10(x0, x1, x2, x3, k0, k1);
      const float u0 = (float)((x0 >> 8) | 1u) * 0x1p-24f, u1 = (float)((x1 >> 8) | 1u) * 0x1p-24f;
      const float u2 = (float)((x2 >> 8) | 1u) * 0x1p-24f, u3 = (float)((x3 >> 8) | 1u) * 0x1p-24f;
      const float ra = sqrtf(-2.0f * glog(u0, c_logtab)), rb = sqrtf(-2.0f * glog(u2, c_logtab));
      float sa, ca, sb, cb;
      gsincos(6.28318530717958647692f * u1, sa, ca);
      gsincos(6.28318530717958647692f * u3, sb, cb);
      const float g[4] = {ra * ca, ra * sa, rb * cb, rb * sb};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = c0 + j;
        if (c < ld) out[r * ld + c] = (__bf16)(c < cols ? g[j] : 0.0f);
      }
    }
  }
}

struct EpFrame {
  float fx, fy, fz, plane_z, ox, oy, oz, weight;
};

__device__ __forceinline__ float qdecode(const float *__restrict__ q, int K, float scale, float g_max, float g) {
  float t = (g + g_max) * scale;
  t = fminf(fmaxf(t, 0.0f), (float)(K - 1));  // NaN -> 0
  const int k = min((int)t, K - 2);
  const float a = t - (float)k;
  const float q0 = q[k], q1 = q[k + 1];
  return fmaf(a, q1 - q0, q0);
}

__global__ void __launch_bounds__(256) exitplane_decode_kernel(const float *__restrict__ Z, int64_t ldz, int64_t n,
                                                               const float *__restrict__ qtab, int K, float g_max,
                                                               EpFrame f, float *__restrict__ ox, float *__restrict__ oy,
                                                               float *__restrict__ oz, float *__restrict__ ou,
                                                               float *__restrict__ ov, float *__restrict__ ow,
                                                               float *__restrict__ oE, float *__restrict__ owt) {
  const float scale = (float)(K - 1) / (2.0f * g_max);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float *zr = Z + i * ldz;
    const float E = qdecode(qtab, K, scale, g_max, zr[0]);
    const float px = qdecode(qtab + K, K, scale, g_max, zr[1]);
    const float py = qdecode(qtab + 2 * K, K, scale, g_max, zr[2]);
    float lu = qdecode(qtab + 3 * K, K, scale, g_max, zr[3]);
    float lv = qdecode(qtab + 4 * K, K, scale, g_max, zr[4]);
    float lw = qdecode(qtab + 5 * K, K, scale, g_max, zr[5]);
    const float l2 = fmaf(lw, lw, fmaf(lv, lv, lu * lu));
    if (!(l2 > 0.0f)) { lu = 0.0f; lv = 0.0f; lw = 1.0f; }
    else normalize3(lu, lv, lw);
    float nx = px - f.fx, ny = py - f.fy, nz = f.plane_z - f.fz;
    normalize3(nx, ny, nz);
    const float up = sqrtf(fmaf(ny, ny, nx * nx));
    float e1x, e1y, e1z, e2x, e2y;
    if (up > 1e-7f) {
      const float iu = 1.0f / up;
      e1x = (nx * nz) * iu; e1y = (ny * nz) * iu; e1z = -up;
      e2x = -ny * iu; e2y = nx * iu;
    } else {
      e1x = nz; e1y = 0.0f; e1z = 0.0f;
      e2x = 0.0f; e2y = 1.0f;
    }
    float du = fmaf(lw, nx, fmaf(lv, e2x, lu * e1x));
    float dv = fmaf(lw, ny, fmaf(lv, e2y, lu * e1y));
    float dw = fmaf(lw, nz, lu * e1z);
    normalize3(du, dv, dw);
    ox[i] = px + f.ox; oy[i] = py + f.oy; oz[i] = f.plane_z + f.oz;
    ou[i] = du; ov[i] = dv; ow[i] = dw;
    oE[i] = E; owt[i] = f.weight;
  }
}

int grid_for(int64_t n) {
  int dev = 0, n_cu = 256;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
  const int64_t want = (n + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)n_cu * 16));
}

}  // namespace

extern "C" int gh_gaussian_bf16(uint64_t seed, uint64_t first_row, int64_t n_rows, int32_t n_cols, int32_t ld,
                                void *out, void *stream) {
  if (!out || n_rows < 0 || n_cols < 0 || ld < n_cols || ld <= 0) return GH_EINVAL;
  if (n_rows == 0) return GH_OK;
  hipLaunchKernelGGL(gaussian_bf16_kernel, dim3(grid_for(n_rows)), dim3(256), 0, (hipStream_t)stream, (uint32_t)seed,
                     (uint32_t)(seed >> 32), first_row, n_rows, (int)n_cols, (int)ld, (__bf16 *)out);
  return hipGetLastError() == hipSuccess ? GH_OK : GH_EHIP;
}

extern "C" int gh_exitplane_decode(const float *Z, int64_t ldz, int64_t n, const float *qtab, int32_t n_knots,
                                   float g_max, const float focal[3], float plane_z, const float offset[3],
                                   float weight, float *const out[8], void *stream) {
  if (!Z || !qtab || !focal || !offset || !out || n < 0 || ldz < 6 || n_knots < 2 || !(g_max > 0.0f))
    return GH_EINVAL;
  for (int k = 0; k < 8; ++k)
    if (!out[k]) return GH_EINVAL;
  if (n == 0) return GH_OK;
  const EpFrame f{focal[0], focal[1], focal[2], plane_z, offset[0], offset[1], offset[2], weight};
  hipLaunchKernelGGL(exitplane_decode_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, Z, ldz, n, qtab,
                     (int)n_knots, g_max, f, out[0], out[1], out[2], out[3], out[4], out[5], out[6], out[7]);
  return hipGetLastError() == hipSuccess ? GH_OK : GH_EHIP;
}
