// SPECT detection for the learned detector response (ARF) and for its plain-transport
// counterpart.
//
// gh_spect_pairs: photons that left the phantom (exit point p, direction d, vacuum
// outside) are projected onto every detector head of the gantry. Head k sits at angle
// a_k around the z axis through `center`, with outward normal n_k = (cos a_k, sin a_k, 0)
// and tangential axis e_k = (-sin a_k, cos a_k, 0); its collimator face is the plane
// (x - center).n_k = radius. A photon reaches head k when c = d.n_k > cos(theta_max)
// (larger incidence angles are absorbed by the collimator); it crosses the face at
// h = p + t d, t = (radius - (p - center).n_k)/c, and is binned at the crossing of the
// plane crystal_depth further (GateARFActor: distance_to_crystal). In the head frame
// x' = axial (z), y' = tangential, z' = -n_k (towards the source) the direction is
// d' = (d_z, d.e_k, -c); theta = acos(c), phi = atan2(d'_y, d'_x)
// (GateARFTrainingDatasetActor.cpp:99-107 with the plane normal as third axis).
// Every (photon, head) pair is written with a wave-aggregated atomic slot: the NN input
// row (bf16: E, theta, phi scaled by norm[]), the pixel index (head*ny + iy)*nx + ix,
// the weight, and optionally the photon in the frame of a detector model whose hexagonal
// hole lattice (pitch, first lattice vector along x') has a lattice point at
// patch_origin: the entry point is reduced to the lattice cell, so a finite model patch
// stands for the whole periodic collimator.
//
// gh_energy_windows: deposited energy per event -> energy window (0 = none) after a
// Gaussian blur whose width follows one of the digitizer laws (fixed sigma, inverse-square
// root FWHM = resolution sqrt(reference E), linear FWHM/E), noise from Philox4x32-10
// (key = seed, counter = (event lo, event hi, "BLUR", 0)).
//
// gh_arf_accumulate / gh_counts_accumulate: add the window probabilities of the network
// (softmax of its n_win+1 logits, class 0 = not detected) or the simulated window of
// each pair into projection images [head][window][ny][nx] (FP64 atomics), optionally
// the squared contributions too.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>

#include "../../include/gatehip.h"
#include "gh_device.h"

namespace {

using namespace gh;

constexpr uint32_t BLUR_TAG = 0x424C5552u;  // "BLUR"
constexpr int MAX_WIN = 8;

struct PairOut {
  __bf16 *x16;
  int32_t *pix;
  float *wgt;
  float *det[8];
  float *tp;
  unsigned long long *count;
  int64_t capacity;
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__global__ void __launch_bounds__(256) spect_pairs_kernel(const float *__restrict__ x, const float *__restrict__ y,
                                                          const float *__restrict__ z, const float *__restrict__ u,
                                                          const float *__restrict__ v, const float *__restrict__ w,
                                                          const float *__restrict__ E, const float *__restrict__ wt,
                                                          int64_t n, gh_spect_desc D, const float *__restrict__ cs,
                                                          PairOut o) {
  __shared__ int owner_lane[4][64];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const float two_pi = 6.28318530717958647692f;
  const float dalpha = two_pi / (float)D.n_proj;
  const float sin_max = sqrtf(fmaxf(fmaf(-D.cos_theta_max, D.cos_theta_max, 1.0f), 0.0f));
  // candidate heads: |azimuth(d) - a_k| < theta_max, so k0 +- margin covers them; when
  // that window wraps onto itself every head is visited once instead
  int margin = (int)ceilf(asinf(fminf(sin_max, 1.0f)) / dalpha) + 2;
  const bool all_heads = 2 * margin + 1 >= D.n_proj;
  const int jlo = all_heads ? 0 : -margin, jhi = all_heads ? D.n_proj - 1 : margin;
  const float half_w = 0.5f * (float)D.nx * D.spacing[0], half_h = 0.5f * (float)D.ny * D.spacing[1];
  const float a2x = 0.5f * D.pitch, a2y = 0.86602540378443865f * D.pitch;
  // every lane of a wave runs the same rounds so that slots are assigned with wave-wide
  // ballots
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
    const int64_t i = base + threadIdx.x;
    const bool valid = i < n;
    float px = 0.f, py = 0.f, pz = 0.f, du = 0.f, dv = 0.f, dw = 1.f, e = 0.f, ww = 0.f;
    bool cand = false;
    int k0 = 0;
    if (valid) {
      px = x[i] - D.center[0]; py = y[i] - D.center[1]; pz = z[i] - D.center[2];
      du = u[i]; dv = v[i]; dw = w[i]; e = E[i]; ww = wt[i];
      // only directions within theta_max of the transverse plane can reach a head
      cand = fabsf(dw) < sin_max && ww > 0.0f;
      if (cand && !all_heads) {
        const float beta = atan2f(dv, du) - D.angle0;
        k0 = (int)floorf(beta / dalpha + 0.5f);
      }
    }
    // expand the (candidate photon, head) work items of the wave over its 64 lanes: lane
    // l of round r takes item 64 r + l = (candidate ci, head offset j); the candidate's
    // data comes from its owner lane by a cross-lane read
    const uint64_t cmask = __ballot(cand);
    const int n_cand = __popcll(cmask);
    if (n_cand == 0) continue;
    if (cand) owner_lane[threadIdx.x >> 6][mbcnt(cmask)] = (int)lane_id();
    __builtin_amdgcn_wave_barrier();
    const int width = jhi - jlo + 1;
    const int n_items = n_cand * width;
    for (int it0 = 0; it0 < n_items; it0 += 64) {
      const int item = it0 + (int)lane_id();
      const bool act = item < n_items;
      const int ci = act ? item / width : 0;
      const int j = jlo + (act ? item - ci * width : 0);
      const int src = owner_lane[threadIdx.x >> 6][ci];
      const float qx = __shfl(px, src, 64), qy = __shfl(py, src, 64), qz = __shfl(pz, src, 64);
      const float qu = __shfl(du, src, 64), qv = __shfl(dv, src, 64), qw = __shfl(dw, src, 64);
      const float qe = __shfl(e, src, 64), qwt = __shfl(ww, src, 64);
      const int qk0 = __shfl(k0, src, 64);
      bool hit = false;
      int32_t pixel = 0;
      float c = 0.f, dy = 0.f, hu = 0.f, hz = 0.f;
      int k = 0;
      if (act) {
        const float px = qx, py = qy, pz = qz, du = qu, dv = qv, dw = qw;
        const int k0 = qk0;
        k = ((k0 + j) % D.n_proj + D.n_proj) % D.n_proj;
        const float ca = cs[2 * k], sa = cs[2 * k + 1];
        c = fmaf(du, ca, dv * sa);
        if (c > D.cos_theta_max) {
          const float s = fmaf(px, ca, py * sa);
          const float t = (D.radius - s) / c;
          if (t >= 0.0f) {
            dy = fmaf(-du, sa, dv * ca);  // d . e_k
            const float ex = fmaf(t, du, px), ey = fmaf(t, dv, py);
            hu = fmaf(-ex, sa, ey * ca);  // tangential coordinate on the face
            hz = fmaf(t, dw, pz);         // axial coordinate on the face
            const float f = D.crystal_depth / c;
            const float bu = fmaf(f, dy, hu), bz = fmaf(f, dw, hz);
            const float fx = (bu + half_w) / D.spacing[0], fy = (bz + half_h) / D.spacing[1];
            if (fx >= 0.0f && fy >= 0.0f && fx < (float)D.nx && fy < (float)D.ny) {
              const int ix = min((int)fx, D.nx - 1), iy = min((int)fy, D.ny - 1);
              pixel = (k * D.ny + iy) * D.nx + ix;
              hit = true;
            }
          }
        }
      }
      const uint64_t m = __ballot(hit);
      if (m == 0) continue;
      const uint32_t leader = __ffsll((unsigned long long)m) - 1;
      unsigned long long slot0 = 0;
      if (lane_id() == leader) slot0 = atomicAdd(o.count, (unsigned long long)__popcll(m));
      slot0 = __shfl(slot0, leader, 64);
      if (!hit) continue;
      const uint64_t slot = slot0 + mbcnt(m);
      if ((int64_t)slot >= o.capacity) continue;
      const float theta = acosf(fminf(c, 1.0f));
      const float phi = atan2f(dy, qw);
      if (o.x16) {
        __bf16 *row = o.x16 + slot * 8;
        row[0] = (__bf16)fmaf(qe, D.norm[0], D.norm[1]);
        row[1] = (__bf16)fmaf(theta, D.norm[2], D.norm[3]);
        row[2] = (__bf16)fmaf(phi, D.norm[4], D.norm[5]);
#pragma unroll
        for (int q = 3; q < 8; ++q) row[q] = (__bf16)0.0f;
      }
      if (o.tp) { o.tp[2 * slot] = theta; o.tp[2 * slot + 1] = phi; }
      o.pix[slot] = pixel;
      o.wgt[slot] = qwt;
      if (o.det[0]) {
        // head frame: x' axial (hz), y' tangential (hu); reduce to the lattice cell
        float rx = hz, ry = hu;
        if (D.pitch > 0.0f) {
          const float jj = floorf(ry / a2y);
          ry = fmaf(-jj, a2y, ry);
          rx = fmaf(-jj, a2x, rx);
          const float ii = floorf(rx / D.pitch);
          rx = fmaf(-ii, D.pitch, rx);
        }
        o.det[0][slot] = D.patch_origin[0] + rx;
        o.det[1][slot] = D.patch_origin[1] + ry;
        o.det[2][slot] = D.patch_origin[2];
        o.det[3][slot] = qw;
        o.det[4][slot] = dy;
        o.det[5][slot] = -c;
        o.det[6][slot] = qe;
        o.det[7][slot] = qwt;
      }
    }
  }
}

struct Windows {
  float lo[MAX_WIN], hi[MAX_WIN];
  int n;
};

// standard deviation of the blurred energy of an event with energy e (MeV); FWHM laws are
// converted with 1 / (2 sqrt(2 ln 2))
__device__ __forceinline__ float blur_sigma(const gh_blur &B, float e) {
  constexpr float fwhm_to_sigma = 0.42466090014400953f;
  if (B.law == GH_BLUR_GAUSSIAN) return B.sigma;
  if (B.law == GH_BLUR_INVERSE_SQUARE) return (B.resolution * sqrtf(B.reference * e)) * fwhm_to_sigma;
  return ((B.slope * (e - B.reference) + B.resolution) * e) * fwhm_to_sigma;
}

__global__ void __launch_bounds__(256) energy_windows_kernel(const float *__restrict__ edep, int64_t n, uint32_t k0,
                                                             uint32_t k1, uint64_t first, gh_blur B, Windows W,
                                                             int32_t *__restrict__ cls) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float e = edep[i];
    int c = 0;
    if (e > 0.0f) {
      float eb = e;
      if (B.law != GH_BLUR_NONE) {
        const uint64_t id = first + (uint64_t)i;
        uint32_t x0 = (uint32_t)id, x1 = (uint32_t)(id >> 32), x2 = BLUR_TAG, x3 = 0u;
        philox10(x0, x1, x2, x3, k0, k1);
        const float u0 = (float)((x0 >> 8) | 1u) * 0x1p-24f, u1 = (float)((x1 >> 8) | 1u) * 0x1p-24f;
        float sn, csn;
        gsincos(6.28318530717958647692f * u1, sn, csn);
        const float g = sqrtf(-2.0f * glog(u0, c_logtab)) * csn;
        eb = fmaf(blur_sigma(B, e), g, e);
      }
      for (int k = 0; k < W.n; ++k)
        if (eb >= W.lo[k] && eb < W.hi[k]) { c = k + 1; break; }
    }
    cls[i] = c;
  }
}

__device__ __forceinline__ void add_f64(double *p, double v) {
  __builtin_amdgcn_global_atomic_fadd_f64((__attribute__((address_space(1))) double *)p, v);
}

__global__ void __launch_bounds__(256) arf_accumulate_kernel(const float *__restrict__ logits, int64_t ldl, int64_t m,
                                                             const int32_t *__restrict__ pix,
                                                             const float *__restrict__ wgt, int n_win, int plane,
                                                             double *img, double *img2) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < m; r += stride) {
    const float *l = logits + r * ldl;
    float mx = l[0];
    for (int k = 1; k <= n_win; ++k) mx = fmaxf(mx, l[k]);
    float s = 0.0f;
    for (int k = 0; k <= n_win; ++k) s += __expf(l[k] - mx);
    const float inv = wgt[r] / s;
    const int32_t q = pix[r];
    const int32_t head = q / plane, rem = q - head * plane;
    for (int k = 1; k <= n_win; ++k) {
      const double c = (double)(__expf(l[k] - mx) * inv);
      const int64_t idx = ((int64_t)head * n_win + (k - 1)) * plane + rem;
      add_f64(img + idx, c);
      if (img2) add_f64(img2 + idx, c * c);
    }
  }
}

__global__ void __launch_bounds__(256) counts_accumulate_kernel(const int32_t *__restrict__ cls,
                                                                const int32_t *__restrict__ pix,
                                                                const float *__restrict__ wgt, int64_t m, int n_win,
                                                                int plane, double *img, double *img2) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < m; r += stride) {
    const int c = cls[r];
    if (c <= 0 || c > n_win) continue;
    const int32_t q = pix[r];
    const int32_t head = q / plane, rem = q - head * plane;
    const int64_t idx = ((int64_t)head * n_win + (c - 1)) * plane + rem;
    const double wv = (double)wgt[r];
    add_f64(img + idx, wv);
    if (img2) add_f64(img2 + idx, wv * wv);
  }
}

int grid_for(int64_t n) {
  int dev = 0, n_cu = 256;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
  const int64_t want = (n + 255) / 256;
  return (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)n_cu * 16));
}

}  // namespace

extern "C" int gh_spect_pairs(const float *const rec[8], int64_t n, const gh_spect_desc *det, const float *angle_cs,
                              int64_t capacity, void *x16, int32_t *pix, float *wgt, float *const det_rec[8],
                              float *theta_phi, unsigned long long *count, void *stream) {
  if (!rec || !det || !angle_cs || !pix || !wgt || !count || n < 0 || capacity < 0) return GH_EINVAL;
  for (int k = 0; k < 8; ++k)
    if (!rec[k]) return GH_EINVAL;
  if (det->n_proj <= 0 || det->nx <= 0 || det->ny <= 0 || !(det->spacing[0] > 0.0f) || !(det->spacing[1] > 0.0f) ||
      !(det->cos_theta_max > 0.0f && det->cos_theta_max < 1.0f) || !(det->radius > 0.0f) || det->pitch < 0.0f ||
      (int64_t)det->n_proj * det->nx * det->ny > INT32_MAX)
    return GH_EINVAL;
  PairOut o{};
  o.x16 = (__bf16 *)x16; o.pix = pix; o.wgt = wgt; o.tp = theta_phi; o.count = count; o.capacity = capacity;
  if (det_rec) {
    for (int k = 0; k < 8; ++k) {
      if (!det_rec[k]) return GH_EINVAL;
      o.det[k] = det_rec[k];
    }
  }
  if (n == 0) return GH_OK;
  hipLaunchKernelGGL(spect_pairs_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, rec[0], rec[1], rec[2],
                     rec[3], rec[4], rec[5], rec[6], rec[7], n, *det, angle_cs, o);
  return hipGetLastError() == hipSuccess ? GH_OK : GH_EHIP;
}

// parameters of a blurring law are finite and give a non-negative width at the law's
// reference point
static bool blur_ok(const gh_blur &B) {
  switch (B.law) {
    case GH_BLUR_NONE: return true;
    case GH_BLUR_GAUSSIAN: return B.sigma >= 0.0f && std::isfinite(B.sigma);
    case GH_BLUR_INVERSE_SQUARE:
      return B.resolution >= 0.0f && std::isfinite(B.resolution) && B.reference > 0.0f && std::isfinite(B.reference);
    case GH_BLUR_LINEAR:
      return B.resolution >= 0.0f && std::isfinite(B.resolution) && B.reference >= 0.0f &&
             std::isfinite(B.reference) && std::isfinite(B.slope);
    default: return false;
  }
}

extern "C" int gh_energy_windows(const float *edep, int64_t n, uint64_t seed, uint64_t first, const gh_blur *blur,
                                 const float *win, int32_t n_win, int32_t *cls, void *stream) {
  if (!edep || !cls || !win || !blur || n < 0 || n_win <= 0 || n_win > MAX_WIN || !blur_ok(*blur))
    return GH_EINVAL;
  Windows W{};
  W.n = n_win;
  for (int k = 0; k < n_win; ++k) {
    W.lo[k] = win[2 * k];
    W.hi[k] = win[2 * k + 1];
    if (!(W.hi[k] > W.lo[k])) return GH_EINVAL;
  }
  if (n == 0) return GH_OK;
  hipLaunchKernelGGL(energy_windows_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, edep, n,
                     (uint32_t)seed, (uint32_t)(seed >> 32), first, *blur, W, cls);
  return hipGetLastError() == hipSuccess ? GH_OK : GH_EHIP;
}

extern "C" int gh_arf_accumulate(const float *logits, int64_t ldl, int64_t m, const int32_t *pix, const float *wgt,
                                 int32_t n_win, int32_t plane, double *img, double *img2, void *stream) {
  if (!logits || !pix || !wgt || !img || m < 0 || n_win <= 0 || n_win > MAX_WIN || ldl < n_win + 1 || plane <= 0)
    return GH_EINVAL;
  if (m == 0) return GH_OK;
  hipLaunchKernelGGL(arf_accumulate_kernel, dim3(grid_for(m)), dim3(256), 0, (hipStream_t)stream, logits, ldl, m, pix,
                     wgt, (int)n_win, (int)plane, img, img2);
  return hipGetLastError() == hipSuccess ? GH_OK : GH_EHIP;
}

extern "C" int gh_counts_accumulate(const int32_t *cls, const int32_t *pix, const float *wgt, int64_t m, int32_t n_win,
                                    int32_t plane, double *img, double *img2, void *stream) {
  if (!cls || !pix || !wgt || !img || m < 0 || n_win <= 0 || n_win > MAX_WIN || plane <= 0) return GH_EINVAL;
  if (m == 0) return GH_OK;
  hipLaunchKernelGGL(counts_accumulate_kernel, dim3(grid_for(m)), dim3(256), 0, (hipStream_t)stream, cls, pix, wgt, m,
                     (int)n_win, (int)plane, img, img2);
  return hipGetLastError() == hipSuccess ? GH_OK : GH_EHIP;
}
