// Device building blocks of the photon transport: counter-based RNG, FP32 math with a
// fixed operation order (so that a CPU build of the same sequence gives the same bits),
// table lookups and interaction samplers. Everything here is written for wave64 CDNA4.
//
// Numerical contract (also followed by the CPU restatement in oracle/transport_ref.c):
//   * compiled with -ffp-contract=off: every fused multiply-add is an explicit fmaf;
//   * division and sqrtf are IEEE correctly rounded (HIP default for FP32);
//   * log and sin/cos are the series below, never the math library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GH_DEV __device__ __forceinline__

namespace gh {

// ---------------------------------------------------------------- constants
constexpr int KBITS = 6;
constexpr int ESHIFT = 23 - KBITS;                    // 17
constexpr uint32_t EMASK = (1u << ESHIFT) - 1u;
constexpr int BUCKET_BASE = ((127 - 10) << 23) >> ESHIFT;  // 7488: E_LO = 2^-10 MeV
// q^2 grid of the form-factor tables: node 0 at 0, node j >= 1 at the float with bits
// (X_BUCKET_BASE + j - 1) << XSHIFT (8 nodes per octave from 2^-12 A^-2)
constexpr int XSHIFT = 20;
constexpr uint32_t XMASK = (1u << XSHIFT) - 1u;
constexpr int X_BUCKET_BASE = ((127 - 12) << 23) >> XSHIFT;  // 920
constexpr float MEC2 = 0x1.05a1a8p-1f;                // 0.51099895 MeV
constexpr float KPM = 0x1.fac5e8p+8f;                 // photon wave number per MeV [1/Angstrom]
constexpr float EXIT_EPS = 1e-3f;                     // shortest in-box chord that counts as a hit [mm]
constexpr int MAX_REJECT = 1000;                      // cap of every rejection loop

// ---------------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC11).
struct Rng {
  uint32_t k0, k1;      // key = seed
  uint32_t c0, c1, c2;  // history id lo/hi, simulation id
  uint32_t c3;          // block counter
  uint32_t b0, b1, b2, b3;
  int n;                // words left in the buffer (consumed b0 first)
};

GH_DEV void philox10(uint32_t &x0, uint32_t &x1, uint32_t &x2, uint32_t &x3, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * x0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
    // the round's two three-way XORs as one bitop3 each (truth table 0x96 = a ^ b ^ c):
    // the compiler emits two XORs per output otherwise
    const uint32_t y0 = __builtin_amdgcn_bitop3_b32((uint32_t)(p1 >> 32), x1, k0, 0x96);
    const uint32_t y2 = __builtin_amdgcn_bitop3_b32((uint32_t)(p0 >> 32), x3, k1, 0x96);
    x1 = (uint32_t)p1;
    x3 = (uint32_t)p0;
    x0 = y0;
    x2 = y2;
  }
}

GH_DEV void rng_init(Rng &g, uint64_t seed, uint64_t history, uint32_t sim) {
  g.k0 = (uint32_t)seed; g.k1 = (uint32_t)(seed >> 32);
  g.c0 = (uint32_t)history; g.c1 = (uint32_t)(history >> 32); g.c2 = sim; g.c3 = 0;
  g.n = 0;
}

GH_DEV uint32_t rng_next(Rng &g) {
  if (g.n == 0) {
    uint32_t x0 = g.c0, x1 = g.c1, x2 = g.c2, x3 = g.c3;
    philox10(x0, x1, x2, x3, g.k0, g.k1);
    g.b0 = x0; g.b1 = x1; g.b2 = x2; g.b3 = x3;
    g.c3 += 1u;
    g.n = 4;
  }
  const uint32_t r = g.b0;
  g.b0 = g.b1; g.b1 = g.b2; g.b2 = g.b3;
  g.n -= 1;
  return r;
}

// uniform in (0, 1): odd multiples of 2^-24
GH_DEV float u01(Rng &g) { return (float)((rng_next(g) >> 8) | 1u) * 0x1p-24f; }

// ---------------------------------------------------------------- math
// Table of glog (written by gen_log_table.py). x = 2^e m with m in [0.5, 1) (frexp), i =
// the top six mantissa bits of x: entry i covers m in [1/2 + i/128, 1/2 + (i+1)/128) and
// holds rcp_i, the float nearest 1 / c_i with c_i the middle of its interval (c = 1 for the
// last interval, next to 1), and lg_i, the float nearest -log(rcp_i).
static __device__ __constant__ float2 c_logtab[64] = {
    {0x1.fc07fp+0f, -0x1.5ee82ap-1f}, {0x1.f4465ap+0f, -0x1.5707a2p-1f}, {0x1.ecc07cp+0f, -0x1.4f45aap-1f}, {0x1.e573acp+0f, -0x1.47a152p-1f},
    {0x1.de5d6ep+0f, -0x1.4019c2p-1f}, {0x1.d77b66p+0f, -0x1.38ae22p-1f}, {0x1.d0cb58p+0f, -0x1.315da4p-1f}, {0x1.ca4b3p+0f, -0x1.2a2786p-1f},
    {0x1.c3f8fp+0f, -0x1.230b0ep-1f}, {0x1.bdd2b8p+0f, -0x1.1c0784p-1f}, {0x1.b7d6c4p+0f, -0x1.151c4p-1f}, {0x1.b20364p+0f, -0x1.0e4898p-1f},
    {0x1.ac5702p+0f, -0x1.078bfp-1f}, {0x1.a6d01ap+0f, -0x1.00e5aep-1f}, {0x1.a16d4p+0f, -0x1.f4aa8p-2f}, {0x1.9c2d14p+0f, -0x1.e7b42ap-2f},
    {0x1.970e5p+0f, -0x1.dae756p-2f}, {0x1.920fb4p+0f, -0x1.ce42fp-2f}, {0x1.8d3018p+0f, -0x1.c1c604p-2f}, {0x1.886e6p+0f, -0x1.b56fa2p-2f},
    {0x1.83c978p+0f, -0x1.a93ed4p-2f}, {0x1.7f406p+0f, -0x1.9d32cp-2f}, {0x1.7ad22p+0f, -0x1.914a84p-2f}, {0x1.767dcep+0f, -0x1.858556p-2f},
    {0x1.724288p+0f, -0x1.79e266p-2f}, {0x1.6e1f76p+0f, -0x1.6e60ecp-2f}, {0x1.6a13cep+0f, -0x1.630034p-2f}, {0x1.661ec6p+0f, -0x1.57bf74p-2f},
    {0x1.623fa8p+0f, -0x1.4c9e0cp-2f}, {0x1.5e75bcp+0f, -0x1.419b44p-2f}, {0x1.5ac056p+0f, -0x1.36b676p-2f}, {0x1.571ed4p+0f, -0x1.2bef08p-2f},
    {0x1.539094p+0f, -0x1.214456p-2f}, {0x1.501502p+0f, -0x1.16b5cep-2f}, {0x1.4cab88p+0f, -0x1.0c42d6p-2f}, {0x1.49539ep+0f, -0x1.01eae4p-2f},
    {0x1.460cbcp+0f, -0x1.ef5adcp-3f}, {0x1.42d662p+0f, -0x1.db13d8p-3f}, {0x1.3fb014p+0f, -0x1.c6ffbcp-3f}, {0x1.3c995ap+0f, -0x1.b31d84p-3f},
    {0x1.3991c2p+0f, -0x1.9f6c3cp-3f}, {0x1.3698ep+0f, -0x1.8beb04p-3f}, {0x1.33ae46p+0f, -0x1.7898dap-3f}, {0x1.30d19p+0f, -0x1.6574ecp-3f},
    {0x1.2e025cp+0f, -0x1.527e5ep-3f}, {0x1.2b404ap+0f, -0x1.3fb454p-3f}, {0x1.288b02p+0f, -0x1.2d1616p-3f}, {0x1.25e228p+0f, -0x1.1aa2bep-3f},
    {0x1.234568p+0f, -0x1.08598ep-3f}, {0x1.20b47p+0f, -0x1.ec738ep-4f}, {0x1.1e2ef4p+0f, -0x1.c88584p-4f}, {0x1.1bb4a4p+0f, -0x1.a4e764p-4f},
    {0x1.194538p+0f, -0x1.8197e2p-4f}, {0x1.16e068p+0f, -0x1.5e959cp-4f}, {0x1.1485fp+0f, -0x1.3bdf4ep-4f}, {0x1.12358ep+0f, -0x1.1973b6p-4f},
    {0x1.0fef02p+0f, -0x1.eea338p-5f}, {0x1.0db20ap+0f, -0x1.aaef1cp-5f}, {0x1.0b7e6ep+0f, -0x1.67c938p-5f}, {0x1.0953f4p+0f, -0x1.252f4p-5f},
    {0x1.07326p+0f, -0x1.c63d06p-6f}, {0x1.05198p+0f, -0x1.432ab2p-6f}, {0x1.03091cp+0f, -0x1.8244ep-7f}, {0x1p+0f, -0x0p+0f}
};

// log(x) for x > 0 normal: log x = e ln2 + lg_i + log1p(r), r = m rcp_i - 1 (one fma,
// |r| <= 2^-7; exact just below 1), log1p by its series to r^4 (the r^5 term is below
// 1e-11). No division. Relative error below 2e-7 on (0, 1], the arguments it is used for
// (logarithms of uniform variates); above 1 the result keeps an absolute error of a few
// 1e-8 (ln 2 - ln 2 cancels next to 1). tab: c_logtab or a copy in LDS.
GH_DEV float glog(float x, const float2 *tab) {
  const uint32_t i = (__float_as_uint(x) >> 17) & 63u;
  const float m = __builtin_amdgcn_frexp_mantf(x);
  const float ef = (float)__builtin_amdgcn_frexp_expf(x);
  const float2 t = tab[i];
  const float r = fmaf(m, t.x, -1.0f);
  float p = fmaf(r, -0.25f, 0x1.555556p-2f);
  p = fmaf(r, p, -0.5f);
  p = fmaf(r, p, 1.0f);
  const float l1 = fmaf(r, p, t.y);
  return fmaf(ef, 0x1.62ep-1f, fmaf(ef, 0x1.0bfbe8p-15f, l1));
}

GH_DEV void gsincos(float x, float &sn, float &cs) {
  // |x| <= 16 pi. Cody-Waite reduction by pi/2, Taylor polynomials on [-pi/4, pi/4].
  const float k = rintf(x * 0x1.45f306p-1f);
  const int q = (int)k;
  float r = fmaf(-k, 0x1.921fp+0f, x);
  r = fmaf(-k, 0x1.6a88p-17f, r);
  r = fmaf(-k, 0x1.0b4612p-34f, r);
  const float r2 = r * r;
  float t = fmaf(r2, 0x1.71de3ap-19f, -0x1.a01a02p-13f);
  t = fmaf(r2, t, 0x1.111112p-7f);
  t = fmaf(r2, t, -0x1.555556p-3f);
  const float sp = fmaf(r * r2, t, r);
  float c = fmaf(r2, -0x1.27e4fcp-22f, 0x1.a01a02p-16f);
  c = fmaf(r2, c, -0x1.6c16c2p-10f);
  c = fmaf(r2, c, 0x1.555556p-5f);
  const float cp = fmaf(r2 * r2, c, fmaf(-0.5f, r2, 1.0f));
  switch (q & 3) {
    case 0: sn = sp; cs = cp; break;
    case 1: sn = cp; cs = -sp; break;
    case 2: sn = -sp; cs = -cp; break;
    default: sn = -cp; cs = sp; break;
  }
}

GH_DEV void normalize3(float &x, float &y, float &z) {
  const float n2 = fmaf(z, z, fmaf(y, y, x * x));
  const float inv = 1.0f / sqrtf(n2);
  x = x * inv; y = y * inv; z = z * inv;
}

// (cos phi, sin phi) for phi = 2 pi u, u uniform in (0, 1): one draw and no rejection
// loop (a rejection loop runs as long as its slowest lane in a wave)
GH_DEV void azimuth(Rng &g, float &cp, float &sp) {
  gsincos(6.28318530717958647692f * u01(g), sp, cp);
}

// new direction after scattering by polar cosine c and azimuth (cp, sp)
GH_DEV void rotate(float &u, float &v, float &w, float c, float cp, float sp) {
  const float st = sqrtf(fmaxf(fmaf(-c, c, 1.0f), 0.0f));
  float nu, nv, nw;
  if (fabsf(w) < 0.99999f) {
    const float tmp = sqrtf(fmaf(-w, w, 1.0f));
    const float f = st / tmp;
    nu = fmaf(c, u, f * fmaf(u * w, cp, -(v * sp)));
    nv = fmaf(c, v, f * fmaf(v * w, cp, u * sp));
    nw = fmaf(c, w, -((st * cp) * tmp));
  } else {
    nu = st * cp;
    nv = st * sp;
    nw = (w > 0.0f) ? c : -c;
  }
  normalize3(nu, nv, nw);
  u = nu; v = nv; w = nw;
}

// largest i in [0, n-2] with a[i] <= x (a ascending)
GH_DEV int bsearch_le(const float *a, int n, float x) {
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

// Guide tables (C. Chen and H. Asau, "On generating random variates from an empirical
// distribution", AIIE Transactions 6 (1974) 163-166): for a non-decreasing table and
// 2^k equal bins of [0, 1), guide[b] is the search result at the bin's lower edge, so a
// search for a value in bin b only has to look between guide[b] and guide[b + 1] and
// returns exactly what the search over the whole table returns.
constexpr int SPEC_GUIDE = 256;    // bins of the source-spectrum CDF guide
constexpr int RAY_GUIDE = 1024;    // bins of the Rayleigh int F^2 guide (per material)

// largest i in [lo, hi-1] with a[i] <= x, given a[lo] <= x and (a[hi] > x or hi = n-1)
GH_DEV int bsearch_le_range(const float *a, int lo, int hi, float x) {
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

// bsearch_le(a, n, x) for x >= 0 with guide[b] = bsearch_le(a, n, b / RAY_GUIDE)
GH_DEV int bsearch_le_guided(const float *a, const int32_t *guide, int n, float x) {
  const int b = min((int)(x * (float)RAY_GUIDE), RAY_GUIDE - 1);
  return bsearch_le_range(a, guide[b], min(guide[b + 1] + 1, n - 1), x);
}

// smallest k in [lo, hi] with cdf[k] > u, given cdf[lo-1] <= u (or lo = 0) and cdf[hi] > u
GH_DEV int cdf_index_range(const float *cdf, int lo, int hi, float u) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cdf[mid] > u) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// cdf_index(cdf, n, u) for u in (0, 1) with guide[b] = cdf_index(cdf, n, b / SPEC_GUIDE)
GH_DEV int cdf_index_guided(const float *cdf, const int32_t *guide, float u) {
  const int b = (int)(u * (float)SPEC_GUIDE);
  return cdf_index_range(cdf, guide[b], guide[b + 1], u);
}

// smallest k in [0, n-1] with cdf[k] > u (cdf ascending, cdf[n-1] = 1 > u)
GH_DEV int cdf_index(const float *cdf, int n, float u) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cdf[mid] > u) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// ---------------------------------------------------------------- voxel index
// Index of the cell of a grid (lower corner c, spacing s > 0, inv ~ 1/s) that holds the
// point p: floor((p - c) / s) evaluated exactly for the real values of the FP32 inputs.
// With the image origin o = c + s/2 at the centre of voxel 0 this is ITK's
// TransformPhysicalPointToIndex, RoundHalfIntegerUp((p - o) / s): a point on a face
// belongs to the upper voxel. floorf((p - c) * inv) alone is off by one for points
// within about an ulp of a face (the rounding of 1/s and of the product); the estimate
// is corrected with the exact residual p - c - k s: TwoSum gives p - c = d + e exactly,
// and fmaf(-k, s, d) is exact wherever it is small enough to decide the sign.
GH_DEV int cell_index(float p, float c, float s, float inv) {
  const float d = p - c;
  const float bv = d - p;
  const float av = d - bv;
  const float e = (p - av) + (-c - bv);
  float k = floorf(d * inv);
  if (fmaf(-k, s, d) + e < 0.0f) {
    k -= 1.0f;
  } else if (fmaf(-(k + 1.0f), s, d) + e >= 0.0f) {
    k += 1.0f;
  }
  return (int)k;
}

// ---------------------------------------------------------------- energy grid
struct ECache {
  int b;         // bucket
  float frac;    // position inside the bucket
  float maj22;   // majorant at E times 2^-22 (the accept test's uniform without its scaling)
  float inv_maj;
};

GH_DEV void ecache_set(ECache &c, float E, const float *maj) {
  const uint32_t bits = __float_as_uint(E);
  c.b = (int)(bits >> ESHIFT) - BUCKET_BASE;
  c.frac = (float)(bits & EMASK) * 0x1p-17f;
  const float m0 = maj[c.b], m1 = maj[c.b + 1];
  const float m = fmaf(c.frac, m1 - m0, m0);
  c.maj22 = m * 0x1p-22f;
  c.inv_maj = 1.0f / m;
}

GH_DEV float lerp_node(const float *row, const ECache &c) {
  const float v0 = row[c.b], v1 = row[c.b + 1];
  return fmaf(c.frac, v1 - v0, v0);
}

// value of material (rho, row) at E: interpolation between its node values, each the FP32
// product rho * row[node] (with rho = 1, the row itself); a table of those products gives
// the same values through lerp_node
GH_DEV float lerp_node_scaled(const float *row, float rho, const ECache &c) {
  const float v0 = rho * row[c.b], v1 = rho * row[c.b + 1];
  return fmaf(c.frac, v1 - v0, v0);
}

// value of a table on the q^2 grid at x >= 0: the interval comes from the bits of x
// (no search), the position inside it from the low mantissa bits; clamped at the top
GH_DEV float qgrid_value(const float *ys, int n, float x) {
  int j;
  float f;
  if (x < 0x1p-12f) {
    j = 0;
    f = x * 0x1p12f;
  } else {
    const uint32_t bits = __float_as_uint(x);
    j = (int)(bits >> XSHIFT) - X_BUCKET_BASE + 1;
    f = (float)(bits & XMASK) * 0x1p-20f;
  }
  if (j > n - 2) { j = n - 2; f = 1.0f; }
  const float y0 = ys[j], y1 = ys[j + 1];
  return fmaf(f, y1 - y0, y0);
}

// material m: its attenuation is rho * (table row `row` of the mu tables); rho is the
// density when materials share composition rows, else 1 (a row per material)
struct MatRow {
  float rho;
  int row;
};

struct Phys {
  const MatRow *mat;     // [M]
  const float *mu;       // [R][NB+1] table rows (global copy)
  const int32_t *rguide; // [R][RAY_GUIDE+1] guide of rcdf
  const float *maj;      // [NB+1]
  const float *pcpr;     // [R][NB+1][2] (sigma_C/mu, (sigma_C+sigma_R)/mu) interleaved
  const float *muen;     // [R][NB+1]
  const float *rx;       // [NX]
  const float *rcdf;     // [R][NX]
  const float *sinc;     // [R][NX]
  const float *inv_rho;  // [M]
  int n_mat, n_rows, nb, nx;
  int shared_rows;       // 1: materials share composition rows (mat[m] = {density, row})
  int mu_expanded;       // launch: the staged mu table holds a row of node products per material
  float e_cut;
};

// Table pointers a kernel reads through (LDS copies or the global arrays)
struct Tabs {
  const float *mu, *maj, *pcpr, *rx, *rcdf, *sinc;
  const int32_t *rguide;
};

// branching ratios at a collision: (sigma_C/mu, (sigma_C+sigma_R)/mu) interpolated at E
GH_DEV void branching(const Tabs &T, int nb1, int lab, const ECache &c, float &pc, float &pr) {
  const float2 a = *(const float2 *)(T.pcpr + 2 * (lab * nb1 + c.b));
  const float2 b = *(const float2 *)(T.pcpr + 2 * (lab * nb1 + c.b + 1));
  pc = fmaf(c.frac, b.x - a.x, a.x);
  pr = fmaf(c.frac, b.y - a.y, a.y);
}

// One try of the Compton sampler below (Kahn's method, then acceptance with S(q)/Z): the
// transport kernel runs a bounded number of tries per pass and resumes a photon whose
// tries were all rejected in a later pass, so that a wave does not wait for its slowest
// lane. The draws of a photon are the same either way. Returns true if accepted.
GH_DEV bool compton_try(Rng &g, const Tabs &T, int nx, int lab, float E, float &q, float &c) {
  const float alpha = E / MEC2;
  const float a2 = alpha + alpha;
  const float onep = 1.0f + a2;
  const float thr = onep / (9.0f + a2);
  const float r1 = u01(g), r2 = u01(g), r3 = u01(g);
  const float t = fmaf(a2, r2, 1.0f);
  const bool low = r1 <= thr;
  q = low ? t : onep / t;
  const float iq = 1.0f / q;
  c = 1.0f - (q - 1.0f) / alpha;
  const bool acc = low ? (r3 <= 4.0f * (iq - iq * iq)) : (r3 <= 0.5f * fmaf(c, c, iq));
  if (!acc) return false;
  const float kk = E * KPM;
  const float two_k2 = (2.0f * kk) * kk;
  const float x = two_k2 * (1.0f - c);
  const float s = qgrid_value(T.sinc + lab * nx, nx, x);
  return u01(g) <= s;
}

// One try of the Rayleigh sampler below: returns true if accepted; c is the try's cosine.
GH_DEV bool rayleigh_try(Rng &g, const Tabs &T, int nx, int lab, float E, float &c) {
  const float kk = E * KPM;
  const float two_k2 = (2.0f * kk) * kk;
  const float xmax = two_k2 + two_k2;
  const float *G = T.rcdf + lab * nx;
  const float gmax = qgrid_value(G, nx, xmax);
  const float t = u01(g) * gmax;
  const int j = bsearch_le_guided(G, T.rguide + lab * (RAY_GUIDE + 1), nx, t);
  const float g0 = G[j], g1 = G[j + 1];
  const float dg = g1 - g0;
  const float x0 = T.rx[j], x1 = T.rx[j + 1];
  const float x = (dg > 0.0f) ? fmaf((t - g0) / dg, x1 - x0, x0) : x0;
  c = fmaxf(1.0f - x / two_k2, -1.0f);
  return u01(g) <= 0.5f * fmaf(c, c, 1.0f);
}

// Klein-Nishina by Kahn's rejection method (H. Kahn, "Applications of Monte Carlo",
// RM-1237-AEC, 1956) followed by acceptance with S(q)/Z. Returns q = E/E' and cos(theta).
GH_DEV void sample_compton(Rng &g, const Tabs &T, int nx, int lab, float E, float &q_out, float &c_out,
                           uint32_t &rejections) {
  const float alpha = E / MEC2;
  const float a2 = alpha + alpha;
  const float onep = 1.0f + a2;
  const float thr = onep / (9.0f + a2);
  const float kk = E * KPM;
  const float two_k2 = (2.0f * kk) * kk;
  const float *srow = T.sinc + lab * nx;
  float q = 1.0f, c = 1.0f;
  int it = 0;
  for (; it < MAX_REJECT; ++it) {
    const float r1 = u01(g), r2 = u01(g), r3 = u01(g);
    // both branches of Kahn's method as selects over shared terms (a wave would run both
    // branches anyway): q = t or (1 + 2 alpha) / t, then 1/q and cos theta once
    const float t = fmaf(a2, r2, 1.0f);
    const bool low = r1 <= thr;
    q = low ? t : onep / t;
    const float iq = 1.0f / q;
    c = 1.0f - (q - 1.0f) / alpha;
    const bool acc = low ? (r3 <= 4.0f * (iq - iq * iq)) : (r3 <= 0.5f * fmaf(c, c, iq));
    if (!acc) { rejections++; continue; }
    const float x = two_k2 * (1.0f - c);
    const float s = qgrid_value(srow, nx, x);
    if (u01(g) <= s) break;
    rejections++;
  }
  if (it == MAX_REJECT) { q = 1.0f; c = 1.0f; }
  q_out = q;
  c_out = fmaxf(c, -1.0f);
}

// Rayleigh: q^2 from the tabulated int F^2 (inverse by linear interpolation), then
// acceptance with (1 + cos^2)/2.
GH_DEV float sample_rayleigh(Rng &g, const Tabs &T, int nx, int lab, float E) {
  const float kk = E * KPM;
  const float two_k2 = (2.0f * kk) * kk;
  const float xmax = two_k2 + two_k2;
  const float *G = T.rcdf + lab * nx;
  const float gmax = qgrid_value(G, nx, xmax);
  float c = 1.0f;
  for (int it = 0; it < MAX_REJECT; ++it) {
    const float t = u01(g) * gmax;
    const int j = bsearch_le_guided(G, T.rguide + lab * (RAY_GUIDE + 1), nx, t);
    const float g0 = G[j], g1 = G[j + 1];
    const float dg = g1 - g0;
    const float x0 = T.rx[j], x1 = T.rx[j + 1];
    const float x = (dg > 0.0f) ? fmaf((t - g0) / dg, x1 - x0, x0) : x0;
    c = fmaxf(1.0f - x / two_k2, -1.0f);
    if (u01(g) <= 0.5f * fmaf(c, c, 1.0f)) break;
  }
  return c;
}

}  // namespace gh
