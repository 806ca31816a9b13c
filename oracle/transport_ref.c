/* transport_ref.c -- CPU restatement of the photon transport + voxel dose path.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the oracle the GPU library is checked
 * against (and the CPU baseline that bench.py times). Nothing in the product links,
 * loads or calls it.
 *
 * It restates, history by history, the algorithm the HIP kernel runs: Philox4x32-10
 * streams keyed per history, source sampling (GateGenericSource.cpp:312-664,
 * GateSPSEneDistribution.cpp:16-159, GatePhaseSpaceSource.cpp:142-228), Woodcock
 * tracking through the voxel labels (the material lookup of
 * GateImageNestedParameterisation::ComputeMaterial, GateImageNestedParameterisation.cpp:47-75),
 * Compton / Rayleigh / photoelectric interactions with local deposition of electron
 * energy, and the dose tally of GateDoseActor::SteppingAction / ScoreSquaredValue
 * (GateDoseActor.cpp:180-243, 359-397): edep = sum of deposits, edep^2 = sum over
 * histories of the squared per-history voxel sum.
 *
 * Numerical contract shared with the kernel: FP32 arithmetic, compiled with
 * -ffp-contract=off, every fused multiply-add spelled fmaf, correctly rounded / and
 * sqrtf, log and sin/cos from the series below. With identical inputs each history
 * then takes the same path on both sides, so integer counters agree exactly and FP64
 * tallies agree up to the order of the FP64 additions.
 *
 * The squared tally is computed two ways: the 4-slot per-history cache the kernel
 * uses (FIFO eviction of the oldest voxel when a history touches a fifth voxel) and,
 * in `exact_edep2`, the exact per-history voxel sums (no slot limit).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/gatehip.h"

/* ------------------------------------------------------------------ Philox */
void ref_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint64_t a = (uint64_t)0xD2511F53u * (uint64_t)x0;
    uint64_t b = (uint64_t)0xCD9E8D57u * (uint64_t)x2;
    uint32_t n0 = (uint32_t)(b >> 32) ^ x1 ^ k0;
    uint32_t n1 = (uint32_t)b;
    uint32_t n2 = (uint32_t)(a >> 32) ^ x3 ^ k1;
    uint32_t n3 = (uint32_t)a;
    x0 = n0; x1 = n1; x2 = n2; x3 = n3;
  }
  out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

/* 21-bit fields of one step-stream block w[0..3] for tentative step k of three: step 0
 * takes the low 21 bits of w0 (distance) and w1 (accept), step 1 those of w2 and w3,
 * step 2 the top 11 bits of w0 | w2 (distance) and of w1 | w3 (accept), the w2 / w3 part
 * above with its top bit dropped. The uniform is then (2 b + 1) 2^-22. */
static void step_fields(const uint32_t w[4], int k, uint32_t *bd, uint32_t *ba) {
  const uint32_t m21 = 0x1FFFFFu;
  if (k < 2) {
    *bd = w[2 * k] & m21;
    *ba = w[2 * k + 1] & m21;
  } else {
    *bd = ((w[0] >> 21) | (w[2] >> 21 << 11)) & m21;
    *ba = ((w[1] >> 21) | (w[3] >> 21 << 11)) & m21;
  }
}

typedef struct {
  uint32_t key[2];
  uint32_t ctr[4];
  uint32_t buf[4];
  int pos; /* next word of buf to hand out; 4 = empty */
} stream_t;

static void stream_open(stream_t *s, uint64_t seed, uint64_t history, uint32_t sim) {
  s->key[0] = (uint32_t)seed;
  s->key[1] = (uint32_t)(seed >> 32);
  s->ctr[0] = (uint32_t)history;
  s->ctr[1] = (uint32_t)(history >> 32);
  s->ctr[2] = sim;
  s->ctr[3] = 0;
  s->pos = 4;
}

static uint32_t stream_word(stream_t *s) {
  if (s->pos == 4) {
    ref_philox4x32_10(s->ctr, s->key, s->buf);
    s->ctr[3]++;
    s->pos = 0;
  }
  return s->buf[s->pos++];
}

static float uni(stream_t *s) { return (float)((stream_word(s) >> 8) | 1u) * 0x1p-24f; }

/* words [0, n) of the stream of one history, for the RNG parity test */
void ref_stream_words(uint64_t seed, uint64_t history, uint32_t sim, uint32_t *out, int n) {
  stream_t s;
  stream_open(&s, seed, history, sim);
  for (int i = 0; i < n; ++i) out[i] = stream_word(&s);
}

/* ------------------------------------------------------------------ math */
static uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

/* log(x), x > 0 normal: x = 2^e m with m in [0.5, 1) (frexp), i = the top six mantissa
 * bits. The table entry (rcp_i, lg_i) -- rcp_i the float nearest 1/c_i, c_i the middle of
 * entry i's interval [1/2 + i/128, 1/2 + (i+1)/128) (1 for the last one, next to 1), lg_i
 * the float nearest -log(rcp_i); literals from opengate_amd/csrc/gen_log_table.py -- gives
 * log x = e ln2 + lg_i + log1p(r) with r = m rcp_i - 1 and
 * log1p(r) = r (1 - r/2 + r^2/3 - r^3/4). Accurate to 2e-7 relative on (0, 1]. */
static const float LOG_TAB[64][2] = {
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

float ref_log(float x) {
  const uint32_t i = (fbits(x) >> 17) & 63u;
  int e;
  const float m = frexpf(x, &e);
  const float ef = (float)e;
  const float *t = LOG_TAB[i];
  const float r = fmaf(m, t[0], -1.0f);
  float p = fmaf(r, -0.25f, 0x1.555556p-2f);
  p = fmaf(r, p, -0.5f);
  p = fmaf(r, p, 1.0f);
  const float l1 = fmaf(r, p, t[1]);
  return fmaf(ef, 0x1.62ep-1f, fmaf(ef, 0x1.0bfbe8p-15f, l1));
}

void ref_sincos(float x, float *sn, float *cs) {
  float k = rintf(x * 0x1.45f306p-1f);
  int q = (int)k;
  float r = fmaf(-k, 0x1.921fp+0f, x);
  r = fmaf(-k, 0x1.6a88p-17f, r);
  r = fmaf(-k, 0x1.0b4612p-34f, r);
  float r2 = r * r;
  float t = fmaf(r2, 0x1.71de3ap-19f, -0x1.a01a02p-13f);
  t = fmaf(r2, t, 0x1.111112p-7f);
  t = fmaf(r2, t, -0x1.555556p-3f);
  float sp = fmaf(r * r2, t, r);
  float c = fmaf(r2, -0x1.27e4fcp-22f, 0x1.a01a02p-16f);
  c = fmaf(r2, c, -0x1.6c16c2p-10f);
  c = fmaf(r2, c, 0x1.555556p-5f);
  float cp = fmaf(r2 * r2, c, fmaf(-0.5f, r2, 1.0f));
  switch (q & 3) {
    case 0: *sn = sp; *cs = cp; break;
    case 1: *sn = cp; *cs = -sp; break;
    case 2: *sn = -sp; *cs = -cp; break;
    default: *sn = -cp; *cs = sp; break;
  }
}

static void unit(float v[3]) {
  float n2 = fmaf(v[2], v[2], fmaf(v[1], v[1], v[0] * v[0]));
  float inv = 1.0f / sqrtf(n2);
  v[0] = v[0] * inv;
  v[1] = v[1] * inv;
  v[2] = v[2] * inv;
}

#define CAP 1000
#define MEC2 0x1.05a1a8p-1f
#define KPM 0x1.fac5e8p+8f

/* (cos phi, sin phi) for phi = 2 pi u: the kernel's azimuth (gh_device.h) */
static void ref_azimuth(stream_t *s, float *cp, float *sp) {
  ref_sincos(6.28318530717958647692f * uni(s), sp, cp);
}

static void scatter_dir(float d[3], float c, float cp, float sp) {
  float st = sqrtf(fmaxf(fmaf(-c, c, 1.0f), 0.0f));
  float n[3];
  if (fabsf(d[2]) < 0.99999f) {
    float tmp = sqrtf(fmaf(-d[2], d[2], 1.0f));
    float f = st / tmp;
    n[0] = fmaf(c, d[0], f * fmaf(d[0] * d[2], cp, -(d[1] * sp)));
    n[1] = fmaf(c, d[1], f * fmaf(d[1] * d[2], cp, d[0] * sp));
    n[2] = fmaf(c, d[2], -((st * cp) * tmp));
  } else {
    n[0] = st * cp;
    n[1] = st * sp;
    n[2] = (d[2] > 0.0f) ? c : -c;
  }
  unit(n);
  d[0] = n[0]; d[1] = n[1]; d[2] = n[2];
}

/* last i in [0, n-2] with a[i] <= x */
static int floor_index(const float *a, int n, float x) {
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

/* first k with cdf[k] > u */
static int first_above(const float *cdf, int n, float u) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (cdf[mid] > u) hi = mid; else lo = mid + 1;
  }
  return lo;
}

/* value of a table on the q^2 grid at x >= 0 (node 0 at 0, node j >= 1 at the float
 * with bits (920 + j - 1) << 20): interval from the bits of x, position from the low
 * 20 mantissa bits, clamped at the top node */
static float qgrid(const float *ys, int n, float x) {
  int j;
  float f;
  if (x < 0x1p-12f) {
    j = 0;
    f = x * 0x1p12f;
  } else {
    uint32_t b;
    memcpy(&b, &x, 4);
    j = (int)(b >> 20) - 920 + 1;
    f = (float)(b & 0xFFFFFu) * 0x1p-20f;
  }
  if (j > n - 2) {
    j = n - 2;
    f = 1.0f;
  }
  return fmaf(f, ys[j + 1] - ys[j], ys[j]);
}

/* ------------------------------------------------------------------ physics view */
typedef struct {
  const gh_physics_desc *d;
  int nb1;
  /* per material: 1/density (dose), the factor of its table row (density with shared
   * compositions, 1 with per-material tables) and that row (composition or material) */
  float *inv_rho, *rho;
  int *row;
} phys_t;

static void phys_init(phys_t *P, const gh_physics_desc *d) {
  const int M = d->n_materials;
  P->d = d;
  P->nb1 = d->n_buckets + 1;
  P->inv_rho = (float *)malloc((size_t)M * sizeof(float));
  P->rho = (float *)malloc((size_t)M * sizeof(float));
  P->row = (int *)malloc((size_t)M * sizeof(int));
  for (int i = 0; i < M; ++i) {
    P->inv_rho[i] = d->density[i] > 0.0f ? 1.0f / d->density[i] : 0.0f;
    P->rho[i] = d->n_compositions > 0 ? d->density[i] : 1.0f;
    P->row[i] = d->n_compositions > 0 ? d->composition[i] : i;
  }
}

static void phys_free(phys_t *P) {
  free(P->inv_rho);
  free(P->rho);
  free(P->row);
}

typedef struct {
  int bucket;
  float frac, maj, inv_maj;
} eidx_t;

static void energy_index(const phys_t *P, float E, eidx_t *x) {
  uint32_t b = fbits(E);
  x->bucket = (int)(b >> 17) - 7488;
  x->frac = (float)(b & 0x1FFFFu) * 0x1p-17f;
  const float *m = P->d->mu_majorant;
  x->maj = fmaf(x->frac, m[x->bucket + 1] - m[x->bucket], m[x->bucket]);
  x->inv_maj = 1.0f / x->maj;
}

static float node_interp(const float *row, const eidx_t *x) {
  return fmaf(x->frac, row[x->bucket + 1] - row[x->bucket], row[x->bucket]);
}

/* value of material m at the energy of x from table `tab`: interpolation between the
 * material's node values, each the FP32 product of its factor and its row's node (factor 1
 * with one row per material: the row itself) */
static float material_node_interp(const phys_t *P, const float *tab, int m, const eidx_t *x) {
  const float *row = tab + (size_t)P->row[m] * P->nb1;
  const float v0 = P->rho[m] * row[x->bucket], v1 = P->rho[m] * row[x->bucket + 1];
  return fmaf(x->frac, v1 - v0, v0);
}

/* attenuation coefficient of material m at the energy of x */
static float material_mu(const phys_t *P, int m, const eidx_t *x) {
  return material_node_interp(P, P->d->mu_total, m, x);
}

/* Kahn's method for Klein-Nishina, then acceptance with S(q)/Z */
static void compton_sample(stream_t *s, const phys_t *P, int m, float E, float *q_out, float *c_out,
                           uint64_t *rej) {
  const gh_physics_desc *d = P->d;
  float alpha = E / MEC2;
  float a2 = alpha + alpha;
  float onep = 1.0f + a2;
  float thr = onep / (9.0f + a2);
  float kk = E * KPM;
  float two_k2 = (2.0f * kk) * kk;
  const float *srow = d->incoh_s + (size_t)m * d->n_x_nodes;
  float q = 1.0f, c = 1.0f;
  int it;
  for (it = 0; it < CAP; ++it) {
    float r1 = uni(s), r2 = uni(s), r3 = uni(s);
    int ok;
    if (r1 <= thr) {
      q = fmaf(a2, r2, 1.0f);
      float iq = 1.0f / q;
      ok = r3 <= 4.0f * (iq - iq * iq);
    } else {
      q = onep / fmaf(a2, r2, 1.0f);
      float cc = 1.0f - (q - 1.0f) / alpha;
      ok = r3 <= 0.5f * fmaf(cc, cc, 1.0f / q);
    }
    if (!ok) { (*rej)++; continue; }
    c = 1.0f - (q - 1.0f) / alpha;
    float x = two_k2 * (1.0f - c);
    float sv = qgrid(srow, d->n_x_nodes, x);
    if (uni(s) <= sv) break;
    (*rej)++;
  }
  if (it == CAP) { q = 1.0f; c = 1.0f; }
  *q_out = q;
  *c_out = fmaxf(c, -1.0f);
}

static float rayleigh_sample(stream_t *s, const phys_t *P, int m, float E) {
  const gh_physics_desc *d = P->d;
  float kk = E * KPM;
  float two_k2 = (2.0f * kk) * kk;
  float xmax = two_k2 + two_k2;
  const float *G = d->ray_cdf + (size_t)m * d->n_x_nodes;
  const float *X = d->ray_x;
  float gmax = qgrid(G, d->n_x_nodes, xmax);
  float c = 1.0f;
  for (int it = 0; it < CAP; ++it) {
    float t = uni(s) * gmax;
    int j = floor_index(G, d->n_x_nodes, t);
    float dg = G[j + 1] - G[j];
    float x = (dg > 0.0f) ? fmaf((t - G[j]) / dg, X[j + 1] - X[j], X[j]) : X[j];
    c = fmaxf(1.0f - x / two_k2, -1.0f);
    if (uni(s) <= 0.5f * fmaf(c, c, 1.0f)) break;
  }
  return c;
}

/* ------------------------------------------------------------------ tally */
typedef struct {
  double *edep, *edep2, *dose, *dose2, *counts;
  double *exact_edep2;   /* exact per-history squared sums, may be NULL */
  /* per-history scratch for the exact squared tally */
  int *hv;
  double *hs;
  int hn, hcap;
} tally_t;

typedef struct {
  int v[4];
  float s[4], dsum[4];
  uint32_t c[4];
  int n;
} slots_t;

typedef struct {
  uint64_t c[12];
  double e[4];
} stats_t;

enum { K_HIST = 0, K_STEPS, K_DEP, K_COMPT, K_RAYL, K_PHOTO, K_ESC, K_WORLD, K_CUT, K_EVICT, K_REC, K_REJ };
enum { EN_DEP = 0, EN_OUT, EN_ESC, EN_WORLD };

static void write_slot(tally_t *t, int v, float s, float ds, uint32_t c, stats_t *st) {
  t->edep[v] += (double)s;
  if (t->edep2) t->edep2[v] += (double)s * (double)s;
  if (t->counts) t->counts[v] += (double)c;
  if (t->dose) t->dose[v] += (double)ds;
  if (t->dose2) t->dose2[v] += (double)ds * (double)ds;
  st->c[K_DEP]++;
  st->e[EN_DEP] += (double)s;
}

static void exact_add(tally_t *t, int v, float x) {
  if (!t->exact_edep2) return;
  for (int i = 0; i < t->hn; ++i)
    if (t->hv[i] == v) { t->hs[i] += (double)x; return; }
  if (t->hn == t->hcap) {
    t->hcap = t->hcap ? 2 * t->hcap : 16;
    t->hv = (int *)realloc(t->hv, t->hcap * sizeof(int));
    t->hs = (double *)realloc(t->hs, t->hcap * sizeof(double));
  }
  t->hv[t->hn] = v;
  t->hs[t->hn] = (double)x;
  t->hn++;
}

static void history_end(tally_t *t, slots_t *sl, stats_t *st) {
  for (int k = 0; k < sl->n; ++k) write_slot(t, sl->v[k], sl->s[k], sl->dsum[k], sl->c[k], st);
  sl->n = 0;
  if (t->exact_edep2) {
    for (int i = 0; i < t->hn; ++i) t->exact_edep2[t->hv[i]] += t->hs[i] * t->hs[i];
    t->hn = 0;
  }
}

static void score(tally_t *t, slots_t *sl, int v, float x, float dx, stats_t *st) {
  if (v < 0) { st->e[EN_OUT] += (double)x; return; }
  exact_add(t, v, x);
  for (int k = 0; k < sl->n; ++k) {
    if (sl->v[k] == v) {
      sl->s[k] = sl->s[k] + x;
      sl->dsum[k] = sl->dsum[k] + dx;
      sl->c[k] += 1u;
      return;
    }
  }
  if (sl->n == 4) {
    write_slot(t, sl->v[0], sl->s[0], sl->dsum[0], sl->c[0], st);
    for (int k = 0; k < 3; ++k) {
      sl->v[k] = sl->v[k + 1]; sl->s[k] = sl->s[k + 1]; sl->dsum[k] = sl->dsum[k + 1]; sl->c[k] = sl->c[k + 1];
    }
    sl->n = 3;
    st->c[K_EVICT]++;
  }
  sl->v[sl->n] = v; sl->s[sl->n] = x; sl->dsum[sl->n] = dx; sl->c[sl->n] = 1u;
  sl->n++;
}

/* ------------------------------------------------------------------ geometry */
typedef struct {
  const gh_phantom_desc *ph;
  const gh_source_desc *src;
  const gh_tally_desc *tal;
  float inv_s[3], L[3], tinv[3], d_ct, d_phi;
  float s2_min, s2_d;   /* cosine law: sin^2 of theta_min, sin^2 theta_max - sin^2 theta_min */
  float axis[3];        /* surface sphere: unit axis of the tangent frame */
  int aligned;
} sim_t;

/* per-source constants of the direction samplers, computed in FP64 and rounded once */
static void sim_source_init(sim_t *S, const gh_source_desc *src) {
  S->d_ct = src->cos_theta_min - src->cos_theta_max;
  S->d_phi = src->phi_max - src->phi_min;
  const double c0 = src->cos_theta_min, c1 = src->cos_theta_max;
  const double s0 = 1.0 - c0 * c0, s1 = 1.0 - c1 * c1;
  S->s2_min = (float)s0;
  S->s2_d = (float)(s1 - s0);
  const double a0 = src->surface_axis[0], a1 = src->surface_axis[1], a2 = src->surface_axis[2];
  const double an = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
  S->axis[0] = an > 0.0 ? (float)(a0 / an) : 0.0f;
  S->axis[1] = an > 0.0 ? (float)(a1 / an) : 0.0f;
  S->axis[2] = an > 0.0 ? (float)(a2 / an) : 1.0f;
}

static int label_at(const sim_t *S, int v) {
  return S->ph->label_bytes == 2 ? (int)((const uint16_t *)S->ph->labels)[v] : (int)((const uint8_t *)S->ph->labels)[v];
}

static int voxel_of(const sim_t *S, const float p[3]) {
  int ix = (int)floorf(p[0] * S->inv_s[0]);
  int iy = (int)floorf(p[1] * S->inv_s[1]);
  int iz = (int)floorf(p[2] * S->inv_s[2]);
  if ((unsigned)ix >= (unsigned)S->ph->nx || (unsigned)iy >= (unsigned)S->ph->ny || (unsigned)iz >= (unsigned)S->ph->nz)
    return -1;
  return ix + S->ph->nx * (iy + S->ph->ny * iz);
}

/* floor((p - c) / s) for the real values of FP32 p, c, s: the ITK index rule
 * RoundHalfIntegerUp((p - o) / s) with the origin o = c + s/2 at the centre of voxel 0
 * (GateDoseActor.cpp:144-178, TransformPhysicalPointToIndex). Estimate from the rounded
 * quotient, corrected by the sign of the exact residual p - c - k s (TwoSum of p - c,
 * then an fmaf that is exact where the residual is small). */
int ref_cell_index(float p, float c, float s, float inv) {
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

/* phantom voxel of a birth point by the exact rule (the Woodcock steps use voxel_of) */
static int birth_voxel(const sim_t *S, const float p[3]) {
  int ix = ref_cell_index(p[0], 0.0f, S->ph->spacing[0], S->inv_s[0]);
  int iy = ref_cell_index(p[1], 0.0f, S->ph->spacing[1], S->inv_s[1]);
  int iz = ref_cell_index(p[2], 0.0f, S->ph->spacing[2], S->inv_s[2]);
  if ((unsigned)ix >= (unsigned)S->ph->nx || (unsigned)iy >= (unsigned)S->ph->ny || (unsigned)iz >= (unsigned)S->ph->nz)
    return -1;
  return ix + S->ph->nx * (iy + S->ph->ny * iz);
}

/* tally voxel of a deposit at p, -1 outside; an aligned tally is the phantom grid */
static int tally_index(const sim_t *S, const float p[3]) {
  const gh_tally_desc *t = S->tal;
  int idx[3];
  for (int i = 0; i < 3; ++i) {
    const float c = S->aligned ? 0.0f : t->corner[i];
    const float sp = S->aligned ? S->ph->spacing[i] : t->spacing[i];
    idx[i] = ref_cell_index(p[i], c, sp, S->tinv[i]);
  }
  if ((unsigned)idx[0] >= (unsigned)t->nx || (unsigned)idx[1] >= (unsigned)t->ny || (unsigned)idx[2] >= (unsigned)t->nz)
    return -1;
  return idx[0] + t->nx * (idx[1] + t->ny * idx[2]);
}

static void box_hit(const sim_t *S, const float p[3], const float d[3], float *tnear, float *tfar, int *face) {
  *tnear = -INFINITY;
  *tfar = INFINITY;
  *face = -1;
  for (int i = 0; i < 3; ++i) {
    if (d[i] != 0.0f) {
      float inv = 1.0f / d[i];
      float t0 = (-p[i]) * inv;
      float t1 = (S->L[i] - p[i]) * inv;
      float lo = fminf(t0, t1), hi = fmaxf(t0, t1);
      if (lo > *tnear) *tnear = lo;
      if (hi < *tfar) { *tfar = hi; *face = 2 * i + (d[i] > 0.0f ? 1 : 0); }
    } else if (p[i] < 0.0f || p[i] > S->L[i]) {
      *tnear = INFINITY;
    }
  }
}

static float world_distance(const sim_t *S, const float p[3], const float d[3]) {
  float t = INFINITY;
  for (int i = 0; i < 3; ++i) {
    float inv = 1.0f / d[i];
    if (d[i] > 0.0f) t = fminf(t, (S->ph->world_hi[i] - p[i]) * inv);
    else if (d[i] < 0.0f) t = fminf(t, (S->ph->world_lo[i] - p[i]) * inv);
  }
  return fmaxf(t, 0.0f);
}

/* ------------------------------------------------------------------ source */
typedef struct {
  float p[3], d[3], E, w;
} photon_t;

/* d = d0 x' + d1 y' + d2 n in the tangent frame of a surface point with outward unit
   normal n (G4SPSPosDistribution::GeneratePointsOnSurface, sphere): x' = unit(A x n),
   y' = unit(x' x n); a zero vector stays zero */
static void unit_or_zero(float v[3]) {
  const float n2 = fmaf(v[2], v[2], fmaf(v[1], v[1], v[0] * v[0]));
  if (n2 > 0.0f) {
    const float inv = 1.0f / sqrtf(n2);
    v[0] = v[0] * inv; v[1] = v[1] * inv; v[2] = v[2] * inv;
  }
}

static void tangent_frame(const float A[3], const float n[3], float d[3]) {
  float xp[3] = {fmaf(A[1], n[2], -(A[2] * n[1])), fmaf(A[2], n[0], -(A[0] * n[2])), fmaf(A[0], n[1], -(A[1] * n[0]))};
  unit_or_zero(xp);
  float yp[3] = {fmaf(xp[1], n[2], -(xp[2] * n[1])), fmaf(xp[2], n[0], -(xp[0] * n[2])),
                 fmaf(xp[0], n[1], -(xp[1] * n[0]))};
  unit_or_zero(yp);
  const float d0 = d[0], d1 = d[1], d2 = d[2];
  for (int i = 0; i < 3; ++i) d[i] = fmaf(d0, xp[i], fmaf(d1, yp[i], d2 * n[i]));
}

/* two independent N(0,1) values by Box-Muller: radius sqrt(-2 log u), angle 2 pi u' */
static void normal_pair(stream_t *rs, float *n1, float *n2) {
  const float r = sqrtf(-2.0f * ref_log(uni(rs)));
  float cp, sp;
  ref_azimuth(rs, &cp, &sp);
  *n1 = r * cp;
  *n2 = r * sp;
}

static void emit(const sim_t *S, stream_t *rs, uint64_t h, photon_t *ph) {
  const gh_source_desc *src = S->src;
  float lp[3] = {0.0f, 0.0f, 0.0f}, ld[3] = {0.0f, 0.0f, 1.0f};
  float E = src->energy_mono, w = src->weight;
  float nrm[3] = {0.0f, 0.0f, 1.0f}; /* surface sphere: outward unit normal at the point */
  if (src->pos_type == GH_POS_PHSP) {
    int64_t row = (src->phsp_start + (int64_t)h) % src->phsp_n;
    for (int k = 0; k < 3; ++k) {
      lp[k] = src->phsp[k][row];
      ld[k] = src->phsp[3 + k][row];
    }
    E = src->phsp[6][row];
    w = src->phsp[7][row];
  } else {
    if (src->pos_type == GH_POS_BOX) {
      for (int k = 0; k < 3; ++k) lp[k] = (uni(rs) - 0.5f) * src->pos_param[k];
    } else if (src->pos_type == GH_POS_SPHERE || src->pos_type == GH_POS_ELLIPSOID) {
      float a, b, c, r2;
      int tries = 0;
      do {
        a = fmaf(2.0f, uni(rs), -1.0f);
        b = fmaf(2.0f, uni(rs), -1.0f);
        c = fmaf(2.0f, uni(rs), -1.0f);
        r2 = fmaf(c, c, fmaf(b, b, a * a));
      } while (r2 > 1.0f && ++tries < CAP);
      int ell = src->pos_type == GH_POS_ELLIPSOID;
      lp[0] = src->pos_param[0] * a;
      lp[1] = src->pos_param[ell ? 1 : 0] * b;
      lp[2] = src->pos_param[ell ? 2 : 0] * c;
    } else if (src->pos_type == GH_POS_VOXELS) {
      /* Walker alias table (GateSPSVoxelsPosDistribution::VGenerateOne samples the same
         distribution with three nested CDF searches) */
      const int64_t i = (int64_t)(((uint64_t)stream_word(rs) * (uint64_t)src->vox_n) >> 32);
      const int64_t j = (uni(rs) < src->vox_prob[i]) ? i : (int64_t)src->vox_alias[i];
      const int v = src->vox_index[j];
      const int nx = src->vox_size[0], ny = src->vox_size[1];
      const int ix = v % nx, iyz = v / nx;
      const int iy = iyz % ny, iz = iyz / ny;
      lp[0] = ((float)ix + uni(rs)) * src->vox_spacing[0];
      lp[1] = ((float)iy + uni(rs)) * src->vox_spacing[1];
      lp[2] = ((float)iz + uni(rs)) * src->vox_spacing[2];
    } else if (src->pos_type == GH_POS_DISC || src->pos_type == GH_POS_CYLINDER) {
      float a, b, r2;
      int tries = 0;
      do {
        a = fmaf(2.0f, uni(rs), -1.0f);
        b = fmaf(2.0f, uni(rs), -1.0f);
        r2 = fmaf(b, b, a * a);
      } while (r2 > 1.0f && ++tries < CAP);
      lp[0] = src->pos_param[0] * a;
      lp[1] = src->pos_param[0] * b;
      if (src->pos_type == GH_POS_CYLINDER) lp[2] = fmaf(2.0f, uni(rs), -1.0f) * src->pos_param[1];
      if (src->pos_type == GH_POS_DISC && (src->pos_sigma[0] > 0.0f || src->pos_sigma[1] > 0.0f)) {
        /* G4SPSPosDistribution "Beam" circle: the uniform point plus a Gaussian spread */
        float n1, n2;
        normal_pair(rs, &n1, &n2);
        lp[0] = fmaf(src->pos_sigma[0], n1, lp[0]);
        lp[1] = fmaf(src->pos_sigma[1], n2, lp[1]);
      }
    } else if (src->pos_type == GH_POS_SURFACE_SPHERE) {
      /* G4SPSPosDistribution "Surface" sphere: cos(theta) uniform in [-1, 1], phi uniform */
      const float ct = fmaf(2.0f, uni(rs), -1.0f);
      const float st = sqrtf(fmaxf(fmaf(-ct, ct, 1.0f), 0.0f));
      float cp, sp;
      ref_azimuth(rs, &cp, &sp);
      nrm[0] = st * cp;
      nrm[1] = st * sp;
      nrm[2] = ct;
      lp[0] = src->pos_param[0] * nrm[0];
      lp[1] = src->pos_param[0] * nrm[1];
      lp[2] = src->pos_param[0] * nrm[2];
    }
    if (src->dir_type == GH_DIR_ISO || src->dir_type == GH_DIR_COS) {
      /* G4SPSAngDistribution::GenerateIsotropicFlux: cos t uniform in [cos t_max, cos t_min];
         GenerateCosineLawFlux: sin^2 t uniform in [sin^2 t_min, sin^2 t_max] */
      float ct, st;
      if (src->dir_type == GH_DIR_ISO) {
        ct = fmaf(-uni(rs), S->d_ct, src->cos_theta_min);
        st = sqrtf(fmaxf(fmaf(-ct, ct, 1.0f), 0.0f));
      } else {
        const float s2 = fmaf(uni(rs), S->s2_d, S->s2_min);
        st = sqrtf(s2);
        ct = sqrtf(fmaxf(1.0f - s2, 0.0f));
      }
      float cp, sp;
      if (src->phi_full) {
        ref_azimuth(rs, &cp, &sp);
      } else {
        float phi = fmaf(uni(rs), S->d_phi, src->phi_min);
        ref_sincos(phi, &sp, &cp);
      }
      ld[0] = -(st * cp);
      ld[1] = -(st * sp);
      ld[2] = -ct;
      if (src->pos_type == GH_POS_SURFACE_SPHERE) tangent_frame(S->axis, nrm, ld);
    } else if (src->dir_type == GH_DIR_BEAM2D) {
      /* G4SPSAngDistribution::GenerateBeamFlux ("beam2d"): angles (a, b) Gaussian, polar
         angle |(a, b)|, azimuth the angle of (a, b) (0 when both are 0) */
      float n1, n2;
      normal_pair(rs, &n1, &n2);
      const float a = src->dir_sigma[0] * n1, b = src->dir_sigma[1] * n2;
      const float t = sqrtf(fmaf(b, b, a * a));
      float cp = 1.0f, sp = 0.0f;
      if (t > 0.0f) {
        cp = a / t;
        sp = b / t;
      }
      float sn, cs;
      ref_sincos(t, &sn, &cs);
      ld[0] = -(sn * cp);
      ld[1] = -(sn * sp);
      ld[2] = -cs;
    } else if (src->dir_type == GH_DIR_HISTOGRAM) {
      /* G4SPSAngDistribution::GenerateUserDefFlux: theta, then phi, each from its table */
      int k = first_above(src->theta_cdf, src->n_theta, uni(rs));
      const float t = fmaf(uni(rs), src->theta_width[k], src->theta_value[k]);
      k = first_above(src->phi_cdf, src->n_phi, uni(rs));
      const float p = fmaf(uni(rs), src->phi_width[k], src->phi_value[k]);
      float st, ct, sp, cp;
      ref_sincos(t, &st, &ct);
      ref_sincos(p, &sp, &cp);
      ld[0] = -(st * cp);
      ld[1] = -(st * sp);
      ld[2] = -ct;
    } else if (src->dir_type == GH_DIR_MOMENTUM) {
      ld[0] = src->momentum[0]; ld[1] = src->momentum[1]; ld[2] = src->momentum[2];
    }
    if (src->ene_type == GH_ENE_LINES) {
      E = src->energy_value[first_above(src->energy_cdf, src->n_energy, uni(rs))];
    } else if (src->ene_type == GH_ENE_HISTOGRAM) {
      int k = first_above(src->energy_cdf, src->n_energy, uni(rs));
      E = fmaf(uni(rs), src->energy_width[k], src->energy_value[k]);
    } else if (src->ene_type == GH_ENE_RANGE) {
      E = fmaf(uni(rs), src->energy_max - src->energy_mono, src->energy_mono);
    } else if (src->ene_type == GH_ENE_GAUSS) {
      /* G4SPSEneDistribution "Gauss": mean + sigma N(0,1), negative energies 0; the tail is
         cut at mean + 8 sigma */
      float n1, n2;
      normal_pair(rs, &n1, &n2);
      const float top = src->energy_mono + 8.0f * src->energy_sigma;
      E = fminf(fmaxf(fmaf(src->energy_sigma, n1, src->energy_mono), 0.0f), top);
    }
    if (src->weight_sigma >= 0.0f) {
      /* Gaussian history weight (GateGenericSource.cpp:247-259) */
      float n1, n2;
      normal_pair(rs, &n1, &n2);
      w = fmaf(src->weight_sigma, n1, src->weight);
    }
  }
  const float *M = src->pos_mat;
  for (int i = 0; i < 3; ++i) ph->p[i] = fmaf(M[3 * i], lp[0], fmaf(M[3 * i + 1], lp[1], fmaf(M[3 * i + 2], lp[2], M[9 + i])));
  if (src->pos_type == GH_POS_PHSP && src->phsp_iso) {
    float ct = fmaf(2.0f, uni(rs), -1.0f);
    float st = sqrtf(fmaxf(fmaf(-ct, ct, 1.0f), 0.0f));
    float cp, sp;
    ref_azimuth(rs, &cp, &sp);
    ph->d[0] = st * cp; ph->d[1] = st * sp; ph->d[2] = ct;
  } else if (src->dir_type == GH_DIR_FOCUSED && src->pos_type != GH_POS_PHSP) {
    for (int i = 0; i < 3; ++i) ph->d[i] = src->focus[i] - ph->p[i];
    unit(ph->d);
  } else {
    const float *D = src->dir_mat;
    for (int i = 0; i < 3; ++i) ph->d[i] = fmaf(D[3 * i], ld[0], fmaf(D[3 * i + 1], ld[1], D[3 * i + 2] * ld[2]));
    unit(ph->d);
  }
  ph->E = E;
  ph->w = w;
}

/* ------------------------------------------------------------------ one history */
typedef struct {
  float *rec[8];
  uint64_t rec_cap, rec_n;
  float *event_edep;
} outputs_t;

static void run_history(const sim_t *S, const phys_t *P, uint64_t seed, uint64_t first, uint32_t sim_id, uint64_t h,
                        tally_t *t, stats_t *st, outputs_t *o) {
  const gh_physics_desc *d = P->d;
  stream_t rs;
  stream_open(&rs, seed, first + h, sim_id);
  photon_t ph;
  emit(S, &rs, h, &ph);
  st->c[K_HIST]++;
  slots_t sl;
  sl.n = 0;
  t->hn = 0;
  float ev = 0.0f;
  const int det = S->tal->detector_label_min;
  /* TLE: Woodcock collision estimator of kerma (see the kernel's TLE variant) */
  const int tle = (S->tal->flags & GH_TALLY_TLE) != 0;
  const float tle_emin = S->tal->tle_energy_min, tle_emax = S->tal->tle_energy_max;
  const int bvox = birth_voxel(S, ph.p);
  int in_box = bvox >= 0;
  /* After the birth (source sampling, and for a photon born outside the phantom its
   * first world step) the event stream continues at the next unused Philox block: the
   * words left in the current block are dropped (the GPU kernel hands a born photon to a
   * lane with only the block counter). */
  int birth = !in_box;
  if (ph.E < d->e_cut || (tle && ph.E <= tle_emin)) {
    if (in_box) {
      int lab = label_at(S, bvox);
      float x = ph.E * ph.w;
      score(t, &sl, tally_index(S, ph.p), x, x * P->inv_rho[lab], st);
      if (det >= 0 && lab >= det) ev += x;
      history_end(t, &sl, st);
    } else {
      st->e[EN_WORLD] += (double)(ph.E * ph.w);
    }
    st->c[K_CUT]++;
    if (o->event_edep) o->event_edep[h] = ev;
    return;
  }
  if (in_box) rs.pos = 4;
  eidx_t ex;
  energy_index(P, ph.E, &ex);
  const int wm = S->ph->world_material;
  /* Woodcock steps draw from a separate "step stream" of the history: one Philox
   * block (counter word 3 = 0x80000000 | step block) serves three tentative steps
   * through 21-bit fields (step_fields); what is left of a block when a step collides
   * or leaves the box is unused.
   * Everything else (source, interactions, world) uses the event stream rs. */
  const uint32_t hid = (uint32_t)(first + h), hid_hi = (uint32_t)((first + h) >> 32);
  uint32_t step_block = 0;
  for (;;) {
    if (in_box) {
      uint32_t w[4];
      const uint32_t sctr[4] = {hid, hid_hi, sim_id, 0x80000000u | step_block};
      ref_philox4x32_10(sctr, rs.key, w);
      step_block++;
      int outcome = 0; /* 0: both steps virtual, 1: real collision, 2: leaves the box */
      int vox = -1, lab = 0;
      for (int k = 0; k < 3; ++k) {
        uint32_t bd, ba;
        step_fields(w, k, &bd, &ba);
        float step = -ref_log((float)((bd << 1) | 1u) * 0x1p-22f) * ex.inv_maj;
        float q[3];
        for (int i = 0; i < 3; ++i) q[i] = fmaf(step, ph.d[i], ph.p[i]);
        vox = voxel_of(S, q);
        if (vox < 0) { outcome = 2; break; }
        ph.p[0] = q[0]; ph.p[1] = q[1]; ph.p[2] = q[2];
        st->c[K_STEPS]++;
        lab = label_at(S, vox);
        float mu = material_mu(P, lab, &ex);
        if (tle && ph.E <= tle_emax) {
          const float x = (ph.E * ph.w) * (material_node_interp(P, d->mu_en, lab, &ex) * ex.inv_maj);
          score(t, &sl, tally_index(S, ph.p), x, x * P->inv_rho[lab], st);
        }
        if ((float)((ba << 1) | 1u) * 0x1p-22f * ex.maj < mu) { outcome = 1; break; }
      }
      if (outcome == 0) continue;
      if (outcome == 1) {
        float r = uni(&rs);
        float pc = node_interp(d->p_compton + (size_t)P->row[lab] * P->nb1, &ex);
        float pr = node_interp(d->p_rayleigh + (size_t)P->row[lab] * P->nb1, &ex);
        int tv = tally_index(S, ph.p);
        float irho = P->inv_rho[lab];
        int is_det = det >= 0 && lab >= det;
        const int tle_on = tle && ph.E <= tle_emax;
        if (r < pc) {
          st->c[K_COMPT]++;
          float qq, c;
          compton_sample(&rs, P, P->row[lab], ph.E, &qq, &c, &st->c[K_REJ]);
          float e1 = ph.E / qq;
          float x = (ph.E - e1) * ph.w;
          if (!tle_on) {
            score(t, &sl, tv, x, x * irho, st);
            if (is_det) ev += x;
          }
          float cp, sp;
          ref_azimuth(&rs, &cp, &sp);
          scatter_dir(ph.d, c, cp, sp);
          ph.E = e1;
          if (ph.E < d->e_cut || (tle && ph.E <= tle_emin)) {
            float x2 = ph.E * ph.w;
            score(t, &sl, tv, x2, x2 * irho, st);
            if (is_det) ev += x2;
            st->c[K_CUT]++;
            break;
          }
          energy_index(P, ph.E, &ex);
        } else if (r < pr) {
          st->c[K_RAYL]++;
          float c = rayleigh_sample(&rs, P, P->row[lab], ph.E);
          float cp, sp;
          ref_azimuth(&rs, &cp, &sp);
          scatter_dir(ph.d, c, cp, sp);
        } else {
          st->c[K_PHOTO]++;
          float x = ph.E * ph.w;
          if (!tle_on) {
            score(t, &sl, tv, x, x * irho, st);
            if (is_det) ev += x;
          }
          break;
        }
      } else {
        float tn, tf;
        int face;
        box_hit(S, ph.p, ph.d, &tn, &tf, &face);
        float tt = fmaxf(tf, 0.0f);
        float e[3];
        for (int i = 0; i < 3; ++i) e[i] = fmaf(tt, ph.d[i], ph.p[i]);
        if (S->tal->record_face >= 0 && (face == S->tal->record_face || S->tal->record_face == GH_FACE_ALL)) {
          if (o->rec[0] && o->rec_n < o->rec_cap) {
            uint64_t k = o->rec_n;
            for (int i = 0; i < 3; ++i) { o->rec[i][k] = e[i]; o->rec[3 + i][k] = ph.d[i]; }
            o->rec[6][k] = ph.E;
            o->rec[7][k] = ph.w;
          }
          o->rec_n++;
          st->c[K_REC]++;
          st->e[EN_ESC] += (double)(ph.E * ph.w);
          break;
        }
        if (wm < 0) {
          st->c[K_ESC]++;
          st->e[EN_ESC] += (double)(ph.E * ph.w);
          break;
        }
        ph.p[0] = e[0]; ph.p[1] = e[1]; ph.p[2] = e[2];
        in_box = 0;
      }
    } else {
      float tn, tf;
      int face;
      box_hit(S, ph.p, ph.d, &tn, &tf, &face);
      int hit = (tn < tf) && (tf > 1e-3f);
      float tin = fmaxf(tn, 0.0f);
      if (wm < 0) {
        if (!hit) {
          st->c[K_ESC]++;
          st->e[EN_ESC] += (double)(ph.E * ph.w);
          break;
        }
        for (int i = 0; i < 3; ++i) ph.p[i] = fmaf(tin, ph.d[i], ph.p[i]);
        in_box = 1;
        if (birth) { rs.pos = 4; birth = 0; }
        continue;
      }
      float muw = material_mu(P, wm, &ex);
      float tw = world_distance(S, ph.p, ph.d);
      float dist = (muw > 0.0f) ? (-ref_log(uni(&rs)) / muw) : INFINITY;
      if (birth) { rs.pos = 4; birth = 0; }
      if (hit && dist >= tin && tin < tw) {
        for (int i = 0; i < 3; ++i) ph.p[i] = fmaf(tin, ph.d[i], ph.p[i]);
        in_box = 1;
        continue;
      }
      if (dist >= tw) {
        st->c[K_ESC]++;
        st->e[EN_ESC] += (double)(ph.E * ph.w);
        break;
      }
      for (int i = 0; i < 3; ++i) ph.p[i] = fmaf(dist, ph.d[i], ph.p[i]);
      st->c[K_WORLD]++;
      float r = uni(&rs);
      float pc = node_interp(d->p_compton + (size_t)P->row[wm] * P->nb1, &ex);
      float pr = node_interp(d->p_rayleigh + (size_t)P->row[wm] * P->nb1, &ex);
      if (r < pc) {
        float qq, c;
        compton_sample(&rs, P, P->row[wm], ph.E, &qq, &c, &st->c[K_REJ]);
        float e1 = ph.E / qq;
        st->e[EN_WORLD] += (double)((ph.E - e1) * ph.w);
        float cp, sp;
        ref_azimuth(&rs, &cp, &sp);
        scatter_dir(ph.d, c, cp, sp);
        ph.E = e1;
        if (ph.E < d->e_cut) {
          st->e[EN_WORLD] += (double)(ph.E * ph.w);
          st->c[K_CUT]++;
          break;
        }
        energy_index(P, ph.E, &ex);
      } else if (r < pr) {
        float c = rayleigh_sample(&rs, P, P->row[wm], ph.E);
        float cp, sp;
        ref_azimuth(&rs, &cp, &sp);
        scatter_dir(ph.d, c, cp, sp);
      } else {
        st->e[EN_WORLD] += (double)(ph.E * ph.w);
        break;
      }
    }
  }
  history_end(t, &sl, st);
  if (o->event_edep) o->event_edep[h] = ev;
}

/* ------------------------------------------------------------------ driver */
typedef struct {
  const sim_t *S;
  const phys_t *P;
  uint64_t seed, first, h0, h1;
  uint32_t sim_id;
  size_t ntal;
  int flags, want_exact;
  tally_t t;
  stats_t st;
  outputs_t o;
} work_t;

static void *worker(void *arg) {
  work_t *w = (work_t *)arg;
  for (uint64_t h = w->h0; h < w->h1; ++h) run_history(w->S, w->P, w->seed, w->first, w->sim_id, h, &w->t, &w->st, &w->o);
  return NULL;
}

/* Runs histories [0, n) of one simulation (sim_id keys the streams) on n_threads
 * threads and adds into the caller's tallies (edep required; others may be NULL).
 * Recording of exit-face photons and per-history detector energy need n_threads = 1.
 * counters[12] and energies[4] are added to. Returns 0 on success. */
int ref_run(const gh_physics_desc *phys, const gh_phantom_desc *ph, const gh_source_desc *src,
            const gh_tally_desc *tal, uint32_t sim_id, uint64_t n, uint64_t first, uint64_t seed, int n_threads,
            double *edep, double *edep2, double *dose, double *dose2, double *counts, double *exact_edep2,
            float *const rec[8], uint64_t rec_cap, uint64_t *rec_n, float *event_edep, uint64_t counters[12],
            double energies[4]) {
  if (!phys || !ph || !src || !tal || !edep || n_threads < 1) return 1;
  if ((rec && rec[0]) || event_edep) n_threads = 1;
  phys_t P;
  phys_init(&P, phys);
  sim_t S;
  S.ph = ph;
  S.src = src;
  S.tal = tal;
  int nn[3] = {ph->nx, ph->ny, ph->nz};
  S.aligned = (tal->flags & GH_TALLY_ALIGNED) != 0;
  for (int i = 0; i < 3; ++i) {
    S.inv_s[i] = 1.0f / ph->spacing[i];
    S.L[i] = (float)((double)nn[i] * (double)ph->spacing[i]);
    S.tinv[i] = 1.0f / (S.aligned ? ph->spacing[i] : tal->spacing[i]);
  }
  sim_source_init(&S, src);
  size_t ntal = (size_t)tal->nx * tal->ny * tal->nz;
  work_t *W = (work_t *)calloc((size_t)n_threads, sizeof(work_t));
  pthread_t *T = (pthread_t *)calloc((size_t)n_threads, sizeof(pthread_t));
  for (int k = 0; k < n_threads; ++k) {
    work_t *w = &W[k];
    w->S = &S;
    w->P = &P;
    w->seed = seed;
    w->first = first;
    w->sim_id = sim_id;
    w->h0 = n * (uint64_t)k / (uint64_t)n_threads;
    w->h1 = n * (uint64_t)(k + 1) / (uint64_t)n_threads;
    if (n_threads == 1) {
      w->t.edep = edep; w->t.edep2 = edep2; w->t.dose = dose; w->t.dose2 = dose2; w->t.counts = counts;
      w->t.exact_edep2 = exact_edep2;
      if (rec) for (int i = 0; i < 8; ++i) w->o.rec[i] = rec[i];
      w->o.rec_cap = rec_cap;
      w->o.event_edep = event_edep;
    } else {
      w->t.edep = (double *)calloc(ntal, sizeof(double));
      w->t.edep2 = edep2 ? (double *)calloc(ntal, sizeof(double)) : NULL;
      w->t.dose = dose ? (double *)calloc(ntal, sizeof(double)) : NULL;
      w->t.dose2 = dose2 ? (double *)calloc(ntal, sizeof(double)) : NULL;
      w->t.counts = counts ? (double *)calloc(ntal, sizeof(double)) : NULL;
      w->t.exact_edep2 = exact_edep2 ? (double *)calloc(ntal, sizeof(double)) : NULL;
    }
  }
  if (n_threads == 1) {
    worker(&W[0]);
  } else {
    for (int k = 0; k < n_threads; ++k) pthread_create(&T[k], NULL, worker, &W[k]);
    for (int k = 0; k < n_threads; ++k) pthread_join(T[k], NULL);
  }
  for (int k = 0; k < n_threads; ++k) {
    work_t *w = &W[k];
    if (n_threads > 1) {
      double *dst[6] = {edep, edep2, dose, dose2, counts, exact_edep2};
      double *srcs[6] = {w->t.edep, w->t.edep2, w->t.dose, w->t.dose2, w->t.counts, w->t.exact_edep2};
      for (int a = 0; a < 6; ++a) {
        if (!dst[a] || !srcs[a]) continue;
        for (size_t i = 0; i < ntal; ++i) dst[a][i] += srcs[a][i];
        free(srcs[a]);
      }
    }
    free(w->t.hv);
    free(w->t.hs);
    if (counters) for (int i = 0; i < 12; ++i) counters[i] += w->st.c[i];
    if (energies) for (int i = 0; i < 4; ++i) energies[i] += w->st.e[i];
    if (rec_n) *rec_n += w->o.rec_n;
  }
  free(W);
  free(T);
  phys_free(&P);
  return 0;
}

/* HU -> labels: out[i] = lut[#{k : bins[k] <= hu[i]}] (numpy.digitize, right=False) */
void ref_hu_to_labels(const int16_t *hu, int64_t n, const double *bins, int n_bins, const uint16_t *lut, uint16_t *out) {
  for (int64_t i = 0; i < n; ++i) {
    int k = 0;
    while (k < n_bins && bins[k] <= (double)hu[i]) ++k;
    out[i] = lut[k];
  }
}

/* Sampler hooks for the distribution tests: n draws of the Compton (q = E/E', cos)
 * and Rayleigh (cos) samplers of table row m (a material, or a composition when the
 * materials share compositions) at energy E, stream (seed, history i). */
void ref_sample_compton(const gh_physics_desc *phys, int m, float E, uint64_t n, uint64_t seed, float *q_out,
                        float *c_out) {
  phys_t P;
  P.d = phys;
  P.nb1 = phys->n_buckets + 1;
  for (uint64_t i = 0; i < n; ++i) {
    stream_t s;
    stream_open(&s, seed, i, 0);
    uint64_t rej = 0;
    compton_sample(&s, &P, m, E, &q_out[i], &c_out[i], &rej);
  }
}

void ref_sample_rayleigh(const gh_physics_desc *phys, int m, float E, uint64_t n, uint64_t seed, float *c_out) {
  phys_t P;
  P.d = phys;
  P.nb1 = phys->n_buckets + 1;
  for (uint64_t i = 0; i < n; ++i) {
    stream_t s;
    stream_open(&s, seed, i, 0);
    c_out[i] = rayleigh_sample(&s, &P, m, E);
  }
}

/* primaries of histories first .. first + n - 1 as the source emits them (before any
   transport): position, direction, energy and weight into out[0..7] */
void ref_sample_source(const gh_source_desc *src, uint64_t n, uint64_t first, uint64_t seed, uint32_t sim_id,
                       float *const out[8]) {
  sim_t S;
  memset(&S, 0, sizeof(S));
  S.src = src;
  sim_source_init(&S, src);
  for (uint64_t i = 0; i < n; ++i) {
    stream_t s;
    stream_open(&s, seed, first + i, sim_id);
    photon_t ph;
    emit(&S, &s, i, &ph);
    for (int k = 0; k < 3; ++k) {
      out[k][i] = ph.p[k];
      out[3 + k][i] = ph.d[k];
    }
    out[6][i] = ph.E;
    out[7][i] = ph.w;
  }
}

/* attenuation coefficient of material m at E, as the transport sees it */
float ref_mu_at(const gh_physics_desc *phys, int m, float E) {
  phys_t P;
  phys_init(&P, phys);
  eidx_t x;
  energy_index(&P, E, &x);
  const float mu = material_mu(&P, m, &x);
  phys_free(&P);
  return mu;
}

/* ------------------------------------------------------------------ learned exit-plane source
 * Restatements of csrc/exitplane.hip (generator noise and decoder) in the same FP32
 * operation order, so that the GPU output is bit-identical. */

/* N(0,1) noise as float (the GPU rounds it to bf16): row r, columns 0..ld-1 (zero from n_cols) */
void ref_gaussian(uint64_t seed, uint64_t first, int64_t n, int cols, int ld, float *out) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int64_t r = 0; r < n; ++r) {
    const uint64_t id = first + (uint64_t)r;
    for (int c0 = 0; c0 < ld; c0 += 4) {
      const uint32_t ctr[4] = {(uint32_t)id, (uint32_t)(id >> 32), 0x4E4F4953u, (uint32_t)(c0 >> 2)};
      uint32_t x[4];
      ref_philox4x32_10(ctr, key, x);
      float u[4];
      for (int j = 0; j < 4; ++j) u[j] = (float)((x[j] >> 8) | 1u) * 0x1p-24f;
      const float ra = sqrtf(-2.0f * ref_log(u[0])), rb = sqrtf(-2.0f * ref_log(u[2]));
      float sa, ca, sb, cb;
      ref_sincos(6.28318530717958647692f * u[1], &sa, &ca);
      ref_sincos(6.28318530717958647692f * u[3], &sb, &cb);
      const float g[4] = {ra * ca, ra * sa, rb * cb, rb * sb};
      for (int j = 0; j < 4; ++j) {
        const int c = c0 + j;
        if (c < ld) out[r * ld + c] = c < cols ? g[j] : 0.0f;
      }
    }
  }
}

static float qdecode_ref(const float *q, int K, float scale, float g_max, float g) {
  float t = (g + g_max) * scale;
  t = fminf(fmaxf(t, 0.0f), (float)(K - 1));
  int k = (int)t;
  if (k > K - 2) k = K - 2;
  const float a = t - (float)k;
  return fmaf(a, q[k + 1] - q[k], q[k]);
}

static void normalize3_ref(float *x, float *y, float *z) {
  const float n2 = fmaf(*z, *z, fmaf(*y, *y, *x * *x));
  const float inv = 1.0f / sqrtf(n2);
  *x = *x * inv; *y = *y * inv; *z = *z * inv;
}

void ref_exitplane_decode(const float *Z, int64_t ldz, int64_t n, const float *qtab, int K, float g_max,
                          const float focal[3], float plane_z, const float offset[3], float weight, float *const out[8]) {
  const float scale = (float)(K - 1) / (2.0f * g_max);
  for (int64_t i = 0; i < n; ++i) {
    const float *zr = Z + i * ldz;
    const float E = qdecode_ref(qtab, K, scale, g_max, zr[0]);
    const float px = qdecode_ref(qtab + K, K, scale, g_max, zr[1]);
    const float py = qdecode_ref(qtab + 2 * K, K, scale, g_max, zr[2]);
    float lu = qdecode_ref(qtab + 3 * K, K, scale, g_max, zr[3]);
    float lv = qdecode_ref(qtab + 4 * K, K, scale, g_max, zr[4]);
    float lw = qdecode_ref(qtab + 5 * K, K, scale, g_max, zr[5]);
    const float l2 = fmaf(lw, lw, fmaf(lv, lv, lu * lu));
    if (!(l2 > 0.0f)) { lu = 0.0f; lv = 0.0f; lw = 1.0f; }
    else normalize3_ref(&lu, &lv, &lw);
    float nx = px - focal[0], ny = py - focal[1], nz = plane_z - focal[2];
    normalize3_ref(&nx, &ny, &nz);
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
    normalize3_ref(&du, &dv, &dw);
    out[0][i] = px + offset[0]; out[1][i] = py + offset[1]; out[2][i] = plane_z + offset[2];
    out[3][i] = du; out[4][i] = dv; out[5][i] = dw;
    out[6][i] = E; out[7][i] = weight;
  }
}

/* ------------------------------------------------------------------ SPECT heads
 * Restatements of csrc/arf.hip. ref_spect_pairs visits every head of every photon in
 * order (the GPU visits a window of candidate heads and writes pairs in atomic-slot
 * order): the two produce the same set of pairs with the same FP32 values except
 * theta / phi, which both take from the platform's acosf / atan2f. */
void ref_spect_pairs(const float *const rec[8], int64_t n, const gh_spect_desc *D, const float *cs, int64_t capacity,
                     float *x3, int32_t *pix, float *wgt, float *const det[8], float *tp, uint64_t *count) {
  const float sin_max = sqrtf(fmaxf(fmaf(-D->cos_theta_max, D->cos_theta_max, 1.0f), 0.0f));
  const float half_w = 0.5f * (float)D->nx * D->spacing[0], half_h = 0.5f * (float)D->ny * D->spacing[1];
  const float a2x = 0.5f * D->pitch, a2y = 0.86602540378443865f * D->pitch;
  uint64_t cnt = 0;
  for (int64_t i = 0; i < n; ++i) {
    const float px = rec[0][i] - D->center[0], py = rec[1][i] - D->center[1], pz = rec[2][i] - D->center[2];
    const float du = rec[3][i], dv = rec[4][i], dw = rec[5][i], e = rec[6][i], ww = rec[7][i];
    if (!(fabsf(dw) < sin_max && ww > 0.0f)) continue;
    for (int k = 0; k < D->n_proj; ++k) {
      const float ca = cs[2 * k], sa = cs[2 * k + 1];
      const float c = fmaf(du, ca, dv * sa);
      if (!(c > D->cos_theta_max)) continue;
      const float s = fmaf(px, ca, py * sa);
      const float t = (D->radius - s) / c;
      if (!(t >= 0.0f)) continue;
      const float dy = fmaf(-du, sa, dv * ca);
      const float ex = fmaf(t, du, px), ey = fmaf(t, dv, py);
      const float hu = fmaf(-ex, sa, ey * ca);
      const float hz = fmaf(t, dw, pz);
      const float f = D->crystal_depth / c;
      const float bu = fmaf(f, dy, hu), bz = fmaf(f, dw, hz);
      const float fx = (bu + half_w) / D->spacing[0], fy = (bz + half_h) / D->spacing[1];
      if (!(fx >= 0.0f && fy >= 0.0f && fx < (float)D->nx && fy < (float)D->ny)) continue;
      int ix = (int)fx, iy = (int)fy;
      if (ix > D->nx - 1) ix = D->nx - 1;
      if (iy > D->ny - 1) iy = D->ny - 1;
      const uint64_t slot = cnt++;
      if ((int64_t)slot >= capacity) continue;
      const float theta = acosf(fminf(c, 1.0f));
      const float phi = atan2f(dy, dw);
      if (x3) {
        x3[3 * slot] = fmaf(e, D->norm[0], D->norm[1]);
        x3[3 * slot + 1] = fmaf(theta, D->norm[2], D->norm[3]);
        x3[3 * slot + 2] = fmaf(phi, D->norm[4], D->norm[5]);
      }
      if (tp) { tp[2 * slot] = theta; tp[2 * slot + 1] = phi; }
      pix[slot] = (k * D->ny + iy) * D->nx + ix;
      wgt[slot] = ww;
      if (det) {
        float rx = hz, ry = hu;
        if (D->pitch > 0.0f) {
          const float jj = floorf(ry / a2y);
          ry = fmaf(-jj, a2y, ry);
          rx = fmaf(-jj, a2x, rx);
          const float ii = floorf(rx / D->pitch);
          rx = fmaf(-ii, D->pitch, rx);
        }
        det[0][slot] = D->patch_origin[0] + rx;
        det[1][slot] = D->patch_origin[1] + ry;
        det[2][slot] = D->patch_origin[2];
        det[3][slot] = dw;
        det[4][slot] = dy;
        det[5][slot] = -c;
        det[6][slot] = e;
        det[7][slot] = ww;
      }
    }
  }
  *count = cnt;
}

/* width of the blurred energy (laws of GateDigitizerBlurringActor.cpp:75-94): law 1 fixed
 * sigma, 2 FWHM = resolution sqrt(reference e), 3 FWHM = (slope (e - reference) + resolution) e */
static float ref_blur_sigma(int law, float sigma, float resolution, float reference, float slope, float e) {
  const float fwhm_to_sigma = 0.42466090014400953f;
  if (law == 1) return sigma;
  if (law == 2) return (resolution * sqrtf(reference * e)) * fwhm_to_sigma;
  return ((slope * (e - reference) + resolution) * e) * fwhm_to_sigma;
}

void ref_energy_windows(const float *edep, int64_t n, uint64_t seed, uint64_t first, int law, float sigma,
                        float resolution, float reference, float slope, const float *win, int n_win, int32_t *cls) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int64_t i = 0; i < n; ++i) {
    const float e = edep[i];
    int c = 0;
    if (e > 0.0f) {
      float eb = e;
      if (law != 0) {
        const uint64_t id = first + (uint64_t)i;
        const uint32_t ctr[4] = {(uint32_t)id, (uint32_t)(id >> 32), 0x424C5552u, 0u};
        uint32_t x[4];
        ref_philox4x32_10(ctr, key, x);
        const float u0 = (float)((x[0] >> 8) | 1u) * 0x1p-24f, u1 = (float)((x[1] >> 8) | 1u) * 0x1p-24f;
        float sn, csn;
        ref_sincos(6.28318530717958647692f * u1, &sn, &csn);
        const float g = sqrtf(-2.0f * ref_log(u0)) * csn;
        eb = fmaf(ref_blur_sigma(law, sigma, resolution, reference, slope, e), g, e);
      }
      for (int k = 0; k < n_win; ++k)
        if (eb >= win[2 * k] && eb < win[2 * k + 1]) { c = k + 1; break; }
    }
    cls[i] = c;
  }
}
