// MI355X photon transport + voxel dose: persistent Woodcock kernel and the C-ABI.
//
// One lane owns one photon history at a time. A wave keeps its 64 lanes busy by
// refilling lanes whose photon died: every `refill_interval` Woodcock iterations the
// wave takes a ballot of dead lanes, hands them consecutive history ids from its
// current work chunk by a prefix count (mbcnt) over the ballot mask, and grabs a new
// chunk from a global queue with one atomic when the chunk runs dry. The work queue
// spans every simulation of the launch, so one launch runs many simulations
// (batched form) and a single simulation is the case n_sims = 1.
//
// Per step a photon reads one label byte (gather from HBM/L2) and the interpolated
// mu of that material from LDS, where the per-material mu tables and the majorant are
// staged at kernel start. Deposits go through a 4-slot per-lane cache that merges
// deposits of the same history in the same voxel in FP32 and flushes each (history,
// voxel) sum into FP64 edep / edep^2 with hardware FP64 atomics: the squared image is
// the per-history sum squared, as GateDoseActor::ScoreSquaredValue
// (GateDoseActor.cpp:359-381) computes it.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "gh_device.h"
#include "../../include/gatehip.h"

using namespace gh;

namespace {

struct SimDev {
  // phantom
  const uint32_t *labels;  // 2^lb-bit labels packed in 32-bit words (lb = 2, 3, 4: 4-, 8-, 16-bit)
  int lb;
  int nx, ny, nz, world_mat;
  float inv_s[3];
  float ts_ph[3];  // voxel spacing
  float L[3];
  float wlo[3], whi[3];
  // source
  int pos_type, dir_type, ene_type, n_energy;
  float pos_param[4];
  float pos_mat[12];
  float dir_mat[9];
  float ct_min, d_ct, phi_min, d_phi;
  int phi_full, phsp_iso;
  float mom[3], focus[3];
  float e_mono, e_max, weight;
  float pos_sigma[2], e_sigma, w_sigma;
  float s2_min, s2_d;      // cosine law: sin^2 theta_min, sin^2 theta_max - sin^2 theta_min
  float axis[3];           // surface sphere: unit axis of the tangent frame
  float dir_sigma[2];      // beam2d
  int n_theta, n_phi;      // histogram direction tables
  const float *th_val, *th_width, *th_cdf, *ph_val, *ph_width, *ph_cdf;
  const float *e_val, *e_width, *e_cdf;
  const int32_t *e_guide;  // [SPEC_GUIDE+1] guide of e_cdf
  int64_t phsp_n, phsp_start;
  const float *phsp[8];
  int64_t vox_n;
  const float *vox_prob;
  const int32_t *vox_alias, *vox_index;
  int vox_nx, vox_ny;
  float vox_s[3];
  // tally
  double *edep, *edep2, *dose, *dose2, *counts;
  int aligned, tnx, tny, tnz;
  float tcorner[3], tinv[3], ts[3];  // tally grid (aligned: the phantom grid, corner 0)
  // recorder / detector
  int rec_face;
  uint32_t rec_cap;
  unsigned long long *rec_count;
  float *rec[8];
  int det_label_min;
  float *event_edep;
  // track-length (collision) kerma estimator
  int tle;
  float tle_emin, tle_emax;
  // run
  uint64_t n_hist, first_hist;
};

struct Launch {
  const SimDev *sims;
  int n_sims;
  const uint64_t *chunk_offsets;  // [n_sims+1] prefix sum of chunks per sim
  uint64_t total_chunks;
  uint32_t chunk;
  unsigned long long *next_chunk;
  uint64_t seed;
  int max_iter;   // step-phase iterations (2 tentative steps each) between slow phases
  int slow_min;   // lanes waiting for slow work that end a step phase
  int refill_min; // dead lanes that trigger a refill (<= slow_min)
  int scatter_tries; // Compton / Rayleigh sampler tries per slow phase before a photon waits
  Phys P;
  unsigned long long *counters;   // [12]
  double *energies;               // [4]
};

// --------------------------------------------------------------------- wave helpers
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// --------------------------------------------------------------------- hot fields
// Per-lane copy of what every Woodcock step reads (in a batched launch each lane may
// follow another simulation); fields needed only at slow events (box size, world
// material, tally alignment) are read from the simulation descriptor there.
// label of voxel vox from 2^lb-bit labels packed in 32-bit words (voxel v in word
// v >> (5 - lb), bits (v mod 2^(5-lb)) << lb): one global load and one bit-field extract
// for every layout, no branch on it
__device__ __forceinline__ int packed_label(const uint32_t *labels, uint32_t lb, int vox) {
  const auto *g = (const __attribute__((address_space(1))) uint32_t *)labels;
  const uint32_t v = (uint32_t)vox;
  const uint32_t w = g[v >> (5u - lb)];
  return (int)__builtin_amdgcn_ubfe(w, (v & ((1u << (5u - lb)) - 1u)) << lb, 1u << lb);
}

struct Hot {
  const uint32_t *labels;
  int nx, ny, nz;
  float isx, isy, isz;
  uint32_t lb;
  __device__ __forceinline__ void load(const SimDev &s) {
    labels = s.labels; nx = s.nx; ny = s.ny; nz = s.nz;
    isx = s.inv_s[0]; isy = s.inv_s[1]; isz = s.inv_s[2];
    lb = (uint32_t)s.lb;
  }
  __device__ __forceinline__ int label(int vox) const { return packed_label(labels, lb, vox); }
};

// (int)floorf(x) in one instruction (floor and conversion; out-of-range values clamp to the
// int range as the conversion alone does, and any of them lies outside a phantom)
__device__ __forceinline__ int floor_int(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// reciprocal direction: the slab tests of a photon reuse one set of divisions
struct InvDir {
  float x, y, z;
};
__device__ __forceinline__ InvDir inv_dir(float dx, float dy, float dz) {
  InvDir r;
  r.x = 1.0f / dx; r.y = 1.0f / dy; r.z = 1.0f / dz;
  return r;
}

// slab test of one axis of the box [0, L]; inv = 1/d (used only when d != 0)
__device__ __forceinline__ void slab(float p, float d, float inv, float L, int axis, float &tnear, float &tfar,
                                     int &face) {
  if (d != 0.0f) {
    const float t0 = (-p) * inv;
    const float t1 = (L - p) * inv;
    const float tmn = fminf(t0, t1), tmx = fmaxf(t0, t1);
    if (tmn > tnear) tnear = tmn;
    if (tmx < tfar) { tfar = tmx; face = 2 * axis + (d > 0.0f ? 1 : 0); }
  } else if (p < 0.0f || p > L) {
    tnear = INFINITY;
  }
}

// ray vs box [0,L]: entry/exit distance and exit face (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z)
__device__ __forceinline__ void ray_box(float px, float py, float pz, float dx, float dy, float dz,
                                        const InvDir &iv, float Lx, float Ly, float Lz, float &tnear,
                                        float &tfar, int &face) {
  tnear = -INFINITY; tfar = INFINITY; face = -1;
  slab(px, dx, iv.x, Lx, 0, tnear, tfar, face);
  slab(py, dy, iv.y, Ly, 1, tnear, tfar, face);
  slab(pz, dz, iv.z, Lz, 2, tnear, tfar, face);
}

__device__ __forceinline__ float world_axis(float p, float d, float inv, float lo, float hi, float t) {
  if (d > 0.0f) return fminf(t, (hi - p) * inv);
  if (d < 0.0f) return fminf(t, (lo - p) * inv);
  return t;
}

// distance to the world box boundary from inside (iv = 1/direction)
__device__ __forceinline__ float world_exit(const SimDev &s, float px, float py, float pz, float dx,
                                            float dy, float dz, const InvDir &iv) {
  float t = INFINITY;
  t = world_axis(px, dx, iv.x, s.wlo[0], s.whi[0], t);
  t = world_axis(py, dy, iv.y, s.wlo[1], s.whi[1], t);
  t = world_axis(pz, dz, iv.z, s.wlo[2], s.whi[2], t);
  return fmaxf(t, 0.0f);
}

// FP64 add to global memory with the hardware atomic (no return value)
__device__ __forceinline__ void gatomic_add(double *p, double v) {
  (void)__builtin_amdgcn_global_atomic_fadd_f64((__attribute__((address_space(1))) double *)p, v);
}

// --------------------------------------------------------------------- photon
enum { MODE_BOX = 0, MODE_WORLD = 1 };

struct Photon {
  float x, y, z, u, v, w, E, wt;
  int mode;
  uint32_t hist;   // local history index within its simulation
  int sim;
};

// squared-value cache: up to 4 (voxel, sum) entries of the current history, as scalar
// registers (slot k: v_k voxel, s_k FP32 energy sum, d_k dose sum, c_k deposit count)
struct SqCache {
  int v0, v1, v2, v3;
  float s0, s1, s2, s3;
  float d0, d1, d2, d3;
  uint32_t c0, c1, c2, c3;
  int n;
};

// Counters: steps and rejections stay in lane registers; the rare event counters are
// summed per workgroup in LDS with one atomic per wave event (ballot + popcount); the
// energy sums go to a per-lane LDS slot; all are flushed to global memory at kernel end.
enum { C_HIST = 0, C_STEPS, C_DEPOSITS, C_COMPTON, C_RAYLEIGH, C_PHOTO, C_ESCAPED, C_WORLD, C_CUT,
       C_EVICT, C_RECORDED, C_REJ, N_COUNTERS };
enum { E_DEP = 0, E_OUT, E_ESC, E_WORLD, N_ENERGIES };

struct LaneStats {
  uint32_t steps, rej;
  unsigned long long *cnt;  // LDS [N_COUNTERS]
  double *en;               // LDS [N_ENERGIES][BLOCK]: this lane's column
  int stride;               // BLOCK
  __device__ __forceinline__ void inc(int k) {
    const uint64_t m = __ballot(1);
    if (__lane_id() == (uint32_t)(__ffsll((unsigned long long)m) - 1))
      atomicAdd(cnt + k, (unsigned long long)__popcll(m));
  }
  __device__ __forceinline__ void add_e(int k, double v) { unsafeAtomicAdd(en + k * stride, v); }
};

// tally arrays of a simulation, read into registers before the first atomic of a flush
// (a read of the descriptor issued after an atomic would wait for that atomic)
struct TallyOut {
  double *edep, *edep2, *counts, *dose, *dose2;
};
__device__ __forceinline__ TallyOut tally_out(const SimDev &s) {
  return TallyOut{s.edep, s.edep2, s.counts, s.dose, s.dose2};
}

__device__ __forceinline__ void flush_slot(const TallyOut &o, int vox, float sum, float dsum, uint32_t cnt,
                                           LaneStats &st) {
  gatomic_add(o.edep + vox, (double)sum);
  if (o.edep2) gatomic_add(o.edep2 + vox, (double)sum * (double)sum);
  if (o.counts) gatomic_add(o.counts + vox, (double)cnt);
  if (o.dose) gatomic_add(o.dose + vox, (double)dsum);
  if (o.dose2) gatomic_add(o.dose2 + vox, (double)dsum * (double)dsum);
  st.inc(C_DEPOSITS);
  st.add_e(E_DEP, (double)sum);
}

__device__ __forceinline__ void sq_flush_all(const SimDev &s, SqCache &c, LaneStats &st) {
  const TallyOut o = tally_out(s);
  if (c.n > 0) flush_slot(o, c.v0, c.s0, c.d0, c.c0, st);
  if (c.n > 1) flush_slot(o, c.v1, c.s1, c.d1, c.c1, st);
  if (c.n > 2) flush_slot(o, c.v2, c.s2, c.d2, c.c2, st);
  if (c.n > 3) flush_slot(o, c.v3, c.s3, c.d3, c.c3, st);
  c.n = 0;
}

// Adds deposit x (dose dx) of the current history to tally voxel tv. A history's
// deposits in the same voxel are summed in FP32 before the squared value is formed;
// with more than 4 distinct voxels the oldest slot is flushed (counted as an eviction).
__device__ __forceinline__ void deposit(const SimDev &s, SqCache &c, int tv, float x, float dx,
                                        LaneStats &st) {
  if (tv < 0) { st.add_e(E_OUT, (double)x); return; }
  // Slots are updated with value selects, never with one store per slot in separate
  // branches: the compiler merges such stores into a store through a selected slot
  // address, which puts the whole cache in scratch memory.
  const bool h0 = c.n > 0 && c.v0 == tv, h1 = c.n > 1 && c.v1 == tv;
  const bool h2 = c.n > 2 && c.v2 == tv, h3 = c.n > 3 && c.v3 == tv;
  if (h0 || h1 || h2 || h3) {
#define GH_SLOT_ADD(K)                                                     \
    c.s##K = h##K ? c.s##K + x : c.s##K; c.d##K = h##K ? c.d##K + dx : c.d##K; \
    c.c##K += h##K ? 1u : 0u;
    GH_SLOT_ADD(0) GH_SLOT_ADD(1) GH_SLOT_ADD(2) GH_SLOT_ADD(3)
#undef GH_SLOT_ADD
    return;
  }
  if (c.n == 4) {
    flush_slot(tally_out(s), c.v0, c.s0, c.d0, c.c0, st);
    c.v0 = c.v1; c.s0 = c.s1; c.d0 = c.d1; c.c0 = c.c1;
    c.v1 = c.v2; c.s1 = c.s2; c.d1 = c.d2; c.c1 = c.c2;
    c.v2 = c.v3; c.s2 = c.s3; c.d2 = c.d3; c.c2 = c.c3;
    c.n = 3;
    st.inc(C_EVICT);
  }
#define GH_SLOT_SET(K)                                                     \
  {                                                                        \
    const bool t = c.n == K;                                               \
    c.v##K = t ? tv : c.v##K; c.s##K = t ? x : c.s##K;                     \
    c.d##K = t ? dx : c.d##K; c.c##K = t ? 1u : c.c##K;                    \
  }
  GH_SLOT_SET(0) GH_SLOT_SET(1) GH_SLOT_SET(2) GH_SLOT_SET(3)
#undef GH_SLOT_SET
  c.n += 1;
}

// Tally voxel of a deposit at (x,y,z) in the phantom frame, -1 outside the tally grid.
// The index follows the ITK contract of GateDoseActor::GetVoxelPosition
// (GateDoseActor.cpp:144-178, TransformPhysicalPointToIndex) exactly (cell_index), also on
// aligned grids, where the Woodcock step's voxel (a navigation analog) can differ from it
// for points within an ulp of a face.
__device__ __forceinline__ int tally_voxel(const SimDev &s, float x, float y, float z) {
  const int ix = cell_index(x, s.tcorner[0], s.ts[0], s.tinv[0]);
  const int iy = cell_index(y, s.tcorner[1], s.ts[1], s.tinv[1]);
  const int iz = cell_index(z, s.tcorner[2], s.ts[2], s.tinv[2]);
  if ((unsigned)ix >= (unsigned)s.tnx || (unsigned)iy >= (unsigned)s.tny || (unsigned)iz >= (unsigned)s.tnz)
    return -1;
  return ix + s.tnx * (iy + s.tny * iz);
}

// ---------------------------------------------------------------------- source
// two independent N(0,1) values (Box-Muller): radius sqrt(-2 log u), angle 2 pi u'
__device__ __forceinline__ void normal_pair(Rng &g, float &n1, float &n2) {
  const float r = sqrtf(-2.0f * glog(u01(g), c_logtab));
  float cp, sp;
  azimuth(g, cp, sp);
  n1 = r * cp;
  n2 = r * sp;
}

__device__ __forceinline__ void unit_or_zero(float &x, float &y, float &z) {
  const float n2 = fmaf(z, z, fmaf(y, y, x * x));
  if (n2 > 0.0f) {
    const float inv = 1.0f / sqrtf(n2);
    x = x * inv; y = y * inv; z = z * inv;
  }
}

// (u, v, w) <- u x' + v y' + w n in the tangent frame at a surface point with outward unit
// normal n: x' = unit(A x n), y' = unit(x' x n) (G4SPSPosDistribution, sphere surface)
__device__ void tangent_frame(const float *A, float nx, float ny, float nz, float &u, float &v, float &w) {
  float ax = fmaf(A[1], nz, -(A[2] * ny)), ay = fmaf(A[2], nx, -(A[0] * nz)), az = fmaf(A[0], ny, -(A[1] * nx));
  unit_or_zero(ax, ay, az);
  float bx = fmaf(ay, nz, -(az * ny)), by = fmaf(az, nx, -(ax * nz)), bz = fmaf(ax, ny, -(ay * nx));
  unit_or_zero(bx, by, bz);
  const float d0 = u, d1 = v, d2 = w;
  u = fmaf(d0, ax, fmaf(d1, bx, d2 * nx));
  v = fmaf(d0, ay, fmaf(d1, by, d2 * ny));
  w = fmaf(d0, az, fmaf(d1, bz, d2 * nz));
}

// source-spectrum tables of a simulation: its global arrays, or copies in LDS
struct SpecTab {
  const float *val, *width, *cdf;
  const int32_t *guide;
};
constexpr int SPEC_LDS_MAX = 256;  // spectra up to this many bins are staged in LDS (1024-thread forms)

__device__ void sample_source(const SimDev &s, const SpecTab &sp, Rng &g, uint32_t h, Photon &p) {
  float x = 0.0f, y = 0.0f, z = 0.0f, u = 0.0f, v = 0.0f, w = 1.0f, E = s.e_mono, wt = s.weight;
  float nx = 0.0f, ny = 0.0f, nz = 1.0f;  // surface sphere: outward unit normal at the point
  if (s.pos_type == GH_POS_PHSP) {
    const int64_t row = (s.phsp_start + (int64_t)h) % s.phsp_n;
    x = s.phsp[0][row]; y = s.phsp[1][row]; z = s.phsp[2][row];
    u = s.phsp[3][row]; v = s.phsp[4][row]; w = s.phsp[5][row];
    E = s.phsp[6][row]; wt = s.phsp[7][row];
  } else {
    // ---- position (GateGenericSource.cpp:312-384; G4SPSPosDistribution shapes)
    switch (s.pos_type) {
      case GH_POS_BOX:
        x = (u01(g) - 0.5f) * s.pos_param[0];
        y = (u01(g) - 0.5f) * s.pos_param[1];
        z = (u01(g) - 0.5f) * s.pos_param[2];
        break;
      case GH_POS_SPHERE:
      case GH_POS_ELLIPSOID: {
        float a, b, c, r2;
        int it = 0;
        do {
          a = fmaf(2.0f, u01(g), -1.0f);
          b = fmaf(2.0f, u01(g), -1.0f);
          c = fmaf(2.0f, u01(g), -1.0f);
          r2 = fmaf(c, c, fmaf(b, b, a * a));
        } while (r2 > 1.0f && ++it < MAX_REJECT);
        if (s.pos_type == GH_POS_SPHERE) {
          x = s.pos_param[0] * a; y = s.pos_param[0] * b; z = s.pos_param[0] * c;
        } else {
          x = s.pos_param[0] * a; y = s.pos_param[1] * b; z = s.pos_param[2] * c;
        }
        break;
      }
      case GH_POS_VOXELS: {
        // Walker alias table: entry i uniform (32-bit multiply-shift), kept with
        // probability prob[i], else its alias; then uniform inside the voxel
        const int64_t i = (int64_t)(((uint64_t)rng_next(g) * (uint64_t)s.vox_n) >> 32);
        const int64_t j = (u01(g) < s.vox_prob[i]) ? i : (int64_t)s.vox_alias[i];
        const int v = s.vox_index[j];
        const int ix = v % s.vox_nx, iyz = v / s.vox_nx;
        const int iy = iyz % s.vox_ny, iz = iyz / s.vox_ny;
        x = ((float)ix + u01(g)) * s.vox_s[0];
        y = ((float)iy + u01(g)) * s.vox_s[1];
        z = ((float)iz + u01(g)) * s.vox_s[2];
        break;
      }
      case GH_POS_DISC:
      case GH_POS_CYLINDER: {
        float a, b, r2;
        int it = 0;
        do {
          a = fmaf(2.0f, u01(g), -1.0f);
          b = fmaf(2.0f, u01(g), -1.0f);
          r2 = fmaf(b, b, a * a);
        } while (r2 > 1.0f && ++it < MAX_REJECT);
        x = s.pos_param[0] * a; y = s.pos_param[0] * b;
        if (s.pos_type == GH_POS_CYLINDER) z = fmaf(2.0f, u01(g), -1.0f) * s.pos_param[1];
        if (s.pos_type == GH_POS_DISC && (s.pos_sigma[0] > 0.0f || s.pos_sigma[1] > 0.0f)) {
          // "Beam" circle: the uniform point plus a Gaussian spread in x and y
          float n1, n2;
          normal_pair(g, n1, n2);
          x = fmaf(s.pos_sigma[0], n1, x);
          y = fmaf(s.pos_sigma[1], n2, y);
        }
        break;
      }
      case GH_POS_SURFACE_SPHERE: {
        const float ct = fmaf(2.0f, u01(g), -1.0f);
        const float st = sqrtf(fmaxf(fmaf(-ct, ct, 1.0f), 0.0f));
        float cp, sp;
        azimuth(g, cp, sp);
        nx = st * cp; ny = st * sp; nz = ct;
        x = s.pos_param[0] * nx; y = s.pos_param[0] * ny; z = s.pos_param[0] * nz;
        break;
      }
      default: break;
    }
    // ---- direction (G4SPSAngDistribution iso convention: momentum = -(st cp, st sp, ct))
    if (s.dir_type == GH_DIR_ISO || s.dir_type == GH_DIR_COS) {
      // iso: cos theta uniform in [cos theta_max, cos theta_min]; cos: sin^2 theta uniform
      // in [sin^2 theta_min, sin^2 theta_max] (GenerateIsotropicFlux / GenerateCosineLawFlux)
      float ct, st;
      if (s.dir_type == GH_DIR_ISO) {
        ct = fmaf(-u01(g), s.d_ct, s.ct_min);
        st = sqrtf(fmaxf(fmaf(-ct, ct, 1.0f), 0.0f));
      } else {
        const float s2 = fmaf(u01(g), s.s2_d, s.s2_min);
        st = sqrtf(s2);
        ct = sqrtf(fmaxf(1.0f - s2, 0.0f));
      }
      float cp, sp;
      if (s.phi_full) {
        azimuth(g, cp, sp);
      } else {
        const float phi = fmaf(u01(g), s.d_phi, s.phi_min);
        gsincos(phi, sp, cp);
      }
      u = -(st * cp); v = -(st * sp); w = -ct;
      if (s.pos_type == GH_POS_SURFACE_SPHERE) tangent_frame(s.axis, nx, ny, nz, u, v, w);
    } else if (s.dir_type == GH_DIR_BEAM2D) {
      // GenerateBeamFlux ("beam2d"): Gaussian angles (a, b), polar angle |(a, b)|, azimuth
      // the angle of (a, b), 0 when both are 0
      float n1, n2;
      normal_pair(g, n1, n2);
      const float a = s.dir_sigma[0] * n1, b = s.dir_sigma[1] * n2;
      const float t = sqrtf(fmaf(b, b, a * a));
      float cp = 1.0f, sp = 0.0f;
      if (t > 0.0f) {
        cp = a / t;
        sp = b / t;
      }
      float sn, cs;
      gsincos(t, sn, cs);
      u = -(sn * cp); v = -(sn * sp); w = -cs;
    } else if (s.dir_type == GH_DIR_HISTOGRAM) {
      // GenerateUserDefFlux: theta, then phi, each piecewise uniform
      int k = cdf_index(s.th_cdf, s.n_theta, u01(g));
      const float t = fmaf(u01(g), s.th_width[k], s.th_val[k]);
      k = cdf_index(s.ph_cdf, s.n_phi, u01(g));
      const float ph = fmaf(u01(g), s.ph_width[k], s.ph_val[k]);
      float st, ct, sp, cp;
      gsincos(t, st, ct);
      gsincos(ph, sp, cp);
      u = -(st * cp); v = -(st * sp); w = -ct;
    } else if (s.dir_type == GH_DIR_MOMENTUM) {
      u = s.mom[0]; v = s.mom[1]; w = s.mom[2];
    }
    // ---- energy (GateSPSEneDistribution.cpp:16-159)
    if (s.ene_type == GH_ENE_LINES) {
      const int k = cdf_index_guided(sp.cdf, sp.guide, u01(g));
      E = sp.val[k];
    } else if (s.ene_type == GH_ENE_HISTOGRAM) {
      const int k = cdf_index_guided(sp.cdf, sp.guide, u01(g));
      E = fmaf(u01(g), sp.width[k], sp.val[k]);
    } else if (s.ene_type == GH_ENE_RANGE) {
      E = fmaf(u01(g), s.e_max - s.e_mono, s.e_mono);
    } else if (s.ene_type == GH_ENE_GAUSS) {
      float n1, n2;
      normal_pair(g, n1, n2);
      E = fminf(fmaxf(fmaf(s.e_sigma, n1, s.e_mono), 0.0f), s.e_max);
    }
    // Gaussian history weight (GateGenericSource.cpp:247-259)
    if (s.w_sigma >= 0.0f) {
      float n1, n2;
      normal_pair(g, n1, n2);
      wt = fmaf(s.w_sigma, n1, s.weight);
    }
  }
  // source frame -> phantom frame
  const float *M = s.pos_mat;
  p.x = fmaf(M[0], x, fmaf(M[1], y, fmaf(M[2], z, M[9])));
  p.y = fmaf(M[3], x, fmaf(M[4], y, fmaf(M[5], z, M[10])));
  p.z = fmaf(M[6], x, fmaf(M[7], y, fmaf(M[8], z, M[11])));
  if (s.pos_type == GH_POS_PHSP && s.phsp_iso) {
    const float ct = fmaf(2.0f, u01(g), -1.0f);
    const float st = sqrtf(fmaxf(fmaf(-ct, ct, 1.0f), 0.0f));
    float cp, sp;
    azimuth(g, cp, sp);
    p.u = st * cp; p.v = st * sp; p.w = ct;
  } else if (s.dir_type == GH_DIR_FOCUSED && s.pos_type != GH_POS_PHSP) {
    p.u = s.focus[0] - p.x; p.v = s.focus[1] - p.y; p.w = s.focus[2] - p.z;
    normalize3(p.u, p.v, p.w);
  } else {
    const float *D = s.dir_mat;
    p.u = fmaf(D[0], u, fmaf(D[1], v, D[2] * w));
    p.v = fmaf(D[3], u, fmaf(D[4], v, D[5] * w));
    p.w = fmaf(D[6], u, fmaf(D[7], v, D[8] * w));
    normalize3(p.u, p.v, p.w);
  }
  p.E = E;
  p.wt = wt;
}

// ---------------------------------------------------------------------- kernel
// lane modes
// (every mode from M_COLLIDE on is slow work)
enum { M_DEAD = 0, M_STEP = 1, M_COLLIDE = 2, M_EXIT = 3, M_WORLD = 4, M_COMPTON = 5, M_RAYLEIGH = 6 };

// global history id of the lane's current history (event-stream counter words)
__device__ __forceinline__ uint64_t history_of(const Rng &g) { return ((uint64_t)g.c1 << 32) | g.c0; }

// the 21-bit fields of one step-stream Philox block (words w0..w3) for tentative step k of
// three: step 0 takes the low 21 bits of w0 (distance) and w1 (accept), step 1 those of w2
// and w3, step 2 the top 11 bits of w0 | w2 (distance) and of w1 | w3 (accept), the w2 / w3
// part above (its top bit dropped)
template <int K>
__device__ __forceinline__ void step_fields(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint32_t &bd,
                                            uint32_t &ba) {
  constexpr uint32_t M21 = 0x1FFFFFu;
  if (K == 0) { bd = w0 & M21; ba = w1 & M21; }
  else if (K == 1) { bd = w2 & M21; ba = w3 & M21; }
  else { bd = ((w0 >> 21) | (w2 >> 21 << 11)) & M21; ba = ((w1 >> 21) | (w3 >> 21 << 11)) & M21; }
}
// odd multiple of 2^-22 in (0, 1) from a 21-bit field, as an exact integer
__device__ __forceinline__ float odd22(uint32_t b) { return (float)((b << 1) | 1u); }

// Persistent transport kernel. The loop of a wave has three phases, each entered by
// all 64 lanes together (the conditions are ballots, so they are wave-uniform):
//   refill  dead lanes take photons from the wave's birth queue once at least refill_min
//           lanes are dead, or when no lane can step. The queue (64 entries per wave in
//           LDS) is filled by a birth pass of all 64 lanes at once: each lane samples the
//           source for one history of the wave's work chunk, and a photon born outside the
//           phantom is carried to it through the world; photons that stop at birth (below
//           the tracking cut, or missing the phantom) are finished in the pass and never
//           take a lane. So the source sampling runs on full waves however few lanes wait.
//   step    every lane in M_STEP takes up to three Woodcock steps per iteration from one
//           Philox block of its step stream; a real collision (M_COLLIDE) or a box exit
//           (M_EXIT) parks the lane; the phase ends when slow_min lanes are parked;
//   slow    parked lanes process their collision, exit or world event together, using
//           the event stream of their history. The Compton and Rayleigh samplers are
//           rejection methods: a lane runs at most `scatter_tries` tries per slow phase and
//           keeps its remaining tries (M_COMPTON / M_RAYLEIGH) for the next one, so a wave
//           does not run as many tries as its slowest lane needs.
// Batching the slow work this way keeps the wave converged: a collision in one lane no
// longer stalls 63 lanes that only want to step.
// After the birth of a history (source sampling, and the world step to the phantom for a
// photon born outside it) its event stream continues at the next unused Philox block, so
// that a queue entry carries the block counter instead of buffered words.
// LDS layout (floats) of the staged tables; STAGE 1 stages mu + majorant, STAGE 2 also the
// form-factor tables of the samplers (the branching ratios, read once per collision, stay
// in global memory, cached)
struct LdsLayout {
  int mu, maj, mat, sinc, rcdf, rx, total;
};
// expand: the staged mu table has one row per material (node products density x row) instead
// of one per table row, so that a Woodcock step reads it without the (density, row) lookup
__host__ __device__ inline LdsLayout lds_layout(int n_rows, int n_mat, int nb1, int nx, int stage, bool expand) {
  LdsLayout o;
  o.mu = 0;
  o.maj = (expand ? n_mat : n_rows) * nb1;
  o.mat = (o.maj + nb1 + 1) & ~1;  // MatRow pairs, 8-byte aligned
  const int end = o.mat + 2 * n_mat;
  o.sinc = end;
  o.rcdf = o.sinc + n_rows * nx;
  o.rx = o.rcdf + n_rows * nx;
  o.total = stage >= 2 ? o.rx + nx : end;
  if (stage == 0) o.total = 0;
  return o;
}

// birth queue entry: 10 words per photon (x y z u v w E weight, history, meta) in LDS,
// [word][workgroup lane] so that the 64 slots of a wave are consecutive
constexpr int QWORDS = 10;
constexpr uint32_t META_COLLIDE = 1u << 16;  // born outside, interacts in the world first
constexpr int META_SIM_SHIFT = 17;           // simulation id in bits 17..31 (c3 in 0..15)
constexpr int MAX_SIMS = 1 << 15;

// label of voxel `vox` of simulation s (both label layouts)
__device__ __forceinline__ int sim_label(const SimDev &s, int vox) { return packed_label(s.labels, (uint32_t)s.lb, vox); }

// phantom voxel of a point (exact cell index, as the tally's), -1 outside the box
__device__ __forceinline__ int sim_voxel(const SimDev &s, float x, float y, float z) {
  const int ix = cell_index(x, 0.0f, s.ts_ph[0], s.inv_s[0]), iy = cell_index(y, 0.0f, s.ts_ph[1], s.inv_s[1]),
            iz = cell_index(z, 0.0f, s.ts_ph[2], s.inv_s[2]);
  if ((unsigned)ix < (unsigned)s.nx && (unsigned)iy < (unsigned)s.ny && (unsigned)iz < (unsigned)s.nz)
    return ix + s.nx * (iy + s.ny * iz);
  return -1;
}

// TLE: dose by the Woodcock collision estimator of kerma instead of analog deposits
// (the role of GateTLEDoseActor, GateTLEDoseActor.cpp:200-300): every tentative
// collision inside the phantom (real or virtual) of a photon at or below tle_emax scores
// w E mu_en(E, m)/mu_maj(E) in its voxel, whose expectation per unit path is the
// track-length estimate w E mu_en L; the energy such a photon gives to electrons at real
// interactions is not deposited. Photons at or below tle_emin (and below the tracking
// cut) deposit their energy where they are and stop. Single-simulation launches only.
// The batched form keeps a few more per-lane values (its simulation's grid) and the
// kerma-estimator form a few more live values in the step loop; the waves-per-EU bound
// holds both at 128 VGPRs (4 waves per SIMD).
template <bool BATCHED, int STAGE, int BLOCK, bool TLE>
__global__ void __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu((BATCHED || TLE) ? 4 : 1)))
transport_kernel(Launch L) {
  extern __shared__ float lds[];
  const Phys &P = L.P;
  const int NB1 = P.nb + 1;
  Tabs T;
  T.mu = P.mu; T.maj = P.maj; T.pcpr = P.pcpr; T.rx = P.rx; T.rcdf = P.rcdf; T.sinc = P.sinc;
  T.rguide = P.rguide;
  const MatRow *mat_tab = P.mat;
  const bool shared_rows = P.shared_rows != 0;  // uniform over the launch
  // mu by material index (rows per material, or shared rows expanded into LDS): no
  // (density, row) lookup in front of the table read
  const bool mu_by_mat = !shared_rows || (STAGE >= 1 && P.mu_expanded != 0);
  if (STAGE >= 1) {
    // staged by the threads of the launch (a launch may use fewer than BLOCK)
    const int nt = (int)blockDim.x;
    const bool expand = shared_rows && P.mu_expanded != 0;
    const LdsLayout o = lds_layout(P.n_rows, P.n_mat, NB1, P.nx, STAGE, expand);
    if (expand) {
      const int n_mu = P.n_mat * NB1;
      for (int i = threadIdx.x; i < n_mu; i += nt) {
        const int m = i / NB1, j = i - m * NB1;
        const MatRow mr = P.mat[m];
        lds[o.mu + i] = mr.rho * P.mu[mr.row * NB1 + j];
      }
    } else {
      const int n_mu = P.n_rows * NB1;
      for (int i = threadIdx.x; i < n_mu; i += nt) lds[o.mu + i] = P.mu[i];
    }
    for (int i = threadIdx.x; i < NB1; i += nt) lds[o.maj + i] = P.maj[i];
    for (int i = threadIdx.x; i < P.n_mat; i += nt) ((MatRow *)(lds + o.mat))[i] = P.mat[i];
    T.mu = lds + o.mu;
    T.maj = lds + o.maj;
    mat_tab = (const MatRow *)(lds + o.mat);
    if (STAGE >= 2) {
      const int n_x = P.n_rows * P.nx;
      for (int i = threadIdx.x; i < n_x; i += nt) {
        lds[o.sinc + i] = P.sinc[i];
        lds[o.rcdf + i] = P.rcdf[i];
      }
      for (int i = threadIdx.x; i < P.nx; i += nt) lds[o.rx + i] = P.rx[i];
      T.sinc = lds + o.sinc;
      T.rcdf = lds + o.rcdf;
      T.rx = lds + o.rx;
    }
  }
  const float *mu_tab = T.mu;
  const float *maj_tab = T.maj;
  __shared__ unsigned long long s_cnt[N_COUNTERS];
  // energy tallies of the statistics: one FP64 slot per lane and category, so that the
  // lanes of a wave never add to one LDS address at once (64 same-address FP64 atomics
  // serialise in the LDS unit and hold up the table reads of every wave of the CU)
  __shared__ double s_en[N_ENERGIES * BLOCK];
  // birth queues: wave w owns columns [64 w, 64 w + 64)
  __shared__ uint32_t s_q[QWORDS * BLOCK];
  // the logarithm's table (a per-lane index: LDS rather than constant memory)
  __shared__ float2 s_logtab[64];
  if (threadIdx.x < 64) s_logtab[threadIdx.x] = c_logtab[threadIdx.x];
  // the source spectrum of a single-simulation launch, read at every birth: in LDS when it
  // is short (three dependent global reads per birth otherwise)
  __shared__ float s_spec[BLOCK == 1024 ? 3 * SPEC_LDS_MAX : 1];
  __shared__ int32_t s_sguide[BLOCK == 1024 ? SPEC_GUIDE + 1 : 1];
  SpecTab spec0{L.sims[0].e_val, L.sims[0].e_width, L.sims[0].e_cdf, L.sims[0].e_guide};
  if (BLOCK == 1024 && !BATCHED) {
    const SimDev &s0 = L.sims[0];
    if ((s0.ene_type == GH_ENE_LINES || s0.ene_type == GH_ENE_HISTOGRAM) && s0.n_energy <= SPEC_LDS_MAX) {
      for (int i = threadIdx.x; i < s0.n_energy; i += (int)blockDim.x) {
        s_spec[i] = s0.e_val[i];
        s_spec[2 * SPEC_LDS_MAX + i] = s0.e_cdf[i];
        if (s0.e_width) s_spec[SPEC_LDS_MAX + i] = s0.e_width[i];
      }
      for (int i = threadIdx.x; i <= SPEC_GUIDE; i += (int)blockDim.x) s_sguide[i] = s0.e_guide[i];
      spec0 = SpecTab{s_spec, s_spec + SPEC_LDS_MAX, s_spec + 2 * SPEC_LDS_MAX, s_sguide};
    }
  }
  if (threadIdx.x < N_COUNTERS) s_cnt[threadIdx.x] = 0ull;
  for (int k = 0; k < N_ENERGIES; ++k) s_en[k * BLOCK + threadIdx.x] = 0.0;
  __syncthreads();
  LaneStats st;
  st.steps = 0u; st.rej = 0u; st.cnt = s_cnt; st.en = s_en + threadIdx.x; st.stride = BLOCK;
  const uint32_t qbase = threadIdx.x & ~63u;

  Hot hot;
  // descriptor of simulation i (always 0 in the single-simulation forms)
  auto sim = [&](int i) -> const SimDev & { return L.sims[BATCHED ? i : 0]; };
  if (!BATCHED) hot.load(sim(0));
  const float tle_emax = TLE ? sim(0).tle_emax : 0.0f;
  const float tle_emin = TLE ? sim(0).tle_emin : 0.0f;

  Photon ph;
  ph.mode = MODE_BOX; ph.sim = 0; ph.hist = 0;
  ph.x = ph.y = ph.z = ph.u = ph.v = ph.w = ph.E = ph.wt = 0.0f;
  ECache ec;
  ec.b = 0; ec.frac = 0.0f; ec.maj22 = 0x1p-22f; ec.inv_maj = 1.0f;
  Rng rng;
  // the key is the run's seed in every lane from the start, so the compiler sees one
  // wave-uniform key (scalar registers) instead of a per-lane value
  rng_init(rng, L.seed, 0, 0);
  uint32_t sblock = 0;           // step-stream block counter of the current history
  SqCache sq = {};
  float ev_edep = 0.0f;
  int mode = M_DEAD;
  int cvox = 0;                  // voxel of a pending collision (-1: in the world material)
  int clab = 0;                  // its material (kept from the step: no second global read)
  int ntry = 0;                  // rejected tries of a pending Compton / Rayleigh sampling

  // wave-uniform work state
  int cur_sim = 0;
  uint64_t cur_next = 0, cur_end = 0;
  bool have_work = true;
  uint32_t qn = 0;               // photons in this wave's birth queue
  const uint32_t slow_min = (uint32_t)L.slow_min;
  const uint32_t refill_min = (uint32_t)L.refill_min;
  const int scatter_tries = L.scatter_tries;

  // world transport outside the phantom box (homogeneous material or vacuum) of photon p
  // with energy cache e; run in the birth pass for photons born outside the box and in the
  // slow phase for photons that left it. Returns the new mode: M_STEP (entered the box),
  // M_COLLIDE (interaction in the world material, p at the interaction point) or M_DEAD
  // (escaped). The single-simulation forms pass the reciprocal direction; the batched
  // form, which has no registers to spare, recomputes it.
  auto world_step = [&](const SimDev &s, Photon &p, Rng &g, const ECache &e, const InvDir &iv0) -> int {
    auto iv = [&]() { return BATCHED ? inv_dir(p.u, p.v, p.w) : iv0; };
    float tn, tf;
    int face;
    ray_box(p.x, p.y, p.z, p.u, p.v, p.w, iv(), s.L[0], s.L[1], s.L[2], tn, tf, face);
    const bool hit = (tn < tf) && (tf > EXIT_EPS);
    const float tin = fmaxf(tn, 0.0f);
    if (s.world_mat < 0) {
      if (hit) {
        p.x = fmaf(tin, p.u, p.x); p.y = fmaf(tin, p.v, p.y); p.z = fmaf(tin, p.w, p.z);
        return M_STEP;
      }
      st.inc(C_ESCAPED);
      st.add_e(E_ESC, (double)(p.E * p.wt));
      return M_DEAD;
    }
    float muw;
    if (mu_by_mat) {
      muw = lerp_node(mu_tab + s.world_mat * NB1, e);
    } else {
      const MatRow mw = mat_tab[s.world_mat];
      muw = lerp_node_scaled(mu_tab + mw.row * NB1, mw.rho, e);
    }
    const float tw = world_exit(s, p.x, p.y, p.z, p.u, p.v, p.w, iv());
    const float d = (muw > 0.0f) ? (-glog(u01(g), s_logtab) / muw) : INFINITY;
    if (hit && d >= tin && tin < tw) {
      p.x = fmaf(tin, p.u, p.x); p.y = fmaf(tin, p.v, p.y); p.z = fmaf(tin, p.w, p.z);
      return M_STEP;
    }
    if (d >= tw) {
      st.inc(C_ESCAPED);
      st.add_e(E_ESC, (double)(p.E * p.wt));
      return M_DEAD;
    }
    p.x = fmaf(d, p.u, p.x); p.y = fmaf(d, p.v, p.y); p.z = fmaf(d, p.w, p.z);
    st.inc(C_WORLD);
    return M_COLLIDE;
  };

  // Birth pass: all 64 lanes sample one history each from the wave's chunk; photons that
  // still need transport go into the wave's queue (empty when the pass starts).
  auto birth_pass = [&]() {
    while (have_work && cur_next >= cur_end) {
      unsigned long long c = 0;
      if (lane_id() == 0) c = atomicAdd(L.next_chunk, 1ull);
      c = __shfl(c, 0, 64);
      if (c >= L.total_chunks) { have_work = false; break; }
      int s = 0;
      while (s + 1 < L.n_sims && L.chunk_offsets[s + 1] <= c) ++s;
      cur_sim = s;
      const uint64_t local = (c - L.chunk_offsets[s]) * (uint64_t)L.chunk;
      cur_next = local;
      const uint64_t e = local + L.chunk;
      const uint64_t nh = sim(s).n_hist;
      cur_end = e < nh ? e : nh;
    }
    if (!have_work) return;
    const uint64_t left = cur_end - cur_next;
    const uint32_t avail = (uint32_t)(left < 64u ? left : 64u);
    const uint32_t lane = lane_id();
    const SimDev &s = sim(cur_sim);
    int bm = M_DEAD;
    Photon b;
    b.x = b.y = b.z = b.u = b.v = b.w = b.E = b.wt = 0.0f;
    uint32_t bh = 0, bc3 = 0;
    if (lane < avail) {
      bh = (uint32_t)(cur_next + lane);
      Rng g;
      rng_init(g, L.seed, s.first_hist + bh, (uint32_t)(BATCHED ? cur_sim : 0));
      sample_source(s, BATCHED ? SpecTab{s.e_val, s.e_width, s.e_cdf, s.e_guide} : spec0, g, bh, b);
      st.inc(C_HIST);
      const int bvox = sim_voxel(s, b.x, b.y, b.z);
      if (b.E < P.e_cut || (TLE && b.E <= tle_emin)) {
        // below the tracking cut from the start: deposit locally (inside) and stop; the
        // history's only deposit is its (voxel, sum) entry
        float ev = 0.0f;
        if (bvox >= 0) {
          const int lab = sim_label(s, bvox);
          const float x = b.E * b.wt;
          const int tv = tally_voxel(s, b.x, b.y, b.z);
          if (lab >= s.det_label_min && s.det_label_min >= 0) ev = x;
          if (s.event_edep) s.event_edep[bh] = ev;
          if (tv >= 0) flush_slot(tally_out(s), tv, x, x * P.inv_rho[lab], 1u, st);
          else st.add_e(E_OUT, (double)x);
        } else {
          if (s.event_edep) s.event_edep[bh] = ev;
          st.add_e(E_WORLD, (double)(b.E * b.wt));
        }
        st.inc(C_CUT);
      } else if (bvox >= 0) {
        bm = M_STEP;
      } else {
        ECache be;
        ecache_set(be, b.E, maj_tab);
        bm = world_step(s, b, g, be, BATCHED ? InvDir{0.0f, 0.0f, 0.0f} : inv_dir(b.u, b.v, b.w));
        if (bm == M_DEAD && s.event_edep) s.event_edep[bh] = 0.0f;
      }
      bc3 = g.c3;  // the event stream goes on at the next unused block
    }
    cur_next += avail;
    const bool keep = bm != M_DEAD;
    const uint64_t m = __ballot(keep);
    if (keep) {
      const uint32_t col = qbase + qn + mbcnt(m);
      const float f[8] = {b.x, b.y, b.z, b.u, b.v, b.w, b.E, b.wt};
#pragma unroll
      for (int k = 0; k < 8; ++k) s_q[k * BLOCK + col] = __float_as_uint(f[k]);
      s_q[8 * BLOCK + col] = bh;
      s_q[9 * BLOCK + col] = (bc3 & 0xFFFFu) | (bm == M_COLLIDE ? META_COLLIDE : 0u) |
                             ((uint32_t)(BATCHED ? cur_sim : 0) << META_SIM_SHIFT);
    }
    qn += (uint32_t)__popcll(m);
    __builtin_amdgcn_wave_barrier();
  };

  for (;;) {
    // ------------------------------------------------------------- refill
    {
      uint64_t dead = __ballot(mode == M_DEAD);
      const uint32_t n_step = (uint32_t)__popcll(__ballot(mode == M_STEP));
      if (dead != 0ull && (qn > 0u || have_work) &&
          ((uint32_t)__popcll(dead) >= refill_min || n_step == 0u)) {
        while (dead != 0ull && (qn > 0u || have_work)) {
          if (qn == 0u) {
            birth_pass();
            continue;
          }
          const uint32_t nd = (uint32_t)__popcll(dead);
          const uint32_t take = qn < nd ? qn : nd;
          if (mode == M_DEAD) {
            const uint32_t rank = mbcnt(dead);
            if (rank < take) {
              const uint32_t col = qbase + qn - 1u - rank;
              ph.x = __uint_as_float(s_q[0 * BLOCK + col]); ph.y = __uint_as_float(s_q[1 * BLOCK + col]);
              ph.z = __uint_as_float(s_q[2 * BLOCK + col]); ph.u = __uint_as_float(s_q[3 * BLOCK + col]);
              ph.v = __uint_as_float(s_q[4 * BLOCK + col]); ph.w = __uint_as_float(s_q[5 * BLOCK + col]);
              ph.E = __uint_as_float(s_q[6 * BLOCK + col]); ph.wt = __uint_as_float(s_q[7 * BLOCK + col]);
              ph.hist = s_q[8 * BLOCK + col];
              const uint32_t meta = s_q[9 * BLOCK + col];
              ph.sim = BATCHED ? (int)(meta >> META_SIM_SHIFT) : 0;
              const SimDev &s = sim(ph.sim);
              if (BATCHED) hot.load(s);
              rng_init(rng, L.seed, s.first_hist + ph.hist, (uint32_t)ph.sim);
              rng.c3 = meta & 0xFFFFu;
              sblock = 0u;
              ev_edep = 0.0f;
              sq.n = 0;
              ecache_set(ec, ph.E, maj_tab);
              if (meta & META_COLLIDE) {
                mode = M_COLLIDE;
                cvox = -1;
              } else {
                mode = M_STEP;
              }
            }
          }
          qn -= take;
          dead = __ballot(mode == M_DEAD);
        }
      }
      if (!have_work && qn == 0u && __ballot(mode != M_DEAD) == 0ull) break;
    }

    // ------------------------------------------------------------- step phase
    for (int it = 0; it < L.max_iter; ++it) {
      const uint32_t n_step = (uint32_t)__popcll(__ballot(mode == M_STEP));
      const uint32_t n_slow = (uint32_t)__popcll(
          __ballot(mode >= M_COLLIDE || ((have_work || qn > 0u) && mode == M_DEAD)));
      if (n_step == 0u || n_slow >= slow_min) break;
      if (mode == M_STEP) {
        // the event stream's counter words hold the global history id and the simulation
        uint32_t w0 = rng.c0, w1 = rng.c1, w2 = rng.c2, w3 = 0x80000000u | sblock;
        philox10(w0, w1, w2, w3, rng.k0, rng.k1);
        sblock++;
        auto woodcock = [&](auto k_c) {
          if (mode == M_STEP) {
            uint32_t bd, ba;
            step_fields<decltype(k_c)::value>(w0, w1, w2, w3, bd, ba);
            const float d = -glog(odd22(bd) * 0x1p-22f, s_logtab) * ec.inv_maj;
            const float nx = fmaf(d, ph.u, ph.x), ny = fmaf(d, ph.v, ph.y), nz = fmaf(d, ph.w, ph.z);
            const int ix = floor_int(nx * hot.isx), iy = floor_int(ny * hot.isy), iz = floor_int(nz * hot.isz);
            if ((unsigned)ix < (unsigned)hot.nx && (unsigned)iy < (unsigned)hot.ny &&
                (unsigned)iz < (unsigned)hot.nz) {
              ph.x = nx; ph.y = ny; ph.z = nz;
              st.steps++;
              const int vox = ix + hot.nx * (iy + hot.ny * iz);
              const int lab = hot.label(vox);
              // table rows per material (or expanded in LDS): one read; shared composition
              // rows: the material's (density, row) pair first
              float mu;
              if (mu_by_mat) {
                mu = lerp_node(mu_tab + lab * NB1, ec);
              } else {
                const MatRow mr = mat_tab[lab];
                mu = lerp_node_scaled(mu_tab + mr.row * NB1, mr.rho, ec);
              }
              if (TLE && ph.E <= tle_emax) {
                const SimDev &s0 = sim(0);
                const MatRow mr = shared_rows ? mat_tab[lab] : MatRow{1.0f, lab};
                const float x = (ph.E * ph.wt) * (lerp_node_scaled(P.muen + mr.row * NB1, mr.rho, ec) * ec.inv_maj);
                deposit(s0, sq, tally_voxel(s0, ph.x, ph.y, ph.z), x, x * P.inv_rho[lab], st);
              }
              // u01 x majorant < mu, with u01 = odd22(ba) 2^-22 and the 2^-22 folded into
              // maj22: the same single rounding of the same exact product
              if (odd22(ba) * ec.maj22 < mu) {
                mode = M_COLLIDE;
                cvox = vox;
                clab = lab;
              }
            } else {
              mode = M_EXIT;
            }
          }
        };
        woodcock(std::integral_constant<int, 0>{});
        woodcock(std::integral_constant<int, 1>{});
        woodcock(std::integral_constant<int, 2>{});
      }
    }

    // ------------------------------------------------------------- slow phase
    // order: box exit -> world event -> collision (branching) -> scattering samplers. A
    // world interaction becomes a collision with the world material and no tally voxel
    // (cvox = -1), so box and world collisions share one copy of the interaction code.
    if (mode >= M_COLLIDE) {
      const SimDev &s = sim(BATCHED ? (int)rng.c2 : 0);
      bool kill = false;
      // one set of divisions for the box exit and the world transport that follows it
      InvDir iv = {0.0f, 0.0f, 0.0f};
      if (!BATCHED && (mode == M_EXIT || mode == M_WORLD)) iv = inv_dir(ph.u, ph.v, ph.w);
      if (mode == M_EXIT) {
        // leaves the phantom box: exit point from the last position inside
        float tn, tf;
        int face;
        ray_box(ph.x, ph.y, ph.z, ph.u, ph.v, ph.w, BATCHED ? inv_dir(ph.u, ph.v, ph.w) : iv, s.L[0], s.L[1],
                s.L[2], tn, tf, face);
        const float t = fmaxf(tf, 0.0f);
        const float ex = fmaf(t, ph.u, ph.x), ey = fmaf(t, ph.v, ph.y), ez = fmaf(t, ph.w, ph.z);
        if (s.rec_face >= 0 && (face == s.rec_face || s.rec_face == GH_FACE_ALL)) {
          // one slot reservation per simulation among the recording lanes: around a chunk
          // boundary the lanes of a batched wave hold photons of two simulations, each with
          // its own record buffer and counter (one pass in the single-simulation forms)
          const int my_sim = BATCHED ? (int)rng.c2 : 0;
          uint64_t pending = __ballot(1);
          while (pending != 0ull) {
            const uint32_t leader = __ffsll((unsigned long long)pending) - 1;
            const int lsim = BATCHED ? __shfl(my_sim, leader, 64) : 0;
            const uint64_t m = __ballot(my_sim == lsim) & pending;
            unsigned long long base = 0;
            if (lane_id() == leader) base = atomicAdd(s.rec_count, (unsigned long long)__popcll(m));
            base = __shfl(base, leader, 64);
            if (my_sim == lsim) {
              const uint64_t slot = base + mbcnt(m);
              if (slot < s.rec_cap) {
                s.rec[0][slot] = ex; s.rec[1][slot] = ey; s.rec[2][slot] = ez;
                s.rec[3][slot] = ph.u; s.rec[4][slot] = ph.v; s.rec[5][slot] = ph.w;
                s.rec[6][slot] = ph.E; s.rec[7][slot] = ph.wt;
              }
            }
            pending &= ~m;
          }
          st.inc(C_RECORDED);
          st.add_e(E_ESC, (double)(ph.E * ph.wt));
          kill = true;
        } else if (s.world_mat < 0) {
          st.inc(C_ESCAPED);
          st.add_e(E_ESC, (double)(ph.E * ph.wt));
          kill = true;
        } else {
          ph.x = ex; ph.y = ey; ph.z = ez;
          mode = M_WORLD;
        }
      }
      if (mode == M_WORLD && !kill) {
        mode = world_step(s, ph, rng, ec, iv);
        if (mode == M_DEAD) kill = true;
        else if (mode == M_COLLIDE) cvox = -1;
      }
      // interaction: branching, then the process (photoelectric at once, the scattering
      // samplers in bounded tries); `fin` marks a finished interaction this pass
      bool fin = false;
      bool comp = false;
      float x_dep = 0.0f;   // energy given to the medium at this interaction
      const bool in_world = cvox < 0;
      if (mode == M_COLLIDE && !kill) clab = in_world ? s.world_mat : clab;
      const int crow = shared_rows ? mat_tab[clab].row : clab;  // table row of the interaction's material
      if (mode == M_COLLIDE && !kill) {
        const float r = u01(rng);
        float pc, pr;
        branching(T, NB1, crow, ec, pc, pr);
        ntry = 0;
        if (r < pc) {
          if (!in_world) st.inc(C_COMPTON);
          mode = M_COMPTON;
        } else if (r < pr) {
          if (!in_world) st.inc(C_RAYLEIGH);
          mode = M_RAYLEIGH;
        } else {
          if (!in_world) st.inc(C_PHOTO);
          x_dep = (TLE && !in_world && ph.E <= tle_emax) ? 0.0f : ph.E * ph.wt;
          kill = true;
          fin = true;
        }
      }
      // polar cosine of an accepted Compton or Rayleigh scatter; one azimuth-and-rotation
      // block serves both (a wave with lanes of each kind runs it once)
      bool scat = false;
      float c_scat = 1.0f;
      if (mode == M_COMPTON && !fin) {
        float q = 1.0f, c = 1.0f;
        bool done = false;
        for (int t = 0; t < scatter_tries && !done; ++t) {
          if (compton_try(rng, T, P.nx, crow, ph.E, q, c)) {
            done = true;
          } else {
            st.rej++;
            if (++ntry == MAX_REJECT) { q = 1.0f; c = 1.0f; done = true; }
          }
        }
        if (done) {
          const float e1 = ph.E / q;
          // electrons of photons scored by the kerma estimator give nothing locally
          x_dep = (TLE && !in_world && ph.E <= tle_emax) ? 0.0f : (ph.E - e1) * ph.wt;
          ph.E = e1;
          c_scat = fmaxf(c, -1.0f);
          scat = true;
          fin = true;
          comp = true;
        }
      }
      if (mode == M_RAYLEIGH && !fin) {
        float c = 1.0f;
        bool done = false;
        for (int t = 0; t < scatter_tries && !done; ++t) {
          if (rayleigh_try(rng, T, P.nx, crow, ph.E, c) || ++ntry == MAX_REJECT) done = true;
        }
        if (done) {
          c_scat = c;
          scat = true;
          fin = true;
        }
      }
      if (scat) {
        float cp, sp;
        azimuth(rng, cp, sp);
        rotate(ph.u, ph.v, ph.w, c_scat, cp, sp);
      }
      if (fin) {
        if (!kill) mode = in_world ? M_WORLD : M_STEP;
        // local deposition: electron energy, and the photon itself below the tracking cut
        float x_cut = 0.0f;
        if (!kill && (ph.E < P.e_cut || (TLE && !in_world && ph.E <= tle_emin))) {
          x_cut = ph.E * ph.wt;
          st.inc(C_CUT);
          kill = true;
        }
        if (in_world) {
          if (x_dep != 0.0f) st.add_e(E_WORLD, (double)x_dep);
          if (x_cut != 0.0f) st.add_e(E_WORLD, (double)x_cut);
        } else {
          const int tv = tally_voxel(s, ph.x, ph.y, ph.z);
          const float irho = P.inv_rho[clab];
          const bool det = s.det_label_min >= 0 && clab >= s.det_label_min;
          if (x_dep != 0.0f) {
            deposit(s, sq, tv, x_dep, x_dep * irho, st);
            if (det) ev_edep += x_dep;
          }
          if (x_cut != 0.0f) {
            deposit(s, sq, tv, x_cut, x_cut * irho, st);
            if (det) ev_edep += x_cut;
          }
        }
        if (!kill && comp) ecache_set(ec, ph.E, maj_tab);
      }
      if (kill) {
        // the event record goes first: a read of the descriptor placed after the flush's
        // atomics would wait for them
        if (s.event_edep) s.event_edep[history_of(rng) - s.first_hist] = ev_edep;
        sq_flush_all(s, sq, st);
        mode = M_DEAD;
      }
    }
  }

  // ------------------------------------------------------------- counters
  {
    const uint64_t steps = wave_sum<uint64_t>((uint64_t)st.steps);
    const uint64_t rej = wave_sum<uint64_t>((uint64_t)st.rej);
    if (lane_id() == 0) {
      atomicAdd(s_cnt + C_STEPS, (unsigned long long)steps);
      atomicAdd(s_cnt + C_REJ, (unsigned long long)rej);
    }
  }
  __syncthreads();
  if (threadIdx.x < N_COUNTERS && s_cnt[threadIdx.x])
    atomicAdd(L.counters + threadIdx.x, s_cnt[threadIdx.x]);
  for (int k = 0; k < N_ENERGIES; ++k) {
    const double v = wave_sum<double>(s_en[k * BLOCK + threadIdx.x]);
    if (lane_id() == 0 && v != 0.0) unsafeAtomicAdd(L.energies + k, v);
  }
}

// ---------------------------------------------------------------------- phsp check
// flags[0] |= 1 if an energy is outside (0, e_hi) or not finite, |= 2 for a bad weight
__global__ void phsp_check_kernel(const float *__restrict__ E, const float *__restrict__ w, int64_t n, float e_hi,
                                  unsigned int *flags) {
  unsigned int bad = 0u;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float e = E[i];
    if (!(e > 0.0f) || !(e < e_hi)) bad |= 1u;
    if (!(w[i] >= 0.0f) || !isfinite(w[i])) bad |= 2u;
  }
  if (bad) atomicOr(flags, bad);
}

// ---------------------------------------------------------------------- HU -> labels
// label = lut[#{k : bins[k] <= hu}] (numpy.digitize, right=False), bins and lut in LDS
// (dynamic: n_bins doubles, then n_bins + 1 uint16); HU as int16 or as FP64 (any image
// type whose values FP64 holds exactly: integers, float32, float64)
constexpr int HU_MAX_BINS = 4096;
template <typename T>
__global__ void hu_to_labels_kernel(const T *__restrict__ hu, int64_t n, const double *__restrict__ bins, int n_bins,
                                    const uint16_t *__restrict__ lut, uint16_t *__restrict__ out) {
  extern __shared__ double hu_lds[];
  double *sb = hu_lds;
  uint16_t *sl = (uint16_t *)(hu_lds + n_bins);
  for (int i = threadIdx.x; i < n_bins; i += blockDim.x) sb[i] = bins[i];
  for (int i = threadIdx.x; i <= n_bins; i += blockDim.x) sl[i] = lut[i];
  __syncthreads();
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double x = (double)hu[i];
    // number of bins <= x (digitize, right=False); NaN compares false everywhere
    int lo = 0, hi = n_bins;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (sb[mid] <= x) lo = mid + 1; else hi = mid;
    }
    out[i] = sl[lo];
  }
}

// ---------------------------------------------------------------------- host side
#define HIPCHK(expr)                                                          \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);           \
      return GH_EHIP;                                                         \
    }                                                                         \
  } while (0)

#define FAIL(code, msg)   \
  do {                    \
    ctx->err = (msg);     \
    return (code);        \
  } while (0)

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
};

struct SimHost {
  SimDev dev;
  std::vector<void *> owned;         // device allocations of this simulation
  std::vector<void *> source_owned;  // device allocations of its current source
  size_t n_tally = 0;
  int flags = 0;
  int det = -1;
  uint64_t event_cap = 0;
  uint64_t event_n = 0;  // histories of the last run (valid per-history detector energies)
};

}  // namespace

struct gh_ctx {
  int device = 0;
  bool have_muen = false;
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int n_cu = 0, lds_max = 0;
  std::string name;
  // physics
  bool have_phys = false;
  Phys P{};
  std::vector<uint8_t> under_majorant;  // per material: mu <= majorant at every node
  std::vector<void *> phys_owned;
  size_t lds1_bytes = 0, lds2_bytes = 0;    // staged tables (mu by table row)
  size_t lds1e_bytes = 0, lds2e_bytes = 0;  // the same with mu expanded per material
  // simulations
  std::vector<SimHost> sims;
  // run scratch
  SimDev *d_sims = nullptr;
  uint64_t *d_chunk_off = nullptr;
  unsigned long long *d_next = nullptr;
  unsigned long long *d_counters = nullptr;
  double *d_energies = nullptr;
  size_t d_sims_cap = 0;
  gh_launch last_launch{};
  int last_stage = 0;
};

namespace {

template <typename T>
int dev_upload(gh_ctx *ctx, const T *host, size_t n, T **out, std::vector<void *> &owned) {
  *out = nullptr;
  if (n == 0) return GH_OK;
  void *p = nullptr;
  HIPCHK(hipMalloc(&p, n * sizeof(T)));
  owned.push_back(p);
  if (host) HIPCHK(hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice));
  *out = (T *)p;
  return GH_OK;
}

void free_all(std::vector<void *> &v) {
  for (void *p : v) (void)hipFree(p);
  v.clear();
}

}  // namespace

namespace {

int validate_source(gh_ctx *ctx, const gh_source_desc *src) {
  if (src->pos_type < GH_POS_POINT || src->pos_type > GH_POS_SURFACE_SPHERE) FAIL(GH_EINVAL, "invalid pos_type");
  if (!(src->pos_sigma[0] >= 0.0f) || !(src->pos_sigma[1] >= 0.0f) || !std::isfinite(src->pos_sigma[0]) ||
      !std::isfinite(src->pos_sigma[1]))
    FAIL(GH_EINVAL, "pos_sigma must be finite and >= 0");
  if ((src->pos_sigma[0] > 0.0f || src->pos_sigma[1] > 0.0f) && src->pos_type != GH_POS_DISC)
    FAIL(GH_EINVAL, "pos_sigma applies to disc positions only");
  if (!std::isfinite(src->weight_sigma)) FAIL(GH_EINVAL, "weight_sigma must be finite");
  if (src->pos_type != GH_POS_VOXELS && (src->vox_n != 0 || src->vox_prob || src->vox_alias || src->vox_index))
    FAIL(GH_EINVAL, "voxel-source fields are set but pos_type is not GH_POS_VOXELS");
  if (src->pos_type == GH_POS_VOXELS) {
    if (src->vox_n <= 0 || src->vox_n > 0x7FFFFFFF || !src->vox_prob || !src->vox_alias || !src->vox_index)
      FAIL(GH_EINVAL, "voxel source needs a non-empty alias table");
    const int64_t nvox = (int64_t)src->vox_size[0] * src->vox_size[1] * src->vox_size[2];
    if (src->vox_size[0] <= 0 || src->vox_size[1] <= 0 || src->vox_size[2] <= 0 || nvox > 0x7FFFFFFF)
      FAIL(GH_EINVAL, "voxel source: invalid image size");
    for (int k = 0; k < 3; ++k)
      if (!(src->vox_spacing[k] > 0.0f)) FAIL(GH_EINVAL, "voxel source: invalid spacing");
    for (int64_t i = 0; i < src->vox_n; ++i) {
      if (src->vox_index[i] < 0 || src->vox_index[i] >= nvox) FAIL(GH_EINVAL, "voxel source: voxel index out of range");
      if (src->vox_alias[i] < 0 || src->vox_alias[i] >= src->vox_n) FAIL(GH_EINVAL, "voxel source: alias out of range");
      if (!(src->vox_prob[i] >= 0.0f && src->vox_prob[i] <= 1.0f)) FAIL(GH_EINVAL, "voxel source: probability outside [0, 1]");
    }
  }
  if (src->dir_type < GH_DIR_ISO || src->dir_type > GH_DIR_HISTOGRAM) FAIL(GH_EINVAL, "invalid dir_type");
  if (src->dir_type == GH_DIR_BEAM2D)
    for (int k = 0; k < 2; ++k)
      // |polar angle| <= 6 sqrt(2) x 5.8 (the largest Box-Muller normal) < 16 pi, the sine's range
      if (!(src->dir_sigma[k] >= 0.0f && src->dir_sigma[k] <= 6.0f)) FAIL(GH_EINVAL, "dir_sigma must lie in [0, 6] rad");
  if (src->pos_type == GH_POS_SURFACE_SPHERE && (src->dir_type == GH_DIR_ISO || src->dir_type == GH_DIR_COS)) {
    const float *a = src->surface_axis;
    if (!std::isfinite(a[0]) || !std::isfinite(a[1]) || !std::isfinite(a[2]) ||
        (a[0] == 0.0f && a[1] == 0.0f && a[2] == 0.0f))
      FAIL(GH_EINVAL, "surface_axis must be finite and nonzero");
  }
  if (src->dir_type == GH_DIR_HISTOGRAM) {
    // piecewise-uniform angle tables: widths >= 0, cdf nondecreasing with last entry 1,
    // angles within the sine's range
    auto table_ok = [](int n, const float *val, const float *wid, const float *cdf) {
      if (n <= 0 || n > (1 << 20) || !val || !wid || !cdf) return false;
      float prev = 0.0f;
      for (int i = 0; i < n; ++i) {
        const float lo = val[i], hi = val[i] + wid[i];
        if (!(wid[i] >= 0.0f) || !(std::fabs(lo) <= 50.0f) || !(std::fabs(hi) <= 50.0f)) return false;
        if (!(cdf[i] >= prev) || !(cdf[i] <= 1.0f)) return false;
        prev = cdf[i];
      }
      return cdf[n - 1] == 1.0f;
    };
    if (!table_ok(src->n_theta, src->theta_value, src->theta_width, src->theta_cdf))
      FAIL(GH_EINVAL, "histogram direction: invalid theta table");
    if (!table_ok(src->n_phi, src->phi_value, src->phi_width, src->phi_cdf))
      FAIL(GH_EINVAL, "histogram direction: invalid phi table");
  }
  if (src->ene_type < GH_ENE_MONO || src->ene_type > GH_ENE_GAUSS) FAIL(GH_EINVAL, "invalid ene_type");
  if (src->ene_type == GH_ENE_GAUSS && (!(src->energy_sigma >= 0.0f) || !std::isfinite(src->energy_sigma)))
    FAIL(GH_EINVAL, "energy_sigma must be finite and >= 0");
  if ((src->ene_type == GH_ENE_LINES || src->ene_type == GH_ENE_HISTOGRAM) &&
      (src->n_energy <= 0 || !src->energy_value || !src->energy_cdf ||
       (src->ene_type == GH_ENE_HISTOGRAM && !src->energy_width)))
    FAIL(GH_EINVAL, "spectrum arrays missing");
  if (src->pos_type == GH_POS_PHSP && (src->phsp_n <= 0 || src->phsp_start < 0))
    FAIL(GH_EINVAL, "phase-space source needs rows");
  if (src->pos_type == GH_POS_PHSP)
    for (int k = 0; k < 8; ++k)
      if (!src->phsp[k]) FAIL(GH_EINVAL, "phase-space array missing");
  {
    // every source energy must lie on the energy grid: (0, E_hi)
    uint32_t hb = (uint32_t)(BUCKET_BASE + ctx->P.nb) << ESHIFT;
    float e_hi;
    std::memcpy(&e_hi, &hb, 4);
    auto bad = [&](float e) { return !(e > 0.0f) || !(e < e_hi); };
    bool out_of_grid = false;
    if (src->pos_type == GH_POS_PHSP && src->phsp_on_device) {
      // device arrays: checked by a kernel
      unsigned int *flags = nullptr;
      HIPCHK(hipMalloc(&flags, sizeof(unsigned int)));
      unsigned int h = 0;
      hipError_t e = hipMemset(flags, 0, sizeof(unsigned int));
      if (e == hipSuccess) {
        hipLaunchKernelGGL(phsp_check_kernel, dim3(std::max(1, ctx->n_cu) * 4), dim3(256), 0, 0, src->phsp[6],
                           src->phsp[7], src->phsp_n, e_hi, flags);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipMemcpy(&h, flags, sizeof(h), hipMemcpyDeviceToHost);
      (void)hipFree(flags);
      if (e != hipSuccess) FAIL(GH_EHIP, std::string("gh_add_simulation: phase-space check: ") + hipGetErrorString(e));
      if (h & 2u) FAIL(GH_EINVAL, "gh_add_simulation: phase-space weights must be finite and >= 0");
      out_of_grid = (h & 1u) != 0;
    } else if (src->pos_type == GH_POS_PHSP) {
      for (int64_t i = 0; i < src->phsp_n; ++i) {
        const float w = src->phsp[7][i];
        if (!(w >= 0.0f) || !std::isfinite(w)) FAIL(GH_EINVAL, "gh_add_simulation: phase-space weights must be finite and >= 0");
        out_of_grid = out_of_grid || bad(src->phsp[6][i]);
      }
    } else if (src->ene_type == GH_ENE_MONO) {
      out_of_grid = bad(src->energy_mono);
    } else if (src->ene_type == GH_ENE_RANGE) {
      out_of_grid = bad(src->energy_mono) || bad(src->energy_max) || src->energy_max < src->energy_mono;
    } else if (src->ene_type == GH_ENE_GAUSS) {
      // the sampler cuts the tail at mean + 8 sigma: that energy must be on the grid
      out_of_grid = bad(src->energy_mono) || bad(src->energy_mono + 8.0f * src->energy_sigma);
    } else {
      for (int i = 0; i < src->n_energy && !out_of_grid; ++i) {
        const float top = src->ene_type == GH_ENE_HISTOGRAM ? src->energy_value[i] + src->energy_width[i]
                                                            : src->energy_value[i];
        out_of_grid = bad(src->energy_value[i]) || bad(top);
        if (i > 0 && src->energy_cdf[i] < src->energy_cdf[i - 1]) out_of_grid = true;
      }
      if (!out_of_grid && !(src->energy_cdf[src->n_energy - 1] == 1.0f)) {
        ctx->err = "gh_add_simulation: energy_cdf must end at exactly 1";
        return GH_EINVAL;
      }
    }
    if (out_of_grid) FAIL(GH_ERANGE, "gh_add_simulation: source energy outside (0, top of the physics tables)");
  }
  return GH_OK;
}

int upload_source(gh_ctx *ctx, const gh_source_desc *src, SimDev &d, std::vector<void *> &owned) {
  int rc;
  d.e_val = d.e_width = d.e_cdf = nullptr;
  d.e_guide = nullptr;
  d.phsp_n = 0;
  d.phsp_start = 0;
  for (int k = 0; k < 8; ++k) d.phsp[k] = nullptr;
  d.pos_type = src->pos_type; d.dir_type = src->dir_type; d.ene_type = src->ene_type;
  d.n_energy = src->n_energy;
  std::memcpy(d.pos_param, src->pos_param, sizeof(d.pos_param));
  std::memcpy(d.pos_mat, src->pos_mat, sizeof(d.pos_mat));
  std::memcpy(d.dir_mat, src->dir_mat, sizeof(d.dir_mat));
  {
    // cosine law and tangent axis: FP64, rounded once
    const double c0 = src->cos_theta_min, c1 = src->cos_theta_max;
    const double s0 = 1.0 - c0 * c0, s1 = 1.0 - c1 * c1;
    d.s2_min = (float)s0;
    d.s2_d = (float)(s1 - s0);
    const double a0 = src->surface_axis[0], a1 = src->surface_axis[1], a2 = src->surface_axis[2];
    const double an = std::sqrt(a0 * a0 + a1 * a1 + a2 * a2);
    d.axis[0] = an > 0.0 ? (float)(a0 / an) : 0.0f;
    d.axis[1] = an > 0.0 ? (float)(a1 / an) : 0.0f;
    d.axis[2] = an > 0.0 ? (float)(a2 / an) : 1.0f;
    d.dir_sigma[0] = src->dir_sigma[0];
    d.dir_sigma[1] = src->dir_sigma[1];
  }
  d.ct_min = src->cos_theta_min;
  d.d_ct = src->cos_theta_min - src->cos_theta_max;
  d.phi_min = src->phi_min;
  d.d_phi = src->phi_max - src->phi_min;
  d.phi_full = src->phi_full;
  d.phsp_iso = src->phsp_iso;
  std::memcpy(d.mom, src->momentum, sizeof(d.mom));
  std::memcpy(d.focus, src->focus, sizeof(d.focus));
  d.e_mono = src->energy_mono; d.e_max = src->energy_max; d.weight = src->weight;
  if (src->ene_type == GH_ENE_GAUSS) d.e_max = src->energy_mono + 8.0f * src->energy_sigma;
  d.pos_sigma[0] = src->pos_sigma[0]; d.pos_sigma[1] = src->pos_sigma[1];
  d.e_sigma = src->energy_sigma;
  d.w_sigma = src->weight_sigma;
  float *t;
  if (src->ene_type == GH_ENE_LINES || src->ene_type == GH_ENE_HISTOGRAM) {
    if ((rc = dev_upload(ctx, src->energy_value, src->n_energy, &t, owned))) return rc;
    d.e_val = t;
    if ((rc = dev_upload(ctx, src->energy_cdf, src->n_energy, &t, owned))) return rc;
    d.e_cdf = t;
    // guide[b] = smallest k with cdf[k] > b / SPEC_GUIDE (n - 1 at the top edge)
    std::vector<int32_t> guide(SPEC_GUIDE + 1);
    for (int b = 0; b <= SPEC_GUIDE; ++b) {
      const float edge = (float)b / (float)SPEC_GUIDE;
      int k = 0;
      while (k < src->n_energy - 1 && !(src->energy_cdf[k] > edge)) ++k;
      guide[b] = k;
    }
    int32_t *tg;
    if ((rc = dev_upload(ctx, (const int32_t *)guide.data(), guide.size(), &tg, owned))) return rc;
    d.e_guide = tg;
    if (src->ene_type == GH_ENE_HISTOGRAM) {
      if ((rc = dev_upload(ctx, src->energy_width, src->n_energy, &t, owned))) return rc;
      d.e_width = t;
    }
  }
  d.n_theta = d.n_phi = 0;
  d.th_val = d.th_width = d.th_cdf = d.ph_val = d.ph_width = d.ph_cdf = nullptr;
  if (src->dir_type == GH_DIR_HISTOGRAM) {
    d.n_theta = src->n_theta;
    d.n_phi = src->n_phi;
    if ((rc = dev_upload(ctx, src->theta_value, src->n_theta, &t, owned))) return rc;
    d.th_val = t;
    if ((rc = dev_upload(ctx, src->theta_width, src->n_theta, &t, owned))) return rc;
    d.th_width = t;
    if ((rc = dev_upload(ctx, src->theta_cdf, src->n_theta, &t, owned))) return rc;
    d.th_cdf = t;
    if ((rc = dev_upload(ctx, src->phi_value, src->n_phi, &t, owned))) return rc;
    d.ph_val = t;
    if ((rc = dev_upload(ctx, src->phi_width, src->n_phi, &t, owned))) return rc;
    d.ph_width = t;
    if ((rc = dev_upload(ctx, src->phi_cdf, src->n_phi, &t, owned))) return rc;
    d.ph_cdf = t;
  }
  d.vox_n = 0; d.vox_prob = nullptr; d.vox_alias = nullptr; d.vox_index = nullptr;
  if (src->pos_type == GH_POS_VOXELS) {
    d.vox_n = src->vox_n;
    d.vox_nx = src->vox_size[0]; d.vox_ny = src->vox_size[1];
    for (int k = 0; k < 3; ++k) d.vox_s[k] = src->vox_spacing[k];
    if ((rc = dev_upload(ctx, src->vox_prob, (size_t)src->vox_n, &t, owned))) return rc;
    d.vox_prob = t;
    int32_t *ti;
    if ((rc = dev_upload(ctx, src->vox_alias, (size_t)src->vox_n, &ti, owned))) return rc;
    d.vox_alias = ti;
    if ((rc = dev_upload(ctx, src->vox_index, (size_t)src->vox_n, &ti, owned))) return rc;
    d.vox_index = ti;
  }
  if (src->pos_type == GH_POS_PHSP) {
    d.phsp_n = src->phsp_n; d.phsp_start = src->phsp_start;
    for (int k = 0; k < 8; ++k) {
      if (src->phsp_on_device) {
        d.phsp[k] = src->phsp[k];  // caller-owned device memory, used in place
      } else {
        if ((rc = dev_upload(ctx, src->phsp[k], (size_t)src->phsp_n, &t, owned))) return rc;
        d.phsp[k] = t;
      }
    }
  }
  return GH_OK;
}
}  // namespace

extern "C" {

int gh_abi_version(void) { return GH_ABI_VERSION; }

int gh_create(int device, gh_ctx **out) {
  if (!out) return GH_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return GH_EHIP;
  if (device < 0 || device >= n) return GH_EINVAL;
  gh_ctx *ctx = new gh_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
    delete ctx;
    return GH_EHIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    ctx->n_cu = prop.multiProcessorCount;
    ctx->lds_max = (int)prop.sharedMemPerBlock;
    ctx->name = prop.gcnArchName;
  }
  *out = ctx;
  return GH_OK;
}

void gh_destroy(gh_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &s : ctx->sims) { free_all(s.owned); free_all(s.source_owned); }
  free_all(ctx->phys_owned);
  if (ctx->d_sims) (void)hipFree(ctx->d_sims);
  if (ctx->d_chunk_off) (void)hipFree(ctx->d_chunk_off);
  if (ctx->d_next) (void)hipFree(ctx->d_next);
  if (ctx->d_counters) (void)hipFree(ctx->d_counters);
  if (ctx->d_energies) (void)hipFree(ctx->d_energies);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

// message of the last failed call of a context-free entry point of this host thread (mlp.hip)
extern "C" __attribute__((visibility("hidden"))) const char *gh_free_call_error();

const char *gh_last_error(const gh_ctx *ctx) { return ctx ? ctx->err.c_str() : gh_free_call_error(); }

int gh_device_info(gh_ctx *ctx, int32_t *n_cu, int32_t *lds_bytes, char *name, int32_t name_len) {
  if (!ctx) return GH_EINVAL;
  if (n_cu) *n_cu = ctx->n_cu;
  if (lds_bytes) *lds_bytes = ctx->lds_max;
  if (name && name_len > 0) {
    std::strncpy(name, ctx->name.c_str(), (size_t)name_len - 1);
    name[name_len - 1] = 0;
  }
  return GH_OK;
}

int gh_set_physics(gh_ctx *ctx, const gh_physics_desc *ph) {
  if (!ctx) return GH_EINVAL;
  if (!ph || ph->n_materials <= 0 || ph->n_materials > 65535 || ph->n_buckets <= 0 || ph->n_x_nodes < 2 ||
      ph->n_compositions < 0 || ph->n_compositions > 65535)
    FAIL(GH_EINVAL, "gh_set_physics: invalid sizes (1..65535 materials)");
  if (!ph->density || !ph->mu_total || !ph->mu_majorant || !ph->p_compton || !ph->p_rayleigh ||
      !ph->ray_x || !ph->ray_cdf || !ph->incoh_s)
    FAIL(GH_EINVAL, "gh_set_physics: missing table");
  if (ph->n_compositions > 0 && !ph->composition) FAIL(GH_EINVAL, "gh_set_physics: composition array missing");
  if (ph->n_buckets > 64 * 24) FAIL(GH_ERANGE, "gh_set_physics: energy grid beyond 2^14 MeV");
  if (!(ph->e_cut >= 0x1p-10f) || !std::isfinite(ph->e_cut))
    FAIL(GH_EINVAL, "gh_set_physics: e_cut must be >= 2^-10 MeV (first energy node)");
  const size_t M = ph->n_materials, NB1 = ph->n_buckets + 1, NX = ph->n_x_nodes;
  const size_t R = ph->n_compositions > 0 ? (size_t)ph->n_compositions : M;  // table rows
  // material m: attenuation rho * row; rho = density (shared compositions) or 1
  std::vector<MatRow> mat(M);
  std::vector<uint8_t> under;
  for (size_t m = 0; m < M; ++m) {
    if (!(ph->density[m] >= 0.0f) || !std::isfinite(ph->density[m]))
      FAIL(GH_EINVAL, "gh_set_physics: densities must be finite and >= 0");
    if (ph->n_compositions > 0) {
      if (ph->composition[m] < 0 || ph->composition[m] >= ph->n_compositions)
        FAIL(GH_EINVAL, "gh_set_physics: composition index out of range");
      mat[m] = MatRow{ph->density[m], ph->composition[m]};
    } else {
      mat[m] = MatRow{1.0f, (int)m};
    }
  }
  {
    for (size_t j = 0; j < NB1; ++j) {
      if (!(ph->mu_majorant[j] > 0.0f) || !std::isfinite(ph->mu_majorant[j]))
        FAIL(GH_EINVAL, "gh_set_physics: majorant must be positive and finite");
      for (size_t r = 0; r < R; ++r)
        if (!(ph->mu_total[r * NB1 + j] >= 0.0f)) FAIL(GH_EINVAL, "gh_set_physics: mu_total must be >= 0");
    }
    // which materials lie within the majorant (mu as the kernel forms it, an FP32
    // product, at every node): a phantom may only hold those (checked per simulation)
    under.assign(M, 1);
    for (size_t m = 0; m < M; ++m)
      for (size_t j = 0; j < NB1; ++j)
        if (!(mat[m].rho * ph->mu_total[(size_t)mat[m].row * NB1 + j] <= ph->mu_majorant[j])) {
          under[m] = 0;
          break;
        }
    if (ph->ray_x[0] != 0.0f) FAIL(GH_EINVAL, "gh_set_physics: ray_x[0] must be 0");
    for (size_t i = 1; i < NX; ++i) {
      uint32_t b;
      std::memcpy(&b, &ph->ray_x[i], 4);
      if (b != ((uint32_t)(X_BUCKET_BASE + (int)i - 1) << XSHIFT))
        FAIL(GH_EINVAL, "gh_set_physics: ray_x must be the q^2 bucket grid (node i >= 1 has bits (920+i-1)<<20)");
    }
    for (size_t r = 0; r < R; ++r)
      for (size_t i = 1; i < NX; ++i)
        if (!(ph->ray_cdf[r * NX + i] >= ph->ray_cdf[r * NX + i - 1]))
          FAIL(GH_EINVAL, "gh_set_physics: ray_cdf must be non-decreasing");
  }
  HIPCHK(hipSetDevice(ctx->device));
  if (!ctx->sims.empty()) FAIL(GH_ESTATE, "gh_set_physics: clear simulations first");
  free_all(ctx->phys_owned);
  std::vector<float> inv_rho(M);
  for (size_t i = 0; i < M; ++i) inv_rho[i] = ph->density[i] > 0.0f ? 1.0f / ph->density[i] : 0.0f;
  std::vector<float> muen_zero;
  const float *muen = ph->mu_en;
  ctx->have_muen = muen != nullptr;
  if (!muen) { muen_zero.assign(R * NB1, 0.0f); muen = muen_zero.data(); }
  Phys P{};
  float *t;
  int rc;
  {
    MatRow *dm;
    if ((rc = dev_upload(ctx, (const MatRow *)mat.data(), M, &dm, ctx->phys_owned))) return rc;
    P.mat = dm;
  }
  if ((rc = dev_upload(ctx, ph->mu_total, R * NB1, &t, ctx->phys_owned))) return rc; P.mu = t;
  if ((rc = dev_upload(ctx, ph->mu_majorant, NB1, &t, ctx->phys_owned))) return rc; P.maj = t;
  std::vector<float> pcpr(2 * R * NB1);
  for (size_t i = 0; i < R * NB1; ++i) {
    pcpr[2 * i] = ph->p_compton[i];
    pcpr[2 * i + 1] = ph->p_rayleigh[i];
  }
  if ((rc = dev_upload(ctx, (const float *)pcpr.data(), 2 * R * NB1, &t, ctx->phys_owned))) return rc; P.pcpr = t;
  if ((rc = dev_upload(ctx, muen, R * NB1, &t, ctx->phys_owned))) return rc; P.muen = t;
  if ((rc = dev_upload(ctx, ph->ray_x, NX, &t, ctx->phys_owned))) return rc; P.rx = t;
  if ((rc = dev_upload(ctx, ph->ray_cdf, R * NX, &t, ctx->phys_owned))) return rc; P.rcdf = t;
  {
    // guide[r][b] = largest i <= NX-2 with ray_cdf[r][i] <= b / RAY_GUIDE
    std::vector<int32_t> guide(R * (RAY_GUIDE + 1));
    for (size_t r = 0; r < R; ++r) {
      const float *G = ph->ray_cdf + r * NX;
      int i = 0;
      for (int b = 0; b <= RAY_GUIDE; ++b) {
        const float edge = (float)b / (float)RAY_GUIDE;
        while (i + 1 <= (int)NX - 2 && G[i + 1] <= edge) ++i;
        guide[r * (RAY_GUIDE + 1) + b] = i;
      }
    }
    int32_t *tg;
    if ((rc = dev_upload(ctx, (const int32_t *)guide.data(), guide.size(), &tg, ctx->phys_owned))) return rc;
    P.rguide = tg;
  }
  if ((rc = dev_upload(ctx, ph->incoh_s, R * NX, &t, ctx->phys_owned))) return rc; P.sinc = t;
  if ((rc = dev_upload(ctx, inv_rho.data(), M, &t, ctx->phys_owned))) return rc; P.inv_rho = t;
  P.n_mat = (int)M;
  P.n_rows = (int)R;
  P.shared_rows = ph->n_compositions > 0 ? 1 : 0;
  P.nb = ph->n_buckets;
  P.nx = (int)NX;
  P.e_cut = ph->e_cut;
  ctx->P = P;
  ctx->lds1_bytes = (size_t)lds_layout((int)R, (int)M, (int)NB1, (int)NX, 1, false).total * sizeof(float);
  ctx->lds2_bytes = (size_t)lds_layout((int)R, (int)M, (int)NB1, (int)NX, 2, false).total * sizeof(float);
  ctx->lds1e_bytes = (size_t)lds_layout((int)R, (int)M, (int)NB1, (int)NX, 1, true).total * sizeof(float);
  ctx->lds2e_bytes = (size_t)lds_layout((int)R, (int)M, (int)NB1, (int)NX, 2, true).total * sizeof(float);
  ctx->under_majorant.swap(under);
  ctx->have_phys = true;
  return GH_OK;
}

int gh_clear_simulations(gh_ctx *ctx) {
  if (!ctx) return GH_EINVAL;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (auto &s : ctx->sims) { free_all(s.owned); free_all(s.source_owned); }
  ctx->sims.clear();
  return GH_OK;
}

int gh_add_simulation(gh_ctx *ctx, const gh_phantom_desc *ph, const gh_source_desc *src,
                      const gh_tally_desc *tal, int32_t *sim_id) {
  if (!ctx) return GH_EINVAL;
  if (!ctx->have_phys) FAIL(GH_ESTATE, "gh_add_simulation: call gh_set_physics first");
  if (!ph || !src || !tal) FAIL(GH_EINVAL, "gh_add_simulation: null descriptor");
  if (ph->nx <= 0 || ph->ny <= 0 || ph->nz <= 0 || !ph->labels)
    FAIL(GH_EINVAL, "gh_add_simulation: invalid phantom");
  const int64_t nvox = (int64_t)ph->nx * ph->ny * ph->nz;
  if (nvox >= (1ll << 31)) FAIL(GH_ERANGE, "gh_add_simulation: phantom has 2^31 voxels or more");
  for (int i = 0; i < 3; ++i)
    if (!(ph->spacing[i] > 0.0f)) FAIL(GH_EINVAL, "gh_add_simulation: spacing must be > 0");
  if (ph->world_material >= ctx->P.n_mat) FAIL(GH_EINVAL, "gh_add_simulation: world material out of range");
  if (ph->label_bytes != 1 && ph->label_bytes != 2) FAIL(GH_EINVAL, "gh_add_simulation: label_bytes must be 1 or 2");
  auto label_in = [&](int64_t i) -> uint32_t {
    return ph->label_bytes == 2 ? ((const uint16_t *)ph->labels)[i] : ((const uint8_t *)ph->labels)[i];
  };
  uint32_t max_label = 0;
  std::vector<uint8_t> present((size_t)ctx->P.n_mat + 1, 0);
  for (int64_t i = 0; i < nvox; ++i) {
    const uint32_t l = label_in(i);
    max_label = std::max(max_label, l);
    present[std::min<size_t>(l, (size_t)ctx->P.n_mat)] = 1;
  }
  if (max_label >= (uint32_t)ctx->P.n_mat) FAIL(GH_EINVAL, "gh_add_simulation: label out of range of the material table");
  for (int m = 0; m < ctx->P.n_mat; ++m)
    if (present[m] && !ctx->under_majorant[m]) {
      ctx->err = "gh_add_simulation: material " + std::to_string(m) +
                 " occurs in the phantom but its mu exceeds the majorant of gh_set_physics";
      return GH_EINVAL;
    }
  {
    const int rc_src = validate_source(ctx, src);
    if (rc_src) return rc_src;
  }
  if (tal->nx <= 0 || tal->ny <= 0 || tal->nz <= 0) FAIL(GH_EINVAL, "invalid tally size");
  if ((tal->flags & GH_TALLY_ALIGNED) && (tal->nx != ph->nx || tal->ny != ph->ny || tal->nz != ph->nz))
    FAIL(GH_EINVAL, "aligned tally must have the phantom size");
  if (tal->record_face >= 0 && ph->world_material >= 0)
    FAIL(GH_EINVAL, "exit-face recording needs a vacuum world");
  if (tal->record_face > GH_FACE_ALL) FAIL(GH_EINVAL, "record_face must be -1..6");
  if (tal->flags & GH_TALLY_TLE) {
    if (tal->detector_label_min >= 0) FAIL(GH_EINVAL, "a TLE tally cannot score detector events");
    if (!(tal->tle_energy_min >= 0.0f) || !(tal->tle_energy_max >= 0.0f))
      FAIL(GH_EINVAL, "TLE energy bounds must be >= 0");
    if (!ctx->have_muen) FAIL(GH_ESTATE, "a TLE tally needs mu_en tables in gh_set_physics");
  }
  HIPCHK(hipSetDevice(ctx->device));

  SimHost h;
  SimDev &d = h.dev;
  std::memset(&d, 0, sizeof(d));
  int rc;
  {
    // device layout: the narrowest of 4-, 8- and 16-bit labels that holds every label of
    // the phantom, packed in 32-bit words (at most 16 materials in use: two labels per
    // byte, ct256 4 MiB instead of 8 MiB against 4 MB of L2 per XCD)
    const int lb = max_label < 16u ? 2 : (max_label < 256u ? 3 : 4);
    const int per_word = 32 >> lb;
    std::vector<uint32_t> packed(((size_t)nvox + per_word - 1) / per_word, 0u);
    for (int64_t i = 0; i < nvox; ++i)
      packed[(size_t)i / per_word] |= label_in(i) << (((uint32_t)i % per_word) << lb);
    uint32_t *lab;
    if ((rc = dev_upload(ctx, (const uint32_t *)packed.data(), packed.size(), &lab, h.owned))) {
      free_all(h.owned);
      return rc;
    }
    d.labels = lab;
    d.lb = lb;
  }
  d.nx = ph->nx; d.ny = ph->ny; d.nz = ph->nz;
  d.world_mat = ph->world_material;
  for (int i = 0; i < 3; ++i) {
    d.inv_s[i] = 1.0f / ph->spacing[i];
    d.ts_ph[i] = ph->spacing[i];
    const int n = i == 0 ? ph->nx : (i == 1 ? ph->ny : ph->nz);
    d.L[i] = (float)((double)n * (double)ph->spacing[i]);
    d.wlo[i] = ph->world_lo[i];
    d.whi[i] = ph->world_hi[i];
  }
  if ((rc = upload_source(ctx, src, d, h.source_owned))) { free_all(h.owned); free_all(h.source_owned); return rc; }
  float *t;
  // tally
  const size_t ntal = (size_t)tal->nx * tal->ny * tal->nz;
  h.n_tally = ntal;
  h.flags = tal->flags;
  d.aligned = (tal->flags & GH_TALLY_ALIGNED) ? 1 : 0;
  d.tnx = tal->nx; d.tny = tal->ny; d.tnz = tal->nz;
  for (int i = 0; i < 3; ++i) {
    // an aligned tally is the phantom grid: corner at the origin of the phantom frame
    d.tcorner[i] = d.aligned ? 0.0f : tal->corner[i];
    d.ts[i] = d.aligned ? ph->spacing[i] : tal->spacing[i];
    d.tinv[i] = 1.0f / d.ts[i];
  }
  double *dd;
  if ((rc = dev_upload<double>(ctx, nullptr, ntal, &dd, h.owned))) { free_all(h.owned); return rc; }
  d.edep = dd;
  if (tal->flags & GH_TALLY_EDEP2) {
    if ((rc = dev_upload<double>(ctx, nullptr, ntal, &dd, h.owned))) { free_all(h.owned); return rc; }
    d.edep2 = dd;
  }
  if (tal->flags & GH_TALLY_DOSE) {
    if ((rc = dev_upload<double>(ctx, nullptr, ntal, &dd, h.owned))) { free_all(h.owned); return rc; }
    d.dose = dd;
  }
  if (tal->flags & GH_TALLY_DOSE2) {
    if ((rc = dev_upload<double>(ctx, nullptr, ntal, &dd, h.owned))) { free_all(h.owned); return rc; }
    d.dose2 = dd;
  }
  if (tal->flags & GH_TALLY_COUNTS) {
    if ((rc = dev_upload<double>(ctx, nullptr, ntal, &dd, h.owned))) { free_all(h.owned); return rc; }
    d.counts = dd;
  }
  d.rec_face = tal->record_face;
  d.rec_cap = tal->record_face >= 0 ? tal->record_capacity : 0;
  if (tal->record_face >= 0) {
    unsigned long long *rc_ptr;
    if ((rc = dev_upload<unsigned long long>(ctx, nullptr, 1, &rc_ptr, h.owned))) { free_all(h.owned); return rc; }
    d.rec_count = rc_ptr;
    for (int k = 0; k < 8; ++k) {
      if ((rc = dev_upload<float>(ctx, nullptr, std::max<uint32_t>(d.rec_cap, 1u), &t, h.owned))) { free_all(h.owned); return rc; }
      d.rec[k] = t;
    }
  }
  d.det_label_min = tal->detector_label_min;
  h.det = tal->detector_label_min;
  d.tle = (tal->flags & GH_TALLY_TLE) ? 1 : 0;
  d.tle_emin = tal->tle_energy_min;
  d.tle_emax = tal->tle_energy_max;
  ctx->sims.push_back(std::move(h));
  if (sim_id) *sim_id = (int32_t)ctx->sims.size() - 1;
  return gh_reset_tally(ctx, (int32_t)ctx->sims.size() - 1);
}

int gh_set_source(gh_ctx *ctx, int32_t sim_id, const gh_source_desc *src) {
  if (!ctx) return GH_EINVAL;
  if (sim_id < 0 || sim_id >= (int32_t)ctx->sims.size()) FAIL(GH_EINVAL, "gh_set_source: bad sim id");
  if (!src) FAIL(GH_EINVAL, "gh_set_source: null descriptor");
  int rc = validate_source(ctx, src);
  if (rc) return rc;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  SimHost &h = ctx->sims[sim_id];
  std::vector<void *> fresh;
  if ((rc = upload_source(ctx, src, h.dev, fresh))) { free_all(fresh); return rc; }
  free_all(h.source_owned);
  h.source_owned = std::move(fresh);
  return GH_OK;
}

int gh_reset_tally(gh_ctx *ctx, int32_t sim_id) {
  if (!ctx) return GH_EINVAL;
  if (sim_id < 0 || sim_id >= (int32_t)ctx->sims.size()) FAIL(GH_EINVAL, "gh_reset_tally: bad sim id");
  HIPCHK(hipSetDevice(ctx->device));
  SimHost &h = ctx->sims[sim_id];
  const size_t bytes = h.n_tally * sizeof(double);
  double *arrs[5] = {h.dev.edep, h.dev.edep2, h.dev.dose, h.dev.dose2, h.dev.counts};
  for (double *a : arrs)
    if (a) HIPCHK(hipMemsetAsync(a, 0, bytes, ctx->stream));
  if (h.dev.rec_count) HIPCHK(hipMemsetAsync(h.dev.rec_count, 0, sizeof(unsigned long long), ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return GH_OK;
}

int gh_reset_phsp(gh_ctx *ctx, int32_t sim_id) {
  if (!ctx) return GH_EINVAL;
  if (sim_id < 0 || sim_id >= (int32_t)ctx->sims.size()) FAIL(GH_EINVAL, "gh_reset_phsp: bad sim id");
  HIPCHK(hipSetDevice(ctx->device));
  SimHost &h = ctx->sims[sim_id];
  if (h.dev.rec_count) {
    HIPCHK(hipMemsetAsync(h.dev.rec_count, 0, sizeof(unsigned long long), ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  return GH_OK;
}

int gh_run(gh_ctx *ctx, const uint64_t *n_histories, const uint64_t *first_history, uint64_t seed,
           const gh_launch *launch, gh_counters *out) {
  if (!ctx) return GH_EINVAL;
  if (!ctx->have_phys || ctx->sims.empty()) FAIL(GH_ESTATE, "gh_run: physics and simulations required");
  if (!n_histories) FAIL(GH_EINVAL, "gh_run: n_histories is null");
  HIPCHK(hipSetDevice(ctx->device));
  const int ns = (int)ctx->sims.size();
  gh_launch lp{256, 0, 64, 0, 52, 12, 2};
  if (launch) {
    if (launch->block > 0) lp.block = launch->block;
    if (launch->grid > 0) lp.grid = launch->grid;
    if (launch->refill_interval > 0) lp.refill_interval = launch->refill_interval;
    if (launch->chunk > 0) lp.chunk = launch->chunk;
    if (launch->slow_min > 0) lp.slow_min = launch->slow_min;
    if (launch->refill_min > 0) lp.refill_min = launch->refill_min;
    if (launch->scatter_tries > 0) lp.scatter_tries = launch->scatter_tries;
  }
  if (ns > MAX_SIMS) FAIL(GH_ERANGE, "gh_run: more than 32768 simulations in one launch");
  if (lp.scatter_tries > MAX_REJECT) lp.scatter_tries = MAX_REJECT;

  std::vector<uint64_t> offs(ns + 1, 0);
  std::vector<SimDev> devs(ns);
  for (int s = 0; s < ns; ++s) {
    SimHost &h = ctx->sims[s];
    if (n_histories[s] > 0xFFFFFFFFull) FAIL(GH_ERANGE, "gh_run: more than 2^32-1 histories for one simulation in one launch");
    h.dev.n_hist = n_histories[s];
    h.dev.first_hist = first_history ? first_history[s] : 0;
    if (h.det >= 0) {
      // per-history detector energy buffer sized to this run
      if (h.event_cap < n_histories[s]) {
        float *t;
        int rc;
        if ((rc = dev_upload<float>(ctx, nullptr, (size_t)std::max<uint64_t>(n_histories[s], 1), &t, h.owned))) return rc;
        h.dev.event_edep = t;
        h.event_cap = n_histories[s];
      }
      h.event_n = n_histories[s];
    }
    devs[s] = h.dev;
  }
  if (!ctx->d_sims || ctx->d_sims_cap < (size_t)ns) {
    if (ctx->d_sims) { (void)hipFree(ctx->d_sims); ctx->d_sims = nullptr; }
    if (ctx->d_chunk_off) { (void)hipFree(ctx->d_chunk_off); ctx->d_chunk_off = nullptr; }
    HIPCHK(hipMalloc(&ctx->d_sims, ns * sizeof(SimDev)));
    HIPCHK(hipMalloc(&ctx->d_chunk_off, (ns + 1) * sizeof(uint64_t)));
    ctx->d_sims_cap = ns;
  }
  if (!ctx->d_next) {
    HIPCHK(hipMalloc(&ctx->d_next, sizeof(unsigned long long)));
    HIPCHK(hipMalloc(&ctx->d_counters, 12 * sizeof(unsigned long long)));
    HIPCHK(hipMalloc(&ctx->d_energies, 4 * sizeof(double)));
  }

  Launch L;
  L.sims = ctx->d_sims;
  L.n_sims = ns;
  L.chunk_offsets = ctx->d_chunk_off;
  L.next_chunk = ctx->d_next;
  L.seed = seed;
  L.max_iter = lp.refill_interval;
  L.slow_min = std::max(1, std::min(64, lp.slow_min));
  L.refill_min = std::max(1, std::min(L.slow_min, lp.refill_min));
  L.scatter_tries = lp.scatter_tries;
  L.P = ctx->P;
  L.counters = ctx->d_counters;
  L.energies = ctx->d_energies;
  // staging level and block: all tables in LDS shared by one 1024-thread workgroup per CU
  // when they fit, else mu + majorant in LDS with 256-thread workgroups, else none. The
  // budgets leave room for the kernel's static LDS (per-lane energy slots, birth queues and
  // counters: 72 KiB at 1024 threads, 18 KiB at 256; a short source spectrum: 4 KiB at 1024)
  // within the 160 KiB of a CU; a
  // 256-thread workgroup keeps to 40 KiB so that four of them share a CU.
  auto static_lds = [](size_t block) {
    return (size_t)N_ENERGIES * block * sizeof(double) + (size_t)QWORDS * block * sizeof(uint32_t) +
           N_COUNTERS * sizeof(unsigned long long) + 64 * sizeof(float2) +
           (block == 1024 ? (3 * SPEC_LDS_MAX + SPEC_GUIDE + 1) * sizeof(float) : 0);
  };
  const size_t budget_big = 160 * 1024 - static_lds(1024) - 512;
  const size_t budget_small = 40 * 1024 - static_lds(256) - 256;
  int stage, big;
  if (launch && launch->block > 0) {
    big = lp.block > 256;
    if (big) stage = ctx->lds2_bytes <= budget_big ? 2 : (ctx->lds1_bytes <= budget_big ? 1 : 0);
    else stage = ctx->lds1_bytes <= budget_small ? 1 : 0;
    if (big && stage == 0) FAIL(GH_ERANGE, "gh_run: tables too large for a 1024-thread launch; use block <= 256");
  } else if (ns == 1 && ctx->lds2_bytes <= budget_big) {
    // (the batched kernel needs more registers than a 1024-thread workgroup allows)
    stage = 2; big = 1;
  } else if (ctx->lds1_bytes <= budget_small) {
    stage = 1; big = 0;
  } else if (ns == 1 && ctx->lds1_bytes <= budget_big) {
    stage = 1; big = 1;
  } else {
    stage = 0; big = 0;
  }
  if (big) lp.block = 1024;
  if (lp.block % 64 || lp.block > 1024 || (!big && lp.block > 256))
    FAIL(GH_EINVAL, "gh_run: block must be a multiple of 64, <= 256, or 1024");
  size_t shmem = stage == 2 ? ctx->lds2_bytes : (stage == 1 ? ctx->lds1_bytes : 0);
  // shared composition rows: stage mu per material (node products) when that still fits
  L.P.mu_expanded = 0;
  if (ctx->P.shared_rows && stage >= 1) {
    const size_t e = stage == 2 ? ctx->lds2e_bytes : ctx->lds1e_bytes;
    if (e <= (big ? budget_big : budget_small)) {
      L.P.mu_expanded = 1;
      shmem = e;
    }
  }
  const bool batched = ns > 1;
  bool tle = false;
  for (int s = 0; s < ns; ++s) tle = tle || ctx->sims[s].dev.tle;
  if (tle && batched) FAIL(GH_EINVAL, "gh_run: TLE tallies need a single-simulation launch");
  const void *fn = nullptr;
#define GH_PICK(B, S, BL, TL) (const void *)&transport_kernel<B, S, BL, TL>
  if (!batched && tle) {
    fn = big ? (stage == 2 ? GH_PICK(false, 2, 1024, true) : GH_PICK(false, 1, 1024, true))
             : (stage == 1 ? GH_PICK(false, 1, 256, true) : GH_PICK(false, 0, 256, true));
  } else if (!batched) {
    fn = big ? (stage == 2 ? GH_PICK(false, 2, 1024, false) : GH_PICK(false, 1, 1024, false))
             : (stage == 1 ? GH_PICK(false, 1, 256, false) : GH_PICK(false, 0, 256, false));
  } else {
    fn = big ? (stage == 2 ? GH_PICK(true, 2, 1024, false) : GH_PICK(true, 1, 1024, false))
             : (stage == 1 ? GH_PICK(true, 1, 256, false) : GH_PICK(true, 0, 256, false));
  }
#undef GH_PICK
  if (shmem > 64 * 1024) HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  if (lp.grid <= 0) {
    int per_cu = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, lp.block, shmem));
    lp.grid = std::max(1, ctx->n_cu) * std::max(1, per_cu);
  }
  // work-queue chunk: automatic = about four chunks per wave of the launch (64 to 4096
  // histories, a multiple of 64), so that a short run still spreads over every wave; which
  // wave runs a history does not change its result
  if (lp.chunk <= 0) {
    uint64_t total = 0;
    for (int s = 0; s < ns; ++s) total += n_histories[s];
    const uint64_t waves = (uint64_t)lp.grid * (uint64_t)(lp.block / 64);
    uint64_t c = total / std::max<uint64_t>(1, 4 * waves);
    c = std::min<uint64_t>(4096, std::max<uint64_t>(64, c & ~(uint64_t)63));
    lp.chunk = (int32_t)c;
  }
  for (int s = 0; s < ns; ++s) offs[s + 1] = offs[s] + (n_histories[s] + lp.chunk - 1) / lp.chunk;
  L.total_chunks = offs[ns];
  L.chunk = (uint32_t)lp.chunk;
  HIPCHK(hipMemcpyAsync(ctx->d_sims, devs.data(), ns * sizeof(SimDev), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->d_chunk_off, offs.data(), (ns + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->d_next, 0, sizeof(unsigned long long), ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->d_counters, 0, 12 * sizeof(unsigned long long), ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->d_energies, 0, 4 * sizeof(double), ctx->stream));
  ctx->last_launch = lp;
  ctx->last_stage = stage;
  HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
  if (L.total_chunks > 0) {
    void *args[] = {&L};
    HIPCHK(hipLaunchKernel(fn, dim3(lp.grid), dim3(lp.block), args, shmem, ctx->stream));
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
  unsigned long long c[12];
  double e[4];
  HIPCHK(hipMemcpyAsync(c, ctx->d_counters, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(e, ctx->d_energies, sizeof(e), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (out) {
    out->histories = c[0]; out->steps = c[1]; out->deposits = c[2]; out->compton = c[3];
    out->rayleigh = c[4]; out->photo = c[5]; out->escaped = c[6]; out->world_interactions = c[7];
    out->cut_kills = c[8]; out->sq_evictions = c[9]; out->recorded = c[10]; out->kn_rejections = c[11];
    out->e_deposited = e[0]; out->e_outside_tally = e[1]; out->e_escaped = e[2]; out->e_world = e[3];
    out->kernel_ms = ms;
  }
  return GH_OK;
}

int gh_autotune(gh_ctx *ctx, const uint64_t *n_trial, uint64_t seed, gh_launch *best, double *best_ms) {
  if (!ctx) return GH_EINVAL;
  if (!ctx->have_phys || ctx->sims.empty()) FAIL(GH_ESTATE, "gh_autotune: physics and simulations required");
  if (!n_trial || !best) FAIL(GH_EINVAL, "gh_autotune: null argument");
  HIPCHK(hipSetDevice(ctx->device));
  // snapshot every tally so that trial histories leave no trace
  std::vector<std::pair<double *, double *>> saved;
  std::vector<unsigned long long> rec_counts;
  std::vector<void *> tmp;
  for (auto &h : ctx->sims) {
    double *arrs[5] = {h.dev.edep, h.dev.edep2, h.dev.dose, h.dev.dose2, h.dev.counts};
    for (double *a : arrs) {
      if (!a) continue;
      void *b = nullptr;
      HIPCHK(hipMalloc(&b, h.n_tally * sizeof(double)));
      tmp.push_back(b);
      HIPCHK(hipMemcpyAsync(b, a, h.n_tally * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      saved.push_back({a, (double *)b});
    }
    unsigned long long rc = 0;
    if (h.dev.rec_count) HIPCHK(hipMemcpy(&rc, h.dev.rec_count, sizeof(rc), hipMemcpyDeviceToHost));
    rec_counts.push_back(rc);
  }
  // candidates: workgroup shape (automatic staging / 256 threads) x slow-phase threshold
  // x refill threshold (small, frequent refills keep more lanes stepping)
  struct Cand { int block, slow, refill, iter; };
  std::vector<Cand> cands;
  for (int block : {0, 256})
    for (int slow : {16, 32, 48, 52, 56, 60, 62})
      for (int refill : {8, 16, 24})
        if (refill <= slow) cands.push_back({block, slow, refill, 64});
  double best_t = 1e300;
  gh_launch best_lp{};
  int rc = GH_OK;
  std::vector<uint64_t> first(ctx->sims.size(), 0);
  for (const Cand &c : cands) {
    gh_launch lp{c.block, 0, c.iter, 0, c.slow, c.refill, 0};
    gh_counters out{};
    rc = gh_run(ctx, n_trial, first.data(), seed, &lp, &out);
    if (rc != GH_OK) break;
    if (out.kernel_ms < best_t) {
      best_t = out.kernel_ms;
      best_lp = ctx->last_launch;
      best_lp.block = c.block;
      best_lp.grid = 0;
      best_lp.chunk = 0;  // automatic: sized by the production run's own history count
    }
  }
  for (auto &pr : saved) {
    size_t n = 0;
    for (auto &h : ctx->sims)
      for (double *a : {h.dev.edep, h.dev.edep2, h.dev.dose, h.dev.dose2, h.dev.counts})
        if (a == pr.first) n = h.n_tally;
    (void)hipMemcpyAsync(pr.first, pr.second, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream);
  }
  for (size_t s = 0; s < ctx->sims.size(); ++s)
    if (ctx->sims[s].dev.rec_count)
      (void)hipMemcpyAsync(ctx->sims[s].dev.rec_count, &rec_counts[s], sizeof(unsigned long long),
                           hipMemcpyHostToDevice, ctx->stream);
  (void)hipStreamSynchronize(ctx->stream);
  free_all(tmp);
  // the trials overwrote the per-history detector energies of the last run
  for (auto &h : ctx->sims) h.event_n = 0;
  if (rc != GH_OK) return rc;
  *best = best_lp;
  if (best_ms) *best_ms = best_t;
  return GH_OK;
}

int gh_get_tally(gh_ctx *ctx, int32_t sim_id, double *edep, double *edep2, double *dose, double *dose2,
                 double *counts) {
  if (!ctx) return GH_EINVAL;
  if (sim_id < 0 || sim_id >= (int32_t)ctx->sims.size()) FAIL(GH_EINVAL, "gh_get_tally: bad sim id");
  HIPCHK(hipSetDevice(ctx->device));
  SimHost &h = ctx->sims[sim_id];
  const size_t bytes = h.n_tally * sizeof(double);
  double *dst[5] = {edep, edep2, dose, dose2, counts};
  double *src[5] = {h.dev.edep, h.dev.edep2, h.dev.dose, h.dev.dose2, h.dev.counts};
  const char *names[5] = {"edep", "edep2", "dose", "dose2", "counts"};
  for (int k = 0; k < 5; ++k) {
    if (!dst[k]) continue;
    if (!src[k]) FAIL(GH_ESTATE, std::string("gh_get_tally: ") + names[k] + " not enabled");
    HIPCHK(hipMemcpyAsync(dst[k], src[k], bytes, hipMemcpyDeviceToHost, ctx->stream));
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return GH_OK;
}

int gh_get_phsp(gh_ctx *ctx, int32_t sim_id, float *const out[8], uint64_t capacity, uint64_t *n) {
  if (!ctx) return GH_EINVAL;
  if (sim_id < 0 || sim_id >= (int32_t)ctx->sims.size()) FAIL(GH_EINVAL, "gh_get_phsp: bad sim id");
  SimHost &h = ctx->sims[sim_id];
  if (!h.dev.rec_count) FAIL(GH_ESTATE, "gh_get_phsp: recording not enabled");
  HIPCHK(hipSetDevice(ctx->device));
  unsigned long long cnt = 0;
  HIPCHK(hipMemcpy(&cnt, h.dev.rec_count, sizeof(cnt), hipMemcpyDeviceToHost));
  const uint64_t stored = std::min<uint64_t>(cnt, h.dev.rec_cap);
  if (n) *n = cnt;
  if (out) {
    const uint64_t m = std::min<uint64_t>(stored, capacity);
    for (int k = 0; k < 8; ++k)
      if (out[k] && m) HIPCHK(hipMemcpy(out[k], h.dev.rec[k], m * sizeof(float), hipMemcpyDefault));  // host or device
  }
  return GH_OK;
}

int gh_get_event_edep(gh_ctx *ctx, int32_t sim_id, float *out, uint64_t n) {
  if (!ctx) return GH_EINVAL;
  if (sim_id < 0 || sim_id >= (int32_t)ctx->sims.size()) FAIL(GH_EINVAL, "gh_get_event_edep: bad sim id");
  SimHost &h = ctx->sims[sim_id];
  if (h.det < 0 || !h.dev.event_edep) FAIL(GH_ESTATE, "gh_get_event_edep: detector scoring not enabled");
  if (n > h.event_n) FAIL(GH_ERANGE, "gh_get_event_edep: n beyond the histories of the last run");
  HIPCHK(hipSetDevice(ctx->device));
  if (n) HIPCHK(hipMemcpy(out, h.dev.event_edep, n * sizeof(float), hipMemcpyDefault));  // host or device
  return GH_OK;
}

// Walker alias table (Vose's construction, FP64): entry i is kept with probability
// prob[i], otherwise replaced by alias[i]; P(i) = (prob[i] + sum_{j: alias[j] = i}
// (1 - prob[j])) / n = w[i] / sum(w). Host-only helper (no device needed).
int gh_alias_table(const double *w, int64_t n, float *prob, int32_t *alias) {
  if (!w || !prob || !alias || n <= 0 || n > 0x7FFFFFFF) return GH_EINVAL;
  double total = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    if (!(w[i] >= 0.0) || !std::isfinite(w[i])) return GH_EINVAL;
    total += w[i];
  }
  if (!(total > 0.0)) return GH_EINVAL;
  std::vector<double> p((size_t)n);
  std::vector<int32_t> small, large;
  small.reserve((size_t)n);
  large.reserve((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    p[i] = w[i] * (double)n / total;
    (p[i] < 1.0 ? small : large).push_back((int32_t)i);
  }
  while (!small.empty() && !large.empty()) {
    const int32_t a = small.back(), b = large.back();
    small.pop_back();
    large.pop_back();
    prob[a] = (float)p[a];
    alias[a] = b;
    p[b] = (p[b] + p[a]) - 1.0;
    (p[b] < 1.0 ? small : large).push_back(b);
  }
  for (int32_t i : large) { prob[i] = 1.0f; alias[i] = i; }
  for (int32_t i : small) { prob[i] = 1.0f; alias[i] = i; }  // rounding leftovers
  return GH_OK;
}

extern "C++" {
template <typename T>
static int hu_to_labels_impl(gh_ctx *ctx, const T *hu, int64_t n_voxels, const double *bins, int32_t n_bins,
                             const uint16_t *labels_sorted, uint16_t *out, const char *fname) {
  if (!ctx) return GH_EINVAL;
  if (!hu || !bins || !labels_sorted || !out || n_voxels < 0) FAIL(GH_EINVAL, std::string(fname) + ": null argument");
  if (n_bins < 0 || n_bins > HU_MAX_BINS) FAIL(GH_ERANGE, std::string(fname) + ": at most 4096 bins");
  for (int i = 1; i < n_bins; ++i)
    if (!(bins[i - 1] <= bins[i])) FAIL(GH_EINVAL, std::string(fname) + ": bins must be ascending");
  if (n_voxels == 0) return GH_OK;
  HIPCHK(hipSetDevice(ctx->device));
  std::vector<void *> tmp;
  T *d_hu;
  double *d_bins;
  uint16_t *d_lut, *d_out;
  int rc;
  double dummy = 0.0;
  if ((rc = dev_upload(ctx, hu, (size_t)n_voxels, &d_hu, tmp))) { free_all(tmp); return rc; }
  if ((rc = dev_upload(ctx, n_bins ? bins : &dummy, (size_t)std::max(n_bins, 1), &d_bins, tmp))) { free_all(tmp); return rc; }
  if ((rc = dev_upload(ctx, labels_sorted, (size_t)n_bins + 1, &d_lut, tmp))) { free_all(tmp); return rc; }
  if ((rc = dev_upload<uint16_t>(ctx, nullptr, (size_t)n_voxels, &d_out, tmp))) { free_all(tmp); return rc; }
  const int block = 256;
  const int64_t want = (n_voxels + block - 1) / block;
  const int grid = (int)std::min<int64_t>(want, (int64_t)std::max(ctx->n_cu, 1) * 16);
  const size_t shmem = (size_t)n_bins * sizeof(double) + ((size_t)n_bins + 1) * sizeof(uint16_t);
  hipLaunchKernelGGL(hu_to_labels_kernel<T>, dim3(grid), dim3(block), shmem, ctx->stream, d_hu, n_voxels, d_bins,
                     n_bins, d_lut, d_out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, (size_t)n_voxels * sizeof(uint16_t), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  free_all(tmp);
  if (e != hipSuccess) FAIL(GH_EHIP, std::string(fname) + ": " + hipGetErrorString(e));
  return GH_OK;
}
}  // extern "C++"

int gh_hu_to_labels(gh_ctx *ctx, const int16_t *hu, int64_t n_voxels, const double *bins, int32_t n_bins,
                    const uint16_t *labels_sorted, uint16_t *out) {
  return hu_to_labels_impl(ctx, hu, n_voxels, bins, n_bins, labels_sorted, out, "gh_hu_to_labels");
}

int gh_hu_to_labels_f64(gh_ctx *ctx, const double *hu, int64_t n_voxels, const double *bins, int32_t n_bins,
                        const uint16_t *labels_sorted, uint16_t *out) {
  return hu_to_labels_impl(ctx, hu, n_voxels, bins, n_bins, labels_sorted, out, "gh_hu_to_labels_f64");
}

}  // extern "C"
