/* gatehip.h -- C-ABI of the MI355X photon transport + voxel dose library.
 *
 * This is the boundary a host binding (ctypes here; a pybind11 shim in
 * opengate_core) calls. Only plain C types, pointers and sizes cross it.
 *
 * What each entry point replaces in the reference (OpenGATE 10.1.1):
 *   gh_set_physics      per-material cross sections the Geant4 gamma processes
 *                       use (physics.py:206-438) and the mu tables of
 *                       GateMaterialMuHandler (GateMaterialMuHandler.cpp:159-254)
 *   gh_hu_to_labels     ImageVolume.create_label_image (geometry/volumes.py:1200-1243)
 *   gh_add_simulation   ImageVolume construction + GateImageNestedParameterisation
 *                       (volumes.py:1050-1142, GateImageNestedParameterisation.cpp:19-75),
 *                       GateGenericSource / GatePhaseSpaceSource initialisation
 *                       (GateGenericSource.cpp:312-664, GatePhaseSpaceSource.cpp:28-273),
 *                       GateDoseActor::InitializeCpp / BeginOfRunActionMasterThread
 *                       (GateDoseActor.cpp:54-117)
 *   gh_run              the Geant4 event loop driven by GateSourceManager::StartMasterThread
 *                       (GateSourceManager.cpp:198-244) with GateDoseActor::SteppingAction,
 *                       ScoreSquaredValue and FlushSquaredValues (GateDoseActor.cpp:180-243,
 *                       359-397) and BeginOfEventAction event counting (:139-142)
 *   gh_get_tally        DoseActor.EndOfRunActionMasterThread image fetch
 *                       (actors/doseactors.py:662-715, cpp_edep_image.to_pyarray)
 *   gh_get_phsp         PhaseSpaceActor recording of photons leaving a volume
 *                       (GatePhaseSpaceActor.cpp:119) restricted to one exit face
 *
 * Conventions: every function returns 0 on success and a non-zero GH_E* code on
 * failure; gh_last_error() then holds a message. Host buffers are caller-owned
 * and only read/written during the call. Device memory is owned by the context.
 * A context is used from one host thread at a time. Units: mm, MeV, g/cm3.
 */
#ifndef GATEHIP_H
#define GATEHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GH_ABI_VERSION 6

enum {
  GH_OK = 0,
  GH_EINVAL = 1,   /* invalid argument */
  GH_EHIP = 2,     /* HIP runtime error */
  GH_ESTATE = 3,   /* call out of order (e.g. run before physics) */
  GH_ERANGE = 4    /* size beyond a supported limit */
};

/* position types of gh_source_desc.pos_type (GateGenericSource.cpp:312-384) */
enum { GH_POS_POINT = 0, GH_POS_BOX = 1, GH_POS_SPHERE = 2, GH_POS_ELLIPSOID = 3,
       GH_POS_DISC = 4, GH_POS_CYLINDER = 5, GH_POS_PHSP = 6,
       GH_POS_VOXELS = 7, /* activity image (VoxelSource), Walker alias table over its voxels */
       GH_POS_SURFACE_SPHERE = 8 /* uniform on a sphere's surface (G4SPSPosDistribution "Surface") */ };
/* direction types (GateGenericSource.cpp:386-480; G4SPSAngDistribution). ISO, COS and
 * HISTOGRAM give momentum -(sin t cos p, sin t sin p, cos t); on a GH_POS_SURFACE_SPHERE,
 * ISO and COS take it in the tangent frame at the point (surface_axis below). */
enum { GH_DIR_ISO = 0, GH_DIR_MOMENTUM = 1, GH_DIR_FOCUSED = 2,
       GH_DIR_COS = 3,       /* cosine law: sin^2 t uniform between sin^2 of the theta range */
       GH_DIR_BEAM2D = 4,    /* (a, b) = dir_sigma N(0,1)^2, t = |(a, b)|, p = the angle of (a, b) */
       GH_DIR_HISTOGRAM = 5  /* t and p piecewise uniform ("user" theta and phi histograms) */ };
/* energy types (GateGenericSource.cpp:492-664, GateSPSEneDistribution.cpp:16-159) */
enum { GH_ENE_MONO = 0, GH_ENE_LINES = 1, GH_ENE_HISTOGRAM = 2, GH_ENE_RANGE = 3,
       GH_ENE_GAUSS = 4 /* energy_mono + energy_sigma N(0,1), negative values 0 (G4SPSEneDistribution
                           "Gauss"); tails beyond energy_mono + 8 energy_sigma are cut there */ };
/* tally flags */
enum { GH_TALLY_EDEP2 = 1, GH_TALLY_DOSE = 2, GH_TALLY_DOSE2 = 4, GH_TALLY_COUNTS = 8,
       GH_TALLY_ALIGNED = 16,
       GH_TALLY_TLE = 32 /* track-length (Woodcock collision) kerma estimator, TLEDoseActor */ };
/* record_face value that records photons leaving through any face of the phantom box */
enum { GH_FACE_ALL = 6 };

typedef struct gh_ctx gh_ctx;

/* FP32 tables on the float-bit energy grid (node j has the FP32 energy with bits
 * (bucket_base + j) << 17, bucket_base = 7488, 64 buckets per octave from 2^-10 MeV) and on
 * the q^2 grid ray_x (ascending, ray_x[0] = 0).
 * Two layouts. n_compositions = 0: the tables ("[T]" below) are per material, mu_total and
 * mu_en in 1/mm. n_compositions = C > 0: materials that share an elemental composition
 * share its tables (the Schneider HU conversion makes many materials of one composition at
 * different densities, HounsfieldUnit_to_material, geometry/materials.py:149-230): the
 * tables are per composition, mu_total and mu_en per unit density (1/mm per g/cm3), and
 * material m has mu = density[m] * mu_total[composition[m]] (FP32 product). Up to 65535
 * materials. */
typedef struct gh_physics_desc {
  int32_t n_materials;
  int32_t n_buckets;         /* nodes per table row = n_buckets + 1 */
  int32_t n_x_nodes;
  int32_t n_compositions;    /* 0 or C (see above) */
  const float *density;      /* [M] g/cm3 */
  const int32_t *composition;/* [M] composition of each material (n_compositions > 0), else NULL */
  const float *mu_total;     /* [T][NB+1] */
  const float *mu_majorant;  /* [NB+1]   1/mm, >= the mu of every material that occurs in a phantom
                                (checked by gh_add_simulation), at every node */
  const float *p_compton;    /* [T][NB+1] sigma_C / mu */
  const float *p_rayleigh;   /* [T][NB+1] (sigma_C + sigma_R) / mu */
  const float *mu_en;        /* [T][NB+1] (optional) */
  const float *ray_x;        /* [NX] q^2 nodes, 1/Angstrom^2 */
  const float *ray_cdf;      /* [T][NX] normalised int_0^x F^2 */
  const float *incoh_s;      /* [T][NX] S(q)/Z */
  float e_cut;               /* photons below e_cut deposit locally and stop */
  int32_t reserved1;
} gh_physics_desc;

/* Voxel phantom. Transport runs in the phantom frame: the box spans [0, n*spacing)
 * per axis, labels are x-fastest (index = x + nx*(y + ny*z)). The reference's label image
 * is uint16 (ImageVolume.create_label_image, geometry/volumes.py:1237): label_bytes = 2
 * takes it as it is; 1 takes uint8 labels. */
typedef struct gh_phantom_desc {
  int32_t nx, ny, nz;
  int32_t world_material;    /* material index of the world outside the box, -1 = vacuum */
  float spacing[3];
  float world_lo[3];         /* world box in the phantom frame (ignored for vacuum) */
  float world_hi[3];
  int32_t label_bytes;       /* 1 (uint8) or 2 (uint16) */
  const void *labels;        /* material index per voxel */
} gh_phantom_desc;

typedef struct gh_source_desc {
  int32_t pos_type, dir_type, ene_type, n_energy;
  float pos_param[4];        /* box: size x,y,z; sphere, surface_sphere: r; ellipsoid: a,b,c; disc: r; cylinder: r, half z */
  float pos_mat[12];         /* row-major 3x3 then translation: source frame -> phantom frame */
  float dir_mat[9];          /* row-major 3x3: source direction frame -> phantom frame */
  float cos_theta_min, cos_theta_max, phi_min, phi_max; /* iso: ranges in radians */
  float momentum[3];         /* GH_DIR_MOMENTUM, source direction frame */
  float focus[3];            /* GH_DIR_FOCUSED, phantom frame */
  float energy_mono;         /* mono energy; range: Emin */
  float energy_max;          /* range: Emax */
  float weight;              /* history weight (non-phsp) */
  int32_t phi_full;          /* 1 if phi range is the full [0, 2pi) */
  const float *energy_value; /* lines: energies; histogram: bin lower edges */
  const float *energy_width; /* histogram: bin widths */
  const float *energy_cdf;   /* normalised cumulative weights, last = 1 */
  int64_t phsp_n;            /* GH_POS_PHSP: number of rows */
  int64_t phsp_start;        /* first row used by history 0 */
  int32_t phsp_iso;          /* 1: replace directions by isotropic ones */
  int32_t phsp_on_device;    /* 1: phsp[] are device pointers, used in place: they must stay
                                allocated until the simulation is cleared or its source replaced */
  const float *phsp[8];      /* x, y, z, dx, dy, dz, E, weight (source frame) */
  int64_t vox_n;             /* GH_POS_VOXELS: alias-table entries (voxels with activity) */
  const float *vox_prob;     /* acceptance probability of entry i */
  const int32_t *vox_alias;  /* alias entry of entry i */
  const int32_t *vox_index;  /* voxel x + nx (y + ny z) of entry i */
  int32_t vox_size[3];       /* activity image size (x, y, z) */
  int32_t reserved2;
  float vox_spacing[3];      /* voxel size [mm]; source frame: image corner at the origin */
  float pos_sigma[2];        /* GH_POS_DISC: Gaussian spread added to x and y (GateGenericSource.cpp:362-366) */
  float energy_sigma;        /* GH_ENE_GAUSS: standard deviation */
  float weight_sigma;        /* >= 0: the history weight is weight + weight_sigma N(0,1)
                                (GateGenericSource.cpp:247-259); < 0: the weight itself */
  float dir_sigma[2];        /* GH_DIR_BEAM2D: angular sigma in x and y [rad], each in [0, 6] */
  float surface_axis[3];     /* GH_POS_SURFACE_SPHERE with GH_DIR_ISO / GH_DIR_COS: axis A of the
                                tangent frame at the local outward normal n: x' = unit(A x n),
                                y' = unit(x' x n), z' = n (G4SPSPosDistribution takes A = the
                                source rotation's third column); nonzero, normalised on upload */
  int32_t n_theta, n_phi;    /* GH_DIR_HISTOGRAM: bins of the theta and phi tables */
  const float *theta_value;  /* bin lower edges [rad], within [0, pi] */
  const float *theta_width;  /* bin widths (0: a point mass) */
  const float *theta_cdf;    /* normalised cumulative weights, last = 1 */
  const float *phi_value;    /* the same for phi, within [0, 2 pi] */
  const float *phi_width;
  const float *phi_cdf;
} gh_source_desc;

typedef struct gh_tally_desc {
  int32_t nx, ny, nz;
  int32_t flags;             /* GH_TALLY_* ; ALIGNED: the tally grid is the phantom grid */
  float corner[3];           /* lower corner of the tally grid in the phantom frame */
  float spacing[3];
  int32_t record_face;       /* -1 off, else 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z, 6 (GH_FACE_ALL) any face
                                (vacuum world only) */
  uint32_t record_capacity;  /* max photons recorded */
  int32_t detector_label_min;/* -1 off, else per-history energy in labels >= this value */
  float tle_energy_min;      /* GH_TALLY_TLE: photons at or below stop and deposit locally [MeV] */
  float tle_energy_max;      /* GH_TALLY_TLE: estimator for photons at or below, analog above [MeV] */
} gh_tally_desc;

typedef struct gh_launch {
  int32_t block;             /* threads per workgroup: a multiple of 64 up to 256, or 1024;
                                0 = automatic (one simulation whose tables fit in LDS: all
                                tables in LDS, 1024 threads; else 256 threads) */
  int32_t grid;              /* workgroups; 0 = automatic */
  int32_t refill_interval;   /* max step-phase iterations (2 Woodcock steps each) per loop */
  int32_t chunk;             /* histories per work-queue chunk (0: automatic, about 4 chunks per wave,
                                64..4096) */
  int32_t slow_min;          /* parked lanes (collision / exit / world / dead) that end a step phase (0: 52) */
  int32_t refill_min;        /* dead lanes that trigger a refill (clamped to slow_min; 0: 12) */
  int32_t scatter_tries;     /* Compton / Rayleigh sampler tries of a photon per slow phase before it
                                waits for the next one (0: 2); results do not depend on it */
} gh_launch;

typedef struct gh_counters {
  uint64_t histories, steps, deposits, compton, rayleigh, photo;
  uint64_t escaped, world_interactions, cut_kills, sq_evictions, recorded, kn_rejections;
  double e_deposited, e_outside_tally, e_escaped, e_world;
  double kernel_ms;
} gh_counters;

int gh_abi_version(void);
int gh_create(int device, gh_ctx **out);
void gh_destroy(gh_ctx *ctx);
const char *gh_last_error(const gh_ctx *ctx);

int gh_set_physics(gh_ctx *ctx, const gh_physics_desc *phys);

/* Adds one simulation (phantom + source + tally); returns its id in *sim_id. */
int gh_add_simulation(gh_ctx *ctx, const gh_phantom_desc *ph, const gh_source_desc *src,
                      const gh_tally_desc *tal, int32_t *sim_id);
int gh_clear_simulations(gh_ctx *ctx);
/* Replaces the source of a simulation (tallies are kept): several sources of one
 * simulation are run one after the other into the same tally, as the per-event source
 * selection of GateSourceManager::PrepareNextSource (GateSourceManager.cpp:313-386)
 * sums them. */
int gh_set_source(gh_ctx *ctx, int32_t sim_id, const gh_source_desc *src);
int gh_reset_tally(gh_ctx *ctx, int32_t sim_id);

/* Transports n_histories[s] histories of every simulation s in one persistent launch.
 * History h of simulation s draws from the Philox4x32-10 stream with key
 * (seed lo, seed hi) and counter (first_history[s] + h, sim s, block). */
int gh_run(gh_ctx *ctx, const uint64_t *n_histories, const uint64_t *first_history,
           uint64_t seed, const gh_launch *launch, gh_counters *out);

/* Autotuner (no reference counterpart: the reference has no launch shape to tune).
 * Runs n_trial[s] histories per simulation (RNG seed `seed`, keep it different from the
 * production seed) for each candidate launch (workgroup shape, slow-phase threshold,
 * refill threshold, step iterations), keeps the fastest in *best (kernel ms in *best_ms)
 * and restores every tally and recorder count to its state before the call. Per-history
 * detector energies (gh_get_event_edep) of earlier runs are not kept. */
int gh_autotune(gh_ctx *ctx, const uint64_t *n_trial, uint64_t seed, gh_launch *best, double *best_ms);

/* Copies a tally to host. Any pointer may be NULL. counts/dose images only if enabled. */
int gh_get_tally(gh_ctx *ctx, int32_t sim_id, double *edep, double *edep2, double *dose,
                 double *dose2, double *counts);
/* Recorded exit-face photons: 8 float arrays of length *n (x y z dx dy dz E w, phantom frame).
   out[k] may be host or device pointers (copied with the HIP runtime's default direction). */
int gh_get_phsp(gh_ctx *ctx, int32_t sim_id, float *const out[8], uint64_t capacity, uint64_t *n);
/* Empties the exit-face recorder of a simulation (tallies are kept), so that a long run can
 * be recorded and consumed in history chunks, as the hit collection of a PhaseSpaceActor is
 * emptied into its file every clear_every entries (FillToRootIfNeeded,
 * GatePhaseSpaceActor.cpp:241-247). */
int gh_reset_phsp(gh_ctx *ctx, int32_t sim_id);
/* Per-history detector energy (tally.detector_label_min >= 0), n_histories floats; out may be a
   host or a device pointer. */
int gh_get_event_edep(gh_ctx *ctx, int32_t sim_id, float *out, uint64_t n);

/* HU image -> uint16 labels: out[i] = labels_sorted[#{k : bins[k] <= hu[i]}]
 * (numpy.digitize with ascending bins, right=False). */
int gh_hu_to_labels(gh_ctx *ctx, const int16_t *hu, int64_t n_voxels, const double *bins,
                    int32_t n_bins, const uint16_t *labels_sorted, uint16_t *out);
/* The same rule for images of any other type (create_label_image takes any dtype,
 * volumes.py:1233-1239): values as FP64, which holds integers and float32 exactly. Up to
 * 4096 bins in either form. */
int gh_hu_to_labels_f64(gh_ctx *ctx, const double *hu, int64_t n_voxels, const double *bins,
                        int32_t n_bins, const uint16_t *labels_sorted, uint16_t *out);

/* ---- dense-layer GEMM for the learned models (no reference counterpart: the reference
 * runs its GAN source and ARF networks through torch, gansources.py:512, arfactors.py:388).
 * C[M,N] = A[M,K] * B[K,N] in bf16 with FP32 accumulation on the matrix cores;
 * A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]; C row pitch ldc. Device pointers,
 * `stream` is a hipStream_t (NULL = default stream). Epilogue modes: */
enum { GH_EPI_STORE = 0,      /* C32 = acc */
       GH_EPI_BIAS_RELU = 1,  /* C32 = acc + bias (optional), C16 = relu(acc + bias) */
       GH_EPI_BIAS = 2,       /* C32 = acc + bias and/or C16 = bf16(acc + bias) */
       GH_EPI_DRELU = 3 };    /* C32 = acc * relu'(zprev) (optional), C16 = bf16(same) */
/* In GH_EPI_STORE mode a long reduction over few output tiles (weight gradients) is split
 * across workgroups and the slices are summed in a fixed order (library-owned workspace,
 * allocated on first use, one per device and stream). */
int gh_gemm_bf16(const void *A, int64_t sam, int64_t sak, const void *B, int64_t sbk, int64_t sbn,
                 int32_t M, int32_t N, int32_t K, const float *bias, const float *zprev, int64_t ldz,
                 float *C32, void *C16, int64_t ldc, int32_t mode, void *stream);

/* Allocation-free variant of gh_gemm_bf16 (same arguments, kernels and results). No
 * counterpart in the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It calls no
 * hipMalloc or hipFree: the split-K slices of weight-gradient shapes go to a static device
 * array of GH_GEMM_NOALLOC_WS_BYTES that each GPU holds until the process ends. A split into
 * S slices is only made for fewer than 512 output tiles of 64 x 64 and with S <= 1024 / tiles,
 * so the slices (S x M x N x 4 bytes) never exceed 1024 x 64 x 64 x 4 bytes = 16 MiB, the
 * array's size; a larger request would fail with GH_ERANGE. Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams, other host threads) race.
 * gh_last_error(NULL) returns the message of this host thread's last failed call. */
#define GH_GEMM_NOALLOC_WS_BYTES (16u << 20)
int gh_gemm_bf16_noalloc(const void *A, int64_t sam, int64_t sak, const void *B, int64_t sbk, int64_t sbn,
                         int32_t M, int32_t N, int32_t K, const float *bias, const float *zprev, int64_t ldz,
                         float *C32, void *C16, int64_t ldc, int32_t mode, void *stream);

/* Inference of a whole dense network in one kernel (activations stay in LDS): layers
 * Z = X W_l + b_l, ReLU between layers, FP32 output of the last layer. X: M rows of bf16
 * (pitch ldx, first k_in <= 32 columns used); wt[l]: device bf16 weights transposed,
 * [N_l][K_l] with K contiguous, the first layer padded to K = 32 and the last to N = 16
 * (zeros); bias[l]: device FP32 [N_l]; hidden widths all equal to `hidden` (128 or 256);
 * 2..8 layers; Z: M x ldz FP32, columns 0..n_out-1 (n_out <= 16) written. Same values as
 * the layer-by-layer gh_gemm_bf16 path. (Replaces the torch forward of the generator in
 * GANSourceDefaultGenerator.generator, sources/gansources.py:661-669, and of the ARF
 * network in ARFActor.arf_build_image_from_projected_points, actors/arfactors.py:409-421.) */
int gh_mlp_forward_bf16(const void *X, int64_t ldx, int64_t M, int32_t n_layers, int32_t k_in, int32_t hidden,
                        int32_t n_out, const void *const *wt, const float *const *bias, float *Z, int64_t ldz,
                        void *stream);

/* Training: one Adam step (Kingma & Ba) over all parameter tensors of a network in one
 * launch. No counterpart in the reference's source: its networks come trained from the
 * external gaga / garf packages (loaded in sources/gansources.py and actors/arfactors.py);
 * here the learned exit-plane source and detector response are trained in place. Each tensor is rows x cols FP32,
 * row-major, with its gradient and both moments; entries outside the valid_rows x valid_cols
 * corner are set to zero (GEMM padding); param_bf16, if not NULL, receives bf16(param).
 * decay: L2 weight decay added to the gradient. step = 1, 2, ... (bias correction). */
typedef struct gh_adam_tensor {
  float *param;
  const float *grad;
  float *m, *v;
  void *param_bf16;
  int32_t rows, cols, valid_rows, valid_cols;
  float decay;
  int32_t reserved;
} gh_adam_tensor;
int gh_adam_step(const gh_adam_tensor *tensors, int32_t n_tensors, float lr, float beta1, float beta2, float eps,
                 int32_t step, void *stream);

/* Learned exit-plane source (replaces the generator call of GANSourceDefaultGenerator.generator,
 * sources/gansources.py:644-684, and its copy/move of the keys, :686-757).
 * gh_gaussian_bf16: out[r*ld + c] = N(0,1) noise in bf16 for rows first_row .. first_row+n_rows-1
 * (c < n_cols; columns n_cols..ld-1 are zero). Row r depends only on (seed, r). */
int gh_gaussian_bf16(uint64_t seed, uint64_t first_row, int64_t n_rows, int32_t n_cols, int32_t ld, void *out,
                     void *stream);
/* Generator outputs Z (n rows, pitch ldz >= 6, FP32: E, x, y, u', v', w' in Gaussianised space)
 * -> 8 device float arrays x y z u v w E weight. qtab: 6 x n_knots quantile values at the
 * Gaussian knots -g_max + 2 g_max k/(n_knots-1). (u', v', w') is the direction in the frame whose
 * pole is the ray from focal[3] through (x, y, plane_z); the output position is
 * (x, y, plane_z) + offset[3]. */
int gh_exitplane_decode(const float *Z, int64_t ldz, int64_t n, const float *qtab, int32_t n_knots, float g_max,
                        const float focal[3], float plane_z, const float offset[3], float weight,
                        float *const out[8], void *stream);

/* SPECT heads for the learned collimator-detector response (ARFActor, actors/arfactors.py:140-463:
 * apply / arf_build_image_from_projected_points, and GateARFActor::SteppingAction,
 * GateARFActor.cpp:60-105) and for the training set of ARFTrainingDatasetActor
 * (GateARFTrainingDatasetActor.cpp:85-140: E, theta = acos(d_normal), phi = atan2(d_y, d_x),
 * window id + 1 with 0 = not detected). */
typedef struct gh_spect_desc {
  float center[3];        /* point on the gantry axis (axis = +z), phantom frame [mm] */
  float radius;           /* axis to collimator face [mm] */
  int32_t n_proj;         /* heads (projections), head k at angle0 + 2 pi k / n_proj */
  int32_t nx, ny;         /* image columns (tangential), rows (axial) */
  float angle0;           /* [rad] */
  float spacing[2];       /* pixel size (tangential, axial) [mm] */
  float cos_theta_max;    /* photons with d.normal <= cos_theta_max do not reach the crystal */
  float crystal_depth;    /* collimator face -> plane where pixels are assigned [mm] */
  float pitch;            /* hexagonal hole lattice pitch of the detector model [mm] (0: none) */
  float patch_origin[3];  /* detector-model frame: lattice point where reduced entries land */
  float norm[6];          /* NN input row = (E n0 + n1, theta n2 + n3, phi n4 + n5, 0...) */
} gh_spect_desc;

/* Photons that left the phantom (8 device arrays x y z u v w E weight, phantom frame, vacuum
 * outside) -> (photon, head) pairs: NN input rows x16 (capacity x 8 bf16, may be NULL), pixel
 * index (head*ny + iy)*nx + ix, weight, optionally (theta, phi) and the photon in the detector
 * model frame (det_rec, may be NULL). angle_cs: device array of n_proj (cos, sin) pairs.
 * *count (device, zeroed by the caller) receives the number of pairs; rows beyond capacity are
 * counted but not written. */
int gh_spect_pairs(const float *const rec[8], int64_t n, const gh_spect_desc *det, const float *angle_cs,
                   int64_t capacity, void *x16, int32_t *pix, float *wgt, float *const det_rec[8],
                   float *theta_phi, unsigned long long *count, void *stream);
/* Energy blurring laws of GateDigitizerBlurringActor (GateDigitizerBlurringActor.cpp:75-94;
 * parameters as set by DigitizerBlurringActor.set_param_*, actors/digitizers.py:524-570):
 *   GAUSSIAN        sigma fixed [MeV]
 *   INVERSE_SQUARE  FWHM = resolution sqrt(reference E)       (FWHM/E = resolution at reference)
 *   LINEAR          FWHM = (slope (E - reference) + resolution) E
 * with sigma = FWHM / (2 sqrt(2 ln 2)). NONE leaves the energy as it is. */
#define GH_BLUR_NONE 0
#define GH_BLUR_GAUSSIAN 1
#define GH_BLUR_INVERSE_SQUARE 2
#define GH_BLUR_LINEAR 3
typedef struct {
  int32_t law;        /* GH_BLUR_* */
  float sigma;        /* GAUSSIAN [MeV] */
  float resolution;   /* INVERSE_SQUARE, LINEAR: FWHM / E at the reference energy */
  float reference;    /* INVERSE_SQUARE, LINEAR: reference energy [MeV] */
  float slope;        /* LINEAR [1/MeV] */
} gh_blur;

/* Deposited energy per event (device) -> window class (0 none, k = 1..n_win): the energy is
 * blurred by `blur`, then the first window k with lo <= E < hi is taken
 * (GateDigitizerEnergyWindowsActor::ApplyThreshold, GateDigitizerEnergyWindowsActor.cpp:117-131;
 * with overlapping windows the caller runs one window per call: the blurred energy of an
 * event is the same in every call). win: host array of n_win (lo, hi) pairs in MeV. Events
 * with no energy get class 0. Event i draws its noise from the key (seed, first + i). */
int gh_energy_windows(const float *edep, int64_t n, uint64_t seed, uint64_t first, const gh_blur *blur,
                      const float *win, int32_t n_win, int32_t *cls, void *stream);
/* Projection images [head][window][ny][nx] (FP64, device) += weight x softmax(logits)[1..n_win]
 * per pair (ARF), or += weight at the simulated window (plain); img2 (optional) += squares. */
int gh_arf_accumulate(const float *logits, int64_t ldl, int64_t m, const int32_t *pix, const float *wgt,
                      int32_t n_win, int32_t plane, double *img, double *img2, void *stream);
int gh_counts_accumulate(const int32_t *cls, const int32_t *pix, const float *wgt, int64_t m, int32_t n_win,
                         int32_t plane, double *img, double *img2, void *stream);

/* Walker alias table of n non-negative weights (host arrays; no device needed): the voxel
 * source samples entry i uniformly, keeps it with probability prob[i], else takes alias[i]
 * (GateSPSVoxelsPosDistribution::VGenerateOne, GateSPSVoxelsPosDistribution.cpp:24-52,
 * samples the same distribution with nested CDF searches). */
int gh_alias_table(const double *w, int64_t n, float *prob, int32_t *alias);

/* Device properties used by the autotuner and reports. */
int gh_device_info(gh_ctx *ctx, int32_t *n_cu, int32_t *lds_bytes, char *name, int32_t name_len);

#ifdef __cplusplus
}
#endif
#endif
