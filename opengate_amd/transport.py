"""Host-side descriptors of one simulation and the GPU engine handle.

`PhantomSpec`, `SourceSpec` and `TallySpec` hold numpy buffers and produce the C
descriptors of include/gatehip.h; `GpuEngine` owns a gh_ctx. The same specs feed the
CPU oracle in tests, so both sides always see identical FP32 inputs.
"""
import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _abi as A


def _f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


@dataclass
class PhantomSpec:
    """Label volume in the phantom frame: box [0, n*spacing), labels[z, y, x]. Labels are
    uint16, as the reference's label image (geometry/volumes.py:1237), or uint8."""

    labels: np.ndarray  # uint8 or uint16 (nz, ny, nx)
    spacing: tuple  # mm (sx, sy, sz)
    world_material: int = -1  # -1: vacuum outside the box
    world_lo: tuple = (0.0, 0.0, 0.0)  # world box in the phantom frame
    world_hi: tuple = (0.0, 0.0, 0.0)

    def __post_init__(self):
        lab = np.asarray(self.labels)
        if lab.dtype != np.uint16 and lab.dtype != np.uint8:
            if lab.size and (lab.min() < 0 or lab.max() > 65535):
                raise ValueError("labels must lie in 0..65535")
            lab = lab.astype(np.uint16 if lab.size and lab.max() > 255 else np.uint8)
        self.labels = np.ascontiguousarray(lab)
        assert self.labels.ndim == 3

    @property
    def shape_xyz(self):
        nz, ny, nx = self.labels.shape
        return nx, ny, nz

    def desc(self):
        nx, ny, nz = self.shape_xyz
        d = A.PhantomDesc()
        d.nx, d.ny, d.nz = nx, ny, nz
        d.world_material = int(self.world_material)
        d.spacing = _f3(self.spacing)
        d.world_lo = _f3(self.world_lo)
        d.world_hi = _f3(self.world_hi)
        d.label_bytes = self.labels.dtype.itemsize
        d.labels = self.labels.ctypes.data_as(C.c_void_p)
        return d


def spectrum_lines(energies, weights):
    """spectrum_discrete: normalised cumulative weights (GateGenericSource.cpp:544-573)."""
    w = np.cumsum(np.asarray(weights, dtype=np.float64))
    cdf = (w / w[-1]).astype(np.float32)
    cdf[-1] = 1.0
    return np.asarray(energies, dtype=np.float32), None, cdf


def spectrum_histogram(edges, weights):
    """spectrum_histogram with no interpolation: cumulative w_i * dE_i
    (GateGenericSource.cpp:583-609), energy uniform inside the chosen bin."""
    e = np.asarray(edges, dtype=np.float64)
    w = np.asarray(weights, dtype=np.float64)
    if len(e) != len(w) + 1:
        raise ValueError("spectrum_energy_bin_edges must have one more element than spectrum_weights")
    acc = np.cumsum(w * np.diff(e))
    cdf = (acc / acc[-1]).astype(np.float32)
    cdf[-1] = 1.0
    return e[:-1].astype(np.float32), np.diff(e).astype(np.float32), cdf


def angle_histogram(edges, weights, lo, hi):
    """Piecewise-uniform table (lower edges, widths, cdf) of a "user" angle histogram:
    GateGenericSource.cpp:437-466 gives G4SPSAngDistribution the points (edges[0], 0),
    (edges[i], weights[i-1]), whose inverse cumulative is linear between points
    (G4PhysicsVector::GetEnergy), so bin i is uniform on (edges[i], edges[i+1]] with
    probability weights[i]. G4SPSAngDistribution::GenerateUserDefFlux draws again while
    the angle lies outside [lo, hi] ([0, pi] for theta, [0, 2 pi] for phi): the bins are
    clipped to that range with their weights scaled by the part kept, which is the same law."""
    e = np.asarray(edges, dtype=np.float64)
    w = np.asarray(weights, dtype=np.float64)
    if len(e) != len(w) + 1 or len(w) == 0:
        raise ValueError("an angle histogram needs one more angle than weights")
    if np.any(np.diff(e) <= 0):
        raise ValueError("histogram angles must increase strictly")
    if np.any(w < 0) or not np.all(np.isfinite(w)) or not np.all(np.isfinite(e)):
        raise ValueError("histogram weights must be finite and >= 0")
    a, b = np.maximum(e[:-1], lo), np.minimum(e[1:], hi)
    keep = b > a
    frac = np.where(keep, (b - a) / np.diff(e), 0.0)
    wk = (w * frac)[keep]
    if wk.sum() <= 0:
        raise ValueError(f"the angle histogram has no weight inside [{lo}, {hi}]")
    acc = np.cumsum(wk)
    cdf = (acc / acc[-1]).astype(np.float32)
    cdf[-1] = 1.0
    return a[keep].astype(np.float32), (b - a)[keep].astype(np.float32), cdf


def energy_user_histogram(energies, weights):
    """Energy type "histogram" (GateGenericSource.cpp:530-542: G4SPSEneDistribution "User"
    with the points (energies[i], weights[i])). G4SPSEneDistribution::GenerateUserEnergies
    normalises the running sum of the weights and inverts it linearly between points: the
    first weight is a point mass at energies[0], weight i > 0 is spread uniformly on
    (energies[i-1], energies[i]]. Returned as a piecewise-uniform table (lower edges,
    widths, cdf) whose first bin has width 0."""
    e = np.asarray(energies, dtype=np.float64)
    w = np.asarray(weights, dtype=np.float64)
    if len(e) != len(w) or len(e) == 0:
        raise ValueError("histogram_energy and histogram_weight must have the same length")
    if np.any(np.diff(e) <= 0):
        raise ValueError("histogram_energy must increase strictly")
    if np.any(w < 0) or not np.all(np.isfinite(w)) or w.sum() <= 0:
        raise ValueError("histogram_weight must be finite, >= 0 and not all 0")
    lo = np.concatenate([[e[0]], e[:-1]])
    width = np.concatenate([[0.0], np.diff(e)])
    acc = np.cumsum(w)
    cdf = (acc / acc[-1]).astype(np.float32)
    cdf[-1] = 1.0
    return lo.astype(np.float32), width.astype(np.float32), cdf


def voxel_alias_table(activity):
    """(flat voxel indices with activity > 0, prob, alias) for an activity image (z, y, x):
    the Walker alias table of the library (gh_alias_table, a host-only routine)."""
    a = np.ascontiguousarray(activity, dtype=np.float64).ravel()
    if np.any(~np.isfinite(a)) or np.any(a < 0):
        raise ValueError("activity must be finite and >= 0")
    idx = np.flatnonzero(a > 0).astype(np.int32)
    if len(idx) == 0:
        raise ValueError("activity image is empty")
    w = np.ascontiguousarray(a[idx])
    prob = np.zeros(len(idx), np.float32)
    alias = np.zeros(len(idx), np.int32)
    rc = A.load_library().gh_alias_table(A.dptr(w), len(idx), A.fptr(prob), alias.ctypes.data_as(C.POINTER(C.c_int32)))
    if rc != 0:
        raise ValueError(f"gh_alias_table failed ({rc})")
    return idx, prob, alias


@dataclass
class SourceSpec:
    pos_type: str = "point"
    pos_param: tuple = (0.0, 0.0, 0.0, 0.0)
    pos_rot: np.ndarray = field(default_factory=lambda: np.eye(3))  # source frame -> phantom frame
    pos_trans: tuple = (0.0, 0.0, 0.0)
    dir_type: str = "iso"
    dir_rot: np.ndarray = field(default_factory=lambda: np.eye(3))
    theta: tuple = (0.0, np.pi)
    phi: tuple = (0.0, 2 * np.pi)
    momentum: tuple = (0.0, 0.0, 1.0)
    focus: tuple = (0.0, 0.0, 0.0)  # phantom frame
    ene_type: str = "mono"
    energy: float = 0.1405  # mono / range min (MeV)
    energy_max: float = 0.0
    energy_value: Optional[np.ndarray] = None
    energy_width: Optional[np.ndarray] = None
    energy_cdf: Optional[np.ndarray] = None
    weight: float = 1.0
    phsp: Optional[list] = None  # 8 float32 arrays x y z dx dy dz E w (source frame), numpy or device tensors
    phsp_start: int = 0
    phsp_iso: bool = False
    activity: Optional[np.ndarray] = None  # voxels: activity image (z, y, x), image corner at the source origin
    activity_spacing: tuple = (1.0, 1.0, 1.0)  # voxels: voxel size (x, y, z) mm
    pos_sigma: tuple = (0.0, 0.0)  # disc: Gaussian spread added to x and y (mm)
    energy_sigma: float = 0.0  # gauss: standard deviation (MeV)
    weight_sigma: float = -1.0  # >= 0: Gaussian history weight around `weight`
    dir_sigma: tuple = (0.0, 0.0)  # beam2d: angular sigma in x and y (rad)
    surface_axis: tuple = (0.0, 0.0, 1.0)  # surface_sphere iso / cos: axis of the tangent frame
    theta_table: Optional[tuple] = None  # histogram direction: (lower edges, widths, cdf) of theta
    phi_table: Optional[tuple] = None  # ... and of phi

    def _phsp_on_device(self):
        return self.phsp is not None and hasattr(self.phsp[0], "data_ptr")

    def max_energy(self):
        if self.pos_type == "phsp":
            if self._phsp_on_device():
                return float(self.phsp[6].max()) if self.phsp[6].numel() else 0.0
            return float(np.max(self.phsp[6])) if len(self.phsp[6]) else 0.0
        if self.ene_type == "mono":
            return float(self.energy)
        if self.ene_type == "range":
            return float(self.energy_max)
        if self.ene_type == "gauss":
            # the sampler cuts the Gaussian tail at 8 standard deviations
            return float(self.energy) + 8.0 * float(self.energy_sigma)
        if self.ene_type == "spectrum_histogram":
            return float(self.energy_value[-1] + self.energy_width[-1])
        return float(np.max(self.energy_value))

    def desc(self):
        d = A.SourceDesc()
        d.pos_type = A.POS[self.pos_type]
        d.dir_type = A.DIR[self.dir_type]
        d.ene_type = A.ENE[self.ene_type]
        pp = list(self.pos_param) + [0.0] * (4 - len(self.pos_param))
        d.pos_param = (C.c_float * 4)(*[float(x) for x in pp[:4]])
        R = np.asarray(self.pos_rot, dtype=np.float64).reshape(3, 3)
        d.pos_mat = (C.c_float * 12)(*[float(x) for x in R.ravel()], *[float(x) for x in self.pos_trans])
        D = np.asarray(self.dir_rot, dtype=np.float64).reshape(3, 3)
        d.dir_mat = (C.c_float * 9)(*[float(x) for x in D.ravel()])
        d.cos_theta_min = float(np.cos(self.theta[0]))
        d.cos_theta_max = float(np.cos(self.theta[1]))
        d.phi_min = float(self.phi[0])
        d.phi_max = float(self.phi[1])
        d.phi_full = int(abs((self.phi[1] - self.phi[0]) - 2 * np.pi) < 1e-12)
        d.momentum = _f3(self.momentum)
        d.focus = _f3(self.focus)
        d.energy_mono = float(self.energy)
        d.energy_max = float(self.energy_max)
        d.weight = float(self.weight)
        d.pos_sigma = (C.c_float * 2)(float(self.pos_sigma[0]), float(self.pos_sigma[1]))
        d.energy_sigma = float(self.energy_sigma)
        d.weight_sigma = float(self.weight_sigma)
        d.dir_sigma = (C.c_float * 2)(float(self.dir_sigma[0]), float(self.dir_sigma[1]))
        d.surface_axis = _f3(self.surface_axis)
        if self.dir_type == "histogram":
            if self.theta_table is None or self.phi_table is None:
                raise ValueError("a histogram direction needs theta_table and phi_table")
            self._ang = [np.ascontiguousarray(a, dtype=np.float32) for a in (*self.theta_table, *self.phi_table)]
            d.n_theta, d.n_phi = len(self._ang[0]), len(self._ang[3])
            d.theta_value, d.theta_width, d.theta_cdf, d.phi_value, d.phi_width, d.phi_cdf = (
                A.fptr(a) for a in self._ang)
        if self.ene_type in ("spectrum_discrete", "spectrum_histogram"):
            self._ev = np.ascontiguousarray(self.energy_value, dtype=np.float32)
            self._ec = np.ascontiguousarray(self.energy_cdf, dtype=np.float32)
            d.n_energy = len(self._ev)
            d.energy_value = A.fptr(self._ev)
            d.energy_cdf = A.fptr(self._ec)
            if self.ene_type == "spectrum_histogram":
                self._ew = np.ascontiguousarray(self.energy_width, dtype=np.float32)
                d.energy_width = A.fptr(self._ew)
        if self.pos_type == "voxels":
            if self.activity is None or self.activity.ndim != 3:
                raise ValueError("a voxel source needs a 3-D activity image (z, y, x)")
            if not hasattr(self, "_vox") or self._vox[0] is not self.activity:
                self._vox = (self.activity,) + voxel_alias_table(self.activity)
            _, idx, prob, alias = self._vox
            nz, ny, nx = self.activity.shape
            d.vox_n = len(idx)
            d.vox_prob = A.fptr(prob)
            d.vox_alias = alias.ctypes.data_as(C.POINTER(C.c_int32))
            d.vox_index = idx.ctypes.data_as(C.POINTER(C.c_int32))
            d.vox_size = (C.c_int32 * 3)(nx, ny, nz)
            d.vox_spacing = _f3(self.activity_spacing)
        if self.pos_type == "phsp":
            if self._phsp_on_device():
                self._ph = [a.contiguous().float() for a in self.phsp]
                d.phsp_n = int(self._ph[0].numel())
                d.phsp_on_device = 1
                d.phsp = (A.FP * 8)(*[C.cast(C.c_void_p(a.data_ptr()), A.FP) for a in self._ph])
            else:
                self._ph = [np.ascontiguousarray(a, dtype=np.float32) for a in self.phsp]
                d.phsp_n = len(self._ph[0])
                d.phsp = (A.FP * 8)(*[A.fptr(a) for a in self._ph])
            d.phsp_start = int(self.phsp_start)
            d.phsp_iso = int(bool(self.phsp_iso))
        return d


@dataclass
class TallySpec:
    size: tuple  # (nx, ny, nz)
    spacing: tuple = (1.0, 1.0, 1.0)
    corner: tuple = (0.0, 0.0, 0.0)  # phantom frame
    aligned: bool = True
    edep2: bool = True
    dose: bool = False  # per-deposit dose (needed only when not aligned)
    dose2: bool = False
    counts: bool = False
    record_face: int = -1
    record_capacity: int = 0
    detector_label_min: int = -1
    tle: bool = False  # track-length (Woodcock collision) kerma estimator instead of analog deposits
    tle_energy_min: float = 0.0  # MeV
    tle_energy_max: float = float("inf")  # MeV

    def flags(self):
        f = 0
        if self.edep2:
            f |= A.TALLY_EDEP2
        if self.dose:
            f |= A.TALLY_DOSE
        if self.dose2:
            f |= A.TALLY_DOSE2
        if self.counts:
            f |= A.TALLY_COUNTS
        if self.aligned:
            f |= A.TALLY_ALIGNED
        if self.tle:
            f |= A.TALLY_TLE
        return f

    @property
    def n_voxels(self):
        return int(np.prod(self.size))

    def desc(self):
        d = A.TallyDesc()
        d.nx, d.ny, d.nz = [int(x) for x in self.size]
        d.flags = self.flags()
        d.corner = _f3(self.corner)
        d.spacing = _f3(self.spacing)
        d.record_face = int(self.record_face)
        d.record_capacity = int(self.record_capacity)
        d.detector_label_min = int(self.detector_label_min)
        d.tle_energy_min = float(self.tle_energy_min)
        d.tle_energy_max = float(min(self.tle_energy_max, 3.0e38))
        return d


def physics_desc(tables):
    """C descriptor for an opengate_amd.physics.xs.PhysicsTables (keeps arrays alive on the desc)."""
    d = A.PhysicsDesc()
    d.n_materials = tables.n_materials
    d.n_buckets = tables.n_buckets
    d.n_x_nodes = len(tables.ray_x)
    keep = []

    def f(a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        keep.append(a)
        return A.fptr(a)

    d.density = f(tables.density)
    comp = getattr(tables, "composition", None)
    if comp is not None:
        c = np.ascontiguousarray(comp, dtype=np.int32)
        keep.append(c)
        d.n_compositions = int(tables.n_compositions)
        d.composition = c.ctypes.data_as(C.POINTER(C.c_int32))
    d.mu_total = f(tables.mu_total)
    d.mu_majorant = f(tables.mu_majorant)
    d.p_compton = f(tables.p_compton)
    d.p_rayleigh = f(tables.p_rayleigh)
    d.mu_en = f(tables.mu_en)
    d.ray_x = f(tables.ray_x)
    d.ray_cdf = f(tables.ray_cdf)
    d.incoh_s = f(tables.incoh_s)
    d.e_cut = float(tables.e_cut)
    d._keep = keep
    return d


class GhError(RuntimeError):
    pass


class GpuEngine:
    """Owns one gh_ctx on one device."""

    def __init__(self, device=0):
        self.lib = A.load_library()
        h = C.c_void_p()
        rc = self.lib.gh_create(int(device), C.byref(h))
        if rc != A.GH_OK:
            raise GhError(f"gh_create(device={device}) failed with code {rc} (no usable HIP device?)")
        self.ctx = h
        self.device = int(device)
        self.sims = []
        self.tables = None

    def _check(self, rc, what):
        if rc != A.GH_OK:
            msg = self.lib.gh_last_error(self.ctx).decode()
            raise GhError(f"{what}: {msg} (code {rc})")

    def close(self):
        if self.ctx:
            self.lib.gh_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        n_cu, lds = C.c_int32(), C.c_int32()
        name = C.create_string_buffer(64)
        self._check(self.lib.gh_device_info(self.ctx, C.byref(n_cu), C.byref(lds), name, 64), "gh_device_info")
        return {"n_cu": n_cu.value, "lds_bytes": lds.value, "arch": name.value.decode()}

    def set_physics(self, tables):
        self._pd = physics_desc(tables)
        self._check(self.lib.gh_set_physics(self.ctx, C.byref(self._pd)), "gh_set_physics")
        self.tables = tables

    def clear(self):
        self._check(self.lib.gh_clear_simulations(self.ctx), "gh_clear_simulations")
        self.sims = []

    def add_simulation(self, phantom: PhantomSpec, source: SourceSpec, tally: TallySpec):
        if self.tables is not None and source.max_energy() >= self.tables.e_max:
            raise GhError(f"source energy {source.max_energy()} MeV beyond the physics tables ({self.tables.e_max} MeV)")
        pd, sd, td = phantom.desc(), source.desc(), tally.desc()
        sid = C.c_int32()
        self._check(self.lib.gh_add_simulation(self.ctx, C.byref(pd), C.byref(sd), C.byref(td), C.byref(sid)),
                    "gh_add_simulation")
        self.sims.append((phantom, source, tally))
        return sid.value

    def set_source(self, sim_id, source: SourceSpec):
        if self.tables is not None and source.max_energy() >= self.tables.e_max:
            raise GhError(f"source energy {source.max_energy()} MeV beyond the physics tables ({self.tables.e_max} MeV)")
        sd = source.desc()
        self._check(self.lib.gh_set_source(self.ctx, int(sim_id), C.byref(sd)), "gh_set_source")
        ph, _, tal = self.sims[sim_id]
        self.sims[sim_id] = (ph, source, tal)

    def reset_tally(self, sim_id):
        self._check(self.lib.gh_reset_tally(self.ctx, int(sim_id)), "gh_reset_tally")

    def run(self, n_histories, seed, first_history=None, block=0, grid=0, refill_interval=64, chunk=0,
            slow_min=48, refill_min=12, scatter_tries=0):
        n = np.ascontiguousarray(np.atleast_1d(n_histories), dtype=np.uint64)
        if len(n) != len(self.sims):
            raise GhError("one history count per simulation is required")
        f = np.zeros_like(n) if first_history is None else np.ascontiguousarray(np.atleast_1d(first_history), dtype=np.uint64)
        lp = A.Launch(int(block), int(grid), int(refill_interval), int(chunk), int(slow_min), int(refill_min),
                      int(scatter_tries))
        out = A.Counters()
        self._check(self.lib.gh_run(self.ctx, n.ctypes.data_as(C.POINTER(C.c_uint64)),
                                    f.ctypes.data_as(C.POINTER(C.c_uint64)), C.c_uint64(int(seed)),
                                    C.byref(lp), C.byref(out)), "gh_run")
        return out.as_dict()

    def autotune(self, n_trial, seed=0x5EED_A070):
        """Best launch parameters for the current simulations (tallies are left untouched).
        Returns (kwargs for run(), kernel ms of the best trial)."""
        n = np.ascontiguousarray(np.atleast_1d(n_trial), dtype=np.uint64)
        if len(n) != len(self.sims):
            raise GhError("one trial history count per simulation is required")
        best = A.Launch()
        ms = C.c_double()
        self._check(self.lib.gh_autotune(self.ctx, n.ctypes.data_as(C.POINTER(C.c_uint64)), C.c_uint64(int(seed)),
                                         C.byref(best), C.byref(ms)), "gh_autotune")
        kw = {"block": best.block, "grid": best.grid, "refill_interval": best.refill_interval, "chunk": best.chunk,
              "slow_min": best.slow_min, "refill_min": best.refill_min, "scatter_tries": best.scatter_tries}
        return kw, ms.value

    def get_tally(self, sim_id=0):
        _, _, tal = self.sims[sim_id]
        nz_ny_nx = tuple(reversed([int(x) for x in tal.size]))
        res = {"edep": np.zeros(nz_ny_nx)}
        if tal.edep2:
            res["edep2"] = np.zeros(nz_ny_nx)
        if tal.dose:
            res["dose"] = np.zeros(nz_ny_nx)
        if tal.dose2:
            res["dose2"] = np.zeros(nz_ny_nx)
        if tal.counts:
            res["counts"] = np.zeros(nz_ny_nx)
        self._check(self.lib.gh_get_tally(self.ctx, int(sim_id), A.dptr(res["edep"]), A.dptr(res.get("edep2")),
                                          A.dptr(res.get("dose")), A.dptr(res.get("dose2")),
                                          A.dptr(res.get("counts"))), "gh_get_tally")
        return res

    def get_phsp(self, sim_id=0, device=False):
        """Recorded exit-face photons: 8 arrays (x, y, z, u, v, w, E, weight) and the number of
        photons that reached the face (may exceed the capacity). device=True returns device
        tensors (copied on the GPU) instead of numpy arrays."""
        _, _, tal = self.sims[sim_id]
        n = C.c_uint64()
        self._check(self.lib.gh_get_phsp(self.ctx, int(sim_id), (A.FP * 8)(), 0, C.byref(n)), "gh_get_phsp")
        m = min(int(n.value), int(tal.record_capacity))
        if device:
            import torch

            arrs = [torch.empty(m, dtype=torch.float32, device=f"cuda:{self.device}") for _ in range(8)]
            ptrs = [C.cast(C.c_void_p(a.data_ptr()), A.FP) for a in arrs]
        else:
            arrs = [np.zeros(m, dtype=np.float32) for _ in range(8)]
            ptrs = [A.fptr(a) for a in arrs]
        self._check(self.lib.gh_get_phsp(self.ctx, int(sim_id), (A.FP * 8)(*ptrs), m, C.byref(n)), "gh_get_phsp")
        return arrs, int(n.value)

    def reset_phsp(self, sim_id=0):
        """Empty the exit-face recorder (tallies are kept)."""
        self._check(self.lib.gh_reset_phsp(self.ctx, int(sim_id)), "gh_reset_phsp")

    def get_event_edep(self, sim_id, n, device=False):
        """Energy deposited in the detector labels by each of the first n histories of the
        last run (numpy, or a device tensor with device=True)."""
        if device:
            import torch

            out = torch.empty(int(n), dtype=torch.float32, device=f"cuda:{self.device}")
            ptr = C.cast(C.c_void_p(out.data_ptr()), A.FP)
        else:
            out = np.zeros(int(n), dtype=np.float32)
            ptr = A.fptr(out)
        self._check(self.lib.gh_get_event_edep(self.ctx, int(sim_id), ptr, int(n)), "gh_get_event_edep")
        return out

    def hu_to_labels(self, hu, bins, labels_sorted):
        """label = labels_sorted[digitize(hu, bins)] on the GPU: int16 images as they are,
        any other type as FP64 (exact for integers and float32)."""
        hu = np.asarray(hu)
        bins = np.ascontiguousarray(bins, dtype=np.float64)
        lut = np.ascontiguousarray(labels_sorted, dtype=np.uint16)
        if len(lut) != len(bins) + 1:
            raise GhError("labels_sorted must have len(bins)+1 entries")
        out = np.zeros(hu.shape, dtype=np.uint16)
        outp = out.ctypes.data_as(C.POINTER(C.c_uint16))
        lutp = lut.ctypes.data_as(C.POINTER(C.c_uint16))
        if hu.dtype == np.int16:
            hu = np.ascontiguousarray(hu)
            self._check(self.lib.gh_hu_to_labels(self.ctx, hu.ctypes.data_as(C.POINTER(C.c_int16)), hu.size,
                                                 A.dptr(bins), len(bins), lutp, outp), "gh_hu_to_labels")
        else:
            if hu.dtype.kind not in "iuf" or (hu.dtype.kind in "iu" and hu.dtype.itemsize > 4):
                raise TypeError(f"HU image of type {hu.dtype}: integers up to 32 bits or floats")
            h = np.ascontiguousarray(hu, dtype=np.float64)
            self._check(self.lib.gh_hu_to_labels_f64(self.ctx, A.dptr(h), h.size, A.dptr(bins), len(bins), lutp,
                                                     outp), "gh_hu_to_labels_f64")
        return out
