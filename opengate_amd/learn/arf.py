"""Learned collimator-detector response (ARF) for SPECT projections.

A 4-layer dense network (width 128, BF16 weights, FP32 accumulation) maps a photon
(E, theta, phi) arriving at a collimator face to the probabilities of being detected in
each energy window (class 0 = not detected), as ARFActor does with a garf network
(actors/arfactors.py:140-463; training set layout of GateARFTrainingDatasetActor,
GateARFTrainingDatasetActor.cpp:85-140). Everything on the GPU runs in the library's
kernels: the detector model transport that labels the training set (gh_run with a
device phase-space source and per-event crystal energy), the energy blur and windows
(gh_energy_windows), the network (gh_gemm_bf16, cross-entropy gradient from PyTorch
tensors), the photon-to-head pairing (gh_spect_pairs) and the image accumulation
(gh_arf_accumulate / gh_counts_accumulate).

The plain counterpart of the learned response transports every (photon, head) pair
through the voxelised detector model (hexagonal-hole lead collimator over a NaI layer)
at its true position in the hole lattice, and bins the simulated window the same way.
"""
import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from .. import _abi as A
from .mlp import MLP

SIZES = [3, 128, 128, 128, 4]
# Tc-99m triple energy window (MeV): lower scatter, photopeak 140.5 keV +-10 %, upper scatter
WINDOWS = ((0.114, 0.126), (0.126, 0.154), (0.154, 0.166))
E_RANGE = (0.010, 0.160)
# materials of the head model, by label: gap, collimator, crystal
HEAD_MATERIALS = ("G4_AIR", "G4_Pb", "G4_SODIUM_IODIDE")


@dataclass(frozen=True)
class Blur:
    """Energy blurring law of the digitizer (GateDigitizerBlurringActor.cpp:75-94): "none",
    "Gaussian" (fixed sigma, MeV), "InverseSquare" (FWHM = resolution sqrt(reference E)) or
    "Linear" (FWHM = (slope (E - reference) + resolution) E)."""

    law: str = "InverseSquare"
    sigma: float = 0.0
    resolution: float = 0.095
    reference: float = 0.1405
    slope: float = 0.0

    def desc(self):
        if self.law not in A.BLUR_LAWS:
            raise ValueError(f"unknown blurring law {self.law!r} (one of {tuple(A.BLUR_LAWS)})")
        return A.Blur(A.BLUR_LAWS[self.law], float(self.sigma), float(self.resolution), float(self.reference),
                      float(self.slope))


# NaI at 140.5 keV: 9.5 % FWHM, FWHM ~ sqrt(E)
BLUR = Blur()


def windows_overlap(windows):
    """True when two half-open windows [lo, hi) share energies."""
    w = sorted((float(lo), float(hi)) for lo, hi in windows)
    return any(w[k + 1][0] < w[k][1] for k in range(len(w) - 1))


def _torch():
    import torch

    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p()


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


# ---------------------------------------------------------------------- detector model
@dataclass
class HeadModel:
    """Voxelised detector head: hexagonal holes (flat-to-flat `hole`, lattice `pitch`, first
    lattice vector along x) in a lead block of `length`, over an air gap and a NaI layer.
    Frame: box [0, patch) x [0, patch) x [0, nz*voxel); photons come from +z."""

    hole: float = 1.5
    pitch: float = 1.7
    length: float = 35.0
    gap: float = 1.0
    crystal: float = 9.5
    voxel: float = 0.1
    patch: float = 40.0

    @property
    def face_z(self):
        return self.crystal + self.gap + self.length

    @property
    def crystal_depth(self):
        """collimator face -> middle of the crystal"""
        return self.length + self.gap + 0.5 * self.crystal

    @property
    def origin(self):
        """lattice point at the patch centre, just above the collimator face"""
        return (0.5 * self.patch, 0.5 * self.patch, self.face_z + 0.01)

    def hole_mask(self):
        """(ny, nx) bool: voxel centre inside a hexagonal hole of the lattice through the
        patch centre (hexagon = |p.n_i| <= hole/2 for the normals at 0, 60, 120 deg)."""
        n = int(round(self.patch / self.voxel))
        c = (np.arange(n) + 0.5) * self.voxel - 0.5 * self.patch
        X, Y = np.meshgrid(c, c, indexing="xy")
        a2x, a2y = 0.5 * self.pitch, 0.5 * math.sqrt(3.0) * self.pitch
        j0 = np.round(Y / a2y)
        inside = np.zeros_like(X, dtype=bool)
        normals = [(1.0, 0.0), (0.5, 0.5 * math.sqrt(3.0)), (-0.5, 0.5 * math.sqrt(3.0))]
        for dj in (-1, 0, 1):
            j = j0 + dj
            i0 = np.round((X - j * a2x) / self.pitch)
            for di in (-1, 0, 1):
                i = i0 + di
                rx = X - i * self.pitch - j * a2x
                ry = Y - j * a2y
                m = np.ones_like(inside)
                for nx_, ny_ in normals:
                    m &= np.abs(rx * nx_ + ry * ny_) <= 0.5 * self.hole
                inside |= m
        return inside

    def labels(self):
        """uint8 (nz, ny, nx): 0 air, 1 lead, 2 NaI"""
        n = int(round(self.patch / self.voxel))
        nzc = int(round(self.crystal / self.voxel))
        nzg = int(round(self.gap / self.voxel))
        nzl = int(round(self.length / self.voxel))
        lab = np.zeros((nzc + nzg + nzl, n, n), np.uint8)
        lab[:nzc] = 2
        col = np.where(self.hole_mask(), 0, 1).astype(np.uint8)
        lab[nzc + nzg:] = col[None]
        return lab

    def simulation(self, capacity_events=0, e_max=None, e_cut=None):
        """(tables, phantom, tally) of the head model; the tally scores the per-event NaI energy.
        The physics tables cover E_RANGE and e_max (MeV) if larger; photons below e_cut (MeV,
        None = the lowest table energy) deposit where they are."""
        from ..physics import xs
        from ..physics.materials import MaterialDatabase
        from ..transport import PhantomSpec, TallySpec

        db = MaterialDatabase()
        top = max(E_RANGE[1], float(e_max or 0.0))
        tables = xs.build_tables([db.get(n) for n in HEAD_MATERIALS], top * 1.001, e_cut=e_cut)
        ph = PhantomSpec(self.labels(), (self.voxel,) * 3, world_material=-1)
        tal = TallySpec(size=(1, 1, 1), spacing=(self.patch, self.patch, self.face_z), aligned=False, edep2=False,
                        detector_label_min=2)
        return tables, ph, tal


# ---------------------------------------------------------------------- gantry
@dataclass
class Heads:
    """n_proj detector heads around the z axis through `center` (phantom frame)."""

    center: tuple
    radius: float
    n_proj: int = 128
    nx: int = 128
    ny: int = 128
    spacing: tuple = (4.8, 4.8)
    angle0: float = 0.0
    theta_max: float = math.radians(10.0)
    model: HeadModel = None
    e_range: tuple = E_RANGE

    def __post_init__(self):
        if self.model is None:
            self.model = HeadModel()

    @property
    def plane(self):
        return self.nx * self.ny

    def desc(self):
        d = A.SpectDesc()
        d.center = (C.c_float * 3)(*[float(v) for v in self.center])
        d.radius = float(self.radius)
        d.n_proj, d.nx, d.ny = int(self.n_proj), int(self.nx), int(self.ny)
        d.angle0 = float(self.angle0)
        d.spacing = (C.c_float * 2)(*[float(v) for v in self.spacing])
        d.cos_theta_max = float(math.cos(self.theta_max))
        d.crystal_depth = float(self.model.crystal_depth)
        d.pitch = float(self.model.pitch)
        d.patch_origin = (C.c_float * 3)(*[float(v) for v in self.model.origin])
        d.norm = (C.c_float * 6)(*[float(v) for v in input_norm(self.theta_max, self.e_range)])
        return d

    def angle_table(self):
        a = self.angle0 + 2.0 * np.pi * np.arange(self.n_proj) / self.n_proj
        return np.stack([np.cos(a), np.sin(a)], 1).astype(np.float32)


def input_norm(theta_max, e_range=E_RANGE):
    """NN input scaling: E -> [-1, 1] over e_range, theta -> [-1, 1] over [0, theta_max], phi / pi."""
    lo, hi = e_range
    return (2.0 / (hi - lo), -(hi + lo) / (hi - lo), 2.0 / theta_max, -1.0, 1.0 / math.pi, 0.0)


class PairBuffer:
    """Device buffers of gh_spect_pairs for up to `capacity` pairs."""

    def __init__(self, capacity, device, det=False, x16=True):
        torch = _torch()
        self.capacity = int(capacity)
        self.x16 = torch.empty(self.capacity, 8, device=device, dtype=torch.bfloat16) if x16 else None
        self.pix = torch.empty(self.capacity, device=device, dtype=torch.int32)
        self.wgt = torch.empty(self.capacity, device=device, dtype=torch.float32)
        self.det = [torch.empty(self.capacity, device=device) for _ in range(8)] if det else None
        self.count = torch.zeros(1, device=device, dtype=torch.int64)


def spect_pairs(lib, rec, n, heads, cs, buf):
    """Pairs of the first n photons of rec (8 device arrays) into buf; returns the count."""
    buf.count.zero_()
    det = (A.FP * 8)(*[C.cast(_ptr(a), A.FP) for a in buf.det]) if buf.det is not None else None
    rc = lib.gh_spect_pairs((A.FP * 8)(*[C.cast(_ptr(a), A.FP) for a in rec]), int(n), C.byref(heads.desc()),
                            _ptr(cs), buf.capacity, _ptr(buf.x16), _ptr(buf.pix), _ptr(buf.wgt), det, C.c_void_p(),
                            _ptr(buf.count), _stream())
    if rc != 0:
        raise RuntimeError(f"gh_spect_pairs failed ({rc})")
    m = int(buf.count.item())
    if m > buf.capacity:
        raise RuntimeError(f"gh_spect_pairs: {m} pairs exceed the buffer capacity {buf.capacity}")
    return m


def energy_windows(lib, edep, n, seed, first, cls, windows=WINDOWS, blur=BLUR):
    """Window class (device int32, 0 = none) of n event energies after the blurring law."""
    win = (C.c_float * (2 * len(windows)))(*[float(v) for w in windows for v in w])
    rc = lib.gh_energy_windows(_ptr(edep), int(n), int(seed) & (2**64 - 1), int(first), C.byref(blur.desc()), win,
                               len(windows), _ptr(cls), _stream())
    if rc != 0:
        raise RuntimeError(f"gh_energy_windows failed ({rc})")
    return cls


# ---------------------------------------------------------------------- network
class ArfModel:
    """ARF network and its training on labelled detector-model photons."""

    def __init__(self, theta_max=math.radians(10.0), device="cuda", seed=0, sizes=None, windows=WINDOWS,
                 e_range=E_RANGE):
        self.theta_max = float(theta_max)
        self.e_range = (float(e_range[0]), float(e_range[1]))
        self.windows = tuple(tuple(w) for w in windows)
        self.sizes = list(sizes or SIZES)
        if self.sizes[-1] != len(self.windows) + 1:
            raise ValueError("network output = number of windows + 1 (not detected)")
        self.mlp = MLP(self.sizes, device=device, seed=seed)
        self.dev = self.mlp.dev
        self.lib = self.mlp.lib

    @property
    def n_win(self):
        return len(self.windows)

    def inputs(self, E, theta, phi):
        """NN input rows (bf16, padded to 8) from float tensors."""
        torch = _torch()
        n = input_norm(self.theta_max, self.e_range)
        X = torch.zeros(E.shape[0], self.mlp.psizes[0], device=self.dev, dtype=torch.bfloat16)
        X[:, 0] = (E * n[0] + n[1]).to(torch.bfloat16)
        X[:, 1] = (theta * n[2] + n[3]).to(torch.bfloat16)
        X[:, 2] = (phi * n[4] + n[5]).to(torch.bfloat16)
        return X

    def fit(self, X, cls, steps=3000, batch=16384, lr=2e-3, lr_min=5e-5, seed=1, log_every=0, class_weight=None):
        """Cross-entropy training on rows X (bf16 inputs) with classes cls (int, 0..n_win).
        class_weight (one per class, optional) weights the rows of each class in the loss: a
        training set whose undetected rows were kept with probability 1/r (Russian roulette of
        GateARFTrainingDatasetActor.cpp:117-121) is trained with class_weight[0] = r."""
        torch = _torch()
        N = X.shape[0]
        g = torch.Generator(device=self.dev).manual_seed(int(seed))
        cls = cls.long()
        cw = None
        if class_weight is not None:
            cw = torch.as_tensor(class_weight, dtype=torch.float32, device=self.dev)
            if cw.numel() != self.n_win + 1 or bool((cw < 0).any()):
                raise ValueError("class_weight needs one non-negative weight per class")
        losses = torch.empty(steps, device=self.dev)
        for it in range(steps):
            self.mlp.lr = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * it / max(1, steps)))
            idx = torch.randint(0, N, (batch,), device=self.dev, generator=g)
            xb = X[idx].contiguous()
            z = self.mlp.forward(xb)
            logp = torch.log_softmax(z, 1)
            yb = cls[idx]
            grad = logp.exp()
            grad[torch.arange(batch, device=self.dev), yb] -= 1.0
            if cw is None:
                losses[it] = -logp.gather(1, yb[:, None]).mean()
                grad /= batch
            else:
                w = cw[yb]
                ws = w.sum().clamp_min(1e-30)
                losses[it] = -(w * logp.gather(1, yb[:, None])[:, 0]).sum() / ws
                grad *= (w / ws)[:, None]
            dW, db = self.mlp.backward(grad)
            self.mlp.step(dW, db)
            if log_every and (it % log_every == 0 or it == steps - 1):
                print(f"[arf fit] step {it} loss {float(losses[it]):.5f}", flush=True)
        self.history = losses.cpu().numpy()
        return self.history

    def probabilities(self, X):
        torch = _torch()
        return torch.softmax(self.mlp.forward(X, keep=False), 1)

    def project(self, buf, m, img, img2=None, chunk=1 << 21, bufs=None):
        """Add the window probabilities of the m pairs in buf into img [n_proj*n_win, ny, nx] (FP64)."""
        torch = _torch()
        plane = img.shape[-1] * img.shape[-2]
        c = min(int(chunk), max(m, 1))
        Z = torch.empty(c, self.mlp.psizes[-1], device=self.dev)
        fused = self.mlp.fused_ok()
        if bufs is None and not fused:
            bufs = [torch.empty(c, max(self.mlp.psizes[1:-1]), device=self.dev, dtype=torch.bfloat16)
                    for _ in range(2)]
        for s in range(0, m, c):
            k = min(c, m - s)
            if fused:
                self.mlp.infer_fused(buf.x16[s:s + k], Z[:k])
            else:
                self.mlp.infer(buf.x16[s:s + k], Z[:k], bufs)
            rc = self.lib.gh_arf_accumulate(_ptr(Z), Z.shape[1], k, C.c_void_p(buf.pix.data_ptr() + 4 * s),
                                            C.c_void_p(buf.wgt.data_ptr() + 4 * s), self.n_win, plane, _ptr(img),
                                            _ptr(img2), _stream())
            if rc != 0:
                raise RuntimeError(f"gh_arf_accumulate failed ({rc})")

    def flops_per_pair(self):
        return self.mlp.flops_forward(1)

    def save(self, path):
        torch = _torch()
        d = self.mlp.state_dict()
        torch.save({"sizes": torch.tensor(self.sizes), "W": d["W"], "b": d["b"],
                    "theta_max": torch.tensor(self.theta_max, dtype=torch.float64),
                    "windows": torch.tensor(self.windows, dtype=torch.float64),
                    "e_range": torch.tensor(self.e_range, dtype=torch.float64)}, path)

    @classmethod
    def load(cls, path, device="cuda"):
        torch = _torch()
        d = torch.load(path, map_location="cpu", weights_only=True)
        e_range = tuple(d["e_range"].tolist()) if "e_range" in d else E_RANGE
        m = cls(float(d["theta_max"]), device=device, sizes=d["sizes"].tolist(),
                windows=[tuple(w) for w in d["windows"].tolist()], e_range=e_range)
        m.mlp.load_state_dict({"W": d["W"], "b": d["b"]})
        return m


# ---------------------------------------------------------------------- detector transport
class HeadSimulator:
    """Plain transport through the head model: labels photons with their energy window."""

    def __init__(self, model=None, device=0, e_max=None, e_cut=None):
        from ..transport import GpuEngine

        self.model = model or HeadModel()
        self.engine = GpuEngine(device)
        self.tables, self.phantom, self.tally = self.model.simulation(e_max=e_max, e_cut=e_cut)
        self.engine.set_physics(self.tables)
        self.sid = None

    def event_energy(self, rec, n, seed, first=0):
        """Transport n detector-frame photons (8 device arrays, head-model frame) and return
        the energy each leaves in the crystal (device float32, MeV): the sum over the event's
        crystal hits that an adder digitizer forms (GateDigitizerAdderActor)."""
        from ..transport import SourceSpec

        src = SourceSpec(pos_type="phsp", phsp=[a[:n] for a in rec])
        if self.sid is None:
            self.sid = self.engine.add_simulation(self.phantom, src, self.tally)
        else:
            self.engine.set_source(self.sid, src)
        self.engine.run([n], seed, first_history=[first])
        return self.engine.get_event_edep(self.sid, n, device=True)

    def windows(self, rec, n, seed, first=0, cls=None, windows=WINDOWS, blur=BLUR):
        """Window class of each of n detector-frame photons (device int32, 0 = none)."""
        torch = _torch()
        edep = self.event_energy(rec, n, seed, first)
        if cls is None:
            cls = torch.empty(n, device=edep.device, dtype=torch.int32)
        energy_windows(self.engine.lib, edep, n, seed, first, cls, windows, blur)
        return cls

    def photons(self, a, b, E, theta, phi, weight=None):
        """Detector-frame photons (8 device arrays) entering the collimator face at the lattice-
        cell point (a, b) in [0, 1)^2 (fractions of the cell spanned by the first lattice vector
        and the row offset, as gh_spect_pairs reduces entry points) with energy E and direction
        (theta, phi) relative to the inward face normal (radians)."""
        torch = _torch()
        m = self.model
        ox, oy, oz = m.origin
        x = ox + a * m.pitch
        y = oy + b * 0.5 * math.sqrt(3.0) * m.pitch
        ct = torch.cos(theta)
        st = torch.sqrt(torch.clamp(1.0 - ct * ct, min=0.0))
        w = torch.ones_like(x) if weight is None else weight
        return [x, y, torch.full_like(x, oz), st * torch.cos(phi), st * torch.sin(phi), -ct, E, w]

    def training_photons(self, n, seed, theta_max, device="cuda", theta_power=2.0, e_focus=(0.5, 0.100)):
        """n photons uniformly over the lattice cell that gh_spect_pairs reduces entries to
        ([0, pitch) x [0, pitch sqrt(3)/2) from the lattice point at the patch centre), E
        theta = theta_max u^theta_power (u uniform), phi uniform in (-pi, pi], E uniform in
        E_RANGE for a fraction 1 - e_focus[0] of the photons and uniform in
        [e_focus[1], E_RANGE[1]] for the others: (rec, E, theta, phi). The network learns the
        detection probability given (E, theta, phi), so the input distribution only decides
        where it is trained most: with theta_power = 2 about half of the photons fall inside
        the ~2.5 deg acceptance of the holes, and the energies of unscattered and once-
        scattered 140 keV photons get half of the samples."""
        torch = _torch()
        g = torch.Generator(device=device).manual_seed(int(seed))
        a = torch.rand(n, device=device, generator=g)
        b = torch.rand(n, device=device, generator=g)
        lo = torch.where(torch.rand(n, device=device, generator=g) < e_focus[0], e_focus[1], E_RANGE[0])
        E = lo + (E_RANGE[1] - lo) * torch.rand(n, device=device, generator=g)
        theta = theta_max * torch.rand(n, device=device, generator=g) ** theta_power
        phi = (2.0 * torch.rand(n, device=device, generator=g) - 1.0) * math.pi
        return self.photons(a, b, E, theta, phi), E, theta, phi

    def close(self):
        self.engine.close()


# ---------------------------------------------------------------------- projections
def project(rec, n, heads, arf=None, head_sim=None, seed=0, chunk=1 << 22, squared=True, pairs_per_photon=3.0,
            windows=None, blur=BLUR, out=None):
    """SPECT projections of n photons that left a phantom (8 device arrays x y z u v w E
    weight, phantom frame, vacuum outside), for the gantry `heads`: learned (arf, an
    ArfModel, whose windows are those it was trained on) and/or plain (head_sim, a
    HeadSimulator, with the energy `windows` (default: the ArfModel's, else WINDOWS) after
    the blurring law `blur`; overlapping windows each count an event that falls in them, as
    the channels of DigitizerEnergyWindowsActor do). Images are FP64 device tensors
    [n_proj * n_win, ny, nx] (projection-major, as ARFActor's and DigitizerProjectionActor's
    outputs), with squared sums when `squared`. Passing the dict of a previous call as `out`
    adds to its images and continues its pair numbering (the noise streams of the plain
    path), so a long run can be projected in photon chunks. Returns the dict with the
    images, the pair count and the time spent in each stage (s)."""
    import time

    torch = _torch()
    lib = (arf or head_sim.engine).lib
    dev = rec[0].device
    if arf is not None and (abs(heads.theta_max - arf.theta_max) > 1e-9 or tuple(heads.e_range) != arf.e_range):
        raise ValueError("the gantry's acceptance angle and energy range must be those the ARF network was trained "
                         f"with (theta_max {arf.theta_max}, e_range {arf.e_range})")
    if windows is None:
        windows = arf.windows if arf is not None else WINDOWS
    windows = tuple(tuple(float(v) for v in w) for w in windows)
    if arf is not None and head_sim is not None and len(windows) != arf.n_win:
        raise ValueError("learned and plain projections in one call need the same number of windows")
    n_win = arf.n_win if arf is not None else len(windows)
    shape = (heads.n_proj * n_win, heads.ny, heads.nx)
    if out is None:
        out = {"pairs": 0, "t_pairs": 0.0, "t_arf": 0.0, "t_plain": 0.0}
        for name, on in (("arf", arf is not None), ("plain", head_sim is not None)):
            if on:
                out[name] = torch.zeros(shape, device=dev, dtype=torch.float64)
                out[name + "2"] = torch.zeros(shape, device=dev, dtype=torch.float64) if squared else None
    overlap = head_sim is not None and windows_overlap(windows)
    cs = torch.from_numpy(heads.angle_table()).to(dev)
    step = int(chunk)
    buf = PairBuffer(int(step * pairs_per_photon) + 1024, dev, det=head_sim is not None, x16=arf is not None)
    bufs = None
    if arf is not None and not arf.mlp.fused_ok():
        c = min(1 << 21, buf.capacity)
        bufs = [torch.empty(c, max(arf.mlp.psizes[1:-1]), device=dev, dtype=torch.bfloat16) for _ in range(2)]
    s = 0
    while s < n:
        k = min(step, n - s)
        torch.cuda.synchronize()
        t = time.perf_counter()
        m = spect_pairs_checked(lib, [a[s:s + k] for a in rec], k, heads, cs, buf)
        torch.cuda.synchronize()
        out["t_pairs"] += time.perf_counter() - t
        if m is None:  # more pairs than the buffer holds: smaller photon chunks
            step = max(1, step // 2)
            continue
        if arf is not None:
            t = time.perf_counter()
            arf.project(buf, m, out["arf"], out.get("arf2"), bufs=bufs)
            torch.cuda.synchronize()
            out["t_arf"] += time.perf_counter() - t
        if head_sim is not None and m:
            t = time.perf_counter()
            edep = head_sim.event_energy(buf.det, m, seed, first=out["pairs"])
            cls = torch.empty(m, device=dev, dtype=torch.int32)
            if not overlap:
                energy_windows(lib, edep, m, seed, out["pairs"], cls, windows, blur)
                _counts(lib, cls, buf, m, n_win, heads.plane, out["plain"], out.get("plain2"))
            else:
                # one window per pass: an event's blurred energy is the same in every pass
                tmp = torch.zeros((heads.n_proj, heads.ny, heads.nx), device=dev, dtype=torch.float64)
                tmp2 = torch.zeros_like(tmp) if out.get("plain2") is not None else None
                for c, w in enumerate(windows):
                    energy_windows(lib, edep, m, seed, out["pairs"], cls, (w,), blur)
                    tmp.zero_()
                    if tmp2 is not None:
                        tmp2.zero_()
                    _counts(lib, cls, buf, m, 1, heads.plane, tmp, tmp2)
                    out["plain"].view(heads.n_proj, n_win, heads.ny, heads.nx)[:, c] += tmp
                    if tmp2 is not None:
                        out["plain2"].view(heads.n_proj, n_win, heads.ny, heads.nx)[:, c] += tmp2
            torch.cuda.synchronize()
            out["t_plain"] += time.perf_counter() - t
        out["pairs"] += m
        s += k
    return out


def _counts(lib, cls, buf, m, n_win, plane, img, img2):
    rc = lib.gh_counts_accumulate(_ptr(cls), _ptr(buf.pix), _ptr(buf.wgt), m, n_win, plane, _ptr(img), _ptr(img2),
                                  _stream())
    if rc != 0:
        raise RuntimeError(f"gh_counts_accumulate failed ({rc})")


def spect_pairs_checked(lib, rec, n, heads, cs, buf):
    """gh_spect_pairs; None when the pairs do not fit in buf (nothing usable was written)."""
    try:
        return spect_pairs(lib, rec, n, heads, cs, buf)
    except RuntimeError as e:
        if "exceed the buffer capacity" in str(e):
            return None
        raise


def compare_projections(test, ref, test2=None, ref2=None, min_counts=20.0, block=1):
    """Agreement of two projection sets [n_proj * n_win, ny, nx] (numpy): per-window total
    ratio and, over pixels (block x block pixel sums) with at least min_counts reference
    counts, the mean of (test - ref)^2 / (var_test + var_ref), about 1 when the two agree
    within their noise; also per window."""
    out = {"block": block}
    n_win = len(WINDOWS)
    if block > 1:
        def red(a):
            if a is None:
                return None
            h, w = a.shape[-2] // block * block, a.shape[-1] // block * block
            a = a[..., :h, :w]
            return a.reshape(*a.shape[:-2], h // block, block, w // block, block).sum((-3, -1))

        test, ref, test2, ref2 = red(test), red(ref), red(test2), red(ref2)
    t = test.reshape(-1, n_win, *test.shape[-2:])
    r = ref.reshape(-1, n_win, *ref.shape[-2:])
    out["window_ratio"] = [float(t[:, k].sum() / max(r[:, k].sum(), 1e-300)) for k in range(n_win)]
    out["window_counts_ref"] = [float(r[:, k].sum()) for k in range(n_win)]
    if test2 is not None and ref2 is not None:
        v = test2 + ref2
        m = ref >= min_counts
        out["pixels"] = int(m.sum())
        out["chi2_per_pixel"] = float(np.mean((test[m] - ref[m]) ** 2 / np.maximum(v[m], 1e-300))) if m.any() else None
        t4 = test.reshape(-1, n_win, *test.shape[-2:])
        r4 = ref.reshape(-1, n_win, *ref.shape[-2:])
        v4 = v.reshape(-1, n_win, *v.shape[-2:])
        per = []
        for k in range(n_win):
            mk = r4[:, k] >= min_counts
            per.append(float(np.mean((t4[:, k][mk] - r4[:, k][mk]) ** 2 / np.maximum(v4[:, k][mk], 1e-300)))
                       if mk.any() else None)
        out["chi2_per_pixel_window"] = per
    return out


# ---------------------------------------------------------------------- training from a dataset
def train_arf(dataset, pth_filename=None, steps=4000, batch=16384, lr=2e-3, device="cuda", seed=0, log_every=0):
    """Train an ArfModel on an ARFTrainingDatasetActor dataset (.npz path or dict with E [MeV],
    Theta and Phi [deg], window, channels, acceptance_angle, e_range, russian_roulette):
    the counterpart of training a garf network on the reference's training ROOT file.
    Undetected rows are weighted by the Russian roulette factor. Saves to pth_filename
    when given; returns the model."""
    torch = _torch()
    d = dict(dataset) if isinstance(dataset, dict) else dict(np.load(dataset, allow_pickle=False))
    windows = [tuple(float(v) for v in w) for w in np.asarray(d["channels"]).reshape(-1, 2)]
    model = ArfModel(float(d["acceptance_angle"]), device=device, seed=seed, windows=windows,
                     sizes=SIZES[:-1] + [len(windows) + 1], e_range=tuple(np.asarray(d["e_range"]).tolist()))
    E = torch.as_tensor(np.asarray(d["E"], np.float32), device=model.dev)
    th = torch.as_tensor(np.radians(np.asarray(d["Theta"], np.float64)).astype(np.float32), device=model.dev)
    ph = torch.as_tensor(np.radians(np.asarray(d["Phi"], np.float64)).astype(np.float32), device=model.dev)
    cls = torch.as_tensor(np.asarray(d["window"]).astype(np.int64), device=model.dev)
    if len(E) == 0:
        raise ValueError("empty ARF training dataset")
    if int(cls.min()) < 0 or int(cls.max()) > len(windows):
        raise ValueError("window ids outside 0..number of channels")
    rr = float(d.get("russian_roulette", 1))
    cw = [rr] + [1.0] * len(windows) if rr != 1.0 else None
    model.fit(model.inputs(E, th, ph), cls, steps=steps, batch=min(batch, len(E)), lr=lr, seed=seed + 1,
              log_every=log_every, class_weight=cw)
    if pth_filename is not None:
        model.save(pth_filename)
    return model
