"""Learned exit-plane source: an MLP generator of phase-space photons.

The generator maps 8-D Gaussian noise to 6 values (E, x, y, u', v', w') through a
6-layer dense network (width 256, BF16 weights, FP32 accumulation), evaluated and
trained with the library's own kernels (gh_gaussian_bf16 for the noise, gh_gemm_bf16
for every product of the training forward and backward passes, gh_mlp_forward_bf16 for
generation (all six layers in one kernel), gh_exitplane_decode for the mapping back to
photons). It replaces the upstream transport that produced the
exit-plane photons (the role of GANSource + GANSourceDefaultGenerator,
sources/gansources.py:270-431,457-757).

Representation (what the network is trained on):
  * direction in the focal frame: (u', v', w') are the direction components in the
    frame whose pole is the ray from the focal point (the point source) through the
    exit position, so photons that crossed the collimator without interacting have
    (0, 0, 1) whatever their position;
  * every feature is Gaussianised by its own empirical quantile function: training
    targets are Phi^-1((rank + 1/2)/N); generated values go back through a table of
    the data quantiles at equally spaced Gaussian knots. The aperture edges (sharp
    steps of the position marginals) and the point mass of unscattered directions
    are then carried by the tables, and the network only has to learn the smooth
    dependence between features.

Training loss: sliced 2-Wasserstein distance between a generated batch and a data
batch in the Gaussianised space (random unit projections plus the coordinate axes,
1-D optimal transport by sorting). The loss and its gradient w.r.t. the generator
output are computed with PyTorch tensor operations; the gradient is then pushed
through the network by the hand-written backward GEMMs.
"""
import ctypes as C
import math
import time

import numpy as np

from .. import _abi as A
from .mlp import MLP

SIZES = [8, 256, 256, 256, 256, 256, 6]
N_FEATURES = 6
FEATURES = ("E", "x", "y", "u'", "v'", "w'")


def _torch():
    import torch

    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _fptr(t):
    return C.cast(C.c_void_p(t.data_ptr()), A.FP)


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _as_tensor(a, device):
    torch = _torch()
    if hasattr(a, "data_ptr"):
        return a.to(device)
    return torch.from_numpy(np.ascontiguousarray(a)).to(device)


def focal_frame(nx, ny, nz):
    """Right-handed frame (e1, e2, n) around the unit vector n (torch, any float dtype):
    e1 = (nx nz, ny nz, -up)/up, e2 = (-ny, nx, 0)/up with up = |(nx, ny)|;
    for up <= 1e-7: e1 = (nz, 0, 0), e2 = (0, 1, 0)."""
    torch = _torch()
    up = torch.sqrt(nx * nx + ny * ny)
    ok = up > 1e-7
    iu = torch.where(ok, 1.0 / torch.where(ok, up, torch.ones_like(up)), torch.zeros_like(up))
    e1 = (torch.where(ok, nx * nz * iu, nz), torch.where(ok, ny * nz * iu, torch.zeros_like(up)),
          torch.where(ok, -up, torch.zeros_like(up)))
    e2 = (torch.where(ok, -ny * iu, torch.zeros_like(up)), torch.where(ok, nx * iu, torch.ones_like(up)))
    return e1, e2


def exit_features(phsp, focal, plane_z, device="cuda"):
    """Phase-space records (8 arrays x y z u v w E weight, exit-plane frame) -> (N, 6)
    float32 features (E, x, y, u', v', w') on the device."""
    torch = _torch()
    x, y, u, v, w, E = (_as_tensor(phsp[k], device).double() for k in (0, 1, 3, 4, 5, 6))
    nx, ny = x - focal[0], y - focal[1]
    nz = torch.full_like(x, float(plane_z) - focal[2])
    inv = 1.0 / torch.sqrt(nx * nx + ny * ny + nz * nz)
    nx, ny, nz = nx * inv, ny * inv, nz * inv
    (e1x, e1y, e1z), (e2x, e2y) = focal_frame(nx, ny, nz)
    lu = u * e1x + v * e1y + w * e1z
    lv = u * e2x + v * e2y
    lw = u * nx + v * ny + w * nz
    return torch.stack([E, x, y, lu, lv, lw], 1).float()


class QuantileMap:
    """Per-feature quantile tables at the Gaussian knots g_k = -g_max + 2 g_max k/(K-1)."""

    def __init__(self, qtab, g_max):
        self.qtab = qtab.contiguous().float()  # (D, K) on the device
        self.g_max = float(g_max)

    @property
    def n_knots(self):
        return int(self.qtab.shape[1])

    @classmethod
    def fit(cls, F, n_knots=4096, g_max=5.5):
        torch = _torch()
        N, D = F.shape
        g = torch.linspace(-g_max, g_max, n_knots, device=F.device, dtype=torch.float64)
        lev = torch.special.ndtr(g)
        idx = torch.clamp(torch.round(lev * (N - 1)), 0, N - 1).long()
        qtab = torch.empty(D, n_knots, device=F.device, dtype=torch.float32)
        for j in range(D):
            qtab[j] = torch.sort(F[:, j]).values[idx]
        return cls(qtab, g_max)

    @staticmethod
    def gaussianize(F):
        """Rank-based Gaussian scores of every column: Phi^-1((rank + 1/2)/N)."""
        torch = _torch()
        N, D = F.shape
        G = torch.empty_like(F, dtype=torch.float32)
        ranks = torch.empty(N, device=F.device, dtype=torch.float64)
        ar = (torch.arange(N, device=F.device, dtype=torch.float64) + 0.5) / N
        for j in range(D):
            order = torch.argsort(F[:, j], stable=True)
            ranks[order] = ar
            G[:, j] = torch.special.ndtri(ranks).float()
        return G


class ExitPlaneGenerator:
    """Learned exit-plane source for one collimator geometry (focal point, exit plane z)."""

    def __init__(self, focal, plane_z, device="cuda", seed=0, sizes=None):
        self.focal = tuple(float(v) for v in focal)
        self.plane_z = float(plane_z)
        self.sizes = list(sizes or SIZES)
        self.mlp = MLP(self.sizes, device=device, seed=seed)
        self.dev = self.mlp.dev
        self.lib = self.mlp.lib
        self.qmap = None
        self.history = []

    # ------------------------------------------------------------------ noise
    def noise(self, n, seed, first=0, out=None):
        torch = _torch()
        ld = self.mlp.psizes[0]
        if out is None:
            out = torch.empty(n, ld, device=self.dev, dtype=torch.bfloat16)
        rc = self.lib.gh_gaussian_bf16(int(seed) & (2**64 - 1), int(first), int(n), self.sizes[0], ld, _ptr(out),
                                       _stream())
        if rc != 0:
            raise RuntimeError(f"gh_gaussian_bf16 failed ({rc})")
        return out

    # ------------------------------------------------------------------ training
    def fit(self, phsp, steps=4000, batch=16384, projections=128, lr=1e-3, lr_min=3e-5, seed=1, n_knots=4096,
            log_every=0):
        """Train on recorded exit-plane photons (8 arrays, host or device). Returns the
        loss history (one value per step, on the host)."""
        torch = _torch()
        t0 = time.perf_counter()
        F = exit_features(phsp, self.focal, self.plane_z, self.dev)
        self.qmap = QuantileMap.fit(F, n_knots=n_knots)
        G = QuantileMap.gaussianize(F)
        del F
        N = G.shape[0]
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        gen = torch.Generator(device=self.dev).manual_seed(int(seed))
        eye = torch.eye(N_FEATURES, device=self.dev)
        losses = torch.empty(steps, device=self.dev)
        X = torch.empty(batch, self.mlp.psizes[0], device=self.dev, dtype=torch.bfloat16)
        for it in range(steps):
            self.mlp.lr = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * it / max(1, steps)))
            self.noise(batch, seed * 1_000_003 + 17, first=it * batch, out=X)
            out = self.mlp.forward(X)
            o = out.detach().requires_grad_(True)
            d = G[torch.randint(0, N, (batch,), device=self.dev, generator=gen)]
            th = torch.randn(N_FEATURES, projections, device=self.dev, generator=gen)
            th = torch.cat([th / th.norm(dim=0, keepdim=True), eye], 1)
            loss = ((torch.sort(o @ th, 0).values - torch.sort(d @ th, 0).values) ** 2).mean()
            loss.backward()
            dW, db = self.mlp.backward(o.grad)
            self.mlp.step(dW, db)
            losses[it] = loss.detach()
            if log_every and (it % log_every == 0 or it == steps - 1):
                print(f"[exitplane fit] step {it} loss {float(loss):.5f}", flush=True)
        self.history = losses.cpu().numpy()
        t2 = time.perf_counter()
        # GEMM work of one step: forward, weight gradients, and input gradients below layer 0
        kn = [k * n for k, n in zip(self.sizes[:-1], self.sizes[1:])]
        flops = 2.0 * batch * (2 * sum(kn) + sum(kn[1:]))
        self.fit_stats = {"steps": int(steps), "batch": int(batch), "setup_s": t1 - t0, "loop_s": t2 - t1,
                          "step_ms": 1e3 * (t2 - t1) / max(1, steps), "gemm_flop_per_step": flops,
                          "gemm_tflops_over_loop": flops * steps / max(t2 - t1, 1e-9) / 1e12}
        return self.history

    # ------------------------------------------------------------------ generation
    def generate(self, n, seed, first=0, out=None, chunk=1 << 22, weight=1.0, events=None):
        """n exit-plane photons as 8 device float32 arrays (x y z u v w E weight) in the
        exit-plane frame. Photon i depends only on (seed, first + i). events: optional list
        that receives a (start, end) pair of CUDA events recorded around each network forward
        (on the current stream, where the forward runs)."""
        torch = _torch()
        if self.qmap is None:
            raise RuntimeError("generator not trained or loaded")
        n = int(n)
        if out is None:
            out = [torch.empty(n, device=self.dev, dtype=torch.float32) for _ in range(8)]
        m = min(int(chunk), max(n, 1))
        X = torch.empty(m, self.mlp.psizes[0], device=self.dev, dtype=torch.bfloat16)
        bufs = None
        if not self.mlp.fused_ok():
            bufs = [torch.empty(m, max(self.mlp.psizes[1:-1]), device=self.dev, dtype=torch.bfloat16)
                    for _ in range(2)]
        Z = torch.empty(m, self.mlp.psizes[-1], device=self.dev, dtype=torch.float32)
        focal = (C.c_float * 3)(*self.focal)
        zero = (C.c_float * 3)(0.0, 0.0, 0.0)
        fused = self.mlp.fused_ok()
        for c0 in range(0, n, m):
            k = min(m, n - c0)
            self.noise(k, seed, first=first + c0, out=X[:k])
            if events is not None:
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
            if fused:
                self.mlp.infer_fused(X[:k], Z[:k])
            else:
                self.mlp.infer(X[:k], Z[:k], bufs)
            if events is not None:
                e1.record()
                events.append((e0, e1))
            ptrs = (A.FP * 8)(*[C.cast(C.c_void_p(a.data_ptr() + 4 * c0), A.FP) for a in out])
            rc = self.lib.gh_exitplane_decode(_ptr(Z), Z.shape[1], k, _ptr(self.qmap.qtab), self.qmap.n_knots,
                                              self.qmap.g_max, focal, self.plane_z, zero, float(weight), ptrs,
                                              _stream())
            if rc != 0:
                raise RuntimeError(f"gh_exitplane_decode failed ({rc})")
        return out

    def flops_per_photon(self):
        return self.mlp.flops_forward(1)

    # ------------------------------------------------------------------ files
    def state_dict(self):
        torch = _torch()
        d = self.mlp.state_dict()
        return {"sizes": torch.tensor(self.sizes), "W": d["W"], "b": d["b"], "qtab": self.qmap.qtab.cpu(),
                "g_max": torch.tensor(self.qmap.g_max), "focal": torch.tensor(self.focal),
                "plane_z": torch.tensor(self.plane_z)}

    def save(self, path):
        _torch().save(self.state_dict(), path)

    @classmethod
    def load(cls, path, device="cuda"):
        torch = _torch()
        d = torch.load(path, map_location="cpu", weights_only=True)
        g = cls(d["focal"].tolist(), float(d["plane_z"]), device=device, sizes=d["sizes"].tolist())
        g.mlp.load_state_dict({"W": d["W"], "b": d["b"]})
        g.qmap = QuantileMap(d["qtab"].to(g.dev), float(d["g_max"]))
        return g


# ---------------------------------------------------------------------- comparison
def _np(a, n=None):
    a = a.cpu().numpy() if hasattr(a, "cpu") else np.asarray(a)
    return a[:n] if n is not None else a


def compare_exit_planes(gen, ref, distance, bin_mm=1.0, max_n=4_000_000):
    """Statistics of two exit-plane photon sets (8 arrays each): 1-D Wasserstein distances
    of E, x, y; the fluence projected `distance` mm along the directions (2-D histogram
    relative RMS difference in the field core against its Poisson expectation, core
    ratio, 20-80 % penumbra widths of the x profile)."""
    from scipy.stats import wasserstein_distance as wd

    n = min(len(_np(gen[0])), len(_np(ref[0])), max_n)
    g = [_np(a, n).astype(np.float64) for a in gen]
    r = [_np(a, n).astype(np.float64) for a in ref]
    out = {"n": n}
    for k, name in ((6, "E"), (0, "x"), (1, "y")):
        out[f"w1_{name}"] = float(wd(g[k], r[k]))

    def proj(a):
        t = distance / np.abs(a[5])
        return a[0] + t * a[3], a[1] + t * a[4]

    gx, gy = proj(g)
    rx, ry = proj(r)
    lo = min(np.percentile(rx, 0.5), np.percentile(ry, 0.5)) - 10
    hi = max(np.percentile(rx, 99.5), np.percentile(ry, 99.5)) + 10
    bins = np.arange(lo, hi + bin_mm, bin_mm)
    Hg = np.histogram2d(gx, gy, [bins, bins])[0]
    Hr = np.histogram2d(rx, ry, [bins, bins])[0]
    core = Hr > 0.5 * Hr.max()
    out["core_ratio"] = float(Hg[core].sum() / Hr[core].sum())
    out["core_rel_rms"] = float(np.sqrt(np.mean(((Hg[core] - Hr[core]) / Hr[core]) ** 2)))
    out["core_rel_rms_poisson"] = float(np.sqrt(np.mean((Hg[core] + Hr[core]) / Hr[core] ** 2)))

    def penumbra(h):
        m = h.max()
        i = np.where(h > 0.8 * m)[0]
        j = np.where(h > 0.2 * m)[0]
        return float(bins[i[0]] - bins[j[0]]), float(bins[j[-1]] - bins[i[-1]])

    cy = np.abs(0.5 * (bins[1:] + bins[:-1]) - np.median(ry)) < 5 * bin_mm
    out["penumbra_gen"] = penumbra(Hg[:, cy].sum(1))
    out["penumbra_ref"] = penumbra(Hr[:, cy].sum(1))
    return out
