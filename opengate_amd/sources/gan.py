"""GANSource: a learned exit-plane source with the reference's user parameters.

Reference: opengate/sources/gansources.py:270-431 (GANSource user info: pth_filename,
backward_distance, energy_min/max_threshold, batch_size) and GateGANSource.cpp:147-265
(position = R_global (R_local p + t_local) + t_global, direction = R_global d / |d|,
events whose energy falls outside the thresholds are skipped and the next generated
particle is used). The generator here is an ExitPlaneGenerator file (torch tensors
only, loaded with weights_only=True) instead of a gaga .pth; the photons are generated
on the GPU and handed to the transport as device arrays.
"""
import sys

import numpy as np

from ..base import Box
from .generic import SourceBase


class GANSource(SourceBase):
    user_info_defaults = {
        "pth_filename": (None, {"doc": "generator file written by ExitPlaneGenerator.save"}),
        "particle": ("gamma", {"doc": "only 'gamma' on the photon path"}),
        "batch_size": (1 << 22, {"doc": "photons per generator call"}),
        "backward_distance": (None, {"doc": "move generated photons back by this distance along -direction"}),
        "backward_force": (None, {"doc": "accepted for compatibility: no time is tracked on this path"}),
        "energy_min_threshold": (-1, {"doc": "generated photons below this energy are skipped"}),
        "energy_max_threshold": (sys.float_info.max, {"doc": "generated photons above this energy are skipped"}),
        "position": (Box(translation=[0.0, 0.0, 0.0], rotation=np.eye(3)), {"doc": "local placement"}),
        "seed": (None, {"doc": "generator noise seed (default: the simulation seed)"}),
    }
    source_type = "GANSource"
    runtime_attributes = ("simulation", "_gen")

    def __init__(self, name=None, **kw):
        super().__init__(name, **kw)
        self._gen = None

    def generator(self, device):
        if self._gen is None:
            if not self.pth_filename:
                raise ValueError(f"GANSource '{self.name}': pth_filename is not set")
            from ..learn.exitplane import ExitPlaneGenerator

            self._gen = ExitPlaneGenerator.load(self.pth_filename, device=device)
        return self._gen

    def generate(self, n, seed, device="cuda"):
        """n accepted photons as 8 device arrays (x y z u v w E weight) in the source's
        local frame (before position.rotation / translation)."""
        import torch

        if self.particle != "gamma":
            raise NotImplementedError(f"GANSource '{self.name}': particle '{self.particle}'")
        g = self.generator(device)
        lo = float(self.energy_min_threshold)
        hi = float(self.energy_max_threshold)
        parts = []
        got = 0
        first = 0
        while got < n:
            m = int(min(max(self.batch_size, 1), max(n - got, 1) * (1 if got == 0 else 2)))
            rec = g.generate(m, seed, first=first)
            first += m
            keep = (rec[6] >= lo) & (rec[6] <= hi)
            k = int(keep.sum())
            if k == 0 and first > 100 * max(n, 1):
                raise RuntimeError(f"GANSource '{self.name}': no generated energy inside the thresholds")
            parts.append([a[keep] for a in rec] if k < m else rec)
            got += k
        out = [torch.cat([p[j] for p in parts])[:n].contiguous() for j in range(8)]
        back = self.backward_distance
        if back:
            for j in range(3):
                out[j] = out[j] - float(back) * out[3 + j]
        return out
