"""SPECT gantry: a detector head (parallel hexagonal-hole collimator over a scintillator)
at `number_of_projections` angles around an axis, as the reference builds from its SPECT
head helpers (contrib/spect/ge_discovery_nm670.py add_spect_head, one run per gantry angle
with the head rotated by run_timing_intervals) and scores with DigitizerProjectionActor or
ARFActor.

Here every angle is scored in one pass: a photon that leaves the phantom box travels in
vacuum, so it can reach each head that it points at; the photon is paired with each of
them (gh_spect_pairs) and the pairs go through the learned (ARF) or the transported
(head model) detector response. Geometry, in the world frame:
  * the gantry axis is parallel to world z through `rotation_center`;
  * head k is at angle a_k = start_angle + 2 pi k / number_of_projections, with its
    collimator face at distance `radius` from the axis, facing the axis (outward normal
    (cos a_k, sin a_k, 0));
  * image columns run along the tangential direction (-sin a_k, cos a_k, 0), rows along z;
    pixels are assigned where the photon crosses the plane in the middle of the crystal;
  * photons arriving at more than `acceptance_angle` from the face normal are taken to be
    stopped by the collimator septa.
The head model (transported response and ARF training) is a lead collimator with hexagonal
holes (flat-to-flat `hole_diameter`, lattice pitch hole_diameter + septal_thickness),
`collimator_length` long, an air gap, and a NaI crystal `crystal_thickness` thick, voxelised
at `model_voxel` over a `model_patch` square; entry points are reduced to one lattice cell
of the hole pattern, so the patch stands for the whole periodic collimator.
"""
import math

import numpy as np

from ..units import g4_units
from .volumes import VolumeBase

mm = g4_units.mm
deg = g4_units.deg


class SPECTHeadVolume(VolumeBase):
    user_info_defaults = {
        "rotation_center": ([0.0, 0.0, 0.0], {"doc": "point on the gantry axis (world z direction) [mm]"}),
        "radius": (250.0 * mm, {"doc": "gantry axis to collimator face [mm]"}),
        "number_of_projections": (128, {"doc": "gantry angles over 360 deg (one projection per angle)"}),
        "start_angle": (0.0, {"doc": "angle of the first head [rad]"}),
        "acceptance_angle": (10.0 * deg, {"doc": "largest incidence angle that reaches the crystal [rad]"}),
        "hole_diameter": (1.5 * mm, {"doc": "hexagonal hole, flat to flat [mm]"}),
        "septal_thickness": (0.2 * mm, {"doc": "lead between neighbouring holes [mm]"}),
        "collimator_length": (35.0 * mm, {}),
        "crystal_gap": (1.0 * mm, {"doc": "air between collimator and crystal [mm]"}),
        "crystal_thickness": (9.5 * mm, {}),
        "model_voxel": (0.1 * mm, {"doc": "voxel of the transported head model [mm]"}),
        "model_patch": (40.0 * mm, {"doc": "side of the transported head model [mm]"}),
    }
    volume_type = "SPECTHeadVolume"

    def head_model(self):
        from ..learn.arf import HeadModel

        if not (self.hole_diameter > 0 and self.septal_thickness >= 0 and self.collimator_length > 0
                and self.crystal_thickness > 0 and self.crystal_gap >= 0 and self.model_voxel > 0):
            raise ValueError(f"SPECTHeadVolume {self.name}: collimator and crystal sizes must be positive")
        return HeadModel(hole=float(self.hole_diameter), pitch=float(self.hole_diameter + self.septal_thickness),
                         length=float(self.collimator_length), gap=float(self.crystal_gap),
                         crystal=float(self.crystal_thickness), voxel=float(self.model_voxel),
                         patch=float(self.model_patch))

    def heads(self, center_phantom, size, spacing, e_range, theta_max=None):
        """learn.arf.Heads for an image of size [nx (tangential), ny (axial)] pixels of
        spacing [sx, sy] mm; center_phantom = rotation_center in the phantom frame;
        theta_max overrides the acceptance angle (the value a network was trained with)."""
        from ..learn.arf import Heads

        n = int(self.number_of_projections)
        if n < 1:
            raise ValueError(f"SPECTHeadVolume {self.name}: number_of_projections must be >= 1")
        if not (0.0 < float(self.acceptance_angle) < 0.5 * math.pi):
            raise ValueError(f"SPECTHeadVolume {self.name}: acceptance_angle must be in (0, 90 deg)")
        size = [int(v) for v in size]
        spacing = [float(v) for v in spacing]
        if len(size) != 2 or len(spacing) != 2 or min(size) < 1 or min(spacing) <= 0:
            raise ValueError("projection size and spacing must be 2-vectors of positive values")
        return Heads(center=tuple(float(v) for v in center_phantom), radius=float(self.radius), n_proj=n,
                     nx=size[0], ny=size[1], spacing=(spacing[0], spacing[1]), angle0=float(self.start_angle),
                     theta_max=float(self.acceptance_angle if theta_max is None else theta_max),
                     model=self.head_model(),
                     e_range=tuple(float(v) for v in e_range))

    def head_normals(self):
        """(n_proj, 3) outward face normals in the world frame."""
        n = int(self.number_of_projections)
        a = float(self.start_angle) + 2.0 * np.pi * np.arange(n) / n
        return np.stack([np.cos(a), np.sin(a), np.zeros(n)], 1)
