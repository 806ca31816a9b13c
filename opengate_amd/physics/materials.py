"""Material database: named NIST-style materials and mixtures by mass fraction.

Materials are identified by name, as in Geant4/GATE ("G4_WATER", ...). The
compositions below are the NIST/ICRU mass fractions; densities in g/cm3.
User materials (e.g. those created by the Schneider HU conversion,
opengate/geometry/materials.py:149-230) are added with `add_material_weights`,
the same call the reference uses (`MaterialDatabase.add_material_weights`).
"""
from dataclasses import dataclass, field

import numpy as np

from .elements import z_of


@dataclass
class Material:
    name: str
    density: float  # g/cm3
    elements: list  # atomic numbers
    fractions: np.ndarray  # mass fractions, sum 1
    is_vacuum: bool = False
    extra: dict = field(default_factory=dict)


def _mk(name, density, comp, vacuum=False):
    zs = [z_of(s) for s in comp]
    w = np.array([comp[s] for s in comp], dtype=np.float64)
    w = w / w.sum()
    return Material(name, float(density), zs, w, vacuum)


# Mass fractions: NIST "Compositions of Materials" (ICRU Report 37/44/46 values).
_BUILTIN = {
    "G4_Galactic": (1e-25, {"H": 1.0}, True),
    "G4_AIR": (0.00120479, {"C": 0.000124, "N": 0.755268, "O": 0.231781, "Ar": 0.012827}, False),
    "G4_WATER": (1.0, {"H": 0.111894, "O": 0.888106}, False),
    "G4_BONE_CORTICAL_ICRP": (
        1.92,
        {"H": 0.034, "C": 0.155, "N": 0.042, "O": 0.435, "Na": 0.001, "Mg": 0.002,
         "P": 0.103, "S": 0.003, "Ca": 0.225},
        False,
    ),
    "G4_BONE_COMPACT_ICRU": (
        1.85,
        {"H": 0.063984, "C": 0.278, "N": 0.027, "O": 0.410016, "Mg": 0.002, "P": 0.07,
         "S": 0.002, "Ca": 0.147},
        False,
    ),
    "G4_LUNG_ICRP": (
        1.04,
        {"H": 0.105, "C": 0.083, "N": 0.023, "O": 0.779, "Na": 0.002, "P": 0.001,
         "S": 0.002, "Cl": 0.003, "K": 0.002},
        False,
    ),
    "G4_ADIPOSE_TISSUE_ICRP": (
        0.95,
        {"H": 0.114, "C": 0.598, "N": 0.007, "O": 0.278, "Na": 0.001, "S": 0.001, "Cl": 0.001},
        False,
    ),
    "G4_MUSCLE_SKELETAL_ICRP": (
        1.05,
        {"H": 0.102, "C": 0.143, "N": 0.034, "O": 0.71, "Na": 0.001, "P": 0.002, "S": 0.003,
         "Cl": 0.001, "K": 0.004},
        False,
    ),
    "G4_TISSUE_SOFT_ICRP": (
        1.03,
        {"H": 0.104472, "C": 0.23219, "N": 0.02488, "O": 0.630238, "Na": 0.00113,
         "Mg": 0.00013, "P": 0.00133, "S": 0.00199, "Cl": 0.00134, "K": 0.00199,
         "Ca": 0.00023, "Fe": 0.00005, "Zn": 0.00003},
        False,
    ),
    "G4_Al": (2.699, {"Al": 1.0}, False),
    "G4_Cu": (8.96, {"Cu": 1.0}, False),
    "G4_W": (19.3, {"W": 1.0}, False),
    "G4_Pb": (11.35, {"Pb": 1.0}, False),
    "G4_Ti": (4.54, {"Ti": 1.0}, False),
    "G4_SODIUM_IODIDE": (3.667, {"Na": 0.153373, "I": 0.846627}, False),
    "G4_PLEXIGLASS": (1.19, {"H": 0.080538, "C": 0.599848, "O": 0.319614}, False),
}


class MaterialDatabase:
    """Name -> Material. Mirrors the subset of opengate.geometry.materials.MaterialDatabase
    (materials.py:726) used on the photon path: lookup by name and add_material_weights."""

    def __init__(self):
        self.materials = {}
        for name, (rho, comp, vac) in _BUILTIN.items():
            self.materials[name] = _mk(name, rho, comp, vac)

    def add_material_weights(self, name, elements, weights, density):
        """elements: symbols or names; weights: mass fractions; density in Geant4 units
        (g_cm3 = 6.24150907e18), as MaterialDatabase.add_material_weights of the reference."""
        from ..units import g4_units

        comp = {}
        for e, w in zip(elements, weights):
            comp[e] = comp.get(e, 0.0) + float(w)
        self.materials[name] = _mk(name, float(density) / g4_units.g_cm3, comp)
        return self.materials[name]

    def add_material(self, material: Material):
        self.materials[material.name] = material

    def __contains__(self, name):
        return name in self.materials

    def get(self, name):
        try:
            return self.materials[name]
        except KeyError:
            raise KeyError(f"unknown material '{name}'") from None

    def FindOrBuildMaterial(self, name):
        return self.get(name)
