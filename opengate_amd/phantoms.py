"""Seed-generated phantoms, material tables and spectra of the benchmark workloads.

Workloads (all phantoms from numpy.random.Generator(PCG64(seed))):
  * water64:  64^3 water, 2 mm voxels, three cortical-bone spheres (r 8-15 mm),
              140.5 keV point source at the centre;
  * ct256:    256x256x128 int16 HU, 1 mm voxels: ellipsoidal body (HU 40 +- 20),
              two lungs (HU -800), spine and ribs (HU 700-1200), air (HU -1000),
              mapped to 12 materials by HU intervals; 120 kVp Kramers spectrum
              filtered by 2.5 mm Al in 1 keV bins, 30 deg cone (15 deg half angle)
              from 600 mm;
  * batch64:  uneven cubes 64^3..192^3 (1.5-3 mm) with 1-6 inserts and volume
              sources (140.511 keV, or the 112.95/208.37 keV lines of Lu-177).
"""
import json
from pathlib import Path

import numpy as np

from .physics import xs
from .physics.materials import MaterialDatabase

DATA = Path(__file__).resolve().parent / "data"


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


# ------------------------------------------------------------------ water64
def water64_labels(seed=1, n=64, spacing=2.0):
    """labels[z, y, x]: 0 water, 1 cortical bone (three spheres fully inside)."""
    g = rng(seed)
    lab = np.zeros((n, n, n), np.uint8)
    half = n * spacing / 2
    c = (np.arange(n) + 0.5) * spacing - half  # voxel centres, volume centred at 0
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    spheres = []
    for _ in range(3):
        r = g.uniform(8.0, 15.0)
        ctr = g.uniform(-half + r, half - r, size=3)
        spheres.append((ctr, r))
        inside = (X - ctr[0]) ** 2 + (Y - ctr[1]) ** 2 + (Z - ctr[2]) ** 2 <= r * r
        lab[inside] = 1
    return lab, spheres


# ------------------------------------------------------------------ ct256
# 12 materials by HU interval [lo, hi): (name, base composition, density g/cm3)
CT12_TABLE = [
    (-1100, -900, "CT_Air", "G4_AIR", 0.00120479),
    (-900, -600, "CT_Lung_inflated", "G4_LUNG_ICRP", 0.26),
    (-600, -200, "CT_Lung_dense", "G4_LUNG_ICRP", 0.60),
    (-200, -50, "CT_Adipose", "G4_ADIPOSE_TISSUE_ICRP", 0.95),
    (-50, 20, "CT_SoftTissue_lo", "G4_TISSUE_SOFT_ICRP", 1.00),
    (20, 60, "CT_SoftTissue", "G4_TISSUE_SOFT_ICRP", 1.03),
    (60, 120, "CT_Muscle", "G4_MUSCLE_SKELETAL_ICRP", 1.06),
    (120, 400, "CT_Bone_spongy", "G4_BONE_CORTICAL_ICRP", 1.15),
    (400, 700, "CT_Bone_300", "G4_BONE_CORTICAL_ICRP", 1.35),
    (700, 900, "CT_Bone_500", "G4_BONE_CORTICAL_ICRP", 1.55),
    (900, 1100, "CT_Bone_700", "G4_BONE_CORTICAL_ICRP", 1.72),
    (1100, 3100, "CT_Bone_cortical", "G4_BONE_CORTICAL_ICRP", 1.92),
]


def ct12_materials(db: MaterialDatabase):
    """Registers the 12 CT materials and returns voxel_materials [[lo, hi, name], ...]."""
    vm = []
    for lo, hi, name, base, rho in CT12_TABLE:
        if name not in db:
            b = db.get(base)
            from .physics.elements import symbol_of
            from .units import g4_units

            db.add_material_weights(name, [symbol_of(z) for z in b.elements], list(b.fractions),
                                    rho * g4_units.g_cm3)
        vm.append([lo, hi, name])
    return vm


def ct256_hu(seed=2, shape_xyz=(256, 256, 128), spacing=1.0):
    """int16 HU image hu[z, y, x], volume centred at 0."""
    g = rng(seed)
    nx, ny, nz = shape_xyz
    hu = np.full((nz, ny, nx), -1000, np.int16)
    x = (np.arange(nx) + 0.5) * spacing - nx * spacing / 2
    y = (np.arange(ny) + 0.5) * spacing - ny * spacing / 2
    z = (np.arange(nz) + 0.5) * spacing - nz * spacing / 2
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    body = (X / 120.0) ** 2 + (Y / 90.0) ** 2 + (Z / 62.0) ** 2 <= 1.0
    noise = np.clip(np.rint(g.normal(40.0, 20.0, size=hu.shape)), -120, 110).astype(np.int16)
    hu[body] = noise[body]
    for sx in (-55.0, 55.0):
        lung = ((X - sx) / 35.0) ** 2 + ((Y - 10.0) / 45.0) ** 2 + (Z / 50.0) ** 2 <= 1.0
        hu[lung & body] = -800
    bone_hu = g.integers(700, 1201, size=hu.shape).astype(np.int16)
    spine = ((X / 12.0) ** 2 + ((Y + 55.0) / 14.0) ** 2 <= 1.0) & body
    r_ell = np.sqrt((X / 104.0) ** 2 + (Y / 78.0) ** 2)
    ribs = (np.abs(r_ell - 1.0) < 0.035) & (np.mod(Z + 200.0, 24.0) < 8.0) & (Y > -60.0) & body
    hu[spine | ribs] = bone_hu[spine | ribs]
    return hu


def kramers_spectrum(kvp=0.120, e_min=0.010, bin_width=0.001, al_mm=2.5, db=None):
    """Kramers N(E) ~ (E0 - E)/E times the transmission of al_mm of aluminium
    (NIST mu/rho of Al), on bins of bin_width from e_min to kvp. Returns (edges, weights)."""
    db = db or MaterialDatabase()
    edges = np.round(np.arange(e_min, kvp + bin_width / 2, bin_width), 9)
    mid = 0.5 * (edges[1:] + edges[:-1])
    al = db.get("G4_Al")
    mu_al = xs.mu_rho(al, mid) * al.density  # 1/cm
    w = np.clip(kvp - mid, 0.0, None) / mid * np.exp(-mu_al * al_mm / 10.0)
    return edges, w


# ------------------------------------------------------------------ batch64
def batch_simulations(seed=3, n_sims=64):
    """Descriptions of the batched workload: list of dicts (labels, spacing, source, histories)."""
    g = rng(seed)
    lines = json.loads((DATA / "rad_gamma_lines.json").read_text())
    sims = []
    for k in range(n_sims):
        n = int(g.integers(64, 193))
        sp = float(g.uniform(1.5, 3.0))
        half = n * sp / 2
        c = (np.arange(n) + 0.5) * sp - half
        lab = np.zeros((n, n, n), np.uint8)  # 0 = air outside the water ellipsoid
        Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
        ax = g.uniform(0.6, 0.95, size=3) * half
        body = (X / ax[0]) ** 2 + (Y / ax[1]) ** 2 + (Z / ax[2]) ** 2 <= 1.0
        lab[body] = 1  # water
        for _ in range(int(g.integers(1, 7))):
            kind = int(g.integers(2, 5))  # 2 bone, 3 lung, 4 soft tissue
            a = g.uniform(0.08, 0.3, size=3) * half
            ctr = g.uniform(-0.5, 0.5, size=3) * ax
            ins = ((X - ctr[0]) / a[0]) ** 2 + ((Y - ctr[1]) / a[1]) ** 2 + ((Z - ctr[2]) / a[2]) ** 2 <= 1.0
            lab[ins & body] = kind
        if g.uniform() < 0.5:
            src = {"energies": [0.140511], "weights": [0.885]}
        else:
            lu = lines["Lu177"]
            sel = [i for i, e in enumerate(lu["energies"]) if abs(e - 0.11295) < 2e-4 or abs(e - 0.20837) < 2e-4]
            src = {"energies": [lu["energies"][i] for i in sel], "weights": [lu["weights"][i] for i in sel]}
        src_half = g.uniform(0.2, 0.5, size=3) * ax
        hist = int(round(10 ** g.uniform(6, 8)))
        sims.append({"labels": lab, "spacing": sp, "n": n, "source_half": src_half, "spectrum": src,
                     "histories": hist})
    return sims


BATCH_MATERIALS = ["G4_AIR", "G4_WATER", "G4_BONE_CORTICAL_ICRP", "G4_LUNG_ICRP", "G4_TISSUE_SOFT_ICRP"]
