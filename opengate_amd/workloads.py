"""Benchmark workloads as GPU-engine inputs (phantom, source, tally, tables).

ct256 (the CT-scale workload): 256x256x128 int16 HU phantom at 1 mm mapped to 12
materials by HU intervals on the GPU, world of air (3 m cube), 120 kVp Kramers
spectrum (2.5 mm Al, 1 keV bins) in a cone of 15 deg half angle from 600 mm above
the phantom centre, dose on the phantom grid (edep, edep^2).

water64: 64^3 water with three cortical-bone spheres at 2 mm, 140.5 keV point source
at the centre, vacuum world, dose on the phantom grid.
"""
import numpy as np

from . import phantoms as P
from .physics import xs
from .physics.materials import MaterialDatabase
from .transport import PhantomSpec, SourceSpec, TallySpec, spectrum_histogram, spectrum_lines

CT256_DESC = ("ct256: 256x256x128 HU phantom, 1 mm voxels, 12 materials + air world; 120 kVp Kramers "
              "spectrum (2.5 mm Al, 1 keV bins), 15 deg half-angle cone from 600 mm; dose on the 256x256x128 grid")
WATER64_DESC = "water64: 64^3 water + 3 bone spheres, 2 mm; 140.5 keV point source at the centre; dose 64^3"


def ct256(engine, seed=2):
    """Returns (tables, phantom, source, tally, info). Labels are computed on the GPU."""
    db = MaterialDatabase()
    vm = P.ct12_materials(db)
    hu = P.ct256_hu(seed)
    names = ["CT_Air"] + [r[2] for r in vm if r[2] != "CT_Air"]
    lut = {n: i for i, n in enumerate(names)}
    rows = sorted(vm, key=lambda r: r[0])
    bins = np.array([r[0] for r in rows] + [rows[-1][1]], dtype=np.float64)
    labels_sorted = np.array([0] + [lut[r[2]] for r in rows] + [0], dtype=np.uint16)
    lab = engine.hu_to_labels(hu, bins, labels_sorted)
    mats = [db.get(n) for n in names] + [db.get("G4_AIR")]
    tables = xs.build_tables(mats, 0.1201)
    world = len(mats) - 1
    L = np.array([256.0, 256.0, 128.0])
    half_w = 1500.0
    ph = PhantomSpec(lab.astype(np.uint8), (1.0, 1.0, 1.0), world, tuple(-half_w + L / 2), tuple(half_w + L / 2))
    e, w = P.kramers_spectrum(db=db)
    ev, ew, cdf = spectrum_histogram(e, w)
    src = SourceSpec(pos_trans=(128.0, 128.0, 64.0 + 600.0), theta=(0.0, np.deg2rad(15.0)),
                     ene_type="spectrum_histogram", energy_value=ev, energy_width=ew, energy_cdf=cdf)
    tal = TallySpec(size=(256, 256, 128), spacing=(1.0, 1.0, 1.0))
    return tables, ph, src, tal, {"desc": CT256_DESC, "hu": hu, "names": names}


CT256S_DESC = ("ct256-schneider: the ct256 HU phantom mapped by HounsfieldUnit_to_material (Schneider 2000 tables, "
               "density tolerance {tol} g/cm3: {n} materials of {c} compositions, {bits}-bit labels), air world, "
               "120 kVp cone from 600 mm; dose on the 256x256x128 grid")


def ct256_schneider(engine, tol=0.05, seed=2, share_compositions=True, majorant_over_phantom=True):
    """ct256 with the reference's own HU conversion: the Schneider materials and densities
    tables split into sub-intervals of density width <= tol g/cm3
    (HounsfieldUnit_to_material, geometry/materials.py:149-230); the label image follows the
    ImageVolume label rule (default material G4_AIR = label 0). The Woodcock majorant covers
    the materials of this phantom (or every material of the table when
    majorant_over_phantom is False, for other phantoms). Returns (tables, phantom,
    source, tally, info); labels are computed on the GPU and kept as uint16 when needed."""
    from types import SimpleNamespace
    from pathlib import Path as _P

    from .geometry.materials import HounsfieldUnit_to_material
    from .units import g4_units

    db = MaterialDatabase()
    data = _P(__file__).resolve().parent / "data"
    vm, created = HounsfieldUnit_to_material(SimpleNamespace(volume_manager=SimpleNamespace(material_database=db)),
                                             tol * g4_units.g_cm3, str(data / "Schneider2000MaterialsTable.txt"),
                                             str(data / "Schneider2000DensitiesTable.txt"))
    hu = P.ct256_hu(seed)
    names = ["G4_AIR"] + [r[2] for r in vm]
    lut = {n: i for i, n in enumerate(names)}
    rows = sorted(vm, key=lambda r: r[0])
    bins = np.array([r[0] for r in rows] + [rows[-1][1]], dtype=np.float64)
    labels_sorted = np.array([0] + [lut[r[2]] for r in rows] + [0], dtype=np.uint16)
    lab = engine.hu_to_labels(hu, bins, labels_sorted)
    mats = [db.get(n) for n in names]
    present = np.flatnonzero(np.bincount(np.asarray(lab).ravel(), minlength=len(names)))
    tables = xs.build_tables(mats, 0.1201, share_compositions=share_compositions,
                             majorant_materials=present if majorant_over_phantom else None)
    L = np.array([256.0, 256.0, 128.0])
    half_w = 1500.0
    ph = PhantomSpec(lab if len(names) > 256 else lab.astype(np.uint8), (1.0, 1.0, 1.0), 0,
                     tuple(-half_w + L / 2), tuple(half_w + L / 2))
    e, wts = P.kramers_spectrum(db=db)
    ev, ew, cdf = spectrum_histogram(e, wts)
    src = SourceSpec(pos_trans=(128.0, 128.0, 64.0 + 600.0), theta=(0.0, np.deg2rad(15.0)),
                     ene_type="spectrum_histogram", energy_value=ev, energy_width=ew, energy_cdf=cdf)
    tal = TallySpec(size=(256, 256, 128), spacing=(1.0, 1.0, 1.0))
    n_rows = tables.n_compositions or tables.n_materials
    bits = 4 if int(lab.max()) < 16 else (8 if int(lab.max()) < 256 else 16)
    desc = CT256S_DESC.format(tol=tol, n=len(names), c=n_rows, bits=bits)
    return tables, ph, src, tal, {"desc": desc, "hu": hu, "names": names, "n_materials": len(names),
                                  "n_rows": n_rows, "label_bits": bits}


def water64(seed=1):
    db = MaterialDatabase()
    lab, spheres = P.water64_labels(seed)
    tables = xs.build_tables([db.get("G4_WATER"), db.get("G4_BONE_CORTICAL_ICRP")], 0.1406)
    ph = PhantomSpec(lab, (2.0, 2.0, 2.0))
    src = SourceSpec(pos_trans=(64.0, 64.0, 64.0), energy=0.1405)
    tal = TallySpec(size=(64, 64, 64), spacing=(2.0, 2.0, 2.0))
    return tables, ph, src, tal, {"desc": WATER64_DESC, "spheres": spheres}


BATCH64_DESC = ("batch64: 64 simulations, cubes of 64^3..192^3 voxels at 1.5-3 mm (water ellipsoid in air with "
                "1-6 bone/lung/soft-tissue inserts), ellipsoid volume sources with the 140.511 keV line or the "
                "112.95/208.37 keV Lu-177 lines, 1e6..1e8 histories each (log-uniform)")


def batch64(seed=3, n_sims=64):
    """Returns (tables, [(phantom, source, tally, histories), ...], info)."""
    db = MaterialDatabase()
    sims = P.batch_simulations(seed, n_sims)
    tables = xs.build_tables([db.get(n) for n in P.BATCH_MATERIALS], 0.2090)
    out = []
    for s in sims:
        n, sp = s["n"], s["spacing"]
        L = n * sp
        ph = PhantomSpec(s["labels"], (sp, sp, sp), world_material=-1)
        ev, _, cdf = spectrum_lines(s["spectrum"]["energies"], s["spectrum"]["weights"])
        src = SourceSpec(pos_type="ellipsoid", pos_param=tuple(s["source_half"]) + (0.0,),
                         pos_trans=(L / 2, L / 2, L / 2), ene_type="spectrum_discrete", energy_value=ev,
                         energy_cdf=cdf)
        tal = TallySpec(size=(n, n, n), spacing=(sp, sp, sp))
        out.append((ph, src, tal, s["histories"]))
    return tables, out, {"desc": BATCH64_DESC}


# ---------------------------------------------------------------------- exit plane
EXITPLANE_DESC = ("exitplane: tungsten block 40 mm thick with a 20x20 mm aperture (seed-generated edge "
                  "roughness, 0.5 mm voxels) under a 2.5 mm Al filter, 64x64 mm; raw 120 kVp Kramers point "
                  "source 100 mm above the filter, 10 deg half-angle cone; photons recorded on the exit face")
COLLIMATOR_SPACING = 0.5
COLLIMATOR_SHAPE = (128, 128, 85)  # x, y, z voxels: W at z in [0, 40) mm, Al at [40, 42.5) mm
COLLIMATOR_SOURCE = (32.0, 32.0, 142.5)  # point source, collimator box frame (mm)
# collimator box frame -> ct256 phantom frame: the source lands where the ct256 cone
# source sits (600 mm above the phantom centre), the aperture axis on the phantom centre
COLLIMATOR_TO_CT256 = (96.0, 96.0, 521.5)


def collimator_labels(seed=4):
    """uint8 labels (z, y, x): 0 air, 1 tungsten, 2 aluminium. The aperture is 20x20 mm
    centred on the source axis; each 5 mm slab of the block moves each aperture edge by
    -1..+1 voxel (seeded), so the block is a seed-generated casting, not an ideal cut."""
    g = np.random.default_rng(seed)
    nx, ny, nz = COLLIMATOR_SHAPE
    sp = COLLIMATOR_SPACING
    lab = np.zeros((nz, ny, nx), np.uint8)
    nw = int(round(40.0 / sp))
    lab[:nw] = 1
    lab[nw:] = 2
    c = (np.arange(nx) + 0.5) * sp
    for z0 in range(0, nw, 10):
        e = g.integers(-1, 2, size=4) * sp  # x-, x+, y-, y+ edge offsets
        inx = (c >= 22.0 + e[0]) & (c < 42.0 + e[1])
        iny = (c >= 22.0 + e[2]) & (c < 42.0 + e[3])
        lab[z0:z0 + 10][:, iny[:, None] & inx[None, :]] = 0
    return lab


def exitplane(seed=4, e_max=0.1201, capacity=0):
    """Collimator stage: returns (tables, phantom, source, tally, info). The tally records
    photons leaving through the -z face (capacity records); the world is vacuum."""
    db = MaterialDatabase()
    mats = [db.get("G4_AIR"), db.get("G4_W"), db.get("G4_Al")]
    tables = xs.build_tables(mats, e_max)
    sp = COLLIMATOR_SPACING
    ph = PhantomSpec(collimator_labels(seed), (sp, sp, sp), world_material=-1)
    e, w = P.kramers_spectrum(al_mm=0.0, db=db)
    ev, ew, cdf = spectrum_histogram(e, w)
    src = SourceSpec(pos_trans=COLLIMATOR_SOURCE, theta=(0.0, np.deg2rad(10.0)), ene_type="spectrum_histogram",
                     energy_value=ev, energy_width=ew, energy_cdf=cdf)
    tal = TallySpec(size=(8, 8, 8), spacing=(8.0, 8.0, 85 * sp / 8), edep2=False, aligned=False, record_face=4,
                    record_capacity=int(capacity))
    return tables, ph, src, tal, {"desc": EXITPLANE_DESC, "names": [m.name for m in mats]}


SPECT16_DESC = ("spect16: 16 phantoms of the batch64 kind (cubes of 64^3..192^3 voxels at 1.5-3 mm, water "
                "ellipsoid with bone/lung/soft-tissue inserts) with 140.511 keV ellipsoid sources, {hist:.0e} "
                "histories each in one batched launch; photons leaving each phantom are projected on 128 heads "
                "(128x128 pixels of 4.8 mm, hexagonal-hole lead collimator 1.5/0.2/35 mm over 9.5 mm NaI)")


def spect16(histories=100_000_000, seed=5, n_sims=16):
    """Returns (tables, [(phantom, source, tally, histories, heads_kwargs), ...], info). The
    tallies record every photon that leaves the phantom (vacuum world)."""
    db = MaterialDatabase()
    sims = P.batch_simulations(seed, n_sims)
    tables = xs.build_tables([db.get(n) for n in P.BATCH_MATERIALS], 0.1406)
    out = []
    for s in sims:
        n, sp = s["n"], s["spacing"]
        L = n * sp
        ph = PhantomSpec(s["labels"], (sp, sp, sp), world_material=-1)
        src = SourceSpec(pos_type="ellipsoid", pos_param=tuple(s["source_half"]) + (0.0,),
                         pos_trans=(L / 2, L / 2, L / 2), energy=0.140511)
        tal = TallySpec(size=(n, n, n), spacing=(sp, sp, sp), edep2=False, record_face=6,
                        record_capacity=int(histories))
        heads = {"center": (L / 2, L / 2, L / 2), "radius": L * np.sqrt(2.0) / 2 + 10.0}
        out.append((ph, src, tal, int(histories), heads))
    return tables, out, {"desc": SPECT16_DESC.format(hist=histories)}
