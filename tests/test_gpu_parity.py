"""GPU transport vs the CPU oracle on identical inputs (through the C-ABI).

Every history follows the same FP32 path on both sides, so all integer counters
(histories, Woodcock steps, deposits, interactions by process, escapes, ...) must be
bit-exact, and the FP64 tallies must agree up to the order of FP64 additions.
"""
import numpy as np
import pytest

from oracle import oracle as O
from opengate_amd.transport import PhantomSpec, SourceSpec, TallySpec, spectrum_histogram, spectrum_lines

from helpers import aligned_tally, centre, small_phantom, tables

pytestmark = pytest.mark.gpu

INT_KEYS = ["histories", "steps", "deposits", "compton", "rayleigh", "photo", "escaped", "world_interactions",
            "cut_kills", "sq_evictions", "recorded", "kn_rejections"]
E_KEYS = ["e_deposited", "e_outside_tally", "e_escaped", "e_world"]


def compare(gpu_c, ref_c, gpu_t, ref_t, tol=1e-9):
    for k in INT_KEYS:
        assert gpu_c[k] == ref_c[k], (k, gpu_c[k], ref_c[k])
    for k in E_KEYS:
        assert abs(gpu_c[k] - ref_c[k]) <= tol * max(1.0, abs(ref_c[k])), (k, gpu_c[k], ref_c[k])
    for k, ref in ref_t.items():
        if k in ("exact_edep2", "phsp", "event_edep"):
            continue
        g = gpu_t[k]
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.max(np.abs(g - ref)) <= tol * scale, (k, np.max(np.abs(g - ref)), scale)


def run_both(engine, T, ph, src, tal, n, seed, **launch):
    engine.clear()
    engine.set_physics(T)
    sid = engine.add_simulation(ph, src, tal)
    print(f"[gpu case] n={n} seed={seed} src={src.pos_type}/{src.dir_type}/{src.ene_type} "
          f"phantom={ph.shape_xyz} launch={launch}", flush=True)
    c = engine.run([n], seed, **launch)
    g = engine.get_tally(sid)
    r, rc = O.run(T, ph, src, tal, n, seed)
    return c, g, rc, r


def test_point_source_water_bone(engine):
    T = tables()
    ph = small_phantom(32, 2.0)
    src = SourceSpec(pos_trans=centre(ph), energy=0.1405)
    tal = aligned_tally(ph, counts=True)
    c, g, rc, r = run_both(engine, T, ph, src, tal, 200_000, 1234)
    compare(c, rc, g, r)
    assert c["histories"] == 200_000


def test_box_source_lines_and_launch_shapes(engine):
    T = tables(e_max=0.21)
    ph = small_phantom(24, 3.0)
    ev, ew, cdf = spectrum_lines([0.11295, 0.20837], [0.062, 0.1038])
    src = SourceSpec(pos_type="box", pos_param=(20.0, 30.0, 10.0), pos_trans=centre(ph), ene_type="spectrum_discrete",
                     energy_value=ev, energy_cdf=cdf)
    tal = aligned_tally(ph)
    # block 0: automatic staging (all tables in LDS, one 1024-thread workgroup per CU)
    for launch in ({"block": 256, "refill_interval": 16, "chunk": 4096},
                   {"block": 64, "grid": 7, "refill_interval": 1, "chunk": 100},
                   {"block": 128, "grid": 301, "refill_interval": 64, "chunk": 333},
                   {"block": 0, "slow_min": 52, "refill_min": 16},
                   {"block": 1024, "grid": 5, "refill_interval": 8, "chunk": 1000, "slow_min": 8, "refill_min": 4}):
        c, g, rc, r = run_both(engine, T, ph, src, tal, 150_000, 77, **launch)
        compare(c, rc, g, r)


def test_small_workgroups_stage_every_table_entry(engine):
    """A launch with fewer threads than its kernel's workgroup size (64 of 256) stages the
    whole table: right after a full-size launch has left another table set in LDS, its
    results still equal the oracle's (each launch must write every entry it reads)."""
    ph = small_phantom(24, 3.0)
    src = SourceSpec(pos_trans=centre(ph), energy=0.1405)
    tal = aligned_tally(ph)
    other = tables(("G4_LUNG_ICRP", "G4_BONE_CORTICAL_ICRP"), 0.1405)
    c, g, rc, r = run_both(engine, other, ph, src, tal, 50_000, 5, block=256)
    compare(c, rc, g, r)
    T = tables(("G4_WATER", "G4_BONE_COMPACT_ICRU"), 0.1405)
    for launch in ({"block": 64, "grid": 256}, {"block": 192, "grid": 300}):
        c, g, rc, r = run_both(engine, T, ph, src, tal, 50_000, 6, **launch)
        compare(c, rc, g, r)


def test_cone_from_outside_air_world_histogram_nonaligned(engine):
    T = tables(("G4_WATER", "G4_BONE_CORTICAL_ICRP", "G4_AIR"), 0.125)
    ph = small_phantom(32, 2.0)
    ph.world_material = 2
    ph.world_lo = (-400.0, -400.0, -400.0)
    ph.world_hi = (464.0, 464.0, 464.0)
    e = np.arange(10, 121) * 1e-3
    mid = 0.5 * (e[1:] + e[:-1])
    w = np.clip(0.120 - mid, 0, None) / mid
    ev, ew, cdf = spectrum_histogram(e, w)
    cx, cy, cz = centre(ph)
    # cone of half-angle 15 deg from 300 mm above, pointing to -z (G4 iso convention)
    src = SourceSpec(pos_trans=(cx, cy, cz + 300.0), theta=(0.0, np.deg2rad(15.0)), ene_type="spectrum_histogram",
                     energy_value=ev, energy_width=ew, energy_cdf=cdf)
    tal = TallySpec(size=(20, 20, 20), spacing=(2.5, 2.5, 2.5), corner=(7.0, 7.0, 7.0), aligned=False,
                    dose=True, dose2=True, counts=True)
    c, g, rc, r = run_both(engine, T, ph, src, tal, 100_000, 99)
    compare(c, rc, g, r)
    assert c["world_interactions"] > 0


def test_spectra_with_empty_bins_and_lines(engine):
    """Spectra whose CDF has flat stretches (zero-weight bins and lines, many narrow bins in
    one guide bin) and a histogram of a few hundred bins: the guided CDF search of the
    kernel must pick the same bin as the oracle's plain binary search."""
    T = tables(e_max=0.21)
    ph = small_phantom(24, 3.0)
    tal = aligned_tally(ph)
    # lines: zero weights first, in the middle and last
    ev, ew, cdf = spectrum_lines([0.02, 0.05, 0.08, 0.1405, 0.16, 0.2], [0.0, 1.0, 0.0, 3.0, 1e-6, 0.0])
    src = SourceSpec(pos_trans=centre(ph), ene_type="spectrum_discrete", energy_value=ev, energy_cdf=cdf)
    c, g, rc, r = run_both(engine, T, ph, src, tal, 100_000, 5)
    compare(c, rc, g, r)
    # histogram: 400 bins, a third of them empty, one bin holding most of the weight
    rng = np.random.default_rng(8)
    e = np.linspace(0.010, 0.200, 401)
    w = rng.uniform(0.0, 1.0, 400)
    w[rng.uniform(size=400) < 0.33] = 0.0
    w[0] = 0.0
    w[-1] = 0.0
    w[123] = 500.0
    ev, ew, cdf = spectrum_histogram(e, w)
    src = SourceSpec(pos_trans=centre(ph), ene_type="spectrum_histogram", energy_value=ev, energy_width=ew,
                     energy_cdf=cdf)
    c, g, rc, r = run_both(engine, T, ph, src, tal, 100_000, 6)
    compare(c, rc, g, r)


def test_restricted_phi_and_sphere_source(engine):
    T = tables()
    ph = small_phantom(20, 2.5)
    src = SourceSpec(pos_type="sphere", pos_param=(8.0,), pos_trans=centre(ph), theta=(0.3, 2.0),
                     phi=(0.2, 4.0), energy=0.1)
    tal = aligned_tally(ph)
    c, g, rc, r = run_both(engine, T, ph, src, tal, 100_000, 5)
    compare(c, rc, g, r)


def test_phase_space_source_iso_and_file_directions(engine):
    T = tables()
    ph = small_phantom(24, 2.0)
    rng = np.random.default_rng(3)
    n = 5000
    cx, cy, cz = centre(ph)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    arrs = [np.full(n, cx, np.float32) + rng.uniform(-5, 5, n).astype(np.float32),
            np.full(n, cy, np.float32), np.full(n, cz, np.float32),
            d[:, 0].astype(np.float32), d[:, 1].astype(np.float32), d[:, 2].astype(np.float32),
            rng.uniform(0.02, 0.14, n).astype(np.float32), rng.uniform(0.5, 1.5, n).astype(np.float32)]
    for iso in (False, True):
        src = SourceSpec(pos_type="phsp", phsp=arrs, phsp_start=17, phsp_iso=iso)
        tal = aligned_tally(ph)
        c, g, rc, r = run_both(engine, T, ph, src, tal, 20_000, 11)
        compare(c, rc, g, r)


def test_phase_space_source_on_device(engine):
    """The same records handed over as device arrays (used in place) give the oracle's result;
    an out-of-table energy in device memory is rejected before any transport."""
    import torch

    from opengate_amd.transport import GhError

    T = tables()
    ph = small_phantom(24, 2.0)
    rng = np.random.default_rng(4)
    n = 3000
    cx, cy, cz = centre(ph)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    arrs = [np.full(n, cx, np.float32), np.full(n, cy, np.float32) + rng.uniform(-4, 4, n).astype(np.float32),
            np.full(n, cz, np.float32), d[:, 0].astype(np.float32), d[:, 1].astype(np.float32),
            d[:, 2].astype(np.float32), rng.uniform(0.02, 0.14, n).astype(np.float32),
            rng.uniform(0.5, 1.5, n).astype(np.float32)]
    dev = [torch.from_numpy(a).cuda() for a in arrs]
    tal = aligned_tally(ph)
    engine.clear()
    engine.set_physics(T)
    sid = engine.add_simulation(ph, SourceSpec(pos_type="phsp", phsp=dev, phsp_start=5), tal)
    c = engine.run([10_000], 21)
    g = engine.get_tally(sid)
    r, rc = O.run(T, ph, SourceSpec(pos_type="phsp", phsp=arrs, phsp_start=5), tal, 10_000, 21)
    compare(c, rc, g, r)
    engine.clear()
    engine.set_physics(T)
    for k, v in ((6, float("nan")), (6, -0.01), (7, -1.0), (7, float("inf"))):
        bad = [t.clone() for t in dev]
        bad[k][17] = v
        with pytest.raises(GhError):
            engine.add_simulation(ph, SourceSpec(pos_type="phsp", phsp=bad), tal)


def test_tle_estimator_bit_exact(engine):
    """Track-length (collision) kerma estimator: all photons, then only photons at or
    below 80 keV (analog above), with an energy floor."""
    T = tables()
    ph = small_phantom(32, 2.0)
    src = SourceSpec(pos_trans=centre(ph), energy=0.1405)
    for kw in ({"tle": True}, {"tle": True, "tle_energy_max": 0.08, "tle_energy_min": 0.03}):
        tal = aligned_tally(ph, **kw)
        c, g, rc, r = run_both(engine, T, ph, src, tal, 100_000, 31)
        compare(c, rc, g, r)


@pytest.mark.parametrize("n_mat", [16, 17])
def test_label_packing_odd_sizes(engine, n_mat):
    """16 materials: labels packed two per byte on the GPU; 17: one byte each. Odd phantom
    dimensions (odd voxel count, so the last byte is half used); every label value used."""
    from opengate_amd.physics.materials import MaterialDatabase
    from opengate_amd.physics import xs

    from opengate_amd.units import g4_units

    db = MaterialDatabase()
    for k in range(3):
        db.add_material_weights(f"WATER_{k}", ["H", "O"], [0.111894, 0.888106],
                                (1.05 + 0.05 * k) * g4_units.g / g4_units.cm3)
    names = ["G4_WATER", "G4_BONE_CORTICAL_ICRP", "G4_LUNG_ICRP", "G4_AIR", "G4_TISSUE_SOFT_ICRP", "G4_Al",
             "G4_MUSCLE_SKELETAL_ICRP", "G4_ADIPOSE_TISSUE_ICRP", "G4_BONE_COMPACT_ICRU", "G4_PLEXIGLASS", "G4_Ti",
             "G4_SODIUM_IODIDE", "G4_Cu", "G4_Pb", "WATER_0", "WATER_1", "WATER_2"][:n_mat]
    T = xs.build_tables([db.get(n) for n in names], 0.1405)
    g = np.random.default_rng(n_mat)
    lab = g.integers(0, n_mat, size=(11, 13, 15)).astype(np.uint8)
    lab[-1, -1, -1] = n_mat - 1
    ph = PhantomSpec(lab, (3.0, 2.5, 2.0))
    src = SourceSpec(pos_type="box", pos_param=(20.0, 20.0, 15.0), pos_trans=centre(ph), energy=0.1)
    tal = aligned_tally(ph)
    c, g_, rc, r = run_both(engine, T, ph, src, tal, 60_000, 3)
    compare(c, rc, g_, r)


@pytest.mark.parametrize("launch", [{"block": 0}, {"block": 256}], ids=["auto", "wg256"])
def test_many_materials_partial_and_no_staging(engine, launch):
    """40 materials: the full table set (about 290 KiB) exceeds the LDS of a CU, so the
    automatic launch stages only mu and the majorant, with 1024-thread workgroups; with
    256-thread workgroups even those exceed the per-workgroup budget and every table is
    read from global memory. Both equal the oracle."""
    from opengate_amd.physics.materials import MaterialDatabase
    from opengate_amd.physics import xs
    from opengate_amd.units import g4_units

    db = MaterialDatabase()
    names = ["G4_WATER", "G4_BONE_CORTICAL_ICRP", "G4_LUNG_ICRP", "G4_AIR", "G4_TISSUE_SOFT_ICRP", "G4_Al",
             "G4_MUSCLE_SKELETAL_ICRP", "G4_ADIPOSE_TISSUE_ICRP", "G4_BONE_COMPACT_ICRU", "G4_PLEXIGLASS"]
    for k in range(30):
        db.add_material_weights(f"WATER_{k}", ["H", "O"], [0.111894, 0.888106],
                                (0.9 + 0.02 * k) * g4_units.g / g4_units.cm3)
        names.append(f"WATER_{k}")
    T = xs.build_tables([db.get(n) for n in names], 0.1405)
    nb1, nx, m = T.n_buckets + 1, len(T.ray_x), len(names)
    lds1 = 4 * (m * nb1 + nb1)
    lds2 = lds1 + 4 * (2 * m * nb1 + 2 * m * nx + nx)
    assert lds2 > 160 * 1024 and 36 * 1024 < lds1 < 100 * 1024
    g = np.random.default_rng(40)
    lab = g.integers(0, m, size=(13, 11, 9)).astype(np.uint8)
    ph = PhantomSpec(lab, (2.0, 2.5, 3.0))
    src = SourceSpec(pos_type="box", pos_param=(18.0, 18.0, 18.0), pos_trans=centre(ph), energy=0.1)
    tal = aligned_tally(ph)
    c, g_, rc, r = run_both(engine, T, ph, src, tal, 50_000, 4, **launch)
    compare(c, rc, g_, r)


def test_degenerate_runs(engine):
    """Zero histories, a one-voxel phantom, and a source that never reaches the phantom."""
    T = tables()
    ph = small_phantom(1, 5.0)
    src = SourceSpec(pos_trans=centre(ph), energy=0.1405)
    tal = aligned_tally(ph)
    engine.clear()
    engine.set_physics(T)
    engine.add_simulation(ph, src, tal)
    c0 = engine.run([0], 1)
    assert c0["histories"] == 0 and c0["steps"] == 0
    assert np.all(engine.get_tally(0)["edep"] == 0)
    c, g, rc, r = run_both(engine, T, ph, src, tal, 10_000, 2)
    compare(c, rc, g, r)
    away = SourceSpec(pos_trans=(100.0, 100.0, 100.0), momentum=(1.0, 0.0, 0.0), dir_type="momentum", energy=0.1)
    c, g, rc, r = run_both(engine, T, ph, away, tal, 10_000, 2)
    compare(c, rc, g, r)
    assert c["escaped"] == 10_000 and c["steps"] == 0


def test_voxel_source(engine):
    """Activity-image source (alias table on the GPU): counters and tallies of the oracle."""
    T = tables()
    ph = small_phantom(24, 2.0)
    g = np.random.default_rng(5)
    act = g.gamma(0.8, size=(6, 7, 8))
    act[act < 0.4] = 0.0
    src = SourceSpec(pos_type="voxels", activity=act, activity_spacing=(3.0, 2.5, 2.0), pos_trans=(12.0, 14.0, 16.0),
                     energy=0.1405)
    tal = aligned_tally(ph, counts=True)
    c, g_, rc, r = run_both(engine, T, ph, src, tal, 100_000, 13)
    compare(c, rc, g_, r)


def test_batched_launch_equals_separate_runs(engine):
    T = tables(e_max=0.21)
    specs = []
    for k, (n, sp) in enumerate([(16, 3.0), (24, 2.0), (20, 1.5)]):
        ph = small_phantom(n, sp)
        src = SourceSpec(pos_type="box", pos_param=(10.0, 10.0, 10.0), pos_trans=centre(ph),
                         energy=[0.1405, 0.2, 0.06][k])
        specs.append((ph, src, aligned_tally(ph)))
    engine.clear()
    engine.set_physics(T)
    for s in specs:
        engine.add_simulation(*s)
    hist = [30_000, 50_000, 7_000]
    print(f"[gpu case] batched 3 sims {hist}", flush=True)
    c = engine.run(hist, 2024, first_history=[0, 1000, 5])
    tot = {k: 0 for k in INT_KEYS}
    for sid, ((ph, src, tal), n, first) in enumerate(zip(specs, hist, [0, 1000, 5])):
        r, rc = O.run(T, ph, src, tal, n, 2024, first=first, sim_id=sid)
        g = engine.get_tally(sid)
        for k in ("edep", "edep2"):
            assert np.max(np.abs(g[k] - r[k])) <= 1e-9 * r[k].max()
        for k in INT_KEYS:
            tot[k] += rc[k]
    for k in INT_KEYS:
        assert c[k] == tot[k], k


@pytest.mark.parametrize("launch", [{"block": 256}, {"block": 0}], ids=["wg256", "staged"])
def test_tle_kerma_estimator_launch_shapes(engine, launch):
    """Track-length (collision kerma) dose tally: counters and tallies equal the oracle's for
    the 256-thread form and the staged form (all tables in LDS, 1024-thread workgroups)."""
    T = tables(e_max=0.21)
    ph = small_phantom(24, 2.0)
    src = SourceSpec(pos_trans=centre(ph), energy=0.1405)
    tal = aligned_tally(ph, tle=True, tle_energy_min=0.003, tle_energy_max=0.15)
    c, g, rc, r = run_both(engine, T, ph, src, tal, 60_000, 21, **launch)
    compare(c, rc, g, r)


def test_batched_recording_per_simulation(engine):
    """Exit-face recording in a batched launch: with small chunks the lanes of a wave hold
    photons of different simulations, and each record must land in the buffer of its own
    simulation (same multiset of rows as the oracle run of that simulation)."""
    T = tables(e_max=0.21)
    specs = []
    for k, (n, sp, face) in enumerate([(16, 3.0, 6), (20, 2.0, 5), (12, 2.5, 6)]):
        ph = small_phantom(n, sp)
        src = SourceSpec(pos_type="box", pos_param=(8.0, 8.0, 8.0), pos_trans=centre(ph),
                         energy=[0.1405, 0.2, 0.06][k])
        specs.append((ph, src, aligned_tally(ph, record_face=face, record_capacity=60_000)))
    engine.clear()
    engine.set_physics(T)
    for sp_ in specs:
        engine.add_simulation(*sp_)
    hist = [20_000, 30_000, 9_000]
    first = [0, 500, 7]
    # 8 workgroups of 256 threads (32 waves) over 923 chunks of 64 histories: every wave
    # crosses chunk boundaries, and the waves that cross a simulation boundary mix sims
    print(f"[gpu case] batched recording 3 sims {hist} chunk 64 grid 8", flush=True)
    c = engine.run(hist, 31, first_history=first, block=256, grid=8, chunk=64, refill_interval=4)
    total = 0
    key = lambda a: np.lexsort([a[6], a[2], a[1], a[0]])
    for sid, ((ph, src, tal), n, f) in enumerate(zip(specs, hist, first)):
        arrs, nrec = engine.get_phsp(sid)
        r, rc = O.run(T, ph, src, tal, n, 31, first=f, sim_id=sid, record=True)
        assert nrec == rc["recorded_total"], sid
        g = np.stack(arrs)[:, key(arrs)]
        rr = np.stack([x[:nrec] for x in r["phsp"]])
        rr = rr[:, key(rr)]
        np.testing.assert_array_equal(g, rr)
        total += nrec
    assert c["recorded"] == total


def test_recorder_overflow_keeps_count_and_real_rows(engine):
    """More exit photons than the record capacity: the count stays exact, exactly capacity
    rows are stored, and each stored row is one of the oracle's records (which ones are
    kept depends on the order the waves reserve their slots)."""
    T = tables()
    ph = small_phantom(16, 2.0)
    src = SourceSpec(pos_trans=centre(ph), energy=0.1405)
    cap = 1000
    engine.clear()
    engine.set_physics(T)
    sid = engine.add_simulation(ph, src, aligned_tally(ph, record_face=6, record_capacity=cap))
    n = 20_000
    print("[gpu case] recorder overflow", flush=True)
    c = engine.run([n], 5)
    arrs, nrec = engine.get_phsp(sid)
    r, rc = O.run(T, ph, src, aligned_tally(ph, record_face=6, record_capacity=n), n, 5, record=True)
    assert c["recorded"] == nrec == rc["recorded_total"] > cap
    assert all(len(a) == cap for a in arrs)
    ref_rows = {tuple(row) for row in np.stack([x[:rc["recorded_total"]] for x in r["phsp"]]).T.tolist()}
    got_rows = [tuple(row) for row in np.stack(arrs).T.tolist()]
    assert len(set(got_rows)) == cap and all(row in ref_rows for row in got_rows)


def test_recorder_and_detector_event_energy(engine):
    T = tables()
    ph = small_phantom(20, 2.0)
    src = SourceSpec(pos_trans=centre(ph), energy=0.1405)
    tal = aligned_tally(ph, record_face=5, record_capacity=200_000, detector_label_min=1)
    engine.clear()
    engine.set_physics(T)
    sid = engine.add_simulation(ph, src, tal)
    n = 50_000
    print("[gpu case] recorder + detector", flush=True)
    c = engine.run([n], 8)
    arrs, nrec = engine.get_phsp(sid)
    ev = engine.get_event_edep(sid, n)
    r, rc = O.run(T, ph, src, tal, n, 8, record=True, event=True)
    assert nrec == rc["recorded_total"] == c["recorded"]
    np.testing.assert_array_equal(ev, r["event_edep"])
    # recorded rows: same multiset (GPU order depends on scheduling)
    key = lambda a: np.lexsort([a[6], a[2], a[1], a[0]])
    g = np.stack(arrs)[:, key(arrs)]
    rr = np.stack([x[:nrec] for x in r["phsp"]])
    rr = rr[:, key(rr)]
    np.testing.assert_array_equal(g, rr)


def test_hu_to_labels_bit_exact(engine):
    rng = np.random.default_rng(0)
    hu = rng.integers(-1100, 3100, size=(17, 33, 65)).astype(np.int16)
    hu.ravel()[:4200] = np.arange(-1100, 3100)[:4200]
    vm = [[-1050, -950, "G4_AIR"], [-950, -120, "G4_LUNG_ICRP"], [-120, 19, "G4_ADIPOSE_TISSUE_ICRP"],
          [19, 80, "G4_WATER"], [120, 400.5, "G4_MUSCLE_SKELETAL_ICRP"], [400.5, 3000, "G4_BONE_CORTICAL_ICRP"]]
    ref, lut = O.label_image(hu, "G4_AIR", vm)
    bins, ls = O.label_bins(lut, vm)
    out = engine.hu_to_labels(hu, bins, ls)
    np.testing.assert_array_equal(out, ref)


@pytest.mark.gpu
def test_hu_to_labels_float_images_and_many_bins(engine):
    """Float HU images (any dtype, create_label_image, volumes.py:1233-1239) and tables of
    more than 256 bins, against numpy.digitize itself: values on every bin edge, one ulp
    either side, NaN and infinities."""
    g = np.random.default_rng(5)
    bins = np.unique(np.round(g.uniform(-1100, 3100, size=600), 3)).astype(np.float64)
    lut = g.integers(0, 700, size=len(bins) + 1).astype(np.uint16)
    lut[0] = lut[-1] = 0
    for dt in (np.float32, np.float64, np.int32):
        edges = bins.astype(dt)
        vals = np.concatenate([edges, np.nextafter(edges, -np.inf) if dt != np.int32 else edges - 1,
                               np.nextafter(edges, np.inf) if dt != np.int32 else edges + 1,
                               g.uniform(-1200, 3200, size=5000).astype(dt)])
        if dt != np.int32:
            vals = np.concatenate([vals, np.array([np.nan, np.inf, -np.inf], dtype=dt)])
        print(f"[gpu case] hu_to_labels {np.dtype(dt).name} {len(bins)} bins", flush=True)
        out = engine.hu_to_labels(vals.reshape(1, 1, -1), bins, lut)
        exp = lut[np.digitize(vals.astype(np.float64), bins)]
        np.testing.assert_array_equal(out.ravel(), exp)
