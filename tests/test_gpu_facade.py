"""The drop-in API on the GPU: a reference-style user script (ImageVolume + HU table,
GenericSource, DoseActor, SimulationStatisticsActor) checked against the CPU oracle
run on the very descriptors the facade handed to the C-ABI."""
import numpy as np
import pytest

import opengate_amd as gate
from oracle import oracle as O
from opengate_amd import phantoms as P

pytestmark = pytest.mark.gpu


def water_bone_image():
    lab, _ = P.water64_labels(seed=1, n=32, spacing=2.0)
    hu = np.where(lab == 1, 1200, 0).astype(np.int16)
    return gate.Image(hu, (2.0, 2.0, 2.0), origin=(-31.0, -31.0, -31.0))


def build_sim(tmp_path, seed=42, two_sources=False):
    u = gate.g4_units
    sim = gate.Simulation()
    sim.output_dir = str(tmp_path)
    sim.random_seed = seed
    sim.world.material = "G4_Galactic"
    ct = sim.add_volume("Image", "ct")
    ct.image = water_bone_image()
    ct.material = "G4_AIR"
    ct.voxel_materials = [[-1000, 500, "G4_WATER"], [500, 3000, "G4_BONE_CORTICAL_ICRP"]]
    ct.translation = [5 * u.mm, -3 * u.mm, 2 * u.mm]
    src = sim.add_source("GenericSource", "tc99m")
    src.particle = "gamma"
    src.position.type = "sphere"
    src.position.radius = 4 * u.mm
    src.position.translation = [5 * u.mm, -3 * u.mm, 2 * u.mm]
    src.direction.type = "iso"
    src.energy.type = "mono"
    src.energy.mono = 140.5 * u.keV
    src.n = 40_000
    if two_sources:
        s2 = sim.add_source("GenericSource", "lu")
        s2.position.type = "box"
        s2.position.size = [10 * u.mm, 10 * u.mm, 10 * u.mm]
        s2.energy.type = "spectrum_discrete"
        s2.energy.spectrum_energies = [112.95 * u.keV, 208.37 * u.keV]
        s2.energy.spectrum_weights = [0.062, 0.1038]
        s2.n = 30_000
    dose = sim.add_actor("DoseActor", "dose")
    dose.attached_to = "ct"
    dose.size = "like_image_volume"
    dose.spacing = "like_image_volume"
    dose.edep_uncertainty.active = True
    dose.dose.active = True
    dose.dose_uncertainty.active = True
    dose.output_filename = "out.mhd"
    stats = sim.add_actor("SimulationStatisticsActor", "stats")
    return sim, dose, stats


@pytest.mark.parametrize("two_sources", [False, True])
def test_user_script_matches_oracle(tmp_path, two_sources):
    sim, dose, stats = build_sim(tmp_path, two_sources=two_sources)
    print(f"[gpu case] facade two_sources={two_sources}", flush=True)
    sim.run()
    eng = sim.engine
    tables, ph, tal = eng.inputs
    edep = np.zeros_like(dose.edep.image.array)
    edep2 = np.zeros_like(edep)
    n_tot = 0
    for spec, n, first in eng.executed:
        r, _ = O.run(tables, ph, spec, tal, n, eng.seed, first=first)
        edep += r["edep"]
        edep2 += r["edep2"]
        n_tot += n
    assert stats.counts["events"] == n_tot == dose.number_of_samples
    g = dose.edep.image.array
    assert np.max(np.abs(g - edep)) <= 1e-9 * edep.max()
    assert np.max(np.abs(dose.edep_squared.image.array - edep2)) <= 1e-9 * edep2.max()
    # dose = edep / rho / V in Gy; uncertainty = the reference formula
    rho = tables.density.astype(np.float64)[ph.labels]
    exp_dose = O.dose_gy(edep, rho, 8.0)
    assert np.allclose(dose.dose.image.array, exp_dose, rtol=1e-9, atol=1e-30)
    np.testing.assert_allclose(dose.edep_uncertainty.image.array, O.uncertainty(edep, edep2, n_tot),
                               rtol=1e-6, atol=1e-12)
    # files written with the reference suffixes, origin of the local frame
    img = gate.read_image(tmp_path / "out_edep.mhd")
    np.testing.assert_allclose(img.origin, -32 * 2.0 / 2 + 1.0 + np.array([0.0, 0.0, 0.0]))
    assert (tmp_path / "out_dose_uncertainty.mhd").exists()
    assert stats.pps > 0


def test_ct256_labels_and_transport_parity(engine):
    from opengate_amd import workloads

    print("[gpu case] ct256 labels + 100k histories", flush=True)
    tables, ph, src, tal, info = workloads.ct256(engine)
    vm = P.ct12_materials(__import__("opengate_amd.physics.materials", fromlist=["x"]).MaterialDatabase())
    ref, lut = O.label_image(info["hu"], "CT_Air", vm)
    remap = np.array([info["names"].index(m) for m, _ in sorted(lut.items(), key=lambda kv: kv[1])])
    np.testing.assert_array_equal(ph.labels, remap[ref])
    engine.clear()
    engine.set_physics(tables)
    engine.add_simulation(ph, src, tal)
    c = engine.run([100_000], 5)
    g = engine.get_tally(0)
    r, rc = O.run(tables, ph, src, tal, 100_000, 5)
    for k in ("histories", "steps", "deposits", "compton", "rayleigh", "photo", "escaped", "world_interactions"):
        assert c[k] == rc[k], k
    assert np.max(np.abs(g["edep"] - r["edep"])) <= 1e-9 * r["edep"].max()


def test_uncertainty_goal_stops_early(tmp_path):
    """The run stops at the first check whose mean uncertainty (the reference's rule,
    restated by the oracle from the output images) is at or below the goal; every earlier
    check was above it and set the next one at (u / goal)^2 n overshoot (truncated)."""
    from opengate_amd.actors.dose import next_uncertainty_check

    sim, dose, stats = build_sim(tmp_path)
    sim.sources["tc99m"].n = 2_000_000
    dose.uncertainty_goal = 0.05
    dose.uncertainty_first_check_after_n_events = 20_000
    print("[gpu case] uncertainty goal", flush=True)
    sim.run()
    n = stats.counts["events"]
    assert n < 2_000_000
    checks = sim.engine.goal_checks
    assert checks[0][0] == 20_000 and checks[-1][0] == n
    assert checks[-1][1] <= 0.05 and all(u > 0.05 for _, u in checks[:-1])
    for (n0, u0), (n1, _) in zip(checks[:-1], checks[1:]):
        assert n1 == max(n0 + 1, next_uncertainty_check(u0, 0.05, n0, dose.uncertainty_overshoot_factor_N_events))
    u = O.dose_mean_uncertainty_reference(dose.edep.image.array, dose.edep_squared.image.array, n,
                                          dose.uncertainty_top_voxels_count, dose.uncertainty_voxel_edep_threshold)
    assert u == pytest.approx(checks[-1][1], rel=1e-9)


def test_runs_split_by_run_timing_intervals(tmp_path):
    """Two run intervals with number_of_primaries per run: the histories of both runs
    continue one stream, so the result equals one run of their sum; an activity source
    gives Poisson counts per run, and the statistics actor counts the runs."""
    u = gate.g4_units
    out = []
    for split in (True, False):
        sim, dose, stats = build_sim(tmp_path / ("split" if split else "one"))
        if split:
            sim.run_timing_intervals = [[0, 1 * u.s], [1 * u.s, 3 * u.s]]
            sim.sources["tc99m"].number_of_primaries = [15_000, 25_000]
            sim.sources["tc99m"].n = 0
        print(f"[gpu case] run split {split}", flush=True)
        sim.run()
        out.append((stats.counts["events"], stats.counts["runs"], dose.edep.image.array.copy()))
    assert out[0][0] == out[1][0] == 40_000
    assert out[0][1] == 2 and out[1][1] == 1
    assert np.max(np.abs(out[0][2] - out[1][2])) <= 1e-9 * out[1][2].max()
    sim, dose, stats = build_sim(tmp_path / "act")
    sim.run_timing_intervals = [[0, 1 * u.s], [1 * u.s, 2 * u.s]]
    src = sim.sources["tc99m"]
    src.n = 0
    src.activity = 30_000 * u.Bq
    print("[gpu case] activity per run", flush=True)
    sim.run()
    counts = sim.engine.events_per_source["tc99m"]
    assert len(counts) == 2 and stats.counts["events"] == sum(counts)
    for c in counts:
        assert abs(c - 30_000) < 5 * np.sqrt(30_000)


def test_tle_dose_actor_script_matches_oracle(tmp_path):
    """TLEDoseActor in a user script (threshold type 'energy', energy_min): the facade hands
    a TLE tally to the C-ABI and the GPU result equals the oracle's on the same inputs."""
    u = gate.g4_units
    sim, dose, stats = build_sim(tmp_path)
    del sim.actors["dose"]
    tle = sim.add_actor("TLEDoseActor", "tle")
    tle.attached_to = "ct"
    tle.size = "like_image_volume"
    tle.spacing = "like_image_volume"
    tle.edep_uncertainty.active = True
    tle.tle_threshold_type = "energy"
    tle.tle_threshold = 120 * u.keV
    tle.energy_min = 20 * u.keV
    tle.output_filename = "tle.mhd"
    print("[gpu case] facade TLEDoseActor", flush=True)
    sim.run()
    eng = sim.engine
    tables, ph, tal = eng.inputs
    assert tal.tle and abs(tal.tle_energy_max - 0.12) < 1e-12 and abs(tal.tle_energy_min - 0.02) < 1e-12
    edep = np.zeros_like(tle.edep.image.array)
    for spec, n, first in eng.executed:
        r, _ = O.run(tables, ph, spec, tal, n, eng.seed, first=first)
        edep += r["edep"]
    assert edep.sum() > 0
    assert np.max(np.abs(tle.edep.image.array - edep)) <= 1e-9 * edep.max()
    assert (tmp_path / "tle_edep_uncertainty.mhd").exists()


def test_split_jobs_merge_to_the_single_run(tmp_path):
    """Two jobs with one random_seed and disjoint history ranges (first_history), merged from
    their files, give the single run's edep, edep_squared and uncertainty (test041-mt
    pattern: N-way split == 1-way)."""
    from opengate_amd import jobs as J

    n_total = 200_000
    outs = []
    for j, (first, n) in enumerate([(0, n_total)] + J.split_histories(n_total, 2)):
        d = tmp_path / f"run{j}"
        sim, dose, stats = build_sim(d, seed=99)
        sim.sources["tc99m"].n = n
        sim.first_history = first
        dose.edep_squared.active = True
        print(f"[gpu case] job split {j}: first {first} n {n}", flush=True)
        sim.run()
        assert dose.number_of_samples == n
        outs.append(d)
    merged, n = J.merge_dose_outputs(outs[1:], "out.mhd", tmp_path / "merged")
    whole = gate.read_image(outs[0] / "out_edep.mhd").array
    whole2 = gate.read_image(outs[0] / "out_edep_squared.mhd").array
    assert n == n_total
    assert np.max(np.abs(merged["edep"].array - whole)) <= 1e-9 * whole.max()
    assert np.max(np.abs(merged["edep_squared"].array - whole2)) <= 1e-9 * whole2.max()
    np.testing.assert_allclose(merged["edep_uncertainty"].array,
                               gate.read_image(outs[0] / "out_edep_uncertainty.mhd").array, rtol=1e-6, atol=1e-12)


def test_phase_space_actor_file_feeds_phase_space_source(tmp_path):
    """PhaseSpaceActor writes the photons leaving the +z face (reference branch names);
    a second simulation reads the file with PhaseSpaceSource and transports every photon:
    one history per recorded photon, and its energy bookkeeping closes on the recorded
    sum of E x weight."""
    u = gate.g4_units
    sim = gate.Simulation()
    sim.output_dir = str(tmp_path)
    sim.random_seed = 3
    sim.world.material = "G4_Galactic"
    box = sim.add_volume("Box", "water")
    box.size = [60 * u.mm, 60 * u.mm, 60 * u.mm]
    box.material = "G4_WATER"
    s = sim.add_source("GenericSource", "tc")
    s.energy.mono = 140.5 * u.keV
    s.n = 200_000
    ps = sim.add_actor("PhaseSpaceActor", "exit")
    ps.attached_to = "water"
    ps.exit_face = "+z"
    ps.output_filename = "exit.npz"
    print("[gpu case] PhaseSpaceActor record", flush=True)
    sim.run()
    d = dict(np.load(tmp_path / "exit.npz"))
    n = len(d["KineticEnergy"])
    assert n == ps.number_of_recorded > 1000
    np.testing.assert_allclose(d["PrePositionLocal_Z"], 30.0, atol=1e-4)
    assert np.all(d["PreDirectionLocal_Z"] > 0)
    e_in = float(np.sum(d["KineticEnergy"].astype(np.float64) * d["Weight"].astype(np.float64)))

    sim2 = gate.Simulation()
    sim2.output_dir = str(tmp_path)
    sim2.random_seed = 4
    sim2.world.material = "G4_Galactic"
    b2 = sim2.add_volume("Box", "slab")
    b2.size = [200 * u.mm, 200 * u.mm, 20 * u.mm]
    b2.translation = [0, 0, 50 * u.mm]
    b2.material = "G4_WATER"
    src = sim2.add_source("PhaseSpaceSource", "phsp")
    src.phsp_file = str(tmp_path / "exit.npz")
    src.global_flag = True  # positions were recorded in a box centred on the world origin
    src.n = n
    stats = sim2.add_actor("SimulationStatisticsActor", "stats")
    print("[gpu case] PhaseSpaceSource replay", flush=True)
    c = sim2.run()
    assert stats.counts["events"] == n
    e_out = c["e_deposited"] + c["e_outside_tally"] + c["e_escaped"] + c["e_world"]
    assert e_out == pytest.approx(e_in, rel=1e-6)
    assert c["e_deposited"] > 0.01 * e_in


def test_autotuned_launch_gives_the_same_result(tmp_path):
    """Simulation.launch = "auto" picks launch parameters with gh_autotune; results do not
    depend on the launch shape (integer counters equal, FP64 tallies up to the order of
    additions) and the trial runs leave no trace in the tallies."""
    outs = []
    for j, launch in enumerate(({}, "auto")):
        sim, dose, stats = build_sim(tmp_path / f"l{j}", seed=123)
        sim.sources["tc99m"].n = 300_000
        sim.launch = launch
        print(f"[gpu case] launch {launch!r}", flush=True)
        c = sim.run()
        outs.append((c, dose.edep.image.array.copy(), dose.edep_squared.image.array.copy()))
    (c0, e0, q0), (c1, e1, q1) = outs
    for k in ("histories", "steps", "deposits", "compton", "rayleigh", "photo", "escaped"):
        assert c0[k] == c1[k], k
    assert np.max(np.abs(e0 - e1)) <= 1e-9 * e0.max()
    assert np.max(np.abs(q0 - q1)) <= 1e-9 * q0.max()
