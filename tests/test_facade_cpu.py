"""Host-side logic of the drop-in API (no GPU): units, images, outputs, frames, spectra,
C-ABI exports."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import opengate_amd as gate
from oracle import oracle as O
from opengate_amd import _abi as A
from opengate_amd.actors.dose import insert_suffix_before_extension, variance_or_uncertainty
from opengate_amd.engine import SimulationEngine
from opengate_amd.image import Image, read_image, write_image
from opengate_amd.transport import spectrum_histogram, spectrum_lines

ROOT = Path(__file__).resolve().parent.parent


def test_units():
    u = gate.g4_units
    assert u.mm == 1 and u.cm == 10 and u.m == 1000 and u.keV == 1e-3 and u.MeV == 1
    assert u.Gy == pytest.approx(1e-12, rel=1e-12)
    assert u.g_cm3 == pytest.approx(6.24150907e18, rel=1e-8)
    assert u.deg == pytest.approx(np.pi / 180)
    # dose of 1 MeV in 1 mm3 of water in Gy
    assert 1.0 / (1.0 * u.g_cm3) / (u.Gy * 1.0) == pytest.approx(1.602176634e-7, rel=1e-8)


def test_metaimage_roundtrip(tmp_path):
    a = np.random.default_rng(0).normal(size=(4, 5, 6))
    img = Image(a, (1.5, 2.0, 2.5), (-3.0, 1.0, 7.0))
    for ext in (".mhd", ".mha"):
        p = write_image(img, tmp_path / f"x{ext}")
        b = read_image(p)
        np.testing.assert_array_equal(b.array, a)
        np.testing.assert_allclose(b.spacing, img.spacing)
        np.testing.assert_allclose(b.origin, img.origin)
    h = write_image(Image(np.arange(24, dtype=np.int16).reshape(2, 3, 4)), tmp_path / "h.mhd")
    assert read_image(h).array.dtype == np.int16


def test_uncertainty_matches_oracle_and_closed_form():
    rng = np.random.default_rng(2)
    # per-history deposit X = E0 with probability p, else 0
    E0, p, N = 0.14, 0.3, 10_000
    x = (rng.uniform(size=(N, 5)) < p) * E0
    S, S2 = x.sum(0), (x**2).sum(0)
    u = variance_or_uncertainty(S, S2, N, "uncertainty")
    np.testing.assert_allclose(u, O.uncertainty(S, S2, N))
    # closed form of the relative standard error of the mean of a Bernoulli deposit
    assert np.allclose(u, np.sqrt((1 - p) / (p * N)), rtol=0.05)
    assert np.all(variance_or_uncertainty(S, S2, 1, "uncertainty") == 1)
    assert np.all(variance_or_uncertainty(S, None, N, "uncertainty") == 0)
    z = variance_or_uncertainty(np.zeros(3), np.zeros(3), N, "uncertainty")
    assert np.all(z == 0)


def test_output_names():
    assert insert_suffix_before_extension("test041.mhd", "edep") == Path("test041_edep.mhd")
    assert insert_suffix_before_extension("a/b.nii.gz", "dose") == Path("a/b_dose.nii.gz")
    sim = gate.Simulation()
    sim.output_dir = "/tmp/out"
    d = sim.add_actor("DoseActor", "dose")
    d.output_filename = "x.mhd"
    assert d.edep_uncertainty.get_output_path() == Path("/tmp/out/x_edep_uncertainty.mhd")
    assert d.edep.active and not d.dose.active
    with pytest.raises(AttributeError):
        d.not_a_parameter = 3
    with pytest.raises(ValueError):
        d.hit_type = "bogus"


def test_spectrum_cdfs():
    ev, _, cdf = spectrum_lines([0.1, 0.2, 0.3], [1, 2, 1])
    np.testing.assert_allclose(cdf, [0.25, 0.75, 1.0])
    # first k with u < cdf[k] (GateSPSEneDistribution::IndexForProbability)
    idx = np.searchsorted(cdf, np.array([0.0, 0.2499, 0.25, 0.9]), side="right")
    np.testing.assert_array_equal(idx, [0, 0, 1, 2])
    lo, w, cdf = spectrum_histogram([0.01, 0.02, 0.04], [1.0, 1.0])
    np.testing.assert_allclose(cdf, [1 / 3, 1.0], rtol=1e-7)
    np.testing.assert_allclose(w, [0.01, 0.02], rtol=1e-6)


def _sim_with_image(tmp_path, translation=(10.0, -20.0, 5.0)):
    sim = gate.Simulation()
    mm = gate.g4_units.mm
    img = sim.add_volume("Image", "ct")
    img.image = Image(np.zeros((10, 20, 30), np.int16), (2.0, 2.0, 2.0))
    img.voxel_materials = [[-2000, 5000, "G4_WATER"]]
    img.translation = list(translation)
    src = sim.add_source("GenericSource", "s")
    src.position.type = "point"
    src.position.translation = [10.0, -20.0, 5.0 + 100 * mm]
    src.energy.mono = 0.1405
    src.n = 10
    return sim, img, src


def test_source_frames(tmp_path):
    sim, img, src = _sim_with_image(tmp_path)
    eng = SimulationEngine(sim)
    eng.R, eng.t = sim.volume_manager.global_transform("ct")
    eng.size = img.size_pix
    eng.spacing = img.spacing
    eng.L = eng.size * eng.spacing
    spec = eng.source_spec(src, 10)
    # world point (10, -20, 105) -> phantom frame p - t + L/2 = (30, 20, 110)
    np.testing.assert_allclose(spec.pos_trans, [30.0, 20.0, 110.0])
    # iso cone pointing down the attached frame by default
    assert spec.dir_type == "iso" and spec.theta == (0.0, np.pi)
    src.direction.type = "focused"
    src.direction.focus_point = [10.0, -20.0, 5.0]
    spec = eng.source_spec(src, 10)
    np.testing.assert_allclose(spec.focus, [30.0, 20.0, 10.0])
    src.particle = "e-"
    with pytest.raises(NotImplementedError):
        eng.source_spec(src, 10)


def test_c_abi_exports_every_declared_symbol():
    header = (ROOT / "include" / "gatehip.h").read_text()
    declared = set(re.findall(r"^(?:int|void|const char \*)\s*(gh_\w+)\(", header, re.M))
    assert len(declared) >= 15
    lib = A.load_library()
    for name in declared:
        assert hasattr(lib, name), name
    assert set(A.SIGNATURES) == declared
    assert lib.gh_abi_version() == 6
    # no compute without a GPU: creating a context fails cleanly when no device is visible
    import torch

    if not torch.cuda.is_available():
        h = C.c_void_p()
        assert lib.gh_create(0, C.byref(h)) != 0


def test_tle_dose_actor_user_info():
    import opengate_amd as gate

    u = gate.g4_units
    sim = gate.Simulation()
    tle = sim.add_actor("TLEDoseActor", "tle")
    assert tle.database == "EPDL" and tle.tle_threshold_type == "None"
    assert tle.tle_energy_bounds() == (0.0, float("inf"))
    tle.tle_threshold_type = "energy"
    tle.tle_threshold = 100 * u.keV
    tle.energy_min = 5 * u.keV
    assert tle.tle_energy_bounds() == pytest.approx((0.005, 0.1))
    tle.tle_threshold_type = "max range"
    with pytest.raises(NotImplementedError):
        tle.tle_energy_bounds()
    with pytest.raises(Exception):
        tle.database = "simulated"


def test_uncertainty_goal_rule_matches_the_reference_loop():
    """dose_actor_mean_uncertainty (vectorised) against the oracle's voxel loop of
    GateDoseActor::ComputeMeanUncertainty: per-voxel clamp at 1, negative variances at 0,
    n raised to 2, no qualifying voxel -> 1; and the truncated next-check estimate."""
    from oracle import oracle as O
    from opengate_amd.actors.dose import dose_actor_mean_uncertainty, next_uncertainty_check

    g = np.random.default_rng(4)
    for trial in range(40):
        n = int(g.integers(0, 60))
        e = g.exponential(1.0, size=(5, 6, 7)) * (g.random((5, 6, 7)) < 0.7)
        # squared sums from 1 (a single deposit) to n (spread) times the mean, and some that
        # make the variance negative
        e2 = e * e * g.uniform(0.0, 1.5, size=e.shape) / max(n, 1) * g.integers(1, 4)
        top = int(g.integers(1, 300))
        th = float(g.uniform(0.0, 1.0))
        a = dose_actor_mean_uncertainty(e, e2, n, top, th)
        b = O.dose_mean_uncertainty_reference(e, e2, n, top, th)
        assert a == pytest.approx(b, rel=1e-12, abs=1e-15), trial
    assert dose_actor_mean_uncertainty(np.zeros(8), np.zeros(8), 100, 10, 0.7) == 1.0
    # Bernoulli voxels: k of n events deposit 1 (S = S2 = k, p = k / n):
    # unc = sqrt((p - p^2) / (n - 1)) / p = sqrt((1 - p) / (p (n - 1)))
    n = 400
    k = np.array([100.0, 120.0, 150.0, 400.0])
    p = k / n
    exp = np.mean(np.sqrt((1 - p) / (p * (n - 1))))
    assert dose_actor_mean_uncertainty(k, k, n, 10, 0.1) == pytest.approx(exp, rel=1e-12)
    # one deposit in n events: unc = 1 (the clamp holds it there)
    assert dose_actor_mean_uncertainty(np.ones(3), np.ones(3), n, 10, 0.5) == pytest.approx(1.0)
    assert next_uncertainty_check(0.1, 0.05, 1000, 1.05) == 4200
    assert next_uncertainty_check(0.0999, 0.05, 1001, 1.0) == int((0.0999 / 0.05) * (0.0999 / 0.05) * 1001)


def test_dose_output_geometry_with_rotation(tmp_path):
    """Output origin and direction of VoxelDepositActor._update_output_coordinate_system
    (doseactors.py:170-232), worked by hand: size (4, 2, 2), spacing (2, 1, 3) give the
    centred origin (-3, -0.5, -1.5); rotation Rz(90) maps it to (0.5, -3, -1.5), plus the
    translation (10, 0, 0). The attached volume sits at (5, 0, 0) rotated by Rx(90) inside a
    mother at (0, 0, 100): "global" is Rx (10.5, -3, -1.5) + (5, 0, 100) = (15.5, 1.5, 97)."""
    sim = gate.Simulation()
    mother = sim.add_volume("Box", "mother")
    mother.size = [500.0, 500.0, 500.0]
    mother.translation = [0.0, 0.0, 100.0]
    box = sim.add_volume("Box", "box")
    box.mother = "mother"
    box.size = [100.0, 100.0, 100.0]
    box.translation = [5.0, 0.0, 0.0]
    Rx = np.array([[1.0, 0, 0], [0, 0, -1], [0, 1, 0]])
    Rz = np.array([[0.0, -1, 0], [1, 0, 0], [0, 0, 1]])
    box.rotation = Rx
    dose = sim.add_actor("DoseActor", "dose")
    dose.attached_to = "box"
    dose.size = [4, 2, 2]
    dose.spacing = [2.0, 1.0, 3.0]
    dose.translation = [10.0, 0.0, 0.0]
    dose.rotation = Rz
    gt = sim.volume_manager.global_transform("box")
    size, spacing = np.array([4, 2, 2]), np.array([2.0, 1.0, 3.0])
    expect = {None: ([-3.0, -0.5, -1.5], np.eye(3)), "local": ([10.5, -3.0, -1.5], Rz),
              "global": ([15.5, 1.5, 97.0], Rx @ Rz)}
    for system, (o, d) in expect.items():
        dose.output_coordinate_system = system
        origin, direction = dose.output_geometry(size, spacing, box, gt)
        np.testing.assert_allclose(origin, o, atol=1e-12)
        np.testing.assert_allclose(direction, d, atol=1e-12)
    # the direction reaches the written header (TransformMatrix) and comes back
    dose.output_coordinate_system = "global"
    origin, direction = dose.output_geometry(size, spacing, box, gt)
    z = np.zeros((2, 2, 4))
    dose.store_results(size, spacing, origin, z, None, None, None, None, None, 10, direction)
    img = dose.edep.image
    np.testing.assert_allclose(img.direction, Rx @ Rz, atol=1e-12)
    back = read_image(write_image(img, tmp_path / "rot_edep.mhd"))
    np.testing.assert_allclose(back.direction, Rx @ Rz, atol=1e-12)
    np.testing.assert_allclose(back.origin, [15.5, 1.5, 97.0], atol=1e-12)
