"""VoxelSource (activity image): the library's alias table reproduces the activity exactly;
the oracle samples it (chi^2 through photons below the tracking cut, which deposit at
their emission point); the facade places the image like the reference
(origin -size/2 spacing + translation + spacing/2 in the attached volume)."""
import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O
from opengate_amd.transport import PhantomSpec, SourceSpec, TallySpec, voxel_alias_table

from helpers import tables


def _reconstruct(prob, alias):
    n = len(prob)
    p = prob.astype(np.float64).copy()
    np.add.at(p, alias, 1.0 - prob.astype(np.float64))
    return p / n


def test_alias_table_is_exact():
    g = np.random.default_rng(1)
    act = g.gamma(0.7, size=(7, 9, 11))
    act[act < 0.3] = 0.0
    act[3, 4, 5] = 50.0  # one hot voxel
    idx, prob, alias = voxel_alias_table(act)
    w = act.ravel()[idx]
    np.testing.assert_allclose(_reconstruct(prob, alias), w / w.sum(), rtol=0, atol=2e-7)
    assert np.all((prob >= 0) & (prob <= 1)) and np.all((alias >= 0) & (alias < len(idx)))
    np.testing.assert_array_equal(idx, np.flatnonzero(act.ravel() > 0))
    with pytest.raises(ValueError):
        voxel_alias_table(np.zeros((2, 2, 2)))
    with pytest.raises(ValueError):
        voxel_alias_table(-np.ones((2, 2, 2)))


def _below_cut_run(act, n, seed, threads=4):
    """Photons of 0.5 keV (below the tracking cut) deposit at their emission point: the edep
    tally on a grid twice as fine as the activity image is the emission histogram."""
    T = tables()
    nz, ny, nx = act.shape
    sp = 2.0
    ph = PhantomSpec(np.zeros((2 * nz, 2 * ny, 2 * nx), np.uint8), (sp / 2,) * 3)
    src = SourceSpec(pos_type="voxels", activity=act, activity_spacing=(sp, sp, sp), energy=0.0005)
    tal = TallySpec(size=(2 * nx, 2 * ny, 2 * nz), spacing=(sp / 2,) * 3, edep2=False)
    r, c = O.run(T, ph, src, tal, n, seed, threads=threads)
    return r["edep"] / 0.0005, c


def test_oracle_samples_the_activity():
    g = np.random.default_rng(2)
    act = g.uniform(0, 1, size=(4, 5, 6))
    act[0, 0, :] = 0.0
    act[2, 3, 1] = 8.0
    n = 400_000
    counts, c = _below_cut_run(act, n, 7)
    assert c["cut_kills"] == n
    coarse = counts.reshape(4, 2, 5, 2, 6, 2).sum((1, 3, 5))
    assert np.all(coarse[0, 0, :] == 0)
    expect = n * act / act.sum()
    m = expect > 0
    chi2 = np.sum((coarse[m] - expect[m]) ** 2 / expect[m])
    assert stats.chi2.sf(chi2, m.sum() - 1) > 1e-3
    # uniform inside each voxel: the 8 sub-voxels share a voxel's counts equally
    sub = counts.reshape(4, 2, 5, 2, 6, 2).transpose(0, 2, 4, 1, 3, 5).reshape(-1, 8)
    tot = sub.sum(1)
    e8 = tot[:, None] / 8.0
    k = tot > 0
    chi2 = np.sum((sub[k] - e8[k]) ** 2 / e8[k])
    assert stats.chi2.sf(chi2, 7 * k.sum()) > 1e-3


def test_facade_places_the_activity_image(tmp_path):
    import opengate_amd as gate
    from opengate_amd.engine import SimulationEngine

    u = gate.g4_units
    sim = gate.Simulation()
    sim.world.material = "G4_Galactic"
    ct = sim.add_volume("Box", "ct")
    ct.size = [20 * u.mm, 20 * u.mm, 20 * u.mm]
    ct.material = "G4_WATER"
    ct.translation = [10 * u.mm, 0, 0]
    act = np.zeros((2, 3, 4))
    act[1, 2, 3] = 1.0
    src = sim.add_source("VoxelSource", "v")
    src.image = gate.Image(act, (1.5, 2.0, 2.5))
    src.attached_to = "ct"
    src.position.translation = [1.0, -2.0, 0.5]
    src.energy.mono = 100 * u.keV
    src.n = 10
    eng = SimulationEngine(sim)
    vol = eng.phantom_volume()
    eng.build_geometry(vol)
    spec = eng.source_spec(src, 10)
    assert spec.pos_type == "voxels"
    # the only active voxel (x 3, y 2, z 1) spans [corner + i sp, corner + (i+1) sp] in the
    # attached volume, corner = -size/2 sp + translation; phantom frame = attached + L/2
    sp = np.array([1.5, 2.0, 2.5])
    corner = -np.array([4, 3, 2]) / 2 * sp + np.array([1.0, -2.0, 0.5])
    lo = corner + np.array([3, 2, 1]) * sp + 10.0
    np.testing.assert_allclose(np.asarray(spec.pos_trans) + np.array([3, 2, 1]) * sp, lo, atol=1e-9)
    np.testing.assert_allclose(spec.pos_rot, np.eye(3), atol=1e-12)
