"""The benchmark workloads themselves against the CPU oracle, on the GPU through the C-ABI:
config 1 at its full size (water64: 64^3 water with three seed-placed cortical-bone
spheres at 2 mm, 140.5 keV point source, 1e6 histories) and the full batch64 set (all 64
simulations of 64^3..192^3 voxels with their ellipsoid volume sources and spectra) in one
batched launch, with the history counts scaled to about 2e4 per simulation. Integer
counters must be equal and FP64 tallies equal up to the order of FP64 additions; every
position type and direction type of the source sampler is checked the same way."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from opengate_amd import workloads
from opengate_amd.transport import SourceSpec, angle_histogram, spectrum_lines

from helpers import aligned_tally, centre, small_phantom, tables
from test_gpu_parity import INT_KEYS, compare, run_both

THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.mark.gpu
def test_config1_water64_full_size(engine):
    T, ph, src, tal, info = workloads.water64()
    engine.clear()
    engine.set_physics(T)
    engine.add_simulation(ph, src, tal)
    n = 1_000_000
    print(f"[gpu case] config 1 water64 {n} histories", flush=True)
    c = engine.run([n], 2025)
    g = engine.get_tally(0)
    r, rc = O.run(T, ph, src, tal, n, 2025, threads=THREADS)
    compare(c, rc, g, r)
    assert c["histories"] == n
    # the three bone spheres are in the phantom and receive dose
    assert (ph.labels == 1).sum() > 0 and g["edep"][ph.labels == 1].sum() > 0


def batch64_scaled(per_sim=20_000):
    T, sims, info = workloads.batch64()
    h = np.array([s[3] for s in sims], dtype=np.float64)
    scaled = np.maximum(1000, np.round(h * per_sim / h.mean())).astype(np.int64)
    return T, sims, [int(x) for x in scaled]


def test_batch64_scaled_counts_keep_the_spread():
    _, sims, hist = batch64_scaled()
    assert len(sims) == 64
    assert abs(np.mean(hist) / 20_000 - 1) < 0.05
    assert max(hist) / min(hist) > 10  # the log-uniform spread of the workload survives


@pytest.mark.gpu
def test_config3_batch64_full_set_one_launch(engine):
    T, sims, hist = batch64_scaled()
    engine.clear()
    engine.set_physics(T)
    for ph, src, tal, _ in sims:
        engine.add_simulation(ph, src, tal)
    first = [k * 7919 for k in range(len(sims))]
    print(f"[gpu case] batch64 full set, {sum(hist)} histories in one launch", flush=True)
    c = engine.run(hist, 64, first_history=first)
    tot = {k: 0 for k in INT_KEYS}
    for sid, ((ph, src, tal, _), n, f) in enumerate(zip(sims, hist, first)):
        r, rc = O.run(T, ph, src, tal, n, 64, first=f, sim_id=sid, threads=THREADS)
        g = engine.get_tally(sid)
        for k in ("edep", "edep2"):
            assert np.max(np.abs(g[k] - r[k])) <= 1e-9 * max(r[k].max(), 1e-300), (sid, k)
        for k in INT_KEYS:
            tot[k] += rc[k]
    for k in INT_KEYS:
        assert c[k] == tot[k], k
    assert c["histories"] == sum(hist)


POS_CASES = [("point", ()), ("box", (14.0, 9.0, 20.0)), ("sphere", (9.0,)), ("ellipsoid", (12.0, 6.0, 9.0)),
             ("disc", (11.0,)), ("cylinder", (8.0, 10.0)), ("surface_sphere", (9.0,))]
DIR_CASES = ["iso", "momentum", "focused", "cos", "beam2d", "histogram"]


@pytest.mark.gpu
@pytest.mark.parametrize("dir_type", DIR_CASES)
@pytest.mark.parametrize("pos_type,param", POS_CASES, ids=[p for p, _ in POS_CASES])
def test_every_position_and_direction_type(engine, pos_type, param, dir_type):
    """Every position shape of GateGenericSource (GateGenericSource.cpp:312-384) under each
    direction type, with a rotated source frame, against the oracle bit for bit."""
    T = tables(e_max=0.21)
    ph = small_phantom(24, 2.5)
    a = 0.4
    rot = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]]) @ \
        np.array([[1, 0, 0], [0, np.cos(0.7), -np.sin(0.7)], [0, np.sin(0.7), np.cos(0.7)]])
    ev, _, cdf = spectrum_lines([0.11295, 0.20837], [0.062, 0.1038])
    cx, cy, cz = centre(ph)
    src = SourceSpec(pos_type=pos_type, pos_param=tuple(param) + (0.0,) * (4 - len(param)), pos_rot=rot,
                     pos_trans=(cx + 1.3, cy - 0.7, cz + 2.1), dir_rot=rot, dir_type=dir_type,
                     momentum=(0.3, -0.4, 0.866), focus=(cx - 20.0, cy + 5.0, cz),
                     ene_type="spectrum_discrete", energy_value=ev, energy_cdf=cdf)
    if dir_type in ("iso", "cos"):
        src.theta = (0.2, 2.5) if dir_type == "iso" else (0.1, 1.3)
        src.phi = (0.3, 4.0)
        src.surface_axis = (0.2, -0.4, 0.9)
    elif dir_type == "beam2d":
        src.dir_sigma = (0.05, 0.3)
    elif dir_type == "histogram":
        src.theta_table = angle_histogram([0.2, 0.9, 1.4, 2.9], [1.0, 0.0, 2.0], 0.0, np.pi)
        src.phi_table = angle_histogram([-0.5, 1.0, 2.0, 2.0001, 7.0], [1.0, 3.0, 2.0, 0.5], 0.0, 2 * np.pi)
    tal = aligned_tally(ph, counts=True)
    c, g, rc, r = run_both(engine, T, ph, src, tal, 60_000, 31)
    compare(c, rc, g, r)
    assert c["histories"] == 60_000


SPREAD_CASES = {
    # disc with its Gaussian spread (GateGenericSource.cpp:343-366) and a Gaussian energy
    "disc-spread-gauss": dict(pos_type="disc", pos_param=(11.0, 0, 0, 0), pos_sigma=(3.0, 1.0),
                              ene_type="gauss", energy=0.14, energy_sigma=0.008),
    # Gaussian energy with 16 % of the draws negative (energy 0: stops at birth) and a
    # Gaussian history weight (GateGenericSource.cpp:247-259)
    "sphere-gauss-at-zero-weight": dict(pos_type="surface_sphere", pos_param=(9.0, 0, 0, 0),
                                        ene_type="gauss", energy=0.02, energy_sigma=0.02,
                                        weight=2.0, weight_sigma=0.5),
    # Gaussian weight on a line spectrum: weights of both signs reach the tally
    "box-weight": dict(pos_type="box", pos_param=(14.0, 9.0, 20.0, 0), weight=0.3, weight_sigma=0.3),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", list(SPREAD_CASES))
def test_gaussian_spreads_energies_and_weights(engine, case):
    """Gaussian source spreads (disc sigma_x / sigma_y, energy type gauss, weight_sigma) and
    the surface sphere, with a rotated source frame, against the oracle bit for bit."""
    T = tables(e_max=0.21)
    ph = small_phantom(24, 2.5)
    a = 0.4
    rot = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    cx, cy, cz = centre(ph)
    kw = dict(SPREAD_CASES[case])
    if "ene_type" not in kw:
        ev, _, cdf = spectrum_lines([0.11295, 0.20837], [0.062, 0.1038])
        kw.update(ene_type="spectrum_discrete", energy_value=ev, energy_cdf=cdf)
    src = SourceSpec(pos_rot=rot, pos_trans=(cx + 1.3, cy - 0.7, cz + 2.1), dir_rot=rot, theta=(0.0, 2.0), **kw)
    tal = aligned_tally(ph, counts=True)
    c, g, rc, r = run_both(engine, T, ph, src, tal, 60_000, 37)
    compare(c, rc, g, r)
    assert c["histories"] == 60_000
    if kw.get("energy") == 0.02:
        assert c["cut_kills"] > 0.1 * 60_000  # the energy-0 births stop at once
