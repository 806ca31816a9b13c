"""Many materials and uint16 labels on the GPU against the oracle, bit for bit in the
integer counters and to FP64 summation order in the tallies: the ct256 phantom through the
reference's Schneider HU conversion at 0.05 g/cm3 (75 materials of 27 composition rows,
8-bit labels), a phantom spanning the whole Schneider HU range at 0.01 g/cm3 (287
materials, uint16 labels, 16-bit device layout) with shared rows and with one row per
material (287 rows: too large for LDS, read from global memory), power-of-two densities
where shared and per-material tables are the same, and one batched launch that mixes
4-, 8- and 16-bit label layouts."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from opengate_amd import workloads
from opengate_amd.physics import xs
from opengate_amd.transport import PhantomSpec, SourceSpec, TallySpec, spectrum_histogram

from test_gpu_parity import INT_KEYS, compare, run_both
from test_material_rows import layered_phantom, water_family

THREADS = max(1, min(16, os.cpu_count() or 1))
pytestmark = pytest.mark.gpu


def test_ct256_schneider_75_materials(engine):
    T, ph, src, tal, info = workloads.ct256_schneider(engine, 0.05)
    assert info["n_materials"] == 75 and info["label_bits"] == 8 and T.n_compositions > 0
    engine.clear()
    engine.set_physics(T)
    engine.add_simulation(ph, src, tal)
    n = 300_000
    print(f"[gpu case] ct256-schneider 0.05: {info['n_materials']} materials, {n} histories", flush=True)
    c = engine.run([n], 71)
    g = engine.get_tally(0)
    r, rc = O.run(T, ph, src, tal, n, 71, threads=THREADS)
    compare(c, rc, g, r)


def schneider_ramp(tol, share):
    """A 48x40x36 phantom at 2 mm whose HU values cover the Schneider range [-1050, 3000]
    (every material of the conversion appears), through the HU -> label rule."""
    class _Host:
        def hu_to_labels(self, hu, bins, labels_sorted):
            return labels_sorted[np.digitize(hu, bins)]

    T, _, _, _, info = workloads.ct256_schneider(_Host(), tol, share_compositions=share, majorant_over_phantom=False)
    n = T.n_materials
    g = np.random.default_rng(11)
    lab = np.repeat(g.integers(1, n, size=(36, 1, 1)), 40, axis=1).repeat(48, axis=2)
    lab[:, 20:, :] = g.integers(0, n, size=(36, 20, 48))
    lab[:, :, :3] = n - 1  # the densest material on one face
    ph = PhantomSpec(lab.astype(np.uint16), (2.0, 2.0, 2.0))
    e = np.arange(20, 121) * 1e-3
    mid = 0.5 * (e[1:] + e[:-1])
    ev, ew, cdf = spectrum_histogram(e, np.clip(0.120 - mid, 0, None) / mid)
    src = SourceSpec(pos_type="box", pos_param=(60.0, 50.0, 40.0), pos_trans=(48.0, 40.0, 36.0),
                     ene_type="spectrum_histogram", energy_value=ev, energy_width=ew, energy_cdf=cdf)
    tal = TallySpec(size=(48, 40, 36), spacing=(2.0, 2.0, 2.0), dose=True)
    return T, ph, src, tal


@pytest.mark.parametrize("share", [True, False], ids=["shared-rows", "row-per-material"])
def test_287_materials_uint16_labels(engine, share):
    T, ph, src, tal = schneider_ramp(0.01, share)
    assert T.n_materials == 287 and ph.labels.dtype == np.uint16 and int(ph.labels.max()) == 286
    assert (T.composition is not None) == share
    print(f"[gpu case] 287 materials uint16 labels share={share}", flush=True)
    c, g, rc, r = run_both(engine, T, ph, src, tal, 60_000, 13)
    compare(c, rc, g, r)
    assert c["compton"] > 0 and c["photo"] > 0


def test_power_of_two_densities_shared_equals_per_material(engine):
    mats = water_family()
    ph = PhantomSpec(layered_phantom(len(mats)), (2.5, 2.5, 2.5))
    src = SourceSpec(pos_trans=(20.0, 22.5, 25.0), energy=0.1405)
    tal = TallySpec(size=(16, 18, 20), spacing=(2.5, 2.5, 2.5))
    out = []
    for share in (True, False):
        T = xs.build_tables(mats, 0.15, share_compositions=share)
        print(f"[gpu case] power-of-two densities share={share}", flush=True)
        c, g, rc, r = run_both(engine, T, ph, src, tal, 50_000, 17)
        compare(c, rc, g, r)
        out.append((c, g))
    for k in INT_KEYS:
        assert out[0][0][k] == out[1][0][k], k
    assert np.max(np.abs(out[0][1]["edep"] - out[1][1]["edep"])) <= 1e-9 * out[1][1]["edep"].max()


def test_batched_launch_mixes_label_layouts(engine):
    T, ph16, src16, tal16 = schneider_ramp(0.01, True)
    n = T.n_materials
    lab4 = np.zeros((16, 16, 16), np.uint16)
    lab4[4:12, 4:12, 4:12] = 7
    ph4 = PhantomSpec(lab4, (3.0, 3.0, 3.0))  # labels < 16: 4-bit layout
    lab8 = (np.arange(20 * 20 * 20).reshape(20, 20, 20) % 200).astype(np.uint16)
    ph8 = PhantomSpec(lab8, (2.0, 2.0, 2.0))  # labels < 256: 8-bit layout
    sims = [(ph16, src16, tal16),
            (ph4, SourceSpec(pos_trans=(24.0, 24.0, 24.0), energy=0.1), TallySpec(size=(16, 16, 16), spacing=(3, 3, 3))),
            (ph8, SourceSpec(pos_trans=(20.0, 20.0, 20.0), energy=0.06), TallySpec(size=(20, 20, 20), spacing=(2, 2, 2)))]
    engine.clear()
    engine.set_physics(T)
    for ph, src, tal in sims:
        engine.add_simulation(ph, src, tal)
    hist = [30_000, 20_000, 20_000]
    first = [0, 100_000, 200_000]
    print(f"[gpu case] batched 16/4/8-bit label layouts, {n} materials", flush=True)
    c = engine.run(hist, 23, first_history=first)
    tot = {k: 0 for k in INT_KEYS}
    for sid, ((ph, src, tal), h, f) in enumerate(zip(sims, hist, first)):
        r, rc = O.run(T, ph, src, tal, h, 23, first=f, sim_id=sid, threads=THREADS)
        g = engine.get_tally(sid)
        assert np.max(np.abs(g["edep"] - r["edep"])) <= 1e-9 * max(r["edep"].max(), 1e-300), sid
        for k in INT_KEYS:
            tot[k] += rc[k]
    for k in INT_KEYS:
        assert c[k] == tot[k], k
