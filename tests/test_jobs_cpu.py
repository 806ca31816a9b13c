"""Job split and merge of dose outputs (no GPU): contiguous history ranges, sample-count
sidecars, summed images and the uncertainty recomputed from the merged sums."""
import json

import numpy as np
import pytest

import opengate_amd as gate
from opengate_amd import jobs as J
from opengate_amd.actors.dose import DoseActor, sample_count_path


def test_split_histories():
    assert J.split_histories(10, 3) == [(0, 4), (4, 3), (7, 3)]
    assert J.split_histories(5, 1) == [(0, 5)]
    parts = J.split_histories(1_000_003, 7)
    assert sum(n for _, n in parts) == 1_000_003
    assert all(parts[k + 1][0] == parts[k][0] + parts[k][1] for k in range(6))
    with pytest.raises(ValueError):
        J.split_histories(10, 0)


def write_job(d, edep, edep2, n):
    d.mkdir(parents=True, exist_ok=True)
    sim = gate.Simulation()
    sim.output_dir = str(d)
    a = sim.add_actor("DoseActor", "dose")
    a.output_filename = "dose.mhd"
    a.edep_uncertainty.active = True
    a.edep_squared.active = True  # merging needs the squared sums on disk
    a.store_results(np.array(edep.shape[::-1]), np.array([2.0, 2.0, 2.0]), np.array([-3.0, -3.0, -3.0]), edep,
                    edep2, None, None, None, None, n)
    a.write_outputs()


def test_merge_dose_outputs(tmp_path):
    g = np.random.default_rng(3)
    # per history: deposit 1 MeV in voxel 0 with probability p; the merged uncertainty is
    # the binomial relative error sqrt((1-p)/(p (N-1)) * N/N) of the summed counts
    ns = [1000, 2500, 1500]
    k_tot = 0
    for j, n in enumerate(ns):
        k = int(g.binomial(n, 0.3))
        k_tot += k
        edep = np.zeros((2, 3, 4))
        edep[0, 0, 0] = k
        edep[1, 2, 3] = 0.5 * j
        edep2 = edep.copy()  # unit deposits: per-history squares are the deposits
        edep2[1, 2, 3] = 0.25 * j
        write_job(tmp_path / f"job{j}", edep, edep2, n)
    assert json.loads(sample_count_path(tmp_path / "job0" / "dose_edep.mhd").read_text()) == {"number_of_samples": 1000}
    merged, n = J.merge_dose_outputs([tmp_path / f"job{j}" for j in range(3)], "dose.mhd", tmp_path / "merged")
    assert n == sum(ns)
    assert merged["edep"].array[0, 0, 0] == k_tot
    N, m = float(n), k_tot / float(n)
    expect = np.sqrt((m - m * m) / (N - 1)) / m
    assert merged["edep_uncertainty"].array[0, 0, 0] == pytest.approx(expect, rel=1e-12)
    back = gate.read_image(tmp_path / "merged" / "dose_edep_uncertainty.mhd")
    np.testing.assert_allclose(back.array, merged["edep_uncertainty"].array)
    np.testing.assert_allclose(back.origin, [-3.0, -3.0, -3.0])
    assert json.loads(sample_count_path(tmp_path / "merged" / "dose_edep.mhd").read_text())["number_of_samples"] == n


def test_merge_errors(tmp_path):
    e = np.ones((2, 2, 2))
    write_job(tmp_path / "a", e, e, 10)
    write_job(tmp_path / "b", e, e, 10)
    sample_count_path(tmp_path / "b" / "dose_edep.mhd").unlink()
    with pytest.raises(FileNotFoundError):
        J.merge_dose_outputs([tmp_path / "a", tmp_path / "b"], "dose.mhd", tmp_path / "m")
    write_job(tmp_path / "c", np.ones((2, 2, 3)), np.ones((2, 2, 3)), 10)
    with pytest.raises(ValueError):
        J.merge_dose_outputs([tmp_path / "a", tmp_path / "c"], "dose.mhd", tmp_path / "m")
