"""Pins of the oracle's FP64 helpers against what the mathematics fixes (CPU):

* siddon_path_lengths: chords are positive, lie in the grid, and sum to the length of the
  segment inside the box (slab clipping, analytic); a segment inside the box keeps its
  whole length; a segment along a voxel row crosses one voxel per spacing;
* beer_lambert: exp(-mu L) in a homogeneous box, and exp(-mu1 L1 - mu2 L2) across a
  plane interface at a voxel face, with L1, L2 from the geometry;
* mu_at (the table lookup the transport sees): equal to the table at the grid nodes,
  linear in the bucket fraction between them, and within 0.1 % of the FP64 mixture model;
* dose_gy: 1 MeV in 1 mm^3 of unit-density matter is 1.602176634e-13 J / 1e-6 kg.
"""
import numpy as np
import pytest

from oracle import oracle as O
from opengate_amd.physics import xs

from helpers import DB, tables


def clip_length(p0, p1, L):
    """length of the part of segment p0->p1 inside the box [0, L] (Liang-Barsky, FP64)"""
    p0, p1 = np.asarray(p0, float), np.asarray(p1, float)
    d = p1 - p0
    t0, t1 = 0.0, 1.0
    for i in range(3):
        if d[i] == 0:
            if p0[i] < 0 or p0[i] > L[i]:
                return 0.0
            continue
        a, b = (0 - p0[i]) / d[i], (L[i] - p0[i]) / d[i]
        t0, t1 = max(t0, min(a, b)), min(t1, max(a, b))
    return max(0.0, t1 - t0) * np.linalg.norm(d)


def test_siddon_chords_sum_to_the_clipped_length():
    g = np.random.default_rng(11)
    shape = (9, 7, 5)
    sp = np.array([1.5, 2.0, 3.2])
    L = np.array(shape) * sp
    for _ in range(300):
        p0 = g.uniform(-10, L + 10)
        p1 = g.uniform(-10, L + 10)
        idx, seg = O.siddon_path_lengths(p0, p1, shape, sp)
        assert np.all(seg > 0)
        assert np.all((idx >= 0) & (idx < np.array(shape)))
        assert seg.sum() == pytest.approx(clip_length(p0, p1, L), rel=1e-12, abs=1e-12)


def test_siddon_inside_segment_and_voxel_row():
    shape, sp = (10, 10, 10), np.array([2.0, 2.0, 2.0])
    idx, seg = O.siddon_path_lengths([3.1, 4.2, 5.3], [15.7, 11.1, 6.9], shape, sp)
    assert seg.sum() == pytest.approx(np.linalg.norm([12.6, 6.9, 1.6]), rel=1e-12)
    # along x through voxel centres: one chord of 2 mm per voxel, indices 0..9 in order
    idx, seg = O.siddon_path_lengths([-5.0, 7.0, 9.0], [30.0, 7.0, 9.0], shape, sp)
    assert np.allclose(seg, 2.0) and list(idx[:, 0]) == list(range(10))
    assert np.all(idx[:, 1] == 3) and np.all(idx[:, 2] == 4)


def test_beer_lambert_homogeneous_box():
    g = np.random.default_rng(12)
    shape = (8, 6, 5)
    sp = (2.5, 2.5, 2.5)
    lab = np.zeros(shape[::-1], np.uint8)
    L = np.array(shape) * np.array(sp)
    mu = 0.0153
    for _ in range(50):
        p0, p1 = g.uniform(-5, L + 5), g.uniform(-5, L + 5)
        expect = np.exp(-mu * clip_length(p0, p1, L))
        assert O.beer_lambert(p0, p1, lab, sp, [mu]) == pytest.approx(expect, rel=1e-12)


def test_beer_lambert_two_slabs():
    """Water below z = 10 mm, bone above: a ray from (x, y, -1) to (x', y', 21) crosses
    10/cos mm of each in a 20 mm box."""
    shape, sp = (10, 10, 10), (2.0, 2.0, 2.0)
    lab = np.zeros(shape[::-1], np.uint8)
    lab[5:] = 1
    mu = [0.0154, 0.0301]
    p0 = np.array([4.0, 6.0, -1.0])
    d = np.array([0.3, 0.2, 1.0])
    p1 = p0 + d * 22.0
    c = d[2] / np.linalg.norm(d)
    expect = np.exp(-(mu[0] + mu[1]) * 10.0 / c)
    assert O.beer_lambert(p0, p1, lab, sp, mu) == pytest.approx(expect, rel=1e-12)
    # a transposed label index (x and z swapped) is caught
    assert O.beer_lambert(p0, p1, np.ascontiguousarray(np.swapaxes(lab, 0, 2)), sp, mu) != pytest.approx(expect, rel=1e-3)


def test_mu_at_nodes_and_model():
    T = tables(("G4_WATER", "G4_BONE_CORTICAL_ICRP"), 0.15)
    nodes = xs.bucket_node_energies(T.n_buckets)
    for m in range(2):
        for j in (0, 7, 64, 200, T.n_buckets - 1):
            assert O.mu_at(T, m, nodes[j]) == T.mu_total[m, j]
        # half-way between nodes j and j+1 (in the bucket's mantissa fraction)
        j = 300
        e_mid = np.float32((np.float64(nodes[j]) + np.float64(nodes[j + 1])) / 2)
        v = O.mu_at(T, m, e_mid)
        lo, hi = sorted((T.mu_total[m, j], T.mu_total[m, j + 1]))
        assert lo <= v <= hi
        assert v == pytest.approx(0.5 * (T.mu_total[m, j] + T.mu_total[m, j + 1]), rel=1e-6)
        # against the FP64 mixture model at energies off the grid
        mat = DB.get(T.material_names[m])
        for E in (0.0153, 0.0467, 0.1405):
            model = xs.mu_rho(mat, E)[0] * mat.density / 10.0
            assert O.mu_at(T, m, E) == pytest.approx(model, rel=1e-3)


def test_dose_gy_by_hand():
    # 1 MeV = 1.602176634e-13 J; 1 mm^3 at 1 g/cm^3 = 1e-6 kg
    assert O.dose_gy(np.array([1.0]), np.array([1.0]), 1.0)[0] == pytest.approx(1.602176634e-7, rel=1e-15)
    # 3 MeV in 8 mm^3 of bone (1.92 g/cm^3): 3 * 1.602176634e-13 / (8e-3 cm^3 * 1.92 g/cm^3 * 1e-3 kg/g)
    assert O.dose_gy(np.array([3.0]), np.array([1.92]), 8.0)[0] == pytest.approx(
        3 * 1.602176634e-13 / (8e-3 * 1.92 * 1e-3), rel=1e-15)
    # no matter, no dose
    assert O.dose_gy(np.array([2.0]), np.array([0.0]), 1.0)[0] == 0.0
