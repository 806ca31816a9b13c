"""Voxel index contract (GateDoseActor::GetVoxelPosition, GateDoseActor.cpp:144-178, through
ITK's TransformPhysicalPointToIndex): the C restatement's cell index of an FP32 point
against exact rational arithmetic and against the FP64 ITK formula (oracle.itk_index),
on points at, and one ulp either side of, every voxel face of small grids."""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O

SPACINGS = (0.3, 1.7, 2.5)


def exact_index(p, corner, s):
    """floor((p - c) / s) for the real values of the FP32 inputs."""
    return math.floor((Fraction(float(np.float32(p))) - Fraction(float(np.float32(corner))))
                      / Fraction(float(np.float32(s))))


def face_points(corner, s, n):
    """FP32 points at each face corner + k s (k = 0..n) and one ulp below and above."""
    c, sf = np.float32(corner), np.float32(s)
    out = []
    for k in range(n + 1):
        f = np.float32(c + np.float32(k) * sf)
        out += [np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf))]
    return out


@pytest.mark.parametrize("s", SPACINGS)
@pytest.mark.parametrize("corner", (0.0, 0.37, -12.91))
def test_cell_index_is_exact_floor_on_faces(s, corner):
    for p in face_points(corner, s, 40):
        assert O.cell_index(p, corner, s) == exact_index(p, corner, s), (p, corner, s)


def test_cell_index_is_exact_floor_on_random_points():
    g = np.random.default_rng(5)
    for _ in range(20000):
        s = np.float32(g.uniform(0.05, 5.0))
        c = np.float32(g.uniform(-50, 50))
        p = np.float32(c + g.uniform(-3, 300) * s)
        assert O.cell_index(p, c, s) == exact_index(p, c, s), (p, c, s)


@pytest.mark.parametrize("s", SPACINGS)
def test_cell_index_matches_itk_formula_off_ties(s):
    """ITK's FP64 RoundHalfIntegerUp((p - o) * (1/s)) with o = corner + s/2 (the centre of voxel 0)
    equals the exact floor wherever (p - corner)/s is not an integer; at exact ties (a point on a
    face) the FP64 rounding of ITK's product decides, which only an ITK build could pin."""
    corner = np.float32(0.37)
    n = 30
    ties = agree = 0
    for p in face_points(corner, s, n):
        ex = exact_index(p, corner, s)
        tie = (Fraction(float(p)) - Fraction(float(corner))) % Fraction(float(np.float32(s))) == 0
        origin = float(corner) + float(np.float32(s)) / 2
        idx, _ = O.itk_index([float(p), 0, 0], [origin, 0, 0], [float(np.float32(s)), 1, 1], [n, 1, 1])
        if tie:
            ties += 1
            agree += int(idx[0] == ex)
        else:
            assert idx[0] == ex, (p, s)
        assert O.cell_index(p, corner, s) == ex
    assert ties >= 1


def test_naive_rounded_product_fails_somewhere():
    """The check has teeth: floor((p - c) * fl(1/s)) in FP32, the kernel's former rule, is off
    by one for some of these face points."""
    bad = 0
    for s in SPACINGS:
        sf = np.float32(s)
        inv = np.float32(1.0) / sf
        for p in face_points(0.0, s, 200):
            naive = int(np.floor(np.float32(p) * inv))
            bad += naive != exact_index(p, 0.0, s)
    assert bad > 0
