"""Deposit voxel of the GPU kernel against the ITK index contract (GateDoseActor::GetVoxelPosition,
GateDoseActor.cpp:144-178; GateHelpersImage.txx:60-93), brute force on small grids.

Phase-space photons below the tracking cut deposit their whole energy at their birth point
(the source frame is the phantom frame, so the point reaches the kernel unchanged). The
points sit on, and one ulp either side of, every voxel face of grids with spacings 0.3,
1.7 and 2.5 mm, aligned with the phantom or not. Photon k of a run carries the weight 2^k
and the energy 2^-11 MeV, so each voxel's FP64 sum is a sum of distinct powers of two and
names exactly which photons landed there. The expected voxel is floor((p - corner)/s) in
exact rational arithmetic, which equals the FP64 ITK formula (oracle.itk_index) off exact
ties (checked for every point here; a point on a face, or nearer to it than FP64 resolves,
is a tie, where only an ITK build could say how its rounding falls)."""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from opengate_amd.transport import PhantomSpec, SourceSpec, TallySpec

from helpers import tables

E_DEP = 2.0 ** -11  # below the tracking cut (2^-10 MeV): deposits at birth


def exact_index(p, c, s):
    return math.floor((Fraction(float(np.float32(p))) - Fraction(float(np.float32(c)))) / Fraction(float(np.float32(s))))


def face_coords(c, s, n, off):
    out = []
    for k in range(n + 1):
        f = np.float32(np.float32(c) + np.float32(k) * np.float32(s))
        if off < 0:
            f = np.nextafter(f, np.float32(-np.inf))
        elif off > 0:
            f = np.nextafter(f, np.float32(np.inf))
        out.append(f)
    return out


CASES = [((0.3, 0.3, 0.3), True), ((1.7, 1.7, 1.7), True), ((2.5, 2.5, 2.5), True), ((0.3, 1.7, 2.5), True),
         ((0.3, 0.3, 0.3), False), ((1.7, 1.7, 1.7), False), ((2.5, 2.5, 2.5), False), ((0.3, 1.7, 2.5), False)]


def face_runs(spacing, aligned):
    """(phantom, tally, [(axis, off, source, n, expected edep)], ties) of one grid."""
    size = (7, 5, 4)
    if aligned:
        ph = PhantomSpec(np.zeros(size[::-1], np.uint8), spacing)
        corner = (0.0, 0.0, 0.0)
        tal = TallySpec(size=size, spacing=spacing)
    else:
        # a 24 mm phantom of 2 mm voxels around a tally grid with an odd corner
        ph = PhantomSpec(np.zeros((12, 12, 12), np.uint8), (2.0, 2.0, 2.0))
        corner = (0.37, 2.11, 0.73)
        tal = TallySpec(size=size, spacing=spacing, corner=corner, aligned=False)
    runs = []
    ties = 0
    for axis in range(3):
        for off in (-1, 0, 1):
            pts = []
            for k, fa in enumerate(face_coords(corner[axis], spacing[axis], size[axis], off)):
                p = [np.float32(np.float32(corner[b]) + (np.float32((k % size[b]) + 0.5)) * np.float32(spacing[b]))
                     for b in range(3)]
                p[axis] = fa
                pts.append(p)
            n = len(pts)
            cols = [np.array([p[b] for p in pts], np.float32) for b in range(3)]
            rows = cols + [np.zeros(n, np.float32), np.zeros(n, np.float32), np.ones(n, np.float32),
                           np.full(n, E_DEP, np.float32), (2.0 ** np.arange(n)).astype(np.float32)]
            src = SourceSpec(pos_type="phsp", phsp=rows)
            expect = np.zeros(size[::-1])
            for k, p in enumerate(pts):
                idx = [exact_index(p[b], corner[b], spacing[b]) for b in range(3)]
                # ITK formula in FP64 (origin at the centre of voxel 0) off exact ties
                origin = [float(np.float32(corner[b])) + float(np.float32(spacing[b])) / 2 for b in range(3)]
                itk, _ = O.itk_index([float(x) for x in p], origin, [float(np.float32(x)) for x in spacing], size)
                # a tie: the point lies on a face, or nearer to it than FP64 resolves (the
                # subnormal next to a face at 0)
                q = (Fraction(float(p[axis])) - Fraction(float(np.float32(corner[axis])))) / \
                    Fraction(float(np.float32(spacing[axis])))
                tie = abs(q - round(q)) < Fraction(1, 2 ** 45)
                ties += int(tie)
                if not tie:
                    assert list(itk) == idx, (p, itk, idx)
                if all(0 <= idx[b] < size[b] for b in range(3)):
                    expect[idx[2], idx[1], idx[0]] += E_DEP * 2.0 ** k
            runs.append((axis, off, src, n, expect))
    return ph, tal, runs, ties


@pytest.mark.parametrize("spacing,aligned", CASES)
def test_oracle_deposit_voxel_follows_itk_contract(spacing, aligned):
    T = tables()
    ph, tal, runs, ties = face_runs(spacing, aligned)
    assert ties > 0
    for axis, off, src, n, expect in runs:
        r, c = O.run(T, ph, src, tal, n, 3)
        assert c["cut_kills"] == n
        assert np.array_equal(r["edep"], expect), (axis, off)


@pytest.mark.gpu
@pytest.mark.parametrize("spacing,aligned", CASES)
def test_deposit_voxel_follows_itk_contract(engine, spacing, aligned):
    T = tables()
    ph, tal, runs, _ = face_runs(spacing, aligned)
    engine.clear()
    engine.set_physics(T)
    for k, (axis, off, src, n, expect) in enumerate(runs):
        print(f"[gpu case] index spacing={spacing} aligned={aligned} axis={axis} off={off} n={n}", flush=True)
        if k == 0:
            engine.add_simulation(ph, src, tal)
        else:
            engine.set_source(0, src)
            engine.reset_tally(0)
        c = engine.run([n], 3)
        g = engine.get_tally(0)["edep"]
        assert c["histories"] == n and c["cut_kills"] == n
        assert np.array_equal(g, expect), (axis, off, np.argwhere(g != expect)[:5])
