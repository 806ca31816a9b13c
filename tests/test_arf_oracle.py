"""CPU checks of the SPECT detection path: the head model geometry, the oracle's
photon-to-head pairing against closed forms and an FP64 ray-plane model, and the
energy windows against the normal distribution."""
import math

import numpy as np
import pytest
from scipy import stats

from oracle import arf_ref as R


def _heads(**kw):
    from opengate_amd.learn.arf import Heads

    return Heads(**{"center": (100.0, 100.0, 80.0), "radius": 200.0, "n_proj": 16, "nx": 32, "ny": 24,
                    "spacing": (6.0, 5.0), "angle0": 0.1, **kw})


def test_hole_mask_open_fraction_and_septa():
    from opengate_amd.learn.arf import HeadModel

    m = HeadModel()
    mask = m.hole_mask()
    # hexagon area / cell area = (hole / pitch)^2
    assert abs(mask.mean() - (m.hole / m.pitch) ** 2) < 0.01
    # the lattice point at the patch centre is open, the midpoint to its neighbour along x is lead
    c = mask.shape[0] // 2
    assert mask[c, c] and not mask[c, c + int(round(0.5 * m.pitch / m.voxel))]
    lab = m.labels()
    assert lab.shape == (int(round(m.face_z / m.voxel)), 400, 400)
    assert np.all(lab[: int(m.crystal / m.voxel)] == 2)


def test_photon_along_head_normal_hits_image_centre():
    h = _heads()
    cs = h.angle_table()
    k = 5
    a = h.angle0 + 2 * np.pi * k / h.n_proj
    # photon 3 mm along the tangential axis and 2.5 mm along the gantry axis from the centre,
    # moving along n_k: it lands in the middle of pixel (nx/2, ny/2)
    x0 = 100.0 - 3.0 * math.sin(a)
    y0 = 100.0 + 3.0 * math.cos(a)
    rec = [np.array([x0]), np.array([y0]), np.array([82.5]), np.array([math.cos(a)]), np.array([math.sin(a)]),
           np.array([0.0]), np.array([0.1405]), np.array([2.0])]
    out, n = R.spect_pairs(rec, h.desc(), cs, 16)
    assert n == 1
    iy, ix = divmod(int(out["pix"][0]) - k * h.plane, h.nx)
    assert (ix, iy) == (h.nx // 2, h.ny // 2)
    assert out["tp"][0, 0] < 1e-3 and out["wgt"][0] == 2.0
    assert abs(out["det"][5][0] + 1.0) < 1e-6  # arrives along -z of the head model


def test_pairs_match_fp64_ray_plane_model():
    h = _heads(theta_max=math.radians(12.0))
    cs = h.angle_table()
    g = np.random.default_rng(5)
    n = 4000
    p = g.uniform(60, 140, size=(n, 3))
    d = g.normal(size=(n, 3))
    d[:, 2] *= 0.1
    d /= np.linalg.norm(d, axis=1)[:, None]
    rec = [p[:, 0], p[:, 1], p[:, 2], d[:, 0], d[:, 1], d[:, 2], g.uniform(0.05, 0.15, n), np.ones(n)]
    out, cnt = R.spect_pairs(rec, h.desc(), cs, 50000)
    assert cnt == len(out["pix"]) > 500
    want = []
    c0 = np.asarray(h.center)
    for k in range(h.n_proj):
        ck, sk = cs[k].astype(np.float64)
        nk, ek = np.array([ck, sk, 0.0]), np.array([-sk, ck, 0.0])
        c = d @ nk
        ok = c > math.cos(h.theta_max)
        t = (h.radius - (p - c0) @ nk) / np.where(ok, c, 1.0)
        hit = p + t[:, None] * d
        f = h.model.crystal_depth / np.where(ok, c, 1.0)
        bu = (hit - c0) @ ek + f * (d @ ek)
        bz = hit[:, 2] - c0[2] + f * d[:, 2]
        fx = (bu + h.nx * h.spacing[0] / 2) / h.spacing[0]
        fy = (bz + h.ny * h.spacing[1] / 2) / h.spacing[1]
        ok &= (t >= 0) & (fx >= 0) & (fy >= 0) & (fx < h.nx) & (fy < h.ny)
        want += list((k * h.ny + fy[ok].astype(int)) * h.nx + fx[ok].astype(int))
    got = np.sort(out["pix"])
    want = np.sort(np.array(want))
    # FP32 vs FP64: at most a few pairs land on the other side of a pixel edge
    assert abs(len(got) - len(want)) <= 2
    assert np.mean(np.isin(got, want)) > 0.999


def test_lattice_reduction_moves_by_whole_lattice_vectors():
    h = _heads(theta_max=math.radians(15.0))
    cs = h.angle_table()
    m = h.model
    a2x, a2y = 0.5 * m.pitch, 0.5 * math.sqrt(3) * m.pitch
    g = np.random.default_rng(9)
    c0 = np.asarray(h.center)
    checked = 0
    for _ in range(80):
        p = g.uniform(80, 120, size=3)
        d = g.normal(size=3)
        d[2] *= 0.05
        d /= np.linalg.norm(d)
        rec = [np.array([v]) for v in (*p, *d, 0.1, 1.0)]
        out, cnt = R.spect_pairs(rec, h.desc(), cs, 64)
        for q in range(cnt):
            k = int(out["pix"][q]) // h.plane
            ck, sk = cs[k].astype(np.float64)
            nk, ek = np.array([ck, sk, 0.0]), np.array([-sk, ck, 0.0])
            t = (h.radius - (p - c0) @ nk) / (d @ nk)
            hit = p + t * d - c0
            hz, hu = hit[2], hit @ ek  # head frame: x' axial, y' tangential
            rx = out["det"][0][q] - m.origin[0]
            ry = out["det"][1][q] - m.origin[1]
            assert -1e-4 <= rx < m.pitch + 1e-4 and -1e-4 <= ry < a2y + 1e-4
            j = (hu - ry) / a2y
            i = (hz - rx - round(j) * a2x) / m.pitch
            assert abs(j - round(j)) < 1e-3 and abs(i - round(i)) < 1e-3, (i, j)
            # direction in the head frame and the angles fed to the network
            dz, dy, dn = out["det"][3][q], out["det"][4][q], out["det"][5][q]
            np.testing.assert_allclose([dz, dy, -dn], [d[2], d @ ek, d @ nk], atol=1e-6)
            np.testing.assert_allclose(out["tp"][q], [math.acos(d @ nk), math.atan2(d @ ek, d[2])], atol=1e-3)
            checked += 1
    assert checked > 20


def _window_chi2(cls, e0, sigma, win):
    """chi2 of the window frequencies of events at energy e0 blurred by N(0, sigma)"""
    n = len(cls)
    p = [stats.norm.cdf(hi, e0, sigma) - stats.norm.cdf(lo, e0, sigma) for lo, hi in win]
    p = np.array([1 - sum(p), *p])
    f = np.bincount(cls, minlength=len(win) + 1)
    return float(np.sum((f - n * p) ** 2 / (n * p))), f, n * p


def test_energy_windows_without_and_with_blur():
    win = ((0.114, 0.126), (0.126, 0.154), (0.154, 0.166))
    e = np.array([0.0, 0.11, 0.12, 0.1405, 0.16, 0.2, 0.114, 0.126, 0.166], np.float32)
    # no blur: half-open windows [lo, hi) (GateDigitizerEnergyWindowsActor.cpp:127), first match
    expect = [0, 0, 1, 2, 3, 0, 1, 2, 0]
    np.testing.assert_array_equal(R.energy_windows(e, 1, 0, win, law="none"), expect)
    # InverseSquare with zero resolution draws noise but adds nothing
    np.testing.assert_array_equal(R.energy_windows(e, 1, 0, win, resolution=0.0, reference=0.1405), expect)
    n = 200000
    cls = R.energy_windows(np.full(n, 0.1405, np.float32), 7, 1000, win, resolution=0.095, reference=0.1405)
    chi2, f, fe = _window_chi2(cls, 0.1405, 0.095 * 0.1405 / 2.354820045, win)
    assert chi2 < 30, (f, fe)
    # the noise of event i depends only on (seed, first + i)
    np.testing.assert_array_equal(R.energy_windows(np.full(10, 0.1405, np.float32), 7, 1500, win, resolution=0.095,
                                                   reference=0.1405), cls[500:510])


@pytest.mark.parametrize("law,e0,sigma,kw", [
    # InverseSquare away from the reference: FWHM = R sqrt(E0 E) (GateDigitizerBlurringActor.cpp:80-86)
    ("InverseSquare", 0.120, 0.095 * math.sqrt(0.1405 * 0.120) / 2.354820045,
     {"resolution": 0.095, "reference": 0.1405}),
    # Gaussian: fixed sigma (:75-78)
    ("Gaussian", 0.1405, 0.008, {"sigma": 0.008}),
    # Linear: FWHM = (slope (E - E0) + R) E (:88-94)
    ("Linear", 0.150, (0.4 * (0.150 - 0.1405) + 0.095) * 0.150 / 2.354820045,
     {"resolution": 0.095, "reference": 0.1405, "slope": 0.4}),
])
def test_energy_blur_laws(law, e0, sigma, kw):
    win = ((0.110, 0.126), (0.126, 0.140), (0.140, 0.154), (0.154, 0.170))
    cls = R.energy_windows(np.full(200000, e0, np.float32), 11, 0, win, law=law, **kw)
    chi2, f, fe = _window_chi2(cls, e0, sigma, win)
    assert chi2 < 35, (law, f, fe)
    # a wrong width (10 % off) is rejected by the same test
    chi2_bad, _, _ = _window_chi2(cls, e0, sigma * 1.1, win)
    assert chi2_bad > 100, (law, chi2_bad)


def test_accumulation_models():
    g = np.random.default_rng(2)
    m, n_win, plane, heads = 500, 3, 12, 4
    logits = g.normal(size=(m, 8))
    pix = g.integers(0, heads * plane, m)
    wgt = g.uniform(0.5, 1.5, m)
    img = R.accumulate_arf(logits, pix, wgt, n_win, plane, heads)
    z = logits[:, :4]
    p = np.exp(z) / np.exp(z).sum(1, keepdims=True)
    assert abs(img.sum() - np.sum(wgt * p[:, 1:].sum(1))) < 1e-9
    cls = g.integers(0, 4, m)
    imc = R.accumulate_counts(cls, pix, wgt, n_win, plane, heads)
    assert abs(imc.sum() - wgt[cls > 0].sum()) < 1e-9
    h, r = divmod(int(pix[0]), plane)
    if cls[0] > 0:
        assert imc[(h * n_win + cls[0] - 1) * plane + r] >= wgt[0] - 1e-12
