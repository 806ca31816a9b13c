"""CPU checks of the learned exit-plane source's oracle and host-side feature maps.

The C restatements (noise, decoder) are pinned against closed forms: Box-Muller on
the Philox words computed independently in FP64, normality of the noise, and
decode(encode(record)) == record with a linear quantile table; the product's torch
feature map is checked against the oracle's cross-product formulation.
"""
import math

import numpy as np
import pytest
from scipy import stats

from oracle import exitplane_ref as R
from oracle import oracle as O

FOCAL = (32.0, 32.0, 142.5)


def _records(n, seed=0, scatter=0.1):
    g = np.random.default_rng(seed)
    x = g.uniform(10, 54, n)
    y = g.uniform(10, 54, n)
    d = np.stack([x - FOCAL[0], y - FOCAL[1], np.full(n, -FOCAL[2])], 1)
    d /= np.linalg.norm(d, axis=1)[:, None]
    k = g.uniform(size=n) < scatter
    d[k] += g.normal(scale=0.3, size=(k.sum(), 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    E = g.uniform(0.02, 0.12, n)
    return [x, y, np.zeros(n), d[:, 0], d[:, 1], d[:, 2], E, np.ones(n)]


def test_gaussian_is_box_muller_on_philox_words():
    seed = 0x1234_5678_9ABC
    rows = [0, 1, 12345, 2**33 + 7]
    for r in rows:
        got = R.gaussian(seed, r, 1, 8, 8)[0]
        for blk in range(2):
            w = O.philox4x32_10([r & 0xFFFFFFFF, r >> 32, 0x4E4F4953, blk], [seed & 0xFFFFFFFF, seed >> 32])
            u = [((x >> 8) | 1) * 2.0**-24 for x in w]
            ra, rb = math.sqrt(-2 * math.log(u[0])), math.sqrt(-2 * math.log(u[2]))
            want = [ra * math.cos(2 * math.pi * u[1]), ra * math.sin(2 * math.pi * u[1]),
                    rb * math.cos(2 * math.pi * u[3]), rb * math.sin(2 * math.pi * u[3])]
            np.testing.assert_allclose(got[4 * blk:4 * blk + 4], want, atol=2e-5 * max(1, ra, rb))


def test_gaussian_normal_and_padding():
    a = R.gaussian(7, 1000, 60000, 6, 8)
    assert np.all(a[:, 6:] == 0.0)
    v = a[:, :6]
    for j in range(6):
        assert stats.kstest(v[:, j], "norm").pvalue > 1e-3
    c = np.corrcoef(v.T)
    assert np.max(np.abs(c - np.eye(6))) < 0.02
    # a row depends only on (seed, row): a chunk reproduces the same rows
    np.testing.assert_array_equal(R.gaussian(7, 1500, 10, 6, 8), a[500:510])


def _linear_qtab(F, K=4097, g_max=5.5):
    lo = F.min(0) - 1e-3
    hi = F.max(0) + 1e-3
    kk = np.arange(K) / (K - 1)
    q = lo[:, None] + (hi - lo)[:, None] * kk[None, :]
    Z = -g_max + (F - lo) / (hi - lo) * 2 * g_max
    return q, Z


def test_decode_inverts_encode():
    rec = _records(5000)
    F = R.encode(rec, FOCAL, 0.0)
    q, Z = _linear_qtab(F)
    out = R.decode(Z, q, 5.5, FOCAL, 0.0, offset=(1.0, 2.0, 3.0), weight=0.5)
    np.testing.assert_allclose(out[0], rec[0] + 1.0, atol=2e-4)
    np.testing.assert_allclose(out[1], rec[1] + 2.0, atol=2e-4)
    assert np.all(out[2] == 3.0) and np.all(out[7] == 0.5)
    for k in (3, 4, 5):
        np.testing.assert_allclose(out[k], rec[k], atol=3e-5)
    np.testing.assert_allclose(out[6], rec[6], rtol=1e-5)


def test_decode_unscattered_photons_follow_the_focal_ray():
    rec = _records(2000, scatter=0.0)
    F = R.encode(rec, FOCAL, 0.0)
    np.testing.assert_allclose(F[:, 3:5], 0.0, atol=1e-12)
    np.testing.assert_allclose(F[:, 5], 1.0, atol=1e-12)
    F[:, 3:5] = 0.0
    F[:, 5] = 1.0
    q, Z = _linear_qtab(F)
    out = R.decode(Z, q, 5.5, FOCAL, 0.0)
    ray = np.stack([out[0] - FOCAL[0], out[1] - FOCAL[1], np.full(2000, -FOCAL[2])], 1)
    ray /= np.linalg.norm(ray, axis=1)[:, None]
    d = np.stack(out[3:6], 1)
    assert np.max(np.abs(np.cross(ray, d))) < 2e-6
    assert np.all((ray * d).sum(1) > 0.999999)


def test_decode_on_the_axis_uses_the_limit_frame():
    # exit point straight below the focal point: n = (0, 0, -1), e1 = (nz, 0, 0), e2 = (0, 1, 0)
    q = np.stack([np.linspace(a, b, 3) for a, b in ((0.05, 0.05), (32, 32), (32, 32), (-1, 1), (-1, 1), (-1, 1))])
    Z = np.array([[0, 0, 0, 5.5, 0, 0], [0, 0, 0, 0, 5.5, 0], [0, 0, 0, 0, 0, 5.5], [0, 0, 0, 0, 0, 0]], np.float32)
    out = R.decode(Z, q, 5.5, FOCAL, 0.0)
    d = np.stack(out[3:6], 1)
    np.testing.assert_allclose(d[0], [-1, 0, 0], atol=1e-6)
    np.testing.assert_allclose(d[1], [0, 1, 0], atol=1e-6)
    np.testing.assert_allclose(d[2], [0, 0, -1], atol=1e-6)
    np.testing.assert_allclose(d[3], [0, 0, -1], atol=1e-6)  # zero vector -> the focal ray


def test_product_feature_map_matches_oracle():
    torch = pytest.importorskip("torch")
    from opengate_amd.learn.exitplane import QuantileMap, exit_features

    rec = [a.astype(np.float32) for a in _records(20000, seed=3)]
    F = exit_features(rec, FOCAL, 0.0, device="cpu").double().numpy()
    np.testing.assert_allclose(F, R.encode(rec, FOCAL, 0.0), atol=2e-6)
    Ft = torch.from_numpy(F).float()
    qm = QuantileMap.fit(Ft, n_knots=257, g_max=4.0)
    G = QuantileMap.gaussianize(Ft).numpy()
    for j in range(6):
        # Gaussian scores are monotone in the data and standard normal
        o = np.argsort(F[:, j], kind="stable")
        assert np.all(np.diff(G[o, j]) > 0)
        assert abs(G[:, j].mean()) < 1e-3 and abs(G[:, j].std() - 1) < 2e-3
        # the table holds the data quantiles at Phi(g_k)
        lev = stats.norm.cdf(np.linspace(-4, 4, 257))
        idx = np.clip(np.round(lev * (len(F) - 1)), 0, len(F) - 1).astype(int)
        np.testing.assert_array_equal(qm.qtab[j].numpy(), np.sort(F[:, j]).astype(np.float32)[idx])
