"""SPECT detection on the GPU: photon-to-head pairing and energy windows bit-exact
against the C oracle, the image accumulation against numpy, and the learned detector
response (trained on the voxelised head model) against plain transport of the same
photons through that model."""
import math

import numpy as np
import pytest

from oracle import arf_ref as R
from oracle import mlp_ref as MR

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _heads(**kw):
    from opengate_amd.learn.arf import Heads

    return Heads(**{"center": (100.0, 100.0, 80.0), "radius": 200.0, "n_proj": 16, "nx": 32, "ny": 24,
                    "spacing": (6.0, 5.0), "angle0": 0.1, "theta_max": math.radians(12.0), **kw})


def _photons(n, seed):
    g = np.random.default_rng(seed)
    p = g.uniform(60, 140, size=(n, 3))
    d = g.normal(size=(n, 3))
    d[:, 2] *= 0.1
    d /= np.linalg.norm(d, axis=1)[:, None]
    return [a.astype(np.float32) for a in (p[:, 0], p[:, 1], p[:, 2], d[:, 0], d[:, 1], d[:, 2],
                                           g.uniform(0.02, 0.15, n), g.uniform(0.5, 1.5, n))]


def test_pairs_bit_exact():
    torch = _torch()
    from opengate_amd import _abi as A
    from opengate_amd.learn.arf import PairBuffer, spect_pairs

    lib = A.load_library()
    h = _heads()
    rec = _photons(20000, 3)
    print("[gpu case] gh_spect_pairs 20000 photons x 16 heads", flush=True)
    buf = PairBuffer(100000, "cuda", det=True)
    cs = torch.from_numpy(h.angle_table()).cuda()
    m = spect_pairs(lib, [torch.from_numpy(a).cuda() for a in rec], 20000, h, cs, buf)
    ref, cnt = R.spect_pairs(rec, h.desc(), h.angle_table(), 100000)
    assert m == cnt > 1000
    det = [a[:m].cpu().numpy() for a in buf.det]
    pix = buf.pix[:m].cpu().numpy()
    o = np.lexsort((det[6], det[1], det[0], pix))
    r = np.lexsort((ref["det"][6], ref["det"][1], ref["det"][0], ref["pix"]))
    np.testing.assert_array_equal(pix[o], ref["pix"][r])
    np.testing.assert_array_equal(buf.wgt[:m].cpu().numpy()[o], ref["wgt"][r])
    for k in range(8):
        np.testing.assert_array_equal(det[k][o], ref["det"][k][r])
    x = buf.x16[:m].float().cpu().numpy()[o]
    xr = MR.to_bf16(ref["x3"][r])
    np.testing.assert_array_equal(x[:, 0], xr[:, 0])
    np.testing.assert_allclose(x[:, 1:3], xr[:, 1:3], rtol=2.0**-7, atol=1e-3)
    assert np.all(x[:, 3:] == 0)


@pytest.mark.parametrize("law,kw", [
    ("InverseSquare", {"resolution": 0.095, "reference": 0.1405}),
    ("Gaussian", {"sigma": 0.006}),
    ("Linear", {"resolution": 0.09, "reference": 0.1405, "slope": -0.3}),
    ("none", {}),
])
def test_energy_windows_bit_exact(law, kw):
    torch = _torch()
    from opengate_amd import _abi as A
    from opengate_amd.learn.arf import WINDOWS, Blur, energy_windows

    lib = A.load_library()
    g = np.random.default_rng(4)
    e = g.uniform(0.0, 0.17, 50000).astype(np.float32)
    e[::7] = 0.0
    print(f"[gpu case] gh_energy_windows 50000 {law}", flush=True)
    cls = torch.empty(50000, dtype=torch.int32, device="cuda")
    energy_windows(lib, torch.from_numpy(e).cuda(), 50000, 99, 123, cls, WINDOWS, Blur(law=law, **kw))
    np.testing.assert_array_equal(cls.cpu().numpy(), R.energy_windows(e, 99, 123, WINDOWS, law=law, **kw))


def test_accumulate_kernels():
    torch = _torch()
    from opengate_amd import _abi as A

    lib = A.load_library()
    g = np.random.default_rng(2)
    m, n_win, plane, heads = 20000, 3, 24 * 32, 16
    logits = g.normal(size=(m, 8)).astype(np.float32)
    pix = g.integers(0, heads * plane, m).astype(np.int32)
    wgt = g.uniform(0.5, 1.5, m).astype(np.float32)
    cls = g.integers(0, 4, m).astype(np.int32)
    dv = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    img = torch.zeros(heads * n_win * plane, dtype=torch.float64, device="cuda")
    img2 = torch.zeros_like(img)
    print("[gpu case] accumulate kernels", flush=True)
    L, P, W, Cc = dv(logits), dv(pix), dv(wgt), dv(cls)
    p = lambda t: __import__("ctypes").c_void_p(t.data_ptr())  # noqa: E731
    assert lib.gh_arf_accumulate(p(L), 8, m, p(P), p(W), n_win, plane, p(img), p(img2), None) == 0
    ref = R.accumulate_arf(logits, pix, wgt, n_win, plane, heads)
    np.testing.assert_allclose(img.cpu().numpy(), ref, rtol=1e-5, atol=1e-9)
    img.zero_()
    img2.zero_()
    assert lib.gh_counts_accumulate(p(Cc), p(P), p(W), m, n_win, plane, p(img), p(img2), None) == 0
    ref = R.accumulate_counts(cls, pix, wgt, n_win, plane, heads)
    np.testing.assert_allclose(img.cpu().numpy(), ref, rtol=1e-12, atol=1e-12)
    ref2 = R.accumulate_counts(cls, pix, wgt.astype(np.float64) ** 2, n_win, plane, heads)
    np.testing.assert_allclose(img2.cpu().numpy(), ref2, rtol=1e-6, atol=1e-9)


def _sphere_escapes(engine, n, seed):
    from opengate_amd.physics import xs
    from opengate_amd.physics.materials import MaterialDatabase
    from opengate_amd.transport import PhantomSpec, SourceSpec, TallySpec

    db = MaterialDatabase()
    T = xs.build_tables([db.get("G4_AIR"), db.get("G4_WATER"), db.get("G4_BONE_CORTICAL_ICRP")], 0.1406)
    c = (np.arange(64) + 0.5) * 3.0 - 96.0
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    lab = np.zeros((64, 64, 64), np.uint8)
    lab[X**2 + Y**2 + Z**2 < 80.0**2] = 1
    lab[(X - 30) ** 2 + Y**2 + Z**2 < 20.0**2] = 2
    ph = PhantomSpec(lab, (3.0, 3.0, 3.0), world_material=-1)
    src = SourceSpec(pos_type="sphere", pos_param=(40.0,), pos_trans=(96.0, 96.0, 96.0), energy=0.1405)
    tal = TallySpec(size=(64, 64, 64), spacing=(3.0, 3.0, 3.0), edep2=False, record_face=6, record_capacity=n)
    engine.clear()
    engine.set_physics(T)
    sid = engine.add_simulation(ph, src, tal)
    engine.run([n], seed)
    rec, nrec = engine.get_phsp(sid, device=True)
    return rec, min(nrec, n)


def test_learned_response_matches_plain_projections(engine):
    torch = _torch()
    from opengate_amd.learn.arf import ArfModel, Heads, HeadSimulator, compare_projections, project

    theta_max = math.radians(10.0)
    sim = HeadSimulator()
    n_train = 16_000_000
    print(f"[gpu case] arf training set {n_train} photons through the head model", flush=True)
    rec, E, th, ph = sim.training_photons(n_train, 1, theta_max)
    cls = sim.windows(rec, n_train, 3)
    freq = torch.bincount(cls.long(), minlength=4).cpu().numpy() / n_train
    print("training classes (none, scatter-, peak, scatter+):", freq, flush=True)
    # photopeak: theta within the ~2.5 deg hole acceptance (~1/4 of theta), open fraction
    # (1.5/1.7)^2, E near the line (~1/5 of E), photopeak efficiency of 9.5 mm NaI
    assert 0.01 < freq[2] < 0.2
    arf = ArfModel(theta_max)
    X = arf.inputs(E, th, ph)
    print("[gpu case] arf fit 6000 steps", flush=True)
    hist = arf.fit(X, cls, steps=6000, batch=32768, seed=2)
    assert np.mean(hist[-100:]) < np.mean(hist[:20])
    # calibration on held-out photons: mean predicted probability vs frequency per theta bin
    rec2, E2, th2, ph2 = sim.training_photons(4_000_000, 11, theta_max, theta_power=1.0, e_focus=(1.0, 0.13))
    cls2 = sim.windows(rec2, 4_000_000, 13).long()
    P = arf.probabilities(arf.inputs(E2, th2, ph2)).double()
    peak = (E2 > 0.13).double()
    for lo, hi in ((0.0, 1.0), (1.0, 2.0), (2.0, 3.0), (3.0, 10.0)):
        sel = (th2 >= math.radians(lo)) & (th2 < math.radians(hi)) & (peak > 0)
        k = int(sel.sum())
        f = float((cls2[sel] == 2).double().mean())
        pm = float(P[sel, 2].mean())
        print(f"theta [{lo}, {hi}) deg, E > 130 keV: n {k} peak frequency {f:.4f} predicted {pm:.4f}", flush=True)
        assert abs(pm - f) < 4 * math.sqrt(max(f * (1 - f), 1e-4) / k) + 0.01
    # projections of a water sphere (bone insert, 140.5 keV sphere source)
    print("[gpu case] sphere phantom 1e8 histories, 32 heads", flush=True)
    esc, n_esc = _sphere_escapes(engine, 100_000_000, 21)
    heads = Heads(center=(96.0, 96.0, 96.0), radius=96.0 * math.sqrt(2.0) + 10.0, n_proj=32, nx=64, ny=64,
                  theta_max=theta_max)
    res = project([a[:n_esc] for a in esc], n_esc, heads, arf=arf, head_sim=sim, seed=5)
    cmp = compare_projections(res["arf"].cpu().numpy(), res["plain"].cpu().numpy(), res["arf2"].cpu().numpy(),
                              res["plain2"].cpu().numpy(), block=4)
    print(f"pairs {res['pairs']}  t_pairs {res['t_pairs']:.3f}s t_arf {res['t_arf']:.3f}s "
          f"t_plain {res['t_plain']:.3f}s  {cmp}", flush=True)
    assert 0.97 < cmp["window_ratio"][1] < 1.03  # photopeak
    assert 0.95 < cmp["window_ratio"][0] < 1.05  # lower scatter window
    assert cmp["pixels"] > 200 and cmp["chi2_per_pixel_window"][1] < 2.0
    assert res["t_plain"] > 3 * res["t_arf"]
    sim.close()
