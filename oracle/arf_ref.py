"""Checkers of the SPECT detection path (TEST INFRASTRUCTURE ONLY).

Bindings of ref_spect_pairs / ref_energy_windows (transport_ref.c, FP32 restatements
of csrc/arf.hip) and FP64 numpy models of the image accumulation. Pinned in
tests/test_arf_oracle.py: closed-form head geometry (a photon along a head normal hits
the image centre at theta = 0), an FP64 ray-plane intersection, lattice reduction by
whole lattice vectors, and the blurred window frequencies against the normal CDF.
"""
import ctypes as C

import numpy as np

from . import oracle as O


def _lib():
    L = O.lib()
    if not getattr(L, "_arf_ready", False):
        from opengate_amd import _abi as A

        FP = C.POINTER(C.c_float)
        L.ref_spect_pairs.argtypes = [FP * 8, C.c_int64, C.POINTER(A.SpectDesc), FP, C.c_int64, FP,
                                      C.POINTER(C.c_int32), FP, C.POINTER(FP), FP, C.POINTER(C.c_uint64)]
        L.ref_energy_windows.argtypes = [FP, C.c_int64, C.c_uint64, C.c_uint64, C.c_int, C.c_float, C.c_float,
                                         C.c_float, C.c_float, FP, C.c_int, C.POINTER(C.c_int32)]
        L._arf_ready = True
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def spect_pairs(rec, desc, cs, capacity):
    """Pairs of the photons rec (8 float32 arrays) with the heads of desc (A.SpectDesc);
    cs: float32 (n_proj, 2) cos/sin table. Returns a dict of arrays and the pair count."""
    rec = [np.ascontiguousarray(a, np.float32) for a in rec]
    cs = np.ascontiguousarray(cs, np.float32)
    cap = int(capacity)
    out = {"x3": np.zeros((cap, 3), np.float32), "pix": np.zeros(cap, np.int32), "wgt": np.zeros(cap, np.float32),
           "tp": np.zeros((cap, 2), np.float32), "det": [np.zeros(cap, np.float32) for _ in range(8)]}
    cnt = C.c_uint64(0)
    FP = C.POINTER(C.c_float)
    _lib().ref_spect_pairs((FP * 8)(*[_fp(a) for a in rec]), len(rec[0]), C.byref(desc), _fp(cs), cap,
                           _fp(out["x3"]), out["pix"].ctypes.data_as(C.POINTER(C.c_int32)), _fp(out["wgt"]),
                           (FP * 8)(*[_fp(a) for a in out["det"]]), _fp(out["tp"]), C.byref(cnt))
    n = min(int(cnt.value), cap)
    res = {k: (v[:n] if k != "det" else [a[:n] for a in v]) for k, v in out.items()}
    return res, int(cnt.value)


BLUR_LAWS = {"none": 0, "Gaussian": 1, "InverseSquare": 2, "Linear": 3}


def energy_windows(edep, seed, first, windows, law="InverseSquare", sigma=0.0, resolution=0.0, reference=0.0,
                   slope=0.0):
    """Window class (0 none, k = 1..) of each event energy after the blurring law `law`."""
    edep = np.ascontiguousarray(edep, np.float32)
    win = np.ascontiguousarray(np.asarray(windows, np.float32).reshape(-1, 2))
    cls = np.zeros(len(edep), np.int32)
    _lib().ref_energy_windows(_fp(edep), len(edep), int(seed) & (2**64 - 1), int(first), BLUR_LAWS[law],
                              float(sigma), float(resolution), float(reference), float(slope), _fp(win), len(win),
                              cls.ctypes.data_as(C.POINTER(C.c_int32)))
    return cls


def accumulate_arf(logits, pix, wgt, n_win, plane, n_heads):
    z = np.asarray(logits, np.float64)[:, : n_win + 1]
    p = np.exp(z - z.max(1, keepdims=True))
    p /= p.sum(1, keepdims=True)
    img = np.zeros(n_heads * n_win * plane)
    head, rem = np.divmod(np.asarray(pix, np.int64), plane)
    for k in range(1, n_win + 1):
        np.add.at(img, (head * n_win + k - 1) * plane + rem, p[:, k] * np.asarray(wgt, np.float64))
    return img


def accumulate_counts(cls, pix, wgt, n_win, plane, n_heads):
    img = np.zeros(n_heads * n_win * plane)
    cls = np.asarray(cls, np.int64)
    ok = (cls > 0) & (cls <= n_win)
    head, rem = np.divmod(np.asarray(pix, np.int64)[ok], plane)
    np.add.at(img, (head * n_win + cls[ok] - 1) * plane + rem, np.asarray(wgt, np.float64)[ok])
    return img
