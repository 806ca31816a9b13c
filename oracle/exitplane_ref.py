"""Checkers of the learned exit-plane source (TEST INFRASTRUCTURE ONLY).

* gaussian / decode: bindings of ref_gaussian and ref_exitplane_decode in
  transport_ref.c, FP32 restatements of csrc/exitplane.hip in the same operation
  order (the GPU output must be bit-identical).
* encode: FP64 numpy restatement of the focal-frame features, with the frame built
  from cross products (e2 = z x n / |z x n|, e1 = e2 x n) rather than the closed
  form the library uses, so the two are independent.
Pinned in tests/test_exitplane_oracle.py: Box-Muller closed form on Philox words,
normality (KS), decode(encode(record)) == record, unscattered photons on the focal ray.
"""
import ctypes as C

import numpy as np

from . import oracle as O


def _lib():
    L = O.lib()
    if not getattr(L, "_ep_ready", False):
        FP = C.POINTER(C.c_float)
        L.ref_gaussian.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_int, C.c_int, FP]
        L.ref_exitplane_decode.argtypes = [FP, C.c_int64, C.c_int64, FP, C.c_int, C.c_float, FP, C.c_float, FP,
                                           C.c_float, FP * 8]
        L._ep_ready = True
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def gaussian(seed, first, n, cols, ld):
    out = np.zeros((n, ld), np.float32)
    _lib().ref_gaussian(int(seed) & (2**64 - 1), int(first), int(n), int(cols), int(ld), _fp(out))
    return out


def decode(Z, qtab, g_max, focal, plane_z, offset=(0.0, 0.0, 0.0), weight=1.0):
    Z = np.ascontiguousarray(Z, np.float32)
    qtab = np.ascontiguousarray(qtab, np.float32)
    n = Z.shape[0]
    out = [np.zeros(n, np.float32) for _ in range(8)]
    f = np.asarray(focal, np.float32)
    o = np.asarray(offset, np.float32)
    _lib().ref_exitplane_decode(_fp(Z), Z.shape[1], n, _fp(qtab), qtab.shape[1], float(g_max), _fp(f),
                                float(plane_z), _fp(o), float(weight), (C.POINTER(C.c_float) * 8)(*[_fp(a) for a in out]))
    return out


def encode(rec, focal, plane_z):
    """(N, 6) FP64 features (E, x, y, u', v', w') of records (8 arrays x y z u v w E weight)."""
    x, y = np.asarray(rec[0], np.float64), np.asarray(rec[1], np.float64)
    d = np.stack([np.asarray(rec[k], np.float64) for k in (3, 4, 5)], 1)
    n = np.stack([x - focal[0], y - focal[1], np.full_like(x, plane_z - focal[2])], 1)
    n /= np.linalg.norm(n, axis=1)[:, None]
    e2 = np.cross(np.array([0.0, 0.0, 1.0]), n)
    e2 /= np.linalg.norm(e2, axis=1)[:, None]
    e1 = np.cross(e2, n)
    return np.stack([np.asarray(rec[6], np.float64), x, y, (d * e1).sum(1), (d * e2).sum(1), (d * n).sum(1)], 1)
