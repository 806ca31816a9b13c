"""gh_gemm_bf16_noalloc (the allocation-free GEMM: split-K slices in the library's static
device scratch array) against gh_gemm_bf16 through the C-ABI: equal outputs bit for bit on
two calls in a row, the second with a smaller input, in every epilogue mode and in the
split-K weight-gradient shapes, up to the split that fills the whole scratch array."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPI_STORE, EPI_BIAS_RELU, EPI_BIAS, EPI_DRELU = 0, 1, 2, 3


def _lib():
    from opengate_amd import _abi

    return _abi.load_library(), _abi


def _ptr(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def _call(fn, a, b, M, N, K, mode, c32=None, c16=None, bias=None, zprev=None):
    # A row-major [M, K] (sam = K, sak = 1), B row-major [K, N] (sbk = N, sbn = 1)
    return fn(_ptr(a), K, 1, _ptr(b), N, 1, M, N, K, _ptr(bias), _ptr(zprev), N, _ptr(c32), _ptr(c16), N, mode,
              C.c_void_p(0))


def _inputs(torch, M, N, K, seed):
    g = torch.Generator().manual_seed(seed)
    a = torch.randn(M, K, generator=g).to("cuda", torch.bfloat16)
    b = torch.randn(K, N, generator=g).to("cuda", torch.bfloat16)
    bias = torch.randn(N, generator=g).cuda()
    zp = torch.randn(M, N, generator=g).cuda()
    return a, b, bias, zp


def _both(torch, lib, M, N, K, mode, seed):
    a, b, bias, zp = _inputs(torch, M, N, K, seed)
    outs = []
    for fn in (lib.gh_gemm_bf16, lib.gh_gemm_bf16_noalloc):
        c32 = torch.full((M, N), float("nan"), device="cuda")
        c16 = torch.zeros(M, N, device="cuda", dtype=torch.bfloat16)
        use16 = mode in (EPI_BIAS_RELU, EPI_BIAS, EPI_DRELU)
        rc = _call(fn, a, b, M, N, K, mode, c32=c32, c16=c16 if use16 else None,
                   bias=bias if mode in (EPI_BIAS_RELU, EPI_BIAS) else None,
                   zprev=zp if mode == EPI_DRELU else None)
        assert rc == 0, lib.gh_last_error(None)
        torch.cuda.synchronize()
        outs.append((c32.cpu().numpy(), c16.float().cpu().numpy()))
    (r32, r16), (n32, n16) = outs
    assert np.array_equal(r32, n32, equal_nan=True)
    assert np.array_equal(r16, n16)
    assert not np.isnan(n32).any()
    return n32


@pytest.mark.parametrize("mode", [EPI_STORE, EPI_BIAS_RELU, EPI_BIAS, EPI_DRELU])
def test_noalloc_equals_allocating_entry_point_two_calls(mode):
    import torch

    lib, _ = _lib()
    print(f"[gpu case] gemm noalloc mode {mode}", flush=True)
    big = _both(torch, lib, 200, 136, 320, mode, 11)
    small = _both(torch, lib, 70, 40, 96, mode, 12)
    assert big.shape == (200, 136) and small.shape == (70, 40)


def test_noalloc_split_k_weight_gradient_shapes():
    """Weight-gradient shape (few output tiles, long reduction): the slices go through the
    static scratch; the second call is smaller and must not see the first call's slices."""
    import torch

    lib, _ = _lib()
    print("[gpu case] gemm noalloc split-K 256x256x16384 then 128x64x4096", flush=True)
    big = _both(torch, lib, 256, 256, 16384, EPI_STORE, 21)
    small = _both(torch, lib, 128, 64, 4096, EPI_STORE, 22)
    # against FP64 (bf16 inputs are exact in FP64; FP32 accumulation bound)
    a, b, _, _ = _inputs(torch, 128, 64, 4096, 22)
    A = a.double().cpu().numpy()
    B = b.double().cpu().numpy()
    tol = 1e-5 * (np.abs(A) @ np.abs(B)) + 1e-6
    assert np.all(np.abs(small - A @ B) <= tol)
    assert big.shape == (256, 256)


def test_noalloc_largest_split_fills_the_scratch_exactly():
    """One output tile and a reduction of 2^19: 1024 slices of 64 x 64 floats, the whole
    16 MiB scratch; then a small split-K call."""
    import torch

    lib, abi = _lib()
    M = N = 64
    K = 1 << 19
    S = min((K + 511) // 512, 1024)
    assert S * M * N * 4 == abi.GEMM_NOALLOC_WS_BYTES
    print(f"[gpu case] gemm noalloc {M}x{N}x{K} ({S} slices)", flush=True)
    big = _both(torch, lib, M, N, K, EPI_STORE, 31)
    small = _both(torch, lib, 64, 32, 2048, EPI_STORE, 32)
    assert big.shape == (64, 64) and small.shape == (64, 32)
