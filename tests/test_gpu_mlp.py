"""The BF16 MFMA GEMM and the MLP trained with it, against FP64 references (GPU)."""
import numpy as np
import pytest

from oracle import mlp_ref as R

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (100, 70, 40), (257, 256, 256), (33, 8, 8), (96, 6, 3),
                                   (128, 136, 8192)])
def test_gemm_modes_and_layouts(M, N, K):
    torch = _torch()
    from opengate_amd.learn.mlp import EPI_BIAS, EPI_BIAS_RELU, EPI_DRELU, EPI_STORE, MLP

    print(f"[gpu case] gemm M={M} N={N} K={K}", flush=True)
    net = MLP([8, 8])  # only for the GEMM entry point
    g = torch.Generator().manual_seed(M * 131 + N)
    a = torch.randn(M, K, generator=g).to("cuda", torch.bfloat16)
    bkn = torch.randn(K, N, generator=g).to("cuda", torch.bfloat16)
    bias = torch.randn(N, generator=g).cuda()
    zp = torch.randn(M, N, generator=g).cuda()
    A64 = a.double().cpu().numpy()
    B64 = bkn.double().cpu().numpy()
    ref = A64 @ B64
    tol = 2e-3 * (np.abs(A64) @ np.abs(B64)) + 1e-5
    torch.cuda.synchronize()
    # plain store, A row-major / B row-major (n contiguous)
    c = torch.zeros(M, N, device="cuda")
    net._gemm(a, K, 1, bkn, N, 1, M, N, K, c32=c, ldc=N, mode=EPI_STORE)
    torch.cuda.synchronize()
    assert np.all(np.abs(c.cpu().numpy() - ref) <= tol)
    # transposed operands: A given as A^T (m contiguous), B given as B^T (k contiguous)
    at = a.t().contiguous()
    bt = bkn.t().contiguous()
    c2 = torch.zeros(M, N, device="cuda")
    net._gemm(at, 1, M, bt, 1, K, M, N, K, c32=c2, ldc=N, mode=EPI_STORE)
    torch.cuda.synchronize()
    assert np.all(np.abs(c2.cpu().numpy() - ref) <= tol)
    # fused epilogues
    z = torch.zeros(M, N, device="cuda")
    y = torch.zeros(M, N, device="cuda", dtype=torch.bfloat16)
    net._gemm(a, K, 1, bkn, N, 1, M, N, K, bias=bias, c32=z, c16=y, ldc=N, mode=EPI_BIAS_RELU)
    torch.cuda.synchronize()
    zr = ref + bias.double().cpu().numpy()
    assert np.all(np.abs(z.cpu().numpy() - zr) <= tol)
    assert np.allclose(y.double().cpu().numpy(), R.relu(zr), rtol=1e-2, atol=1e-2)
    net._gemm(a, K, 1, bkn, N, 1, M, N, K, bias=bias, c32=z, ldc=N, mode=EPI_BIAS)
    torch.cuda.synchronize()
    assert np.all(np.abs(z.cpu().numpy() - zr) <= tol)
    net._gemm(a, K, 1, bkn, N, 1, M, N, K, zprev=zp, ldz=N, c32=z, c16=y, ldc=N, mode=EPI_DRELU)
    torch.cuda.synchronize()
    dr = ref * R.drelu(zp.double().cpu().numpy())
    assert np.all(np.abs(z.cpu().numpy() - dr) <= tol * 1.2)


def test_mlp_forward_backward_against_fp64_oracle():
    torch = _torch()
    from opengate_amd.learn.mlp import MLP

    print("[gpu case] mlp fwd/bwd vs oracle", flush=True)
    net = MLP([8, 64, 64, 64, 6], seed=3)
    g = torch.Generator().manual_seed(5)
    x = torch.randn(300, 8, generator=g).cuda()
    out = net.forward(x)
    gout = torch.randn(300, 6, generator=g).cuda()
    dW, db = net.backward(gout)
    torch.cuda.synchronize()
    Ws = [w[:k, :n].double().cpu().numpy() for w, k, n in zip(net.W, net.sizes[:-1], net.sizes[1:])]
    bs = [b[:n].double().cpu().numpy() for b, n in zip(net.b, net.sizes[1:])]
    o_ref, xs, zs = R.forward(Ws, bs, x.double().cpu().numpy())
    assert np.linalg.norm(out.double().cpu().numpy() - o_ref) <= 2e-2 * np.linalg.norm(o_ref)
    dW_ref, db_ref = R.backward(Ws, xs, zs, gout.double().cpu().numpy())
    for l, (k, n) in enumerate(zip(net.sizes[:-1], net.sizes[1:])):
        gw = dW[l][:k, :n].double().cpu().numpy()
        gb = db[l][:n].double().cpu().numpy()
        assert np.linalg.norm(gw - dW_ref[l]) <= 3e-2 * np.linalg.norm(dW_ref[l]), l
        assert np.linalg.norm(gb - db_ref[l]) <= 3e-2 * np.linalg.norm(db_ref[l]), l


def test_mlp_trains_a_regression():
    torch = _torch()
    from opengate_amd.learn.mlp import MLP

    print("[gpu case] mlp training", flush=True)
    net = MLP([8, 128, 128, 4], seed=1, lr=3e-3)
    g = torch.Generator().manual_seed(7)
    W0 = torch.randn(8, 4, generator=g).cuda()
    losses = []
    for it in range(300):
        x = torch.randn(1024, 8, generator=g).cuda()
        t = torch.tanh(x @ W0)
        y = net.forward(x)
        loss = ((y - t) ** 2).mean()
        dW, db = net.backward(2 * (y - t) / y.numel())
        net.step(dW, db)
        losses.append(float(loss))
    assert losses[-1] < 0.1 * losses[0]


def test_fused_adam_step_against_fp64():
    """gh_adam_step (one launch over all tensors) against Adam in FP64 (Kingma & Ba, with
    bias correction and L2 decay on the weights) over three steps: valid corners within
    FP32 rounding, padding exactly zero, the bf16 copy the round-to-nearest-even of the
    parameter."""
    import ctypes as C

    torch = _torch()
    from opengate_amd import _abi as A

    lib = A.load_library()
    print("[gpu case] fused adam", flush=True)
    g = torch.Generator().manual_seed(3)
    shapes = [((40, 24), (33, 20), 0.01), ((1, 24), (1, 20), 0.0), ((24, 8), (24, 5), 0.01)]
    P, G, M_, V_, P16 = [], [], [], [], []
    for (r, c), (vr, vc), _ in shapes:
        p = torch.zeros(r, c, dtype=torch.float64)
        p[:vr, :vc] = torch.randn(vr, vc, generator=g, dtype=torch.float64)
        P.append(p)
        M_.append(torch.zeros(r, c, dtype=torch.float64))
        V_.append(torch.zeros(r, c, dtype=torch.float64))
    dev = [t.float().cuda().contiguous() for t in P]
    dm = [torch.zeros_like(t) for t in dev]
    dv = [torch.zeros_like(t) for t in dev]
    d16 = [torch.empty(t.shape, dtype=torch.bfloat16, device="cuda") for t in dev]
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    for step in range(1, 4):
        grads = []
        for ((r, c), (vr, vc), wd), p in zip(shapes, P):
            gr = torch.randn(r, c, generator=g, dtype=torch.float64)
            gr[vr:, :] = 7.0  # garbage in the padding must not leak
            gr[:, vc:] = 7.0
            grads.append(gr)
        tab = (A.AdamTensor * 3)()
        dg = [x.float().cuda().contiguous() for x in grads]
        for i, ((r, c), (vr, vc), wd) in enumerate(shapes):
            tab[i] = A.AdamTensor(dev[i].data_ptr(), dg[i].data_ptr(), dm[i].data_ptr(), dv[i].data_ptr(),
                                  d16[i].data_ptr() if i != 1 else None, r, c, vr, vc, wd, 0)
        assert lib.gh_adam_step(tab, 3, lr, b1, b2, eps, step, None) == 0
        torch.cuda.synchronize()
        for i, ((r, c), (vr, vc), wd) in enumerate(shapes):
            gg = grads[i][:vr, :vc] + wd * P[i][:vr, :vc]
            M_[i][:vr, :vc] = b1 * M_[i][:vr, :vc] + (1 - b1) * gg
            V_[i][:vr, :vc] = b2 * V_[i][:vr, :vc] + (1 - b2) * gg * gg
            mh = M_[i][:vr, :vc] / (1 - b1**step)
            vh = V_[i][:vr, :vc] / (1 - b2**step)
            P[i][:vr, :vc] -= lr * mh / (vh.sqrt() + eps)
    for i, ((r, c), (vr, vc), wd) in enumerate(shapes):
        got = dev[i].double().cpu()
        np.testing.assert_allclose(got[:vr, :vc].numpy(), P[i][:vr, :vc].numpy(), rtol=2e-5, atol=2e-6)
        assert torch.count_nonzero(got[vr:, :]) == 0 and torch.count_nonzero(got[:, vc:]) == 0
        assert torch.count_nonzero(dm[i].cpu()[vr:, :]) == 0 and torch.count_nonzero(dv[i].cpu()[:, vc:]) == 0
        if i != 1:
            assert torch.equal(d16[i].cpu(), dev[i].cpu().to(torch.bfloat16))
    bad = (A.AdamTensor * 1)(A.AdamTensor(None, None, None, None, None, 1, 1, 1, 1, 0.0, 0))
    assert lib.gh_adam_step(bad, 1, lr, b1, b2, eps, 1, None) != 0
    assert lib.gh_adam_step(tab, 3, lr, 1.0, b2, eps, 1, None) != 0


@pytest.mark.parametrize("sizes,M", [([8, 256, 256, 256, 256, 256, 6], 1000), ([3, 128, 128, 128, 4], 4099),
                                     ([8, 256, 256, 6], 64), ([5, 128, 128, 16], 130),
                                     ([8, 256, 256, 256, 6], 256 * 3 + 17), ([8, 256, 6], 300), ([8, 256, 256, 6], 1)])
def test_fused_forward_equals_layered(sizes, M):
    """gh_mlp_forward_bf16 (all layers in one kernel) gives the layer-by-layer GEMM values
    bit for bit, and both agree with the FP64 model of the bf16 network."""
    torch = _torch()
    from opengate_amd.learn.mlp import MLP

    print(f"[gpu case] fused mlp {sizes} M={M}", flush=True)
    net = MLP(sizes, seed=7)
    for b in net.b:
        b.normal_(0.0, 0.3)
    net._refresh()
    assert net.fused_ok()
    g = torch.Generator().manual_seed(M)
    X = torch.zeros(M, net.psizes[0], dtype=torch.bfloat16)
    X[:, : sizes[0]] = torch.randn(M, sizes[0], generator=g).to(torch.bfloat16)
    X = X.cuda()
    Zl = torch.full((M, net.psizes[-1]), float("nan"), device="cuda")
    bufs = [torch.empty(M, max(net.psizes[1:-1]), device="cuda", dtype=torch.bfloat16) for _ in range(2)]
    net.infer(X, Zl, bufs)
    Zf = torch.full((M, net.psizes[-1]), float("nan"), device="cuda")
    net.infer_fused(X, Zf)
    torch.cuda.synchronize()
    n = sizes[-1]
    np.testing.assert_array_equal(Zf[:, :n].cpu().numpy(), Zl[:, :n].cpu().numpy())
    W = [w.double().cpu().numpy() for w in net.W]
    bb = [b.double().cpu().numpy() for b in net.b]
    zr, _, _ = R.forward(W, bb, X.float().cpu().numpy())
    assert np.max(np.abs(Zf[:, :n].cpu().numpy() - zr[:, :n])) < 0.05 * max(1.0, np.abs(zr).max())
