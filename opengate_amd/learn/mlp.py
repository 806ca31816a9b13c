"""Dense MLP trained and evaluated with the library's BF16 MFMA GEMM (gh_gemm_bf16).

Every matrix product of the forward and backward passes runs in the hand-written
kernel (csrc/mlp.hip) with FP32 accumulation; bias, ReLU and the ReLU derivative are
fused into its epilogues. PyTorch holds the tensors (device memory, the FP32 master
weights, the Adam state and the loss), nothing more.

Layer l: Z_l = X_l W_l + b_l (FP32), X_{l+1} = bf16(relu(Z_l)) for hidden layers,
the last layer is linear. Backward: G_L = dLoss/dZ_L, dW_l = X_l^T G_l,
db_l = sum_m G_l, G_{l-1} = (G_l W_l^T) * relu'(Z_{l-1}).
"""
import ctypes as C

import numpy as np

from .. import _abi as A

EPI_STORE, EPI_BIAS_RELU, EPI_BIAS, EPI_DRELU = 0, 1, 2, 3


def _torch():
    import torch

    return torch


def pad8(n):
    return (n + 7) // 8 * 8


class MLP:
    """sizes = [in, hidden..., out]; inputs/outputs padded to multiples of 8 internally."""

    def __init__(self, sizes, device="cuda", seed=0, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.0):
        torch = _torch()
        self.lib = A.load_library()
        self.sizes = list(sizes)
        self.dev = torch.device(device)
        self.psizes = [pad8(s) for s in sizes]
        g = torch.Generator(device="cpu").manual_seed(seed)
        self.W, self.b = [], []
        for k, n, pk, pn in zip(self.sizes[:-1], self.sizes[1:], self.psizes[:-1], self.psizes[1:]):
            w = torch.zeros(pk, pn)
            w[:k, :n] = torch.randn(k, n, generator=g) * np.sqrt(2.0 / k)
            self.W.append(w.to(self.dev))
            self.b.append(torch.zeros(pn, device=self.dev))
        self.lr, self.betas, self.eps, self.wd = lr, betas, eps, weight_decay
        self.m = [torch.zeros_like(t) for t in self.W + self.b]
        self.v = [torch.zeros_like(t) for t in self.W + self.b]
        self.t = 0
        self._refresh()

    # ------------------------------------------------------------------ helpers
    def _refresh(self):
        """bf16 copies of the weights (kept in place: the optimiser kernel writes them)."""
        torch = _torch()
        if getattr(self, "W16", None) is None or [t.shape for t in self.W16] != [w.shape for w in self.W]:
            self.W16 = [torch.empty(w.shape, device=self.dev, dtype=torch.bfloat16) for w in self.W]
        for t, w in zip(self.W16, self.W):
            t.copy_(w)
        self._fused = None

    def _stream(self):
        return C.c_void_p(_torch().cuda.current_stream().cuda_stream)

    def _gemm(self, a, sam, sak, b, sbk, sbn, M, N, K, bias=None, zprev=None, ldz=0, c32=None, c16=None, ldc=0,
              mode=EPI_STORE):
        p = lambda t: C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p()  # noqa: E731
        rc = self.lib.gh_gemm_bf16(p(a), sam, sak, p(b), sbk, sbn, M, N, K, p(bias), p(zprev), ldz, p(c32), p(c16),
                                   ldc, mode, self._stream())
        if rc != 0:
            raise RuntimeError(f"gh_gemm_bf16 failed ({rc})")

    def n_params(self):
        return sum(int(np.prod(s)) for s in zip(self.sizes[:-1], self.sizes[1:])) + sum(self.sizes[1:])

    def flops_forward(self, batch):
        return 2.0 * batch * sum(k * n for k, n in zip(self.sizes[:-1], self.sizes[1:]))

    # ------------------------------------------------------------------ passes
    def forward(self, x, keep=True):
        """x: (M, in) float tensor on the device, or an already padded contiguous
        (M, in padded) bf16 tensor. Returns (M, out) FP32."""
        torch = _torch()
        M = x.shape[0]
        if x.dtype == torch.bfloat16 and x.shape[1] == self.psizes[0] and x.is_contiguous():
            X = x
        else:
            X = torch.zeros(M, self.psizes[0], device=self.dev, dtype=torch.bfloat16)
            X[:, : self.sizes[0]] = x.to(torch.bfloat16)
        xs, zs = [X], []
        L = len(self.W)
        for l in range(L):
            K, N = self.psizes[l], self.psizes[l + 1]
            Z = torch.empty(M, N, device=self.dev, dtype=torch.float32)
            if l < L - 1:
                Y = torch.empty(M, N, device=self.dev, dtype=torch.bfloat16)
                self._gemm(xs[-1], K, 1, self.W16[l], N, 1, M, N, K, bias=self.b[l], c32=Z, c16=Y, ldc=N,
                           mode=EPI_BIAS_RELU)
                xs.append(Y)
            else:
                self._gemm(xs[-1], K, 1, self.W16[l], N, 1, M, N, K, bias=self.b[l], c32=Z, ldc=N, mode=EPI_BIAS)
            zs.append(Z)
        if keep:
            self._xs, self._zs = xs, zs
        return zs[-1][:, : self.sizes[-1]]

    def infer(self, X, out, bufs):
        """Inference without keeping activations: X (M, in padded) bf16, out (M, out padded)
        FP32, bufs two bf16 buffers of at least (M, widest hidden layer). Hidden layers write
        only their bf16 activations."""
        M = X.shape[0]
        cur = X
        L = len(self.W)
        for l in range(L):
            K, N = self.psizes[l], self.psizes[l + 1]
            if l < L - 1:
                Y = bufs[l % 2].view(-1)[: M * N].view(M, N)
                self._gemm(cur, K, 1, self.W16[l], N, 1, M, N, K, bias=self.b[l], c16=Y, ldc=N, mode=EPI_BIAS_RELU)
                cur = Y
            else:
                self._gemm(cur, K, 1, self.W16[l], N, 1, M, N, K, bias=self.b[l], c32=out, ldc=out.stride(0),
                           mode=EPI_BIAS)
        return out

    # ------------------------------------------------------------------ fused inference
    FUSED_K_IN, FUSED_N_OUT, FUSED_MAX_LAYERS = 32, 16, 8

    def fused_ok(self):
        """gh_mlp_forward_bf16 handles 2..8 layers with equal hidden widths of 128 or 256,
        at most 32 inputs and 16 outputs."""
        hidden = set(self.psizes[1:-1])
        return (2 <= len(self.W) <= self.FUSED_MAX_LAYERS and len(hidden) == 1 and hidden <= {128, 256}
                and self.psizes[0] <= self.FUSED_K_IN and self.psizes[-1] <= self.FUSED_N_OUT)

    def _fused_weights(self):
        """Transposed, padded bf16 weights (first layer K -> 32, last layer N -> 16) and
        biases (last layer -> 16), rebuilt after each optimiser step."""
        torch = _torch()
        if self._fused is None:
            L = len(self.W16)
            wts, bs = [], []
            for l, w16 in enumerate(self.W16):
                K, N = w16.shape
                Kp = self.FUSED_K_IN if l == 0 else K
                Np = self.FUSED_N_OUT if l == L - 1 else N
                t = torch.zeros(Np, Kp, device=self.dev, dtype=torch.bfloat16)
                t[:N, :K] = w16.t()
                wts.append(t.contiguous())
                b = torch.zeros(Np, device=self.dev)
                b[:N] = self.b[l]
                bs.append(b)
            self._fused = (wts, bs)
        return self._fused

    def infer_fused(self, X, out):
        """All layers in one kernel (gh_mlp_forward_bf16): X (M, in padded) bf16 with row
        pitch X.stride(0), out (M, >= out) FP32 with row pitch out.stride(0); only the
        first sizes[-1] columns of out are written. Same values as infer()."""
        wts, bs = self._fused_weights()
        L = len(wts)
        wp = (C.c_void_p * L)(*[t.data_ptr() for t in wts])
        bp = (C.POINTER(C.c_float) * L)(*[C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_float)) for t in bs])
        rc = self.lib.gh_mlp_forward_bf16(C.c_void_p(X.data_ptr()), X.stride(0), X.shape[0], L, self.psizes[0],
                                          self.psizes[1], self.sizes[-1], wp, bp,
                                          C.cast(C.c_void_p(out.data_ptr()), C.POINTER(C.c_float)),
                                          out.stride(0), self._stream())
        if rc != 0:
            raise RuntimeError(f"gh_mlp_forward_bf16 failed ({rc})")
        return out

    def backward(self, grad_out):
        """grad_out: dLoss/dOutput (M, out) FP32. Returns [dW...], [db...] (FP32, padded)."""
        torch = _torch()
        M = grad_out.shape[0]
        L = len(self.W)
        G32 = torch.zeros(M, self.psizes[-1], device=self.dev, dtype=torch.float32)
        G32[:, : self.sizes[-1]] = grad_out
        G16 = G32.to(torch.bfloat16)
        dW, db = [None] * L, [None] * L
        for l in range(L - 1, -1, -1):
            K, N = self.psizes[l], self.psizes[l + 1]
            dw = torch.empty(K, N, device=self.dev, dtype=torch.float32)
            # dW = X^T G : (K x M) (M x N)
            self._gemm(self._xs[l], 1, K, G16, N, 1, K, N, M, c32=dw, ldc=N, mode=EPI_STORE)
            dW[l] = dw
            db[l] = G32.sum(0)
            if l > 0:
                Gp32 = torch.empty(M, K, device=self.dev, dtype=torch.float32)
                Gp16 = torch.empty(M, K, device=self.dev, dtype=torch.bfloat16)
                # G_{l-1} = (G W^T) * relu'(Z_{l-1}) : (M x N) (N x K)
                self._gemm(G16, N, 1, self.W16[l], 1, N, M, K, N, zprev=self._zs[l - 1], ldz=K, c32=Gp32, c16=Gp16,
                           ldc=K, mode=EPI_DRELU)
                G32, G16 = Gp32, Gp16
        return dW, db

    def step(self, dW, db):
        """Adam update of the FP32 master weights and biases in one kernel (gh_adam_step),
        which also keeps the padding at zero and writes the bf16 weight copies."""
        self.t += 1
        params = self.W + self.b
        grads = list(dW) + list(db)
        valid = [(k, n) for k, n in zip(self.sizes[:-1], self.sizes[1:])] + [(1, n) for n in self.sizes[1:]]
        tab = (A.AdamTensor * len(params))()
        for i, (p, g, (vr, vc)) in enumerate(zip(params, grads, valid)):
            if g.shape != p.shape or not g.is_contiguous() or g.dtype != p.dtype:
                raise ValueError("gradient shape / layout differs from its parameter")
            rows, cols = (p.shape[0], p.shape[1]) if p.dim() == 2 else (1, p.shape[0])
            is_w = i < len(self.W)
            tab[i] = A.AdamTensor(p.data_ptr(), g.data_ptr(), self.m[i].data_ptr(), self.v[i].data_ptr(),
                                  self.W16[i].data_ptr() if is_w else None, rows, cols, vr, vc,
                                  float(self.wd) if is_w else 0.0, 0)
        b1, b2 = self.betas
        rc = self.lib.gh_adam_step(tab, len(params), float(self.lr), float(b1), float(b2), float(self.eps), self.t,
                                   self._stream())
        if rc != 0:
            raise RuntimeError(f"gh_adam_step failed ({rc})")
        self._fused = None

    def state_dict(self):
        return {"sizes": self.sizes, "W": [w.cpu() for w in self.W], "b": [b.cpu() for b in self.b]}

    def load_state_dict(self, d):
        self.W = [w.to(self.dev) for w in d["W"]]
        self.b = [b.to(self.dev) for b in d["b"]]
        self._refresh()
