"""FP64 numpy restatement of the dense MLP (TEST INFRASTRUCTURE ONLY).

Mirrors opengate_amd/learn/mlp.py: Z_l = X_l W_l + b_l, X_{l+1} = relu(Z_l), last layer
linear; backward dW_l = X_l^T G_l, db_l = sum G_l, G_{l-1} = (G_l W_l^T) relu'(Z_{l-1}).
With bf16=True the values the GPU rounds to bfloat16 (layer inputs, weights, the
gradients fed to the GEMMs) are rounded the same way (round to nearest even).
Pinned in tests/test_mlp_oracle.py by central finite differences of the FP64 model.
"""
import numpy as np


def to_bf16(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def relu(z):
    return np.maximum(z, 0.0)


def drelu(z):
    return (np.asarray(z) > 0.0).astype(np.float64)


def forward(Ws, bs, x, bf16=True):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    xs, zs = [r(x)], []
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = xs[-1] @ r(W) + b
        zs.append(z)
        if l < len(Ws) - 1:
            xs.append(r(relu(z)))
    return zs[-1], xs, zs


def backward(Ws, xs, zs, g_out, bf16=True):
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    L = len(Ws)
    g32 = np.asarray(g_out, dtype=np.float64)
    dW, db = [None] * L, [None] * L
    for l in range(L - 1, -1, -1):
        g16 = r(g32)
        dW[l] = xs[l].T @ g16
        db[l] = g32.sum(0)
        if l > 0:
            g32 = (g16 @ r(Ws[l]).T) * drelu(zs[l - 1])
    return dW, db
