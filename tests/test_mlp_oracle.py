"""The MLP oracle's backpropagation against central finite differences (CPU)."""
import numpy as np

from oracle import mlp_ref as R


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(0)
    sizes = [5, 7, 6, 3]
    Ws = [rng.normal(size=(a, b)) * 0.5 for a, b in zip(sizes[:-1], sizes[1:])]
    bs = [rng.normal(size=b) * 0.1 for b in sizes[1:]]
    x = rng.normal(size=(4, sizes[0]))
    t = rng.normal(size=(4, sizes[-1]))
    loss = lambda Ws, bs: 0.5 * np.sum((R.forward(Ws, bs, x, bf16=False)[0] - t) ** 2)  # noqa: E731
    out, xs, zs = R.forward(Ws, bs, x, bf16=False)
    # ReLU kinks: no hidden pre-activation near 0, so that the central differences (step
    # 1e-6) stay on one side of every kink
    assert min(np.abs(z).min() for z in zs[:-1]) > 1e-3
    dW, db = R.backward(Ws, xs, zs, out - t, bf16=False)
    h = 1e-6
    for l in range(len(Ws)):
        num = np.zeros_like(Ws[l])
        for idx in np.ndindex(Ws[l].shape):
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][idx] += h
            Wm[l][idx] -= h
            num[idx] = (loss(Wp, bs) - loss(Wm, bs)) / (2 * h)
        assert np.allclose(num, dW[l], rtol=1e-6, atol=1e-8)
        numb = np.zeros_like(bs[l])
        for i in range(len(bs[l])):
            bp = [b.copy() for b in bs]
            bm = [b.copy() for b in bs]
            bp[l][i] += h
            bm[l][i] -= h
            numb[i] = (loss(Ws, bp) - loss(Ws, bm)) / (2 * h)
        assert np.allclose(numb, db[l], rtol=1e-6, atol=1e-8)


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7 + 2**-9, 1.0 + 3 * 2**-8, -2.5], dtype=np.float32)
    # 1+2^-8 is a tie between 1 and 1+2^-7 -> even (1); 1+3*2^-8 ties to 1+2^-6 (even mantissa)
    assert list(R.to_bf16(x)) == [1.0, 1.0, 1.0 + 2**-7, 1.0 + 2**-6, -2.5]
