"""Pins of the oracle's RNG and FP32 math (CPU only).

Philox4x32-10 is checked against the Random123 known-answer vectors (Salmon et al.,
SC11, kat_vectors: philox4x32 10 rounds); log and sin/cos against numpy's FP64
functions (the oracle must be within a few FP32 ulps of the mathematical function).
"""
import numpy as np
import pytest

from oracle import oracle as O

KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    assert tuple(O.philox4x32_10(ctr, key)) == expect


def test_philox_known_answer_fails_when_a_round_is_dropped():
    # the KAT distinguishes 10 rounds from 7 (Random123 philox4x32 7: 5f6fb709 ...)
    assert O.philox4x32_10((0, 0, 0, 0), (0, 0))[0] != 0x5F6FB709


def test_stream_layout():
    seed, hist, sim = 0x0123456789ABCDEF, (5 << 32) + 7, 3
    w = O.stream_words(seed, hist, sim, 12)
    key = (seed & 0xFFFFFFFF, seed >> 32)
    for block in range(3):
        exp = O.philox4x32_10((hist & 0xFFFFFFFF, hist >> 32, sim, block), key)
        assert list(w[4 * block: 4 * block + 4]) == exp
    # different histories / sims give different streams
    assert not np.array_equal(w, O.stream_words(seed, hist + 1, sim, 12))
    assert not np.array_equal(w, O.stream_words(seed, hist, sim + 1, 12))


def test_log_accuracy():
    """The logarithm serves uniform variates, odd multiples of 2^-24 in (0, 1): relative
    error below 2e-7 there (the whole 24-bit grid near 1, where log x -> 0, and random
    points below), and a few 1e-8 absolute above 1."""
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(0, 1, 5000), 2.0 ** rng.uniform(-24, 0, 5000),
                        (2 * np.arange(2 ** 23 - 2000, 2 ** 23) + 1) * 2.0 ** -24]).astype(np.float32)
    x = x[(x > 0) & (x < 1)]
    got = np.array([O.ref_log(v) for v in x])
    ref = np.log(x.astype(np.float64))
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 2e-7
    y = (2.0 ** rng.uniform(0, 10, 5000)).astype(np.float32)
    got = np.array([O.ref_log(v) for v in y])
    assert np.max(np.abs(got - np.log(y.astype(np.float64)))) < 1e-6
    # a dropped e ln2 term or a table entry of the wrong interval fails these
    assert abs(O.ref_log(np.float32(0.25)) - np.log(0.25)) < 1e-7
    assert abs(O.ref_log(np.float32(1 - 2 ** -24)) / np.log1p(-2.0 ** -24) - 1) < 1e-6


def test_sincos_accuracy():
    t = np.random.default_rng(1).uniform(-4 * np.pi, 4 * np.pi, 5000).astype(np.float32)
    sc = np.array([O.ref_sincos(v) for v in t])
    assert np.max(np.abs(sc[:, 0] - np.sin(t.astype(np.float64)))) < 3e-7
    assert np.max(np.abs(sc[:, 1] - np.cos(t.astype(np.float64)))) < 3e-7
