"""The oracle's fft / pfft / RS encode (src/utils.cpp:904-964, 1583-1673;
src/Whir.cpp:158-163), pinned by the DFT definition (O(n^2) evaluation),
the inverse round trip and Horner evaluation of the encoded polynomial."""
import numpy as np
import pytest

import pyoracle as O


def dft_direct(a, inverse=False):
    n = a.shape[0]
    logn = n.bit_length() - 1
    w = O.root_of_unity(logn).reshape(1, 2)
    if inverse:
        w = O.vec_op("inv", w)
    # powers w^(jk) via a table of w^e, e < n
    pw = np.zeros((n, 2), np.uint64)
    pw[0] = [1, 0]
    for e in range(1, n):
        pw[e] = O.vec_op("mul", pw[e - 1:e], w)[0]
    out = np.zeros_like(a)
    for k in range(n):
        idx = (np.arange(n) * k) % n
        terms = O.vec_op("mul", a, pw[idx])
        re = int(np.sum(terms[:, 0].astype(object)) % O.P)
        im = int(np.sum(terms[:, 1].astype(object)) % O.P)
        out[k] = [re, im]
    if inverse:
        ninv = O.vec_op("inv", np.array([[n, 0]], np.uint64))
        out = O.vec_op("mul", out, np.tile(ninv, (n, 1)))
    return out


@pytest.mark.parametrize("logn", [1, 2, 3, 5, 7])
def test_fft_is_the_dft(logn):
    a = O.synth(100 + logn, 0, 1 << logn)
    assert np.array_equal(O.fft(a), dft_direct(a))
    assert np.array_equal(O.fft(a, True), dft_direct(a, True))


@pytest.mark.parametrize("logn", [4, 12, 16])
def test_round_trip(logn):
    a = O.synth(200 + logn, 1, 1 << logn)
    assert np.array_equal(O.fft(O.fft(a), True), a)


def horner(coeffs, x):
    acc = np.zeros((1, 2), np.uint64)
    xx = np.asarray(x, np.uint64).reshape(1, 2)
    for c in coeffs[::-1]:
        acc = O.vec_op("add", O.vec_op("mul", acc, xx), c.reshape(1, 2))
    return acc[0]


@pytest.mark.parametrize("logn,rate", [(6, 4), (8, 2), (9, 8)])
def test_rs_encode_evaluates_polynomial(logn, rate):
    n = 1 << logn
    a = O.synth(300 + logn, 0, n)
    cw = O.rs_encode(a, rate)
    w = O.root_of_unity((n * rate).bit_length() - 1).reshape(1, 2)
    rng = np.random.default_rng(logn)
    for k in rng.integers(0, n * rate, size=6):
        x = np.array([[1, 0]], np.uint64)
        for _ in range(int(k)):
            x = O.vec_op("mul", x, w)
        assert np.array_equal(cw[k], horner(a, x))
