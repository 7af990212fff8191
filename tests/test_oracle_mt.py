"""The oracle's all-cores variants (oracle/oracle_mt.c) against its
single-thread functions, which the other oracle tests pin: identical bytes
for the transform (forward and inverse, both sides of the four-step
threshold), the dup-left and standard trees, the product sumcheck and the
plain proof.  The single-thread transform is itself pinned by the O(n^2)
DFT (test_oracle_ntt.py), so a wrong twist index or transposed step in the
four-step fails here."""
import numpy as np
import pytest

import pyoracle as O


@pytest.mark.parametrize("logn", [3, 11, 12, 13, 15])
@pytest.mark.parametrize("inverse", [False, True])
def test_fft_mt_equals_fft(logn, inverse):
    a = O.synth(0xC05A00F0 + logn, int(inverse), 1 << logn)
    assert np.array_equal(O.fft_mt(a, inverse, threads=4), O.fft(a, inverse))


def test_fft_mt_thread_counts_agree():
    a = O.synth(0xC05A00F1, 0, 1 << 14)
    ref = O.fft(a)
    for th in (1, 3, 8):
        assert np.array_equal(O.fft_mt(a, threads=th), ref)


def test_fft_mt_is_a_dft_at_sampled_points():
    # X[k] = sum_j x_j w^(jk), checked at a few k by Horner (independent of
    # either transform's structure)
    logn = 13
    n = 1 << logn
    x = O.synth(0xC05A00F2, 0, n)
    X = O.fft_mt(x, threads=4)
    w = O.root_of_unity(logn)
    for k in (0, 1, 7, 4095, n - 1):
        wk = O.vec_op("mul", np.array([[1, 0]], np.uint64), np.array([w], np.uint64))[0]
        pt = np.array([1, 0], np.uint64)
        for _ in range(k):
            pt = O.vec_op("mul", pt.reshape(1, 2), wk.reshape(1, 2))[0]
        acc = np.zeros(2, np.uint64)
        for j in range(n - 1, -1, -1):
            acc = O.vec_op("add", O.vec_op("mul", acc.reshape(1, 2), pt.reshape(1, 2)), x[j:j + 1])[0]
        assert np.array_equal(X[k], acc), k


@pytest.mark.parametrize("rate", [2, 4])
def test_rs_encode_mt_equals_rs_encode(rate):
    a = O.synth(0xC05A00F3, rate, 1 << 12)
    assert np.array_equal(O.rs_encode_mt(a, rate, threads=4), O.rs_encode(a, rate))


@pytest.mark.parametrize("dup_left", [True, False])
def test_merkle_mt_equals_merkle(dup_left):
    e = O.synth(0xC05A00F4, 0, 1 << 14)
    assert np.array_equal(O.merkle_commit_mt(e, dup_left, threads=4), O.merkle_commit(e, dup_left))


def test_sumcheck_mt_equals_sumcheck():
    n = 1 << 13
    f, g = O.synth(0xC05A00F5, 0, n), O.synth(0xC05A00F5, 1, n)
    y = O.inner_product(f, g)
    p1, p2 = np.zeros(2, np.uint64), np.zeros(2, np.uint64)
    a = O.product_sumcheck(y, f, g, p1, 0)
    b = O.product_sumcheck_mt(y, f, g, p2, threads=4)
    assert a[0] == 0 and b[0] == 0
    for u, v in zip(a[1:], b[1:]):
        assert np.array_equal(u, v)
    assert np.array_equal(p1, p2)


def test_sumcheck_mt_rejects_wrong_claim():
    n = 1 << 8
    f, g = O.synth(5, 0, n), O.synth(5, 1, n)
    y = O.inner_product(f, g)
    y[1] = (int(y[1]) + 1) % O.P
    rc, *_ = O.product_sumcheck_mt(y, f, g, np.zeros(2, np.uint64), threads=2)
    assert rc == -1


@pytest.mark.parametrize("m", [4, 12, 14])
def test_plain_proof_mt_equals_plain_proof(m):
    n = 1 << m
    f, g = O.synth(0xC05A0001, 0, n), O.synth(0xC05A0001, 1, n)
    prev = np.array([12345, 678], np.uint64)
    exp, _, _ = O.plain_proof(f, g, 4, prev.copy())
    got, ms = O.plain_proof_mt(f, g, 4, prev.copy(), threads=4)
    assert got == exp
    assert ms[3] >= 0


@pytest.mark.parametrize("logn,rate,stop", [(10, 4, 0), (14, 4, 8), (13, 2, 0)])
def test_whir_open_threaded_equals_single_thread(logn, rate, stop):
    n = 1 << logn
    arr, z = O.synth(0xC05A00F6, 0, n), O.synth(0xC05A00F6, 1, logn)
    K = O.whir_claim(arr, z)
    recs = []
    for t in (1, 4):
        O.whir_threads(t)
        try:
            O.libc().srandom(99)
            rc, rec = O.whir_open(arr, z, K, rate, stop)
        finally:
            O.whir_threads(1)
        assert rc == 0
        recs.append(rec)
    assert np.array_equal(recs[0], recs[1])


def test_update_betas_mt_equals_direct_sum():
    n, l = 1 << 10, 5
    b = O.synth(0xC05A00F7, 0, n)
    ch, w = O.synth(0xC05A00F7, 1, l), O.synth(0xC05A00F7, 2, l)
    exp = b.copy()
    for i in range(l):
        pw = np.array([1, 0], np.uint64)
        for j in range(n):
            exp[j] = O.vec_op("add", exp[j:j + 1], O.vec_op("mul", ch[i:i + 1], pw.reshape(1, 2)))[0]
            pw = O.vec_op("mul", pw.reshape(1, 2), w[i:i + 1])[0]
    got = b.copy()
    O.oracle().or_update_betas_mt(O._p(got), n, O._p(np.ascontiguousarray(ch)), O._p(np.ascontiguousarray(w)), l, 3)
    assert np.array_equal(got, exp)
