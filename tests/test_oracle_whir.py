"""The oracle's WHIR opening (src/Whir.cpp:158-467, whir_k = 4).  Pins: the
reference's own verifier checks run inside or_whir_open (every sumcheck
round's H(0)+H(1) = K, and the closing check K = <arr, b> after the
update_betas / K += c_i y_i updates), which only hold if the query folds
y_i (get_RS_multifold_segment_tensor) agree with the RS encoding; a wrong
claim is rejected; the query count follows the reference's formula."""
import math

import numpy as np
import pytest

import pyoracle as O


@pytest.mark.parametrize("logn,rate,stop", [(8, 4, 0), (10, 4, 0), (11, 2, 0), (12, 4, 8)])
def test_whir_open_self_consistent(logn, rate, stop):
    n = 1 << logn
    arr, z = O.synth(900 + logn, 0, n), O.synth(900 + logn, 1, logn)
    K = O.whir_claim(arr, z)
    O.libc().srandom(77 + logn)
    rc, rec = O.whir_open(arr, z, K, rate, stop)
    assert rc == 0
    iters = int(rec[0])
    expect = 0
    cur = n
    while True:
        expect += 1
        cur //= 4
        if cur <= max(4, 1 << stop):
            break
    assert iters == expect


def test_whir_wrong_claim_rejected():
    arr, z = O.synth(950, 0, 1024), O.synth(950, 1, 10)
    K = O.whir_claim(arr, z)
    K[1] = (int(K[1]) + 3) % O.P
    O.libc().srandom(5)
    rc, _ = O.whir_open(arr, z, K, 4)
    assert rc < 0


@pytest.mark.parametrize("rate", [4, 8, 16, 32, 64, 128, 256, 512])
def test_query_count_formula(rate):
    assert O.oracle().or_whir_query_count(rate) == int(-100.0 / math.log2(math.sqrt(1.0 / rate)))


def test_segment_fold_is_folded_polynomial_evaluation():
    # The fold of a whir_k-segment of an RS codeword equals the codeword of the
    # sumcheck-folded coefficient vector at the corresponding point:
    # y(idx) = RS(fold(arr, r0, r1), rate)[idx * 4 / 4 ...] -- checked through
    # the identity y = sum_j f'(j) * w'^(idx*j) with f' the folded vector.
    n, rate = 64, 4
    arr = O.synth(960, 0, n)
    r = O.synth(960, 1, 2)
    cw = O.rs_encode(arr, rate)
    N = n * rate
    one = np.array([[1, 0]], np.uint64)

    def fold(v, c):
        e, o = v[0::2], v[1::2]
        return O.vec_op("add", O.vec_op("mul", np.tile(c.reshape(1, 2), (e.shape[0], 1)), O.vec_op("sub", o, e)), e)
    f2 = fold(fold(arr, r[0]), r[1])
    w = O.root_of_unity(N.bit_length() - 1).reshape(1, 2)
    for idx in (0, 3, 17, N // 4 - 1):
        seg = np.stack([cw[idx + j * N // 4] for j in range(4)])
        y = O.unfe(O.oracle().or_rs_multifold_segment_tensor(O._p(seg), 4, O._p(np.ascontiguousarray(r)), 2, N, idx))
        # w^(4 idx) evaluates f2 (length n/4)
        x = one
        for _ in range(4 * idx):
            x = O.vec_op("mul", x, w)
        acc = np.zeros((1, 2), np.uint64)
        p = one
        for j in range(f2.shape[0]):
            acc = O.vec_op("add", acc, O.vec_op("mul", f2[j:j + 1], p))
            p = O.vec_op("mul", p, x)
        assert np.array_equal(y, acc[0])
