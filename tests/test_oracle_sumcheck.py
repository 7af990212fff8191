"""The oracle's product sumcheck (src/Distributed_Sumcheck.cpp:204-348 at
one rank; tail src/sparse_eval.cpp:140-192), pinned by the round identity
H(0)+H(1) = y, by the final values being the multilinear extensions of
the inputs at the challenges (LSB variable first, src/utils.cpp:1201-1214),
by the reference's compiled linear_poly product (coefficient order), and
by the libc challenge stream of the reference's F::_random()."""
import json
import os

import numpy as np
import pytest

import pyoracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")


def quad_eval(p, x):
    x = np.asarray(x, np.uint64).reshape(1, 2)
    acc = O.vec_op("add", O.vec_op("mul", p[0:1], x), p[1:2])
    return O.vec_op("add", O.vec_op("mul", acc, x), p[2:3])[0]


def mle_brute(v, r):
    # sum_x v[x] * prod_k (bit_k(x) ? r_k : 1 - r_k), bit 0 <-> r[0]
    n = v.shape[0]
    one = np.array([[1, 0]], np.uint64)
    total = np.zeros((1, 2), np.uint64)
    for x in range(n):
        w = one
        for k in range(r.shape[0]):
            rk = r[k:k + 1]
            w = O.vec_op("mul", w, rk if (x >> k) & 1 else O.vec_op("sub", one, rk))
        total = O.vec_op("add", total, O.vec_op("mul", w, v[x:x + 1]))
    return total[0]


@pytest.mark.parametrize("m", [1, 3, 6])
def test_round_identity_and_final_values(m):
    n = 1 << m
    f, g = O.synth(600 + m, 0, n), O.synth(600 + m, 1, n)
    y = O.inner_product(f, g)
    prev = np.zeros(2, np.uint64)
    rc, polys, r, fin = O.product_sumcheck(y, f, g, prev, 0)
    assert rc == 0
    cur = y
    for i in range(m):
        p = polys[i]
        s = O.vec_op("add", quad_eval(p, [0, 0]).reshape(1, 2), quad_eval(p, [1, 0]).reshape(1, 2))[0]
        assert np.array_equal(s, cur)
        cur = quad_eval(p, r[i])
    assert np.array_equal(fin[0], mle_brute(f, r))
    assert np.array_equal(fin[1], mle_brute(g, r))
    # y after the last round is the product of the final values
    assert np.array_equal(cur, O.vec_op("mul", fin[0:1], fin[1:2])[0])
    # the reference's evaluate_vector agrees
    assert np.array_equal(O.evaluate_vector(f, r), fin[0])


def test_challenges_follow_the_transcript():
    m = 5
    f, g = O.synth(700, 0, 32), O.synth(700, 1, 32)
    prev = np.zeros(2, np.uint64)
    _, polys, r, _ = O.product_sumcheck(O.inner_product(f, g), f, g, prev, 0)
    acc = np.zeros(2, np.uint64)
    for i in range(m):
        for c in polys[i]:
            acc = O.hash_F(acc, c)
        assert np.array_equal(acc, r[i])


def test_wrong_claim_detected():
    f, g = O.synth(701, 0, 16), O.synth(701, 1, 16)
    y = O.inner_product(f, g)
    y[1] = (int(y[1]) + 1) % O.P
    rc, *_ = O.product_sumcheck(y, f, g, np.zeros(2, np.uint64), 0)
    assert rc == -1


def test_linear_product_convention_matches_reference_golden():
    g = json.load(open(GOLD))
    for case in g["linear_mul"]:
        l1, l2 = np.array(case["l1"], np.uint64), np.array(case["l2"], np.uint64)
        # (a1 x + b1)(a2 x + b2) = a1a2 x^2 + (a1b2 + b1a2) x + b1b2
        a = O.vec_op("mul", l1[0:1], l2[0:1])[0]
        b = O.vec_op("add", O.vec_op("mul", l1[0:1], l2[1:2]), O.vec_op("mul", l1[1:2], l2[0:1]))[0]
        c = O.vec_op("mul", l1[1:2], l2[1:2])[0]
        assert np.array_equal(np.stack([a, b, c]), np.array(case["quad"], np.uint64))
        assert list(quad_eval(np.array(case["quad"], np.uint64), case["x"])) == case["eval"]


def test_random_element_stream_matches_reference_golden():
    g = json.load(open(GOLD))
    for seed, elems in g["random_elements"].items():
        O.libc().srandom(int(seed))
        got = [list(O.unfe(O.oracle().or_random_element())) for _ in elems]
        assert got == elems


def test_libc_tail_rounds():
    m, off = 6, 4
    f, g = O.synth(702, 0, 64), O.synth(702, 1, 64)
    y = O.inner_product(f, g)
    O.libc().srandom(99)
    rc, polys, r, fin = O.product_sumcheck(y, f, g, np.zeros(2, np.uint64), off)
    assert rc == 0
    O.libc().srandom(99)
    draws = [O.unfe(O.oracle().or_random_element()) for _ in range(4 * off)]
    for i in range(off):
        assert np.array_equal(r[m - off + i], draws[4 * i + 3])
