"""The oracle's GF((2^61-1)^2) arithmetic, pinned three ways: against the
reference's own compiled fieldElement (oracle/_ref and the committed
ref_vectors.json), against Python big-int arithmetic, and by the field's
algebraic identities.  Reference: src/fieldElement.cpp:34-96, 207-210,
238-250; src/utils.cpp:706-717."""
import json
import os

import numpy as np
import pytest

import pyoracle as O

P = O.P
GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.json")


def py_mul(a, b):
    ar, ai = int(a[0]), int(a[1])
    br, bi = int(b[0]), int(b[1])
    return ((ar * br - ai * bi) % P, (ar * bi + ai * br) % P)


def rand_fe(rng, n):
    return rng.integers(0, P, size=(n, 2), dtype=np.uint64)


@pytest.fixture(scope="module")
def gold():
    with open(GOLD) as fh:
        return json.load(fh)


def test_oracle_matches_reference_golden(gold):
    a = np.array(gold["a"], np.uint64)
    b = np.array(gold["b"], np.uint64)
    for name in ("mul", "add", "sub"):
        assert np.array_equal(O.vec_op(name, a, b), np.array(gold[name], np.uint64)), name
    assert np.array_equal(O.vec_op("inv", a), np.array(gold["inv"], np.uint64))


def test_roots_match_reference_golden(gold):
    for logn, r in enumerate(gold["root_of_unity"]):
        assert list(O.root_of_unity(logn)) == r


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")
def test_oracle_matches_reference_library_random():
    rng = np.random.default_rng(11)
    a, b = rand_fe(rng, 20000), rand_fe(rng, 20000)
    R = O.ref()
    for name in ("mul", "add", "sub"):
        out = np.zeros_like(a)
        getattr(R, "ref_" + name)(O._p(a), O._p(b), O._p(out), a.shape[0])
        assert np.array_equal(O.vec_op(name, a, b), out), name


def test_mul_matches_python_bigint():
    rng = np.random.default_rng(12)
    a, b = rand_fe(rng, 2000), rand_fe(rng, 2000)
    got = O.vec_op("mul", a, b)
    for i in range(a.shape[0]):
        assert tuple(int(x) for x in got[i]) == py_mul(a[i], b[i])


def test_identities():
    rng = np.random.default_rng(13)
    a, b, c = rand_fe(rng, 500), rand_fe(rng, 500), rand_fe(rng, 500)
    # canonical range
    assert (O.vec_op("mul", a, b) < P).all()
    # I^2 = -1
    i = np.array([[0, 1]], np.uint64)
    assert O.vec_op("mul", i, i).tolist() == [[P - 1, 0]]
    # distributivity, inverse
    lhs = O.vec_op("mul", a, O.vec_op("add", b, c))
    rhs = O.vec_op("add", O.vec_op("mul", a, b), O.vec_op("mul", a, c))
    assert np.array_equal(lhs, rhs)
    one = np.tile(np.array([[1, 0]], np.uint64), (500, 1))
    assert np.array_equal(O.vec_op("mul", a, O.vec_op("inv", a)), one)
    assert np.array_equal(O.vec_op("sub", O.vec_op("add", a, b), b), a)


@pytest.mark.parametrize("logn", [1, 2, 3, 10, 22, 28, 61])
def test_root_order_exact(logn):
    # root_n has order exactly 2^n: root^(2^(n-1)) = -1
    r = O.root_of_unity(logn).reshape(1, 2)
    x = r
    for _ in range(logn - 1):
        x = O.vec_op("mul", x, x)
    assert x.tolist() == [[P - 1, 0]]
