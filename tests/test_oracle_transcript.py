"""The oracle's Fiat-Shamir chain (src/Fiat_Shamir.cpp:8-56): hash_F is
BLAKE3 of {x.real, x.img, 0, 0, y.real, y.img, 0, 0} with the digest read
at a 7-bit stride.  Pinned by recomputing the digest with BLAKE3 on the
documented block and by the stride's range bound; chain order by folding."""
import numpy as np

import pyoracle as O


def test_hash_F_block_layout_and_packing():
    rng = np.random.default_rng(3)
    for _ in range(50):
        x = rng.integers(0, O.P, size=2, dtype=np.uint64)
        y = rng.integers(0, O.P, size=2, dtype=np.uint64)
        blk = np.array([x[0], x[1], 0, 0, y[0], y[1], 0, 0], np.uint64).tobytes()
        H = O.blake3_64(blk)
        re = sum(H[i] << (7 * i) for i in range(8))
        im = sum(H[8 + i] << (7 * i) for i in range(8))
        assert list(O.hash_F(x, y)) == [re, im]


def test_seven_bit_stride_bound():
    # 255 * sum_i 2^(7i) < 2^57.1: an 8-bit (memcpy) reading would exceed it
    bound = 255 * sum(1 << (7 * i) for i in range(8))
    rng = np.random.default_rng(4)
    for _ in range(200):
        x = rng.integers(0, O.P, size=2, dtype=np.uint64)
        r = O.hash_F(x, x)
        assert int(r[0]) <= bound and int(r[1]) <= bound


def test_hash_to_field_is_a_left_fold():
    e = O.synth(5, 0, 4)
    prev = np.zeros(2, np.uint64)
    out = O.hash_to_field(prev, e)
    acc = np.zeros(2, np.uint64)
    for x in e:
        acc = O.hash_F(acc, x)
    assert np.array_equal(out, acc) and np.array_equal(prev, acc)
