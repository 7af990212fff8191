"""BLAKE3 Merkle commit / open on the GPU against the oracle
(MT_commit_Blake, create_tree_blake, open_tree_blake:
src/merkle_tree.cpp:195-233, 267-299, 320-335)."""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dup_left", [True, False])
@pytest.mark.parametrize("log_elems", [3, 4, 6, 10, 12, 13, 14, 16, 18])
def test_merkle_matches_oracle(ctx, log_elems, dup_left):
    e = O.synth(0xE0 + log_elems, 0, 1 << log_elems)
    print(f"merkle log_elems={log_elems} dup_left={dup_left}", flush=True)
    got = ctx.merkle_commit(e, dup_left)
    exp = O.merkle_commit(e, dup_left)
    assert np.array_equal(got, exp)


def test_merkle_open_matches_oracle(ctx):
    n = 1 << 12
    e = O.synth(0xE7, 0, n)
    host, dev = ctx.merkle_commit(e, True, keep_device=True)
    for pos in [0, 1, 2, 511, 777, n // 4 - 1]:
        assert np.array_equal(ctx.merkle_open(dev, n, pos), O.merkle_open(host, n, pos))
