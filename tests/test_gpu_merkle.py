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


@pytest.mark.parametrize("block_log", [6, 7, 8, 9, 10])
def test_merkle_schedules_build_the_same_tree(block_log):
    """Every Merkle schedule (2^b first-level nodes per workgroup) builds the
    oracle's tree, at sizes where the tree takes one, two and three launches."""
    import cosnark
    c = cosnark.Context(0)
    c.set_merkle_schedule(block_log)
    assert c.merkle_schedule() == block_log
    for log_elems, dup_left in ((8, True), (14, True), (18, False), (22, True)):
        e = O.synth(0xE9 + log_elems, block_log, 1 << log_elems)
        print(f"merkle schedule b={block_log} log_elems={log_elems}", flush=True)
        got = c.merkle_commit(e, dup_left)
        exp = O.merkle_commit_mt(e, dup_left) if log_elems > 18 else O.merkle_commit(e, dup_left)
        assert np.array_equal(got, exp), f"b={block_log} log_elems={log_elems}"
    c.close()


def test_merkle_autotune_keeps_a_valid_schedule():
    import cosnark
    c = cosnark.Context(0)
    b, ms = c.autotune_merkle([1 << 12, 1 << 14, 1 << 16, 1 << 18])
    assert 6 <= b <= 10 and ms > 0 and c.merkle_schedule() == b
    print(f"tuned merkle schedule b={b} ({ms:.4f} ms)", flush=True)
    e = O.synth(0xEF, 0, 1 << 14)
    assert np.array_equal(c.merkle_commit(e, True), O.merkle_commit(e, True))
    for bad in (5, 11):
        with pytest.raises(RuntimeError):
            c.set_merkle_schedule(bad)
    with pytest.raises(RuntimeError):
        c.autotune_merkle([3])
    assert c.merkle_schedule() == b
    c.close()
