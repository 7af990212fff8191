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


@pytest.mark.parametrize("block_log,reg_log", [(8, 0), (9, 1), (9, 0), (6, 0), (10, 2)])
def test_merkle_register_levels_match_oracle(block_log, reg_log):
    """The Merkle schedule's register levels (COSNARK_MK_REG_LOG: 2^RL
    first-level nodes per lane, at least block_log - 8): trees from the
    elements (both parent forms) and from level 2 (a WHIR commit and the
    plain prover's commitment, levels 1 and 2 from the encode) against the
    oracle."""
    import os
    import cosnark
    env = {"COSNARK_MK_BLOCK_LOG": str(block_log), "COSNARK_MK_REG_LOG": str(reg_log),
           "COSNARK_NTT_LEAF_LEVELS": "3"}
    os.environ.update(env)
    try:
        c = cosnark.Context(0)
    finally:
        for k in env:
            del os.environ[k]
    try:
        print(f"merkle register levels b={block_log} rl={reg_log}", flush=True)
        for dup in (True, False):
            e = O.synth(0xE9 + block_log, reg_log, 1 << 18)
            assert np.array_equal(c.merkle_commit(e, dup), O.merkle_commit(e, dup))
        a = O.synth(0xEA, block_log, 1 << 16)
        cw, nodes = c.whir_commit(c.upload(a), 1 << 16, 4)
        ref_cw = O.rs_encode(a, 4)
        got = c.download(nodes, (cosnark.merkle_node_count(4 << 16), 32), np.uint8)
        assert np.array_equal(got, O.merkle_commit(ref_cw, True))
        f, g = O.synth(0xEB, 0, 1 << 16), O.synth(0xEB, 1, 1 << 16)
        df, dg = c.upload(f), c.upload(g)
        proof, cwb, ndb = c.prove_plain_keep(df, dg, 1 << 16, 4, cosnark.Transcript())
        got = c.download(ndb, (cosnark.merkle_node_count(4 << 16), 32), np.uint8)
        assert np.array_equal(got, O.merkle_commit(O.rs_encode(f, 4), True))
    finally:
        c.close()
