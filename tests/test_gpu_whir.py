"""WHIR commit and open (src/Whir.cpp:158-467, whir_k = 4) on the GPU
against the oracle's record, bit for bit."""
import numpy as np
import pytest

import cosnark
import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("logn,rate", [(6, 4), (10, 4), (12, 2)])
def test_whir_commit_matches_oracle(ctx, logn, rate):
    n = 1 << logn
    arr = O.synth(0xC05A0003, 0, n)
    d = ctx.upload(arr)
    cw, nodes = ctx.whir_commit(d, n, rate)
    assert np.array_equal(ctx.download(cw, (n * rate, 2)), O.rs_encode(arr, rate))
    cnt = cosnark.merkle_node_count(n * rate)
    assert np.array_equal(ctx.download(nodes, (cnt, 32), np.uint8), O.merkle_commit(O.rs_encode(arr, rate), True))


@pytest.mark.parametrize("passbits,tile,logn", [(8, 4096, 14), (8, 2048, 14), (6, 4096, 14), (7, 4096, 19)])
def test_whir_commit_fused_tree_levels(ctx, passbits, tile, logn):
    """Codewords of 2^16 and 2^21 points under four NTT plans, the last pass
    writing the leaf digests in each: radix-256 passes with 16- and 8-column
    tiles; 6 + 5 + 5 bits, whose last pass narrows its 64 columns to 32; and
    7 + 7 + 7 bits, whose last pass narrows 128 columns to 16 (at most one
    step-2 item per thread).  Every node must match, since the dup-left root
    alone depends on leaf 0 only."""
    n, rate = 1 << logn, 4
    arr = O.synth(0xC05A0013, passbits, n)
    ctx.set_ntt_plan(logn + 2, passbits, tile)
    try:
        cw, nodes = ctx.whir_commit(ctx.upload(arr), n, rate)
    finally:
        ctx.set_ntt_plan(logn + 2, 0, 0)
    ref_cw = O.rs_encode(arr, rate)
    assert np.array_equal(ctx.download(cw, (n * rate, 2)), ref_cw)
    cnt = cosnark.merkle_node_count(n * rate)
    assert np.array_equal(ctx.download(nodes, (cnt, 32), np.uint8), O.merkle_commit(ref_cw, True))


def test_whir_claim_matches_oracle(ctx):
    n = 1 << 12
    arr, z = O.synth(3, 0, n), O.synth(3, 1, 12)
    assert np.array_equal(ctx.whir_claim(ctx.upload(arr), n, z), O.whir_claim(arr, z))


@pytest.mark.parametrize("logn,rate,stop", [(6, 4, 0), (10, 4, 0), (11, 2, 0), (12, 4, 8), (14, 4, 0), (16, 4, 8),
                                            (18, 4, 8), (18, 2, 8), (20, 4, 8)])
def test_whir_open_matches_oracle(ctx, logn, rate, stop):
    n = 1 << logn
    seed = 0xC05A0003 + logn
    arr, z = O.synth(seed, 0, n), O.synth(seed, 1, logn)
    K = O.whir_claim(arr, z)
    print(f"whir open logn={logn} rate={rate} stop={stop}", flush=True)
    O.libc().srandom(seed)
    rc_o, rec_o = O.whir_open(arr, z, K, rate, stop)
    assert rc_o == 0
    d = ctx.upload(arr)
    cw, nodes = ctx.whir_commit(d, n, rate)
    O.libc().srandom(seed)
    rc, rec = ctx.whir_open(d, n, z, K, rate, stop, cw, nodes)
    assert rc == 0
    assert rec.shape == rec_o.shape
    assert np.array_equal(rec, rec_o)


def test_whir_open_wrong_claim_reports_check(ctx):
    n = 1 << 10
    arr, z = O.synth(11, 0, n), O.synth(11, 1, 10)
    K = O.whir_claim(arr, z)
    K[0] = (int(K[0]) + 1) % O.P
    d = ctx.upload(arr)
    cw, nodes = ctx.whir_commit(d, n, 4)
    O.libc().srandom(1)
    rc, _ = ctx.whir_open(d, n, z, K, 4, 0, cw, nodes)
    assert rc == cosnark.E_CHECK
