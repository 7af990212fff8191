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
    7 + 7 + 7 bits, whose last pass narrows 32 columns to 16 (at most one
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


@pytest.mark.parametrize("logn,rate,passbits", [(14, 4, 7), (14, 8, 8), (16, 4, 8), (16, 8, 8), (15, 2, 5)])
def test_whir_commit_fan_out_with_leaf_digests(logn, rate, passbits):
    """The fan-out first pass (padding by the rate folded into the first
    digit), the last pass narrowing its tile and the leaf digests written by
    that last pass, all in one encode: a context whose fan-out threshold is
    one workgroup (COSNARK_NTT_FAN_MIN=1) takes the fan-out at these small
    sizes, as WHIR's 2^26 and 2^25 encodes do at full size.  Each case's
    plan saves a pass by fanning out: 16 bits at 7 per pass as 9 + 7, 17 at 8
    as 10 + 7, 18 as 10 + 8, 19 as 11 + 8, 16 at 5 as 6 + 5 + 5; the 7-bit
    last passes narrow 32 columns to 16, the 5-bit one 64 to 32.  Every codeword element and tree node
    against the oracle."""
    import os
    n = 1 << logn
    arr = O.synth(0xC05A0014, logn + rate, n)
    old = os.environ.get("COSNARK_NTT_FAN_MIN")
    os.environ["COSNARK_NTT_FAN_MIN"] = "1"
    try:
        c2 = cosnark.Context(0)
    finally:
        if old is None:
            del os.environ["COSNARK_NTT_FAN_MIN"]
        else:
            os.environ["COSNARK_NTT_FAN_MIN"] = old
    try:
        print(f"fan-out commit logn={logn} rate={rate} passbits={passbits}", flush=True)
        c2.set_ntt_plan(logn + rate.bit_length() - 1, passbits, 4096)
        cw, nodes = c2.whir_commit(c2.upload(arr), n, rate)
        ref_cw = O.rs_encode(arr, rate)
        assert np.array_equal(c2.download(cw, (n * rate, 2)), ref_cw)
        cnt = cosnark.merkle_node_count(n * rate)
        assert np.array_equal(c2.download(nodes, (cnt, 32), np.uint8), O.merkle_commit(ref_cw, True))
    finally:
        c2.close()


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


def test_whir_open_2_24_matches_oracle(ctx):
    """The WHIR workload at full size: 2^24 evaluations, blowup 4, whir_k 4,
    folded to 2^8 (8 iterations), the bench's seed; commit and opening
    record bit for bit against the oracle (its encodes, trees and weight
    updates on all cores, the same bytes as single-thread)."""
    logn, rate, stop = 24, 4, 8
    n = 1 << logn
    seed = 0xC05A0003
    arr, z = O.synth(seed, 0, n), O.synth(seed, 1, logn)
    d = ctx.upload(arr)
    K = ctx.whir_claim(d, n, z)
    print("whir 2^24: device commit and open", flush=True)
    cw, nodes = ctx.whir_commit(d, n, rate)
    O.libc().srandom(seed)
    rc, rec = ctx.whir_open(d, n, z, K, rate, stop, cw, nodes)
    assert rc == 0
    print("whir 2^24: oracle", flush=True)
    O.whir_threads(O.max_threads())
    try:
        O.libc().srandom(seed)
        rc_o, rec_o = O.whir_open(arr, z, K, rate, stop)
    finally:
        O.whir_threads(1)
    assert rc_o == 0
    assert rec.shape == rec_o.shape
    assert np.array_equal(rec, rec_o)
    # the claim is <arr, eq(z)>: the oracle's own value
    assert np.array_equal(K, O.whir_claim(arr, z))
    for b in (d, cw, nodes):
        b.free()
