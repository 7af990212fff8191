"""_Whir_Open's prover work on the GPU (cosnark_whir_open_weights,
src/Whir.cpp:469-545) against the oracle: the record, total_r and the
final claim, at fold factors 2 to 16 (the reference's default whir_k is
8), against the codeword's own tree and against the all-zero-leaf tree
that aggregate_commitments leaves; and Whir_Open's multi-claim batching
through the drop-in is covered by tests/dropin."""
import numpy as np
import pytest

import cosnark
import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("logn,rate,whir_k,zero", [(12, 4, 8, False), (14, 4, 8, True), (16, 4, 8, False),
                                                   (16, 2, 2, False), (18, 4, 16, True), (20, 4, 8, False),
                                                   (13, 8, 4, True)])
def test_whir_open_weights_matches_oracle(ctx, logn, rate, whir_k, zero):
    n = 1 << logn
    seed = 0xC05A0060 + logn
    arr, b = O.synth(seed, 0, n), O.synth(seed, 1, n)
    K = O.inner_product(arr, b)
    print(f"whir open weights logn={logn} rate={rate} whir_k={whir_k} zero={zero}", flush=True)
    da, db = ctx.upload(arr), ctx.upload(b)
    cw, nodes = ctx.whir_commit(da, n, rate)
    leaves = n * rate // whir_k if zero else 0
    O.libc().srandom(seed)
    rc, rec, chal, fin = ctx.whir_open_weights(da, n, db, K, cw, n * rate, None if zero else nodes, leaves, rate,
                                               whir_k)
    assert rc == 0
    ref_cw = O.rs_encode(arr, rate)
    ref_nodes = None if zero else O.merkle_commit(ref_cw, True)
    if n >= 1 << 16:
        O.whir_threads(O.max_threads())
    try:
        O.libc().srandom(seed)
        w, orec, ochal, ofin = O.whir_open_weights(arr, b, K, ref_cw, ref_nodes, leaves, rate, whir_k)
    finally:
        O.whir_threads(1)
    assert w > 0
    assert rec.shape == orec.shape and np.array_equal(rec, orec)
    assert np.array_equal(chal, ochal) and np.array_equal(fin, ofin)
    # the caller's weights are read only
    assert np.array_equal(ctx.download(db, (n, 2)), b)
    for x in (da, db, cw, nodes):
        x.free()


def test_whir_open_weights_wrong_claim(ctx):
    n = 1 << 10
    arr, b = O.synth(7, 0, n), O.synth(7, 1, n)
    K = O.inner_product(arr, b)
    K[1] = (int(K[1]) + 1) % O.P
    da, db = ctx.upload(arr), ctx.upload(b)
    cw, nodes = ctx.whir_commit(da, n, 4)
    O.libc().srandom(3)
    rc, *_ = ctx.whir_open_weights(da, n, db, K, cw, 4 * n, nodes, 0, 4, 8)
    assert rc == cosnark.E_CHECK
