"""Batched WHIR (Whir_Batch_Open, src/Whir.cpp:706-899) on the GPU against the
oracle: the whole opening record, bit for bit, for fold factors 2, 4 and 8,
aggregations of equal and smaller arrays (including one at the first step
with kappa = 1 and one at the last scheduled step), and the error behaviour
where the reference reports a failed check."""
import numpy as np
import pytest

import cosnark
import pyoracle as O

pytestmark = pytest.mark.gpu


def _inputs(seed, logs, rate, whir_k=4):
    arrays = [O.synth(seed, 2 * i, 1 << lg) for i, lg in enumerate(logs)]
    zs = [O.synth(seed, 2 * i + 1, lg) for i, lg in enumerate(logs)]
    return O.whir_batch_inputs(arrays, zs, rate, whir_k)


def _gpu(ctx, inp):
    return ctx.whir_batch_open(inp["arrays"], inp["codewords"], inp["trees"], inp["evals"], inp["betas"],
                               inp["rate"], inp["whir_k"])


@pytest.mark.parametrize("logs,rate,whir_k", [([10], 4, 4), ([10, 10], 4, 4), ([12, 8, 6, 6], 4, 4),
                                              ([12, 12, 8, 6, 6], 2, 4), ([12, 9], 4, 8), ([8, 6], 4, 2),
                                              ([8, 2], 4, 4), ([11, 7, 5], 2, 4), ([13, 10, 10, 7], 2, 8),
                                              ([16, 14, 12], 4, 4), ([18, 16], 4, 4), ([19, 12], 4, 8),
                                              ([17, 10], 4, 2)])
def test_batch_open_matches_oracle(ctx, logs, rate, whir_k):
    print(f"whir batch open logs={logs} rate={rate} whir_k={whir_k}", flush=True)
    inp = _inputs(0xC05A0050 + sum(logs), logs, rate, whir_k)
    O.libc().srandom(61 + len(logs))
    rc_o, want = O.whir_batch_open(inp)
    assert rc_o == 0
    O.libc().srandom(61 + len(logs))
    rc, got = _gpu(ctx, inp)
    assert rc == 0, cosnark.lib().cosnark_last_error(ctx.h)
    assert got.shape == want.shape and np.array_equal(got, want)


def test_batch_open_rates_from_library(ctx):
    """Codewords committed on the GPU at cosnark_whir_rates' rates."""
    logs, rate = [12, 10, 8, 8], 4
    inp = _inputs(0xC05A0051, logs, rate)
    rates = cosnark.whir_rates(inp["sizes"], rate)
    assert rates == inp["rates"]
    cws, trees = [], []
    for a, r in zip(inp["arrays"], rates):
        d = ctx.upload(a)
        n = a.shape[0]
        cw = ctx.alloc(16 * n * r)
        nd = ctx.alloc(32 * cosnark.merkle_node_count(n * r))
        ctx._ck(cosnark.lib().cosnark_whir_commit(ctx.h, d.ptr, n, r, cw.ptr, nd.ptr))
        cws.append(ctx.download(cw, (n * r, 2)))
        trees.append(ctx.download(nd, (cosnark.merkle_node_count(n * r), 32), np.uint8))
        d.free(); cw.free(); nd.free()
    for c, t, wc, wt in zip(cws, trees, inp["codewords"], inp["trees"]):
        assert np.array_equal(c, wc) and np.array_equal(t, wt)
    O.libc().srandom(5)
    rc_o, want = O.whir_batch_open(inp)
    O.libc().srandom(5)
    rc, got = ctx.whir_batch_open(inp["arrays"], cws, trees, inp["evals"], inp["betas"], rate)
    assert rc_o == 0 and rc == 0 and np.array_equal(got, want)


def test_batch_open_wrong_claim_reports_check(ctx):
    inp = _inputs(0xC05A0052, [10, 8], 4)
    inp["evals"][0, 0] = (int(inp["evals"][0, 0]) + 1) % O.P
    O.libc().srandom(3)
    rc, _ = _gpu(ctx, inp)
    assert rc == cosnark.E_CHECK


def test_batch_open_rejects_unsorted(ctx):
    inp = _inputs(0xC05A0053, [8, 10], 4)
    O.libc().srandom(3)
    rc, _ = _gpu(ctx, inp)
    assert rc == cosnark.E_ARG


@pytest.mark.parametrize("narr,npts,logn", [(1, 1, 6), (3, 2, 12), (4, 8, 16)])
def test_aggregate_claims_matches_oracle(ctx, narr, npts, logn):
    print(f"aggregate_claims narr={narr} npts={npts} logn={logn}", flush=True)
    n = 1 << logn
    arrays = [O.synth(0xC05A0054 + logn, i, n) for i in range(narr)]
    points = np.stack([O.synth(0xC05A0055 + logn, j, logn) for j in range(npts)])
    O.libc().srandom(23)
    want = O.whir_aggregate_claims(arrays, points)
    O.libc().srandom(23)
    got = ctx.whir_aggregate_claims(arrays, points)
    for g, w in zip(got, want):
        assert np.array_equal(np.asarray(g, np.uint64).reshape(-1), np.asarray(w, np.uint64).reshape(-1))


@pytest.mark.parametrize("m", [1, 5, 12, 20])
def test_precompute_beta_matches_oracle(ctx, m):
    z = O.synth(0xC05A0056, m, m)
    out = ctx.alloc(16 << m)
    ctx._ck(cosnark.lib().cosnark_precompute_beta(ctx.h, cosnark._p(np.ascontiguousarray(z)), m, out.ptr))
    got = ctx.download(out, (1 << m, 2))
    out.free()
    assert np.array_equal(got, O.precompute_beta(z))
