"""Product sumcheck on the GPU against the oracle
(_quadratic_sumcheck, src/Distributed_Sumcheck.cpp:204-348; tail rounds of
quadratic_sumcheck, src/sparse_eval.cpp:140-192)."""
import numpy as np
import pytest

import cosnark
import pyoracle as O

pytestmark = pytest.mark.gpu


def _run_both(ctx, m, seed, offset=0, libc_seed=None, prev0=(0, 0)):
    n = 1 << m
    f, g = O.synth(seed, 0, n), O.synth(seed, 1, n)
    y = O.inner_product(f, g)
    prev = np.array(prev0, np.uint64)
    if libc_seed is not None:
        O.libc().srandom(libc_seed)
    rc_o, polys_o, r_o, fin_o = O.product_sumcheck(y, f, g, prev, offset)
    ts = cosnark.Transcript()
    ts.prev[0], ts.prev[1] = prev0
    if libc_seed is not None:
        O.libc().srandom(libc_seed)
    rc_g, polys_g, r_g, fin_g = ctx.product_sumcheck(f, g, y, ts, offset)
    return (rc_o, polys_o, r_o, fin_o, prev), (rc_g, polys_g, r_g, fin_g, np.array(ts.value(), np.uint64))


@pytest.mark.parametrize("m", [1, 2, 3, 5, 8, 11, 12, 13, 14, 16, 18, 20, 21])
def test_sumcheck_matches_oracle(ctx, m):
    print(f"sumcheck m={m}", flush=True)
    o, g = _run_both(ctx, m, 0x5C00 + m)
    assert o[0] == 0 and g[0] == 0
    for a, b in zip(o[1:], g[1:]):
        assert np.array_equal(a, b)


# (14, 14) and (16, 15): rounds 0 and 1 draw libc challenges on the fused rounds-0-2 path
@pytest.mark.parametrize("m,offset", [(6, 4), (14, 4), (5, 5), (14, 14), (16, 15)])
def test_sumcheck_libc_tail_matches_oracle(ctx, m, offset):
    print(f"sumcheck m={m} offset={offset}", flush=True)
    o, g = _run_both(ctx, m, 0x5D00 + m, offset=offset, libc_seed=1234 + m)
    assert o[0] == 0 and g[0] == 0
    for a, b in zip(o[1:], g[1:]):
        assert np.array_equal(a, b)


def test_sumcheck_nonzero_transcript_start(ctx):
    o, g = _run_both(ctx, 10, 0x5E00, prev0=(12345, 678))
    for a, b in zip(o[1:], g[1:]):
        assert np.array_equal(a, b)


def test_sumcheck_wrong_claim_reports_check(ctx):
    n = 1 << 9
    f, g = O.synth(7, 0, n), O.synth(7, 1, n)
    y = O.inner_product(f, g)
    y[0] = (int(y[0]) + 1) % O.P
    ts = cosnark.Transcript()
    rc, *_ = ctx.product_sumcheck(f, g, y, ts, 0)
    assert rc == cosnark.E_CHECK
