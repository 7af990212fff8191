"""PCS_commit at one rank on the GPU (cosnark_pcs_commit: prepare_data,
Encode, distributed_MT; src/coPCS.cpp:76-250) against the oracle: the
rows, every codeword row and every Base_MT node, at small shares and at
the 2^22 share (q0 = 333, rate 4, rate_offset 64: 18 rows of 2^18 -> 2^20),
where rows 0, 1, 16, 17 and the whole tree are compared (a dup-left column
root depends on rows 0..3 only, so the tree is the oracle's over those)."""
import numpy as np
import pytest

import cosnark
import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("logn,q0,rate,lro", [(10, 3, 4, 6), (14, 100, 4, 6), (16, 333, 2, 4), (12, 30, 8, 0)])
def test_pcs_commit_matches_oracle(ctx, logn, q0, rate, lro):
    n = 1 << logn
    B = n // 16
    nr = 32 * q0 + B
    shares, rnd = O.synth(0xC05A00B0, logn, n), O.synth(0xC05A00B1, logn, nr)
    print(f"pcs_commit n=2^{logn} q0={q0} rate={rate}", flush=True)
    c, rows, cw, nodes = ctx.pcs_commit(ctx.upload(shares), n, ctx.upload(rnd), nr, q0, rate, lro)
    co, rows_o, cw_o, nodes_o = O.pcs_commit(shares, rnd, q0, rate, lro)
    assert c == co
    assert np.array_equal(rows, rows_o) and np.array_equal(cw, cw_o) and np.array_equal(nodes, nodes_o)


def test_pcs_commit_share_2_22(ctx):
    n, q0, rate, lro = 1 << 22, 333, 4, 6
    B = n // 16
    nr = 18 * q0 + B
    shares, rnd = O.synth(0xC05A00B2, 0, n), O.synth(0xC05A00B2, 1, nr)
    print("pcs_commit 2^22 share", flush=True)
    ds, dr = ctx.upload(shares), ctx.upload(rnd)
    c, rows_d, cws_d, nodes_d = ctx.pcs_commit(ds, n, dr, nr, q0, rate, lro, download=False)
    assert c == 18
    co, rows_o = O.prepare_data(shares, rnd, q0)
    assert co == 18
    rows = ctx.download(rows_d, (c, B, 2))
    assert np.array_equal(rows, rows_o[:c])
    cws = ctx.download(cws_d, (c, B * rate, 2))
    pick = [0, 1, 2, 3, 16, 17]
    cw_o = O.copcs_encode(rows_o[pick], rate, lro)
    for j, i in enumerate(pick):
        assert np.array_equal(cws[i], cw_o[j]), f"codeword row {i}"
    nodes = ctx.download(nodes_d, (2 * B * rate - 1, 32), np.uint8)
    assert np.array_equal(nodes, O.distributed_mt(cw_o[:4]))
    for x in (ds, dr, rows_d, cws_d, nodes_d):
        x.free()
