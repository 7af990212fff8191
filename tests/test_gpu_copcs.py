"""coPCS rank-local kernels on the GPU against the oracle, bit for bit:
Encode (src/coPCS.cpp:131-147), distributed_MT (:152-183), the block
sumcheck's two phases (:547-587; src/coSumcheck.cpp:150-213) and
RS_multifold5 (src/coPCS_utils.cpp:386-437).  The commit test runs the
shape of a 2^22-element share: 17 rows of 2^18, blowup 4, rate_offset 64."""
import numpy as np
import pytest

import cosnark
import pyoracle as O

pytestmark = pytest.mark.gpu


def _rows(seed, nrows, L):
    return np.stack([O.synth(seed, i, L) for i in range(nrows)])


@pytest.mark.parametrize("nrows,logL,rate,lro", [(3, 4, 2, 0), (1, 1, 2, 0), (17, 10, 4, 6), (2, 14, 4, 6),
                                                 (32, 6, 8, 1)])
def test_encode_matches_oracle(ctx, nrows, logL, rate, lro):
    print(f"copcs encode nrows={nrows} logL={logL} rate={rate} lro={lro}", flush=True)
    rows = _rows(0xC05A0010 + logL, nrows, 1 << logL)
    assert np.array_equal(ctx.copcs_encode(rows, rate, lro), O.copcs_encode(rows, rate, lro))


@pytest.mark.parametrize("nrows,logLc", [(32, 4), (17, 12), (5, 0), (1, 10), (4, 1), (3, 16)])
def test_distributed_mt_matches_oracle(ctx, nrows, logLc):
    print(f"distributed_mt nrows={nrows} logLc={logLc}", flush=True)
    cw = _rows(0xC05A0011 + logLc, nrows, 1 << logLc)
    assert np.array_equal(ctx.distributed_mt(cw), O.distributed_mt(cw))


def test_distributed_mt_reads_only_rows_0_to_3(ctx):
    cw = _rows(0xC05A0012, 17, 1 << 8)
    a = ctx.distributed_mt(cw)
    cw2 = cw.copy()
    cw2[4:] = O.synth(99, 99, cw2[4:].size // 2).reshape(cw2[4:].shape)
    assert np.array_equal(ctx.distributed_mt(cw2), a)
    cw2[3, 17] = [1, 2]
    assert not np.array_equal(ctx.distributed_mt(cw2), a)


def test_pcs_commit_at_share_shape(ctx):
    """Encode + distributed_MT for a 2^22 share: 16 data rows + 1 mask row of
    2^18, blowup 4, rate_offset 64 -> Base_MT over 2^20 columns."""
    nrows, L, rate, lro = 17, 1 << 18, 4, 6
    print("pcs commit 17 x 2^18", flush=True)
    rows = _rows(0xC05A0002, nrows, L)
    d = ctx.upload(rows)
    cw = ctx.alloc(16 * nrows * L * rate)
    ctx._ck(cosnark.lib().cosnark_copcs_encode(ctx.h, d.ptr, nrows, L, rate, lro, cw.ptr))
    nodes = ctx.alloc(32 * (2 * L * rate - 1))
    ctx._ck(cosnark.lib().cosnark_distributed_mt(ctx.h, cw.ptr, nrows, L * rate, nodes.ptr))
    got_cw = ctx.download(cw, (nrows, L * rate, 2))
    got_nodes = ctx.download(nodes, (2 * L * rate - 1, 32), np.uint8)
    # rows 0..3 fully against the oracle, every row by the degree bound:
    # the codeword is P(w^-1 x) on the 2^20-th roots with deg P < 2^18
    want = O.copcs_encode(rows[:4], rate, lro)
    assert np.array_equal(got_cw[:4], want)
    for i in range(nrows):
        c = ctx.upload(got_cw[i])
        ctx._ck(cosnark.lib().cosnark_ntt(ctx.h, c.ptr, 20, 1))
        coeffs = ctx.download(c, (L * rate, 2))
        c.free()
        assert not coeffs[L:].any()
    assert np.array_equal(got_nodes, O.distributed_mt(want))
    for b in (d, cw, nodes):
        b.free()


@pytest.mark.parametrize("logM,d", [(1, 2), (6, 2), (10, 4), (18, 8), (16, 32), (5, 32), (12, 16), (22, 2),
                                    (12, 64), (14, 128), (13, 256), (8, 256), (18, 256), (11, 2048), (16, 1024)])
def test_phase1_matches_oracle(ctx, logM, d):
    print(f"cosumcheck phase1 logM={logM} d={d}", flush=True)
    M = 1 << logM
    p1, p2 = O.synth(0xC05A0013 + logM, d, M), O.synth(0xC05A0013 + logM, d + 1, M)
    assert np.array_equal(ctx.cosumcheck_phase1(p1, p2, d), O.pcs_sumcheck_phase1(p1, p2, d))


@pytest.mark.parametrize("logM,d", [(1, 2), (10, 4), (18, 8), (16, 32), (12, 64), (14, 128), (13, 4096),
                                    (13, 8192), (20, 2)])
def test_phase2_matches_oracle(ctx, logM, d):
    print(f"cosumcheck phase2 logM={logM} d={d}", flush=True)
    M = 1 << logM
    p1, p2 = O.synth(0xC05A0014 + logM, d, M), O.synth(0xC05A0014 + logM, d + 1, M)
    r = O.synth(0xC05A0014, 7, 1)[0]
    g1, g2 = ctx.cosumcheck_phase2(p1, p2, d, r)
    w1, w2 = O.pcs_sumcheck_phase2(p1, p2, d, r)
    assert np.array_equal(g1, w1) and np.array_equal(g2, w2)


def test_block_sumcheck_rounds_chain(ctx):
    """The PCS opening's rounds (src/coPCS.cpp:1017-1050) on device data:
    sum of even entries = claim, challenge by hash_to_field, next claim
    = sum_k L_k(c) poly[k] over 2d points, fold, repeat with degrees 8, 32, 32."""
    M = 1 << 18
    p1, p2 = O.synth(0xC05A0015, 0, M), O.synth(0xC05A0015, 1, M)
    K = O.inner_product(p1, p2)
    a, b = ctx.upload(p1), ctx.upload(p2)
    ts = cosnark.Transcript()
    prev = np.zeros(2, np.uint64)
    op1, op2 = p1.copy(), p2.copy()
    size = M
    for d in (8, 32, 32):
        poly = ctx.cosumcheck_phase1(None, None, d, dev=(a, b, size))
        assert np.array_equal(poly, O.pcs_sumcheck_phase1(op1, op2, d))
        ev = np.array([int(poly[0::2, 0].astype(object).sum() % O.P), int(poly[0::2, 1].astype(object).sum() % O.P)],
                      np.uint64)
        assert np.array_equal(ev, K)
        c = cosnark.hash_to_field(ts, poly)
        assert np.array_equal(c, O.hash_to_field(prev, poly))
        L2 = O.lagrange_coeff(c, 2 * d)
        K = O.inner_product(L2, poly)
        ctx.cosumcheck_phase2(None, None, d, c, dev=(a, b, size))
        op1, op2 = O.pcs_sumcheck_phase2(op1, op2, d, c)
        size //= d
        assert np.array_equal(ctx.download(a, (size, 2)), op1)
        assert np.array_equal(ctx.download(b, (size, 2)), op2)
    # after the last fold the claim is the inner product of what is left
    assert np.array_equal(K, O.inner_product(op1, op2))
    a.free(); b.free()


@pytest.mark.parametrize("logn,fd", [(1, 2), (8, 2), (10, 4), (18, 8), (20, 32), (5, 32), (12, 16), (14, 64),
                                     (12, 128), (13, 4096), (16, 1024)])
def test_multifold5_matches_oracle(ctx, logn, fd):
    print(f"rs_multifold5 logn={logn} fd={fd}", flush=True)
    n = 1 << logn
    cw, a = O.synth(0xC05A0016 + logn, fd, n), O.synth(0xC05A0016 + logn, fd + 1, fd)
    assert np.array_equal(ctx.rs_multifold5(cw, a), O.rs_multifold5(cw, a))


def test_multifold5_of_a_codeword_at_share_shape(ctx):
    """Folding an RS codeword (2^18 coefficients, blowup 4) by 32 gives the
    codeword of the folded polynomial: a size-independent check at 2^20."""
    m, rate, fd = 1 << 18, 4, 32
    f = O.synth(0xC05A0017, 0, m)
    a = O.synth(0xC05A0017, 1, fd)
    cw = ctx.rs_encode(f, rate)
    out = ctx.rs_multifold5(cw, a)
    g = np.zeros((m // fd, 2), np.uint64)
    for k in range(fd):
        g = O.vec_op("add", g, O.vec_op("mul", f[k::fd], np.tile(a[k], (m // fd, 1))))
    assert np.array_equal(out, ctx.rs_encode(g, rate))


def test_copcs_argument_errors(ctx):
    lib = cosnark.lib()
    buf = ctx.alloc(16 * 64)
    out = ctx.alloc(16 * 64)
    a = np.zeros((64, 2), np.uint64)
    poly = np.zeros((8192, 2), np.uint64)
    assert lib.cosnark_cosumcheck_phase1(ctx.h, buf.ptr, buf.ptr, 64, 3, cosnark._p(poly)) == cosnark.E_ARG
    assert lib.cosnark_cosumcheck_phase1(ctx.h, buf.ptr, buf.ptr, 1 << 13, 4096, cosnark._p(poly)) == cosnark.E_RANGE
    assert lib.cosnark_rs_multifold5(ctx.h, buf.ptr, 64, cosnark._p(a), 4, buf.ptr) == cosnark.E_ARG
    assert lib.cosnark_rs_multifold5(ctx.h, buf.ptr, 64, cosnark._p(a), 128, out.ptr) == cosnark.E_ARG
    assert lib.cosnark_distributed_mt(ctx.h, buf.ptr, 33, 2, out.ptr) == cosnark.E_ARG
    assert lib.cosnark_distributed_mt(ctx.h, buf.ptr, 2, 3, out.ptr) == cosnark.E_ARG
    assert lib.cosnark_copcs_encode(ctx.h, buf.ptr, 1, 1, 2, 0, out.ptr) == cosnark.E_ARG
    buf.free(); out.free()


@pytest.mark.parametrize("logn,rate0,rate1,degree", [(10, 4, 4, 8), (12, 4, 4, 32), (6, 2, 2, 2), (8, 4, 8, 512),
                                                     (3, 1, 2, 1), (18, 4, 4, 8)])
def test_encode_concatenated_matches_oracle(ctx, logn, rate0, rate1, degree):
    print(f"encode_concatenated logn={logn} rates=({rate0},{rate1}) degree={degree}", flush=True)
    data = O.synth(0xC05A0018 + logn, degree, 1 << logn)
    gcp, gC = ctx.encode_concatenated(data, rate0, rate1, degree)
    ocp, oC = O.encode_concatenated(data, rate0, rate1, degree)
    assert np.array_equal(gcp, ocp) and np.array_equal(gC, oC)
    if logn == 10:
        # commit_concatenated: the tree over the rows back to back
        nodes = ctx.merkle_commit(gC.reshape(-1, 2), True)
        assert np.array_equal(nodes, O.merkle_commit(oC.reshape(-1, 2), True))


@pytest.mark.parametrize("logn,fd,l", [(10, 8, 50), (20, 32, 333), (16, 64, 100), (14, 128, 40), (12, 2, 7)])
def test_multifold_segments_match_oracle(ctx, logn, fd, l):
    """delegate_checks (src/coPCS.cpp:740-760): gathered segments and their
    RS_multifold_segment folds at query positions."""
    print(f"multifold_segments logn={logn} fd={fd} l={l}", flush=True)
    n = 1 << logn
    cw, L = O.synth(0xC05A0019 + logn, fd, n), O.synth(0xC05A0019 + logn, fd + 1, fd)
    idx = (O.synth(0xC05A0019 + logn, fd + 2, l)[:, 0] % (n // fd)).astype(np.uint32)
    seg, y = ctx.multifold_segments(cw, L, idx)
    for q in range(l):
        want_seg = cw[int(idx[q])::n // fd][:fd]
        assert np.array_equal(seg[q], want_seg)
        if fd <= 64:
            assert np.array_equal(y[q], O.rs_multifold_segment(want_seg, L, n, int(idx[q])))
    if n <= 1 << 16:
        full = O.rs_multifold5(cw, L)
        assert np.array_equal(y, full[idx.astype(np.int64)])


@pytest.mark.parametrize("nrows,logl", [(16, 20), (1, 10), (31, 12)])
def test_row_combine_matches_oracle(ctx, nrows, logl):
    rows = np.stack([O.synth(0xC05A001A + logl, i, 1 << logl) for i in range(nrows)])
    c = O.synth(0xC05A001A, 99, nrows)
    assert np.array_equal(ctx.row_combine(rows, c), O.row_combine(rows, c))
