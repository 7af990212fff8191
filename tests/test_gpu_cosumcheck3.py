"""The cubic block sumcheck (cubic_coSumcheck_phase1/2, src/coSumcheck.cpp:
9-148) on the GPU against the oracle, bit for bit, including blocks whose p1
part is zero (skipped by the reference); its rounds chained as
cubic_coSumcheck runs them (:355-425); and rank 0's aggregation
(src/MPI_utils.cpp:2604-2785) of per-rank GPU block sums over packed shares.
At 2^22 elements: the even entries against the pointwise closed form, and
additivity over the two halves of the blocks."""
import numpy as np
import pytest

import cosnark
import pyoracle as O

pytestmark = pytest.mark.gpu


def _vecs(seed, M, d, zero_blocks=()):
    v = [O.synth(seed, i, M) for i in range(4)]
    for b in zero_blocks:
        v[0][b * d:(b + 1) * d] = 0
    return v


@pytest.mark.parametrize("logM,d,zb", [(1, 2, ()), (2, 2, (0,)), (6, 2, (3,)), (10, 4, (0, 7)), (12, 8, ()),
                                       (14, 16, (5,)), (12, 32, ()), (5, 32, (0,)), (13, 64, (2,)), (14, 128, ()),
                                       (15, 256, (1,)), (16, 512, (0, 3)), (9, 512, ()), (20, 512, (1, 1025, 2047)), (10, 1024, ()),
                                       (18, 1024, (100,)), (18, 16, ()), (5, 16, (1,)), (22, 16, (9,))])
def test_cubic_phase1_matches_oracle(ctx, logM, d, zb):
    print(f"cosumcheck3 phase1 logM={logM} d={d} zero_blocks={zb}", flush=True)
    M = 1 << logM
    v = _vecs(0xC05A0030 + logM, M, d, zb)
    assert np.array_equal(ctx.cosumcheck3_phase1(*v, d), O.cosumcheck3_phase1(*v, d))


@pytest.mark.parametrize("logM,d", [(1, 2), (10, 4), (14, 8), (12, 64), (14, 128), (13, 1024), (13, 8192), (20, 2)])
def test_cubic_fold_matches_oracle(ctx, logM, d):
    print(f"cosumcheck3 fold logM={logM} d={d}", flush=True)
    M = 1 << logM
    v = _vecs(0xC05A0031 + logM, M, d, (0,) if M // d > 1 else ())
    r = O.synth(0xC05A0031, 9, 1)[0]
    L = O.lagrange_coeff(r, d)
    want = O.cosumcheck3_phase2(*v, d, L)
    for got in (ctx.cosumcheck3_fold(*v, d, r=r), ctx.cosumcheck3_fold(*v, d, L=L)):
        for g, w in zip(got, want):
            assert np.array_equal(g, w)


def _sum_even(H, d):
    return np.array([int(H[0::4, 0].astype(object).sum() % O.P), int(H[0::4, 1].astype(object).sum() % O.P)],
                    np.uint64)


def test_cubic_rounds_chain(ctx):
    """Rounds over degrees 16, 8, 2 on device data: the even entries of each
    round polynomial sum to the claim, the challenge is hash_to_field of the
    polynomial, the next claim is its interpolant at the challenge (the
    4d-point Lagrange weights), and the fold leaves each block's interpolant
    at the challenge."""
    M = 1 << 14
    v = _vecs(0xC05A0032, M, 1)
    bufs = [ctx.upload(x) for x in v]
    ov = [x.copy() for x in v]
    K = None
    ts = cosnark.Transcript()
    prev = np.zeros(2, np.uint64)
    size = M
    for d in (16, 8, 2):
        H = ctx.cosumcheck3_phase1(None, None, None, None, d, dev=(bufs, size))
        assert np.array_equal(H, O.cosumcheck3_phase1(*ov, d))
        if K is not None:
            assert np.array_equal(_sum_even(H, d), K)
        r = cosnark.hash_to_field(ts, H)
        assert np.array_equal(r, O.hash_to_field(prev, H))     # prev is updated in place
        K = O.inner_product(O.lagrange_coeff(r, 4 * d), H)
        ctx.cosumcheck3_fold(None, None, None, None, d, r=r, dev=(bufs, size))
        ov = O.cosumcheck3_phase2(*ov, d, O.lagrange_coeff(r, d))
        size //= d
        for b, w in zip(bufs, ov):
            assert np.array_equal(ctx.download(b, (size, 2)), w)
    # the last claim is the sum of (p1 p2 - p3) b1 over what is left
    t = O.vec_op("mul", O.vec_op("sub", O.vec_op("mul", ov[0], ov[1]), ov[2]), ov[3])
    assert np.array_equal(K, np.array([int(t[:, 0].astype(object).sum() % O.P),
                                       int(t[:, 1].astype(object).sum() % O.P)], np.uint64))
    for b in bufs:
        b.free()


@pytest.mark.parametrize("N,M,d,cubic", [(8, 1 << 12, 16, True), (4, 1 << 10, 8, True), (16, 1 << 11, 4, False)])
def test_aggregated_rank_sums_equal_slot_sums(ctx, N, M, d, cubic):
    """N ranks' GPU block sums over packed shares (k = N/4 secrets per
    polynomial of degree < N/2), combined by the library's rank-0 aggregation,
    equal the weighted sum over the slots of the oracle's block sums over the
    secrets."""
    print(f"aggregation N={N} M={M} d={d} cubic={cubic}", flush=True)
    k, kk = N // 4, N // 2
    ns = 3 if cubic else 2
    secrets = [O.synth(0xC05A0033 + N, i, M * k) for i in range(ns)]
    shares = [O.compute_secret_shares(s, N, k, kk, O.synth(0xC05A0033 + N, 10 + i, M * (kk - k)))
              for i, s in enumerate(secrets)]
    if cubic:
        b1 = O.synth(0xC05A0033 + N, 20, M)
        b2 = O.synth(0xC05A0033 + N, 21, k)
        polys = np.stack([ctx.cosumcheck3_phase1(shares[0][j], shares[1][j], shares[2][j], b1, d) for j in range(N)])
    else:
        b2 = np.tile(np.array([[1, 0]], np.uint64), (k, 1))
        polys = np.stack([ctx.cosumcheck_phase1(shares[0][j], shares[1][j], d) for j in range(N)])
    agg = cosnark.sumcheck_aggr(polys, b2)
    assert np.array_equal(agg, O.sumcheck_aggr(polys, b2))
    want = np.zeros_like(agg)
    for j in range(k):
        sl = [np.ascontiguousarray(s[j::k]) for s in secrets]
        h = O.cosumcheck3_phase1(*sl, b1, d) if cubic else O.pcs_sumcheck_phase1(*sl, d)
        want = O.vec_op("add", want, O.vec_op("mul", h, np.tile(b2[j], (h.shape[0], 1))))
    assert np.array_equal(agg, want)


def _colsum(x):
    """column sums mod p of a (rows, cols) u64 array, exactly"""
    lo = np.sum(x & np.uint64(0xFFFFFFFF), axis=0, dtype=np.uint64)
    hi = np.sum(x >> np.uint64(32), axis=0, dtype=np.uint64)
    return [((int(h) << 32) + int(l)) % O.P for h, l in zip(hi, lo)]


def test_cubic_phase1_at_2_22(ctx):
    M, d = 1 << 22, 512
    print("cosumcheck3 phase1 2^22 d=512", flush=True)
    v = _vecs(0xC05A0034, M, d, (7, 4000))
    H = ctx.cosumcheck3_phase1(*v, d)
    # even entries: pointwise over the live blocks
    t = O.vec_op("mul", O.vec_op("sub", O.vec_op("mul", v[0], v[1]), v[2]), v[3]).reshape(M // d, d, 2)
    t[[7, 4000]] = 0
    assert [int(x) for x in H[0::4, 0]] == _colsum(t[:, :, 0])
    assert [int(x) for x in H[0::4, 1]] == _colsum(t[:, :, 1])
    # additivity over the two halves of the blocks
    h = M // 2
    H1 = ctx.cosumcheck3_phase1(*[x[:h] for x in v], d)
    H2 = ctx.cosumcheck3_phase1(*[x[h:] for x in v], d)
    assert np.array_equal(O.vec_op("add", H1, H2), H)
