"""The oracle's cubic block sumcheck (cubic_coSumcheck_phase1/2,
src/coSumcheck.cpp:9-148) and rank 0's share arithmetic (unpack_secret_share,
compute_secret_shares and sumcheck_aggr's combine, src/MPI_utils.cpp:2604-2785),
and the library's host-only share functions against the oracle.

Pins, each reached by a different route than the oracle's loops:
- phase 1: the interpolant of each block by the DFT definition (Python
  integers), evaluated by Horner at every 4d-th root; the reference's skip of
  blocks whose p1 part is zero, which for the cubic form changes the sum;
- phase 2: the fold is the block interpolant at r (Horner); zero blocks give 0;
- shares: values of a polynomial built by Horner at w_2N^(2j+1) unpack to its
  values at the N-th roots; compute_secret_shares' shares are the group
  interpolant's values at the odd 2N-th roots;
- aggregation: combining the ranks' block sums over shares gives the
  weighted sum, over the packed slots, of the block sums over the secrets
  (the property the protocol relies on), for the quadratic and the cubic form."""
import ctypes

import numpy as np
import pytest

import cosnark
import pyoracle as O

P = O.P


def fmul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def fadd(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def fsub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def fpow(a, e):
    r = (1, 0)
    while e:
        if e & 1:
            r = fmul(r, a)
        a = fmul(a, a)
        e >>= 1
    return r


def finv(a):
    return fpow(a, P * P - 2)


def tup(x):
    return (int(x[0]), int(x[1]))


def root(logn):
    return tup(O.root_of_unity(logn))


def horner(coeffs, x):
    acc = (0, 0)
    for c in coeffs[::-1]:
        acc = fadd(fmul(acc, x), c)
    return acc


def interp_coeffs(vals):
    """coefficients of the degree < n polynomial with vals at the n-th roots (DFT definition)"""
    n = len(vals)
    wi = finv(root(n.bit_length() - 1))
    ni = finv((n, 0))
    out = []
    for m in range(n):
        s = (0, 0)
        for k in range(n):
            s = fadd(s, fmul(tup(vals[k]), fpow(wi, k * m)))
        out.append(fmul(s, ni))
    return out


def arr(vals):
    return np.array([list(v) for v in vals], np.uint64)


def _vecs(seed, M, zero_blocks=(), d=1):
    v = [O.synth(seed, i, M) for i in range(4)]
    for b in zero_blocks:
        v[0][b * d:(b + 1) * d] = 0
    return v


@pytest.mark.parametrize("M,d,zero_blocks", [(16, 2, ()), (32, 4, (1, 5)), (64, 8, (0,)), (8, 8, ())])
def test_cubic_phase1_interpolant_closed_form(M, d, zero_blocks):
    p1, p2, p3, b1 = _vecs(1100 + d, M, zero_blocks, d)
    H = O.cosumcheck3_phase1(p1, p2, p3, b1, d)
    w = root((4 * d).bit_length() - 1)
    want = [(0, 0)] * (4 * d)
    for j in range(M // d):
        if j in zero_blocks:
            continue                    # src/coSumcheck.cpp:22-28
        c = [interp_coeffs(x[j * d:(j + 1) * d]) for x in (p1, p2, p3, b1)]
        for k in range(4 * d):
            x = fpow(w, k)
            e = [horner(ci, x) for ci in c]
            want[k] = fadd(want[k], fmul(fsub(fmul(e[0], e[1]), e[2]), e[3]))
    assert [tup(h) for h in H] == want


def test_cubic_phase1_skip_changes_the_sum():
    M, d = 64, 4
    p1, p2, p3, b1 = _vecs(1200, M, (3,), d)
    H = O.cosumcheck3_phase1(p1, p2, p3, b1, d)
    # even entries: pointwise over the live blocks only
    for i in range(d):
        s = (0, 0)
        for j in range(M // d):
            if j == 3:
                continue
            k = j * d + i
            s = fadd(s, fmul(fsub(fmul(tup(p1[k]), tup(p2[k])), tup(p3[k])), tup(b1[k])))
        assert tup(H[4 * i]) == s
    # counting the zero block would add -p3*b1 there
    k = 3 * d
    assert fmul(tup(p3[k]), tup(b1[k])) != (0, 0)


@pytest.mark.parametrize("M,d", [(32, 4), (64, 16), (16, 2)])
def test_cubic_phase2_is_block_interpolant_at_r(M, d):
    vecs = _vecs(1300 + d, M, (1,), d)
    r = tup(O.synth(1301, 0, 1)[0])
    L = O.lagrange_coeff(r, d)
    out = O.cosumcheck3_phase2(*vecs, d, L)
    for a in range(4):
        for j in range(M // d):
            blk = vecs[a][j * d:(j + 1) * d]
            want = (0, 0) if not blk.any() else horner(interp_coeffs(blk), r)
            assert tup(out[a][j]) == want


@pytest.mark.parametrize("n,kk", [(4, 2), (8, 8), (16, 4), (2, 1), (1, 1)])
def test_unpack_recovers_values_at_roots(n, kk):
    q = [tup(x) for x in O.synth(1400 + n, 0, n)]          # degree < n
    w2 = root(n.bit_length())                               # order 2n
    shares = arr([horner(q, fpow(w2, 2 * j + 1)) for j in range(n)])
    left, v = O.unpack_secret_share(shares, kk)
    wn = root(n.bit_length() - 1)
    vals = [horner(q, fpow(wn, t)) for t in range(n)]
    assert [tup(x) for x in left] == vals
    assert [tup(x) for x in v] == [vals[n * j // kk] for j in range(kk)]


@pytest.mark.parametrize("N,k,kk,pads", [(4, 1, 2, True), (8, 2, 4, True), (8, 2, 4, False), (16, 4, 8, True)])
def test_shares_are_group_interpolant_at_odd_roots(N, k, kk, pads):
    M = 8 * k
    v = O.synth(1500 + N, 0, M)
    pad = O.synth(1500 + N, 1, (M // k) * (kk - k)) if pads else None
    sh = O.compute_secret_shares(v, N, k, kk, pad)
    w2 = root((2 * N).bit_length() - 1)
    for i in range(M // k):
        buf = [tup(v[i * k + j]) for j in range(k)]
        buf += [tup(pad[i * (kk - k) + j]) if pads else (0, 0) for j in range(kk - k)]
        c = interp_coeffs(arr(buf))
        for j in range(N):
            assert tup(sh[j, i]) == horner(c, fpow(w2, 2 * j + 1))
    # unpacking any group's shares at kk = N/2 recovers its buffer
    if kk == N // 2:
        _, back = O.unpack_secret_share(sh[:, 0], kk)
        assert [tup(x) for x in back[:k]] == [tup(v[j]) for j in range(k)]


def _slot(v, k, j):
    return np.ascontiguousarray(v[j::k])


@pytest.mark.parametrize("N,M,d", [(4, 64, 4), (8, 128, 8), (16, 256, 4)])
def test_aggregated_quadratic_equals_sum_over_slots(N, M, d):
    k, kk = N // 4, N // 2
    v1, v2 = O.synth(1600 + N, 0, M * k), O.synth(1600 + N, 1, M * k)
    pads = [O.synth(1600 + N, 2 + t, M * (kk - k)) for t in range(2)]
    s1 = O.compute_secret_shares(v1, N, k, kk, pads[0])
    s2 = O.compute_secret_shares(v2, N, k, kk, pads[1])
    polys = np.stack([O.pcs_sumcheck_phase1(s1[j], s2[j], d) for j in range(N)])
    ones = np.tile(np.array([[1, 0]], np.uint64), (k, 1))
    agg = O.sumcheck_aggr(polys, ones)
    want = [(0, 0)] * (2 * d)
    for j in range(k):
        h = O.pcs_sumcheck_phase1(_slot(v1, k, j), _slot(v2, k, j), d)
        want = [fadd(a, tup(b)) for a, b in zip(want, h)]
    assert [tup(x) for x in agg] == want


@pytest.mark.parametrize("N,M,d", [(4, 64, 4), (8, 64, 2), (16, 128, 8)])
def test_aggregated_cubic_equals_weighted_sum_over_slots(N, M, d):
    k, kk = N // 4, N // 2
    secrets = [O.synth(1700 + N, i, M * k) for i in range(3)]
    shares = [O.compute_secret_shares(s, N, k, kk, O.synth(1700 + N, 10 + i, M * (kk - k)))
              for i, s in enumerate(secrets)]
    b1 = O.synth(1700 + N, 20, M)                           # public: the same on every rank
    b2 = O.synth(1700 + N, 21, k)
    polys = np.stack([O.cosumcheck3_phase1(shares[0][j], shares[1][j], shares[2][j], b1, d) for j in range(N)])
    agg = O.sumcheck_aggr(polys, b2)
    want = [(0, 0)] * (4 * d)
    for j in range(k):
        h = O.cosumcheck3_phase1(*[_slot(s, k, j) for s in secrets], b1, d)
        want = [fadd(a, fmul(tup(b2[j]), tup(x))) for a, x in zip(want, h)]
    assert [tup(x) for x in agg] == want


def test_aggregation_below_four_ranks_is_zero():
    polys = np.stack([O.synth(1800, j, 8) for j in range(2)])
    assert not O.sumcheck_aggr(polys, None).any()


# ---- the library's host-only functions (no GPU needed) ----------------------

def _u64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


@pytest.mark.parametrize("n,kk", [(8, 4), (16, 16), (1, 1), (64, 32)])
def test_library_share_unpack_matches_oracle(n, kk):
    s = O.synth(1900 + n, 0, n)
    left, v = O.unpack_secret_share(s, kk)
    mine = np.ascontiguousarray(s.copy())
    out = np.zeros((kk, 2), np.uint64)
    assert cosnark.lib().cosnark_share_unpack(_u64p(mine), n, kk, _u64p(out)) == 0
    assert np.array_equal(mine, left) and np.array_equal(out, v)


@pytest.mark.parametrize("N,ln", [(4, 8), (8, 64), (32, 4096), (2, 16), (1, 3)])
def test_library_sumcheck_aggr_matches_oracle(N, ln):
    polys = np.stack([O.synth(2000 + N, j, ln) for j in range(N)])
    b = O.synth(2001 + N, 0, max(N // 4, 1))
    want = O.sumcheck_aggr(polys, b)
    out = np.zeros((ln, 2), np.uint64)
    pp = np.ascontiguousarray(polys)
    assert cosnark.lib().cosnark_sumcheck_aggr(_u64p(pp), N, ln, _u64p(b), _u64p(out)) == 0
    assert np.array_equal(out, want)


def test_library_share_functions_reject_bad_sizes():
    x = np.zeros((8, 2), np.uint64)
    o = np.zeros((8, 2), np.uint64)
    assert cosnark.lib().cosnark_share_unpack(_u64p(x), 6, 2, _u64p(o)) != 0
    assert cosnark.lib().cosnark_share_unpack(_u64p(x), 4, 8, _u64p(o)) != 0
    assert cosnark.lib().cosnark_sumcheck_aggr(_u64p(x), 3, 2, _u64p(x), _u64p(o)) != 0
