"""The oracle's coPCS rank-local kernels: Encode (src/coPCS.cpp:131-147),
distributed_MT (:152-183), pcs_sumcheck_phase1/2 (:547-587) and
RS_multifold5 (src/coPCS_utils.cpp:386-437).

Each is pinned by a closed form reached through a different route than the
oracle's own loops: Horner evaluation of the interpolant (Encode, phase 1/2),
pointwise products at the d-th roots (phase 1, even entries), the Lagrange
basis property (phase 2 at r = w^t), a strided sub-transform of the
coefficients (RS_multifold5) and BLAKE3 of the first leaf block (the
dup-left column tree of distributed_MT).  fft itself is pinned by the DFT
definition in test_oracle_ntt.py."""
import numpy as np
import pytest

import pyoracle as O

P = O.P


def fmul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def fadd(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def fpow(a, e):
    r = (1, 0)
    while e:
        if e & 1:
            r = fmul(r, a)
        a = fmul(a, a)
        e >>= 1
    return r


def tup(x):
    return (int(x[0]), int(x[1]))


def horner(coeffs, x):
    acc = (0, 0)
    for c in coeffs[::-1]:
        acc = fadd(fmul(acc, x), tup(c))
    return acc


def root(logn):
    return tup(O.root_of_unity(logn))


def finv(a):
    return tup(O.vec_op("inv", np.array([a], np.uint64))[0])


@pytest.mark.parametrize("nrows,logL,rate,lro", [(3, 4, 2, 0), (2, 5, 4, 0), (1, 3, 8, 2), (32, 2, 2, 1)])
def test_encode_evaluates_shifted_interpolant(nrows, logL, rate, lro):
    L = 1 << logL
    rows = np.stack([O.synth(700 + logL, i, L) for i in range(nrows)])
    out = O.copcs_encode(rows, rate, lro)
    assert out.shape == (nrows, rate * L, 2)
    shift = finv(root(1 + lro + logL))
    zeta = root((rate * L).bit_length() - 1)
    for i in range(nrows):
        coeffs = O.fft(rows[i], inverse=True)
        for k in range(rate * L):
            x = fmul(shift, fpow(zeta, k))
            assert tup(out[i, k]) == horner(coeffs, x)


def _h2(d):
    return O.blake3_64(d + d)


@pytest.mark.parametrize("nrows,logLc", [(32, 4), (5, 3), (1, 5), (4, 0)])
def test_distributed_mt_closed_form(nrows, logLc):
    Lc = 1 << logLc
    cw = np.stack([O.synth(800 + logLc, i, Lc) for i in range(nrows)])
    nodes = O.distributed_mt(cw)
    full = np.zeros((32, Lc, 2), np.uint64)
    full[:nrows] = cw
    # column root: a 32-element dup-left tree keeps only leaf 0 (rows 0..3)
    # and three self-hash levels above it
    for j in range(Lc):
        d = O.blake3_64(full[0:4, j].tobytes())
        for _ in range(3):
            d = _h2(d)
        assert nodes[j].tobytes() == d
        assert O.merkle_commit(full[:, j], True)[-1].tobytes() == d
    # the tree over the column roots: level l node i is the l-fold chain of
    # column root 2^l * i
    off, size, lvl = Lc, Lc // 2, 1
    while size >= 1:
        for i in range(size):
            d = nodes[(1 << lvl) * i].tobytes()
            for _ in range(lvl):
                d = _h2(d)
            assert nodes[off + i].tobytes() == d
        off += size
        size //= 2
        lvl += 1
    assert off == nodes.shape[0]
    # rows beyond nrows are zeros, exactly as an explicit zero row
    assert np.array_equal(O.distributed_mt(full), nodes)


@pytest.mark.parametrize("logM,logd", [(6, 3), (5, 1), (8, 5), (4, 4)])
def test_phase1_products_of_interpolants(logM, logd):
    M, d = 1 << logM, 1 << logd
    data = O.synth(900 + logM, 0, M)
    B = O.synth(900 + logM, 1, M)
    poly = O.pcs_sumcheck_phase1(data, B, d)
    # even entries: the interpolants agree with the blocks at the d-th roots
    for k in range(d):
        acc = (0, 0)
        for j in range(M // d):
            acc = fadd(acc, fmul(tup(data[j * d + k]), tup(B[j * d + k])))
        assert tup(poly[2 * k]) == acc
    # every entry: sum of products of the interpolants at the 2d-th roots
    w2 = root(logd + 1)
    coeffs = [(O.fft(data[j * d:(j + 1) * d], True), O.fft(B[j * d:(j + 1) * d], True))
              for j in range(M // d)]
    for k in range(2 * d):
        x = fpow(w2, k)
        acc = (0, 0)
        for c1, c2 in coeffs:
            acc = fadd(acc, fmul(horner(c1, x), horner(c2, x)))
        assert tup(poly[k]) == acc


@pytest.mark.parametrize("logM,logd", [(6, 3), (5, 1), (7, 5), (4, 4)])
def test_phase2_evaluates_interpolant(logM, logd):
    M, d = 1 << logM, 1 << logd
    data = O.synth(1000 + logM, 0, M)
    B = O.synth(1000 + logM, 1, M)
    r = O.synth(1000 + logM, 2, 1)[0]
    fd, fb = O.pcs_sumcheck_phase2(data, B, d, r)
    for j in range(M // d):
        assert tup(fd[j]) == horner(O.fft(data[j * d:(j + 1) * d], True), tup(r))
        assert tup(fb[j]) == horner(O.fft(B[j * d:(j + 1) * d], True), tup(r))
    # at r = w_d^t the Lagrange weights are the unit vector e_t
    for t in (0, d - 1, d // 2):
        rt = np.array(fpow(root(logd), t), np.uint64)
        L = O.lagrange_coeff(rt, d)
        e = np.zeros((d, 2), np.uint64)
        e[t, 0] = 1
        assert np.array_equal(L, e)
        fd, fb = O.pcs_sumcheck_phase2(data, B, d, rt)
        assert np.array_equal(fd, data[t::d]) and np.array_equal(fb, B[t::d])


@pytest.mark.parametrize("logn,logfd", [(8, 1), (8, 2), (8, 3), (9, 4), (6, 6), (10, 5)])
def test_multifold5_is_a_strided_subtransform(logn, logfd):
    n, fd = 1 << logn, 1 << logfd
    cw = O.synth(1100 + logn, logfd, n)
    a = O.synth(1100 + logn, 100 + logfd, fd)
    out = O.rs_multifold5(cw, a)
    assert out.shape == (n // fd, 2)
    # with p = iNTT_n(cw):  out[i] = sum_k a_k * P_k(w_n^(fd*i)),
    # P_k(z) = sum_s p[fd*s + k] z^s, i.e. a forward NTT of the stride-fd
    # subsequence p[k::fd] of length n/fd
    p = O.fft(cw, True)
    want = [(0, 0)] * (n // fd)
    for k in range(fd):
        sub = np.ascontiguousarray(p[k::fd])
        ev = O.fft(sub) if n // fd > 1 else sub
        for i in range(n // fd):
            want[i] = fadd(want[i], fmul(tup(a[k]), tup(ev[i])))
    assert [tup(x) for x in out] == want
    # the same at a few points by Horner with w_n^(fd*i) directly
    wn = root(logn)
    for i in sorted({0, 1, n // fd - 1} & set(range(n // fd))):
        z = fpow(wn, fd * i)
        acc = (0, 0)
        for k in range(fd):
            acc = fadd(acc, fmul(tup(a[k]), horner(p[k::fd], z)))
        assert tup(out[i]) == acc


def test_multifold5_of_a_codeword_is_the_folded_codeword():
    # folding the RS codeword of f (degree < n/rate) gives the RS codeword
    # of sum_k a_k f_k over the subgroup of order n/fd
    logm, rate, fd = 6, 4, 4
    m = 1 << logm
    f = O.synth(1200, 0, m)
    a = O.synth(1200, 1, fd)
    cw = O.rs_encode(f, rate)
    out = O.rs_multifold5(cw, a)
    g = np.zeros((m // fd, 2), np.uint64)
    for k in range(fd):
        g = O.vec_op("add", g, O.vec_op("mul", f[k::fd], np.tile(a[k], (m // fd, 1))))
    assert np.array_equal(out, O.rs_encode(g, rate))


@pytest.mark.parametrize("logn,rate0,rate1,degree", [(6, 4, 4, 8), (5, 2, 4, 32), (4, 4, 2, 2), (7, 4, 4, 32)])
def test_encode_concatenated_segments_are_codewords(logn, rate0, rate1, degree):
    n = 1 << logn
    data = O.synth(1300 + logn, degree, n)
    cp, C = O.encode_concatenated(data, rate0, rate1, degree)
    assert np.array_equal(cp, O.rs_encode(data, rate0))
    S, D = n * rate0 // degree, degree * rate1
    assert C.shape == (S, D, 2)
    # row j evaluates sum_k cp[j + k S] x^k at the D-th roots of unity
    wD = root(D.bit_length() - 1)
    for j in {0, 1, S - 1}:
        coeffs = cp[j::S][:degree]
        for t in {0, 1, D // 2, D - 1}:
            assert tup(C[j, t]) == horner(coeffs, fpow(wD, t))


@pytest.mark.parametrize("logn,fd", [(8, 2), (8, 8), (10, 32), (6, 4)])
def test_multifold_segment_is_the_fold_at_idx(logn, fd):
    # RS_multifold_segment of the fd-element segment at idx equals RS_multifold5 at idx
    n = 1 << logn
    cw = O.synth(1400 + logn, fd, n)
    L = O.synth(1400 + logn, fd + 1, fd)
    full = O.rs_multifold5(cw, L)
    for idx in {0, 1, 5 % (n // fd), n // fd - 1}:
        seg = cw[idx::n // fd][:fd]
        assert np.array_equal(O.rs_multifold_segment(seg, L, n, idx), full[idx])


def test_row_combine_is_the_weighted_sum():
    rows = np.stack([O.synth(1500, i, 64) for i in range(5)])
    c = O.synth(1501, 0, 5)
    out = O.row_combine(rows, c)
    want = np.zeros((64, 2), np.uint64)
    for i in range(5):
        want = O.vec_op("add", want, O.vec_op("mul", rows[i], np.tile(c[i], (64, 1))))
    assert np.array_equal(out, want)
    # one row: the scaled row; unit weights: the plain sum
    assert np.array_equal(O.row_combine(rows[:1], c[:1]), O.vec_op("mul", rows[0], np.tile(c[0], (64, 1))))
