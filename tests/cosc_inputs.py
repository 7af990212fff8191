"""Well-formed inputs for the coSumcheck drivers over N ranks, built with the
oracle: packed secret shares of seeded secret vectors
(compute_secret_shares, src/MPI_utils.cpp:2742-2766; N/4 secrets and N/4
random pads per share element), per-rank random masks, and the claims K
and K_r that make every check of src/coSumcheck.cpp pass, plus the claims
computed from the secrets directly (the shares folded slot by slot)."""
import numpy as np

import pyoracle as O


def _sum_even(R, stride):
    acc = np.zeros((1, 2), np.uint64)
    for Ri in R:
        for e in Ri[0::stride]:
            acc = O.vec_op("add", acc, e.reshape(1, 2))
    return acc[0]


def _shares(N, secret, seed, stream):
    k, kk = N // 4, N // 2
    pads = O.synth(seed, stream, (secret.shape[0] // k) * (kk - k))
    return O.compute_secret_shares(secret, N, k, kk, pads)


def _fold_slots(secret, N, rounds):
    """The secret vector as its N/4 slots (slot j: secret[k i + j]), each
    folded by the rounds' Lagrange weights: (N/4, remaining, 2)."""
    k = N // 4
    v = [secret[j::k].copy() for j in range(k)]
    for r, d in rounds:
        L = O.lagrange_coeff(r, d)
        out = []
        for x in v:
            y = np.zeros((x.shape[0] // d, 2), np.uint64)
            for t in range(y.shape[0]):
                acc = np.zeros((1, 2), np.uint64)
                prods = O.vec_op("mul", L, x[t * d:(t + 1) * d])
                for p in prods:
                    acc = O.vec_op("add", acc, p.reshape(1, 2))
                y[t] = acc[0]
            out.append(y)
        v = out
    return v


def closing_weights(r, N):
    if N // 4 == 1:
        return np.array([[1, 0]], np.uint64)
    return O.lagrange_coeff(r, N // 4)


def quadratic_inputs(N, M, degrees, seed):
    """N ranks' shares of two secret vectors of (N/4) M elements: (p1, p2
    (N, M, 2), R (N lists of masks), K, K_r, s1, s2)."""
    k = N // 4
    s1, s2 = O.synth(seed, 0, k * M), O.synth(seed, 1, k * M)
    p1, p2 = _shares(N, s1, seed, 2), _shares(N, s2, seed, 3)
    R = [[O.synth(seed, 100 + 10 * m + i, 2 * d) for i, d in enumerate(degrees)] for m in range(N)]
    K = O.inner_product(s1, s2)
    ones = np.zeros((k, 2), np.uint64)
    ones[:, 0] = 1
    S = np.stack([_sum_even(R[m], 2) for m in range(N)]).reshape(N, 1, 2)
    K_r = O.sumcheck_aggr(S, ones)[0]
    return p1, p2, R, K, K_r, s1, s2


def quadratic_expected_claims(s1, s2, N, degrees, r):
    """claims[0..3] from the secrets: slot folds at positions 0 and 1,
    combined with the closing weights."""
    rounds = [(r[1 + i], d) for i, d in enumerate(degrees)]
    L = closing_weights(r[0], N)
    out = []
    for s, pos in ((s1, 0), (s2, 0), (s1, 1), (s2, 1)):
        v = _fold_slots(s, N, rounds)
        acc = np.zeros((1, 2), np.uint64)
        for j in range(N // 4):
            acc = O.vec_op("add", acc, O.vec_op("mul", L[j:j + 1], v[j][pos:pos + 1]))
        out.append(acc[0])
    return np.stack(out)


def cubic_inputs(N, logM, degrees, seed):
    """(p1, p2, p3 (N, M, 2), r1 (logM), r2, R, K, K_r, s1, s2, s3)."""
    k, M = N // 4, 1 << logM
    s1, s2, s3 = O.synth(seed, 0, k * M), O.synth(seed, 1, k * M), O.synth(seed, 2, k * M)
    p1, p2, p3 = _shares(N, s1, seed, 3), _shares(N, s2, seed, 4), _shares(N, s3, seed, 5)
    r1 = O.synth(seed, 6, logM)
    nr2 = max((k - 1).bit_length(), 1)
    r2 = O.synth(seed, 7, nr2)
    b1, b2 = O.precompute_beta(r1), O.precompute_beta(r2)
    R = [[O.synth(seed, 200 + 10 * m + i, 4 * d) for i, d in enumerate(degrees)] for m in range(N)]
    # K = sum over slots j of b2_j sum_i b1_i (s1 s2 - s3)[k i + j]
    t = O.vec_op("sub", O.vec_op("mul", s1, s2), s3)
    w = np.zeros((k * M, 2), np.uint64)
    for j in range(k):
        w[j::k] = O.vec_op("mul", b1, np.repeat(b2[j:j + 1], M, axis=0))
    K = O.inner_product(t, w)
    S = np.stack([_sum_even(R[m], 4) for m in range(N)]).reshape(N, 1, 2)
    K_r = O.sumcheck_aggr(S, b2)[0]
    return p1, p2, p3, r1, r2, R, K, K_r, s1, s2, s3


def cubic_expected_claims(secrets, N, degrees, r):
    rounds = [(r[1 + i], d) for i, d in enumerate(degrees)]
    L = closing_weights(r[0], N)
    out = []
    for s in secrets:
        v = _fold_slots(s, N, rounds)
        acc = np.zeros((1, 2), np.uint64)
        for j in range(N // 4):
            acc = O.vec_op("add", acc, O.vec_op("mul", L[j:j + 1], v[j][0:1]))
        out.append(acc[0])
    return np.stack(out)
