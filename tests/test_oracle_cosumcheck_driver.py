"""The oracle's coSumcheck drivers (or_quadratic_cosumcheck,
or_cubic_cosumcheck: src/coSumcheck.cpp:233-448, all N ranks in lockstep)
pinned by the protocol itself: on packed shares of secret vectors with
per-rank masks, every check the reference prints (each round's sum equals
the running claim, the final check, the cubic's verification 2) holds, and
the claims equal the secret vectors folded slot by slot with the rounds'
Lagrange weights and combined with the closing weights -- computed here
from the secrets, not from the shares.  A wrong claim trips round 0."""
import numpy as np
import pytest

import pyoracle as O
from cosc_inputs import (cubic_expected_claims, cubic_inputs, quadratic_expected_claims, quadratic_inputs)


@pytest.mark.parametrize("N,M,degrees", [(4, 1 << 9, [8, 8, 4]), (8, 1 << 10, [16, 8, 4]), (16, 1 << 8, [32, 4])])
def test_quadratic_driver_checks_and_claims(N, M, degrees):
    p1, p2, R, K, K_r, s1, s2 = quadratic_inputs(N, M, degrees, 0xC05A0070 + N)
    prev = np.zeros(2, np.uint64)
    rc, claims, r, ip, fails = O.quadratic_cosumcheck(list(p1), list(p2), R, K, K_r, degrees, prev)
    assert rc == 0 and fails == 0
    assert np.array_equal(claims[:4], quadratic_expected_claims(s1, s2, N, degrees, r))
    assert ip.shape[0] == sum(2 * d for d in degrees)


def test_quadratic_driver_wrong_claim_trips_round_0():
    N, M, degrees = 4, 1 << 8, [8, 4, 4]
    p1, p2, R, K, K_r, _, _ = quadratic_inputs(N, M, degrees, 0xC05A0071)
    K[0] = (int(K[0]) + 1) % O.P
    rc, *_, fails = O.quadratic_cosumcheck(list(p1), list(p2), R, K, K_r, degrees, np.zeros(2, np.uint64))
    assert rc == 0 and fails & 1


def test_quadratic_driver_rejects_fewer_than_four_ranks():
    N, M = 4, 1 << 6
    p1, p2, R, K, K_r, _, _ = quadratic_inputs(N, M, [4, 4], 3)
    rc, *_ = O.quadratic_cosumcheck(list(p1[:2]), list(p2[:2]), R[:2], K, K_r, [4, 4], np.zeros(2, np.uint64))
    assert rc == -1


@pytest.mark.parametrize("N,logM,degrees", [(4, 9, [8, 16, 4]), (8, 10, [16, 8, 8])])
def test_cubic_driver_checks_and_claims(N, logM, degrees):
    p1, p2, p3, r1, r2, R, K, K_r, s1, s2, s3 = cubic_inputs(N, logM, degrees, 0xC05A0072 + N)
    prev = np.zeros(2, np.uint64)
    rc, claims, r, ip, fails = O.cubic_cosumcheck(list(p1), list(p2), list(p3), r1, r2, R, K, K_r, degrees, prev)
    assert rc == 0 and fails == 0
    assert np.array_equal(claims[:3], cubic_expected_claims((s1, s2, s3), N, degrees, r))
    assert ip.shape[0] == sum(4 * d for d in degrees)
