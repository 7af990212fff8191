"""The oracle's batched WHIR (src/Whir.cpp): get_whir_rates (:18-137), the
fold schedule and aggregation of Whir_Batch_Open (:706-899),
aggregate_arrays (:670-703) and aggregate_commitments (:547-606), with
Whir_Iteration at any fold factor kappa.

Pins: the schedules against hand-derived values; the reference's own checks
run inside the oracle and return an error when they fail: every sumcheck
round, the closing K = <arr, b> (which holds only if every query fold y_i
agrees with the RS encoding and every aggregation combined arrays, weights
and claims consistently), and the closing codeword = RS(arr) (which holds
only if the aggregated codewords are the encodings of the aggregated arrays).
One array with whir_k = 4 reduces to Whir_Open's record; a wrong claim is
rejected; the claims of aggregated arrays are not read (the reference
recomputes them).  The libc draw order and the record layout follow the
reference text: parity partially pinned (the reference does not build here)."""
import numpy as np
import pytest

import pyoracle as O


def _inputs(seed, logs, rate, whir_k=4):
    arrays = [O.synth(seed, 2 * i, 1 << lg) for i, lg in enumerate(logs)]
    zs = [O.synth(seed, 2 * i + 1, lg) for i, lg in enumerate(logs)]
    return O.whir_batch_inputs(arrays, zs, rate, whir_k)


def test_whir_rates_by_hand():
    # 2^12 twice at rate 4; folding 2^12 -> 2^10 (rate 8) -> 2^8 meets 2^8 at
    # rate 16; 2^8 -> 2^6 meets both 2^6 at rate 32
    assert O.whir_rates([1 << 12, 1 << 12, 1 << 8, 1 << 6, 1 << 6], 4, 4) == [4, 4, 16, 32, 32]
    assert O.whir_rates([1 << 10], 2, 4) == [2]
    assert O.whir_rates([1 << 12, 1 << 9], 4, 8) == [4, 8]


def test_kappas_by_hand():
    k, a = O.whir_kappas([1 << 12, 1 << 12, 1 << 8, 1 << 6, 1 << 6], 4)
    assert k == [1, 4, 4, 4, 4, 4]
    assert a == [True, False, True, True, False, False]
    k, a = O.whir_kappas([1 << 8, 1 << 2], 4)
    assert k == [4, 4, 4] and a == [False, False, True]
    k, a = O.whir_kappas([1 << 12, 1 << 9], 8)
    assert k == [8, 8, 8] and a == [True, False, False]


@pytest.mark.parametrize("logs,rate,whir_k", [([10], 4, 4), ([10, 10], 4, 4), ([12, 8, 6, 6], 4, 4),
                                              ([12, 12, 8, 6, 6], 2, 4), ([12, 9], 4, 8), ([8, 6], 4, 2),
                                              ([8, 2], 4, 4), ([11, 7, 5], 2, 4), ([13, 10, 10, 7], 2, 8)])
def test_batch_open_self_consistent(logs, rate, whir_k):
    inp = _inputs(1000 + sum(logs), logs, rate, whir_k)
    O.libc().srandom(31 + len(logs))
    rc, rec = O.whir_batch_open(inp)
    assert rc == 0, rc
    k, a = O.whir_kappas(inp["sizes"], whir_k)
    assert int(rec[0]) == sum(1 for kk, aa in zip(k, a) if not aa or kk != 1) + sum(a)


def test_single_array_is_whir_open():
    n, rate = 1 << 10, 4
    arr, z = O.synth(1100, 0, n), O.synth(1100, 1, 10)
    K = O.whir_claim(arr, z)
    O.libc().srandom(9)
    rc, ref = O.whir_open(arr, z, K, rate)
    assert rc == 0
    inp = O.whir_batch_inputs([arr], [z], rate, 4)
    O.libc().srandom(9)
    rc, rec = O.whir_batch_open(inp)
    assert rc == 0
    assert int(rec[0]) == int(ref[0])
    # the same words once each step's [kind = 0, kappa = 4] header is removed
    pos, rpos = 1, 1
    for it in range(int(ref[0])):
        assert int(rec[pos]) == 0 and int(rec[pos + 1]) == 4
        pos += 2
        # iteration length from the whir_open record: 12 + 4 + 4 + 1 + l + 8l + 2l + 4 l depth + 2
        l = int(ref[rpos + 20])
        depth = (n * rate).bit_length() - 1 - 2 - it    # log2(|cw_it| / 4), |cw| halves each iteration
        ln = 21 + l + 8 * l + 2 * l + 4 * l * depth + 2
        assert np.array_equal(rec[pos:pos + ln], ref[rpos:rpos + ln]), f"iteration {it}"
        pos += ln
        rpos += ln
    assert np.array_equal(rec[pos:], ref[rpos:])


def test_wrong_claim_rejected():
    inp = _inputs(1200, [10, 8], 4)
    inp["evals"][0, 1] = (int(inp["evals"][0, 1]) + 1) % O.P
    O.libc().srandom(3)
    rc, _ = O.whir_batch_open(inp)
    assert rc < 0


def test_aggregated_claims_are_not_read():
    inp = _inputs(1300, [10, 8, 8], 4)
    O.libc().srandom(4)
    rc, rec = O.whir_batch_open(inp)
    assert rc == 0
    inp["evals"][1:] = 12345
    O.libc().srandom(4)
    rc2, rec2 = O.whir_batch_open(inp)
    assert rc2 == 0 and np.array_equal(rec, rec2)


def test_unsorted_input_rejected():
    inp = _inputs(1400, [8, 10], 4)
    O.libc().srandom(5)
    rc, _ = O.whir_batch_open(inp)
    assert rc < 0


@pytest.mark.parametrize("sizes,rate,whir_k", [([1 << 12, 1 << 12, 1 << 8, 1 << 6, 1 << 6], 4, 4), ([1 << 10], 2, 4),
                                               ([1 << 12, 1 << 9], 4, 8), ([1 << 14, 1 << 11, 1 << 5], 2, 2)])
def test_library_whir_rates_match_oracle(sizes, rate, whir_k):
    import cosnark
    assert cosnark.whir_rates(sizes, rate, whir_k) == O.whir_rates(sizes, rate, whir_k)


def test_library_whir_rates_reject_unsorted():
    import cosnark
    with pytest.raises(cosnark.CosnarkError):
        cosnark.whir_rates([1 << 8, 1 << 10], 4, 4)


@pytest.mark.parametrize("narr,npts,logn", [(1, 1, 6), (3, 2, 8), (2, 5, 10)])
def test_aggregate_claims_closed_form(narr, npts, logn):
    n = 1 << logn
    arrays = [O.synth(1500 + logn, i, n) for i in range(narr)]
    points = np.stack([O.synth(1600 + logn, j, logn) for j in range(npts)])
    O.libc().srandom(17)
    arr, b, K, bc, ac = O.whir_aggregate_claims(arrays, points)
    # the challenges: npts for the weights, then narr for the arrays, from libc
    O.libc().srandom(17)
    draws = [O.unfe(O.oracle().or_random_element()) for _ in range(npts + narr)]
    assert [tuple(int(v) for v in x) for x in bc] == [tuple(int(v) for v in d) for d in draws[:npts]]
    assert [tuple(int(v) for v in x) for x in ac] == [tuple(int(v) for v in d) for d in draws[npts:]]
    # K = <arr, b_aggr> = sum_ij bc_j ac_i MLE(A_i)(z_j)
    assert np.array_equal(O.inner_product(arr, b), K)
    want = np.zeros(2, np.uint64)
    for i in range(narr):
        for j in range(npts):
            e = np.array(O.evaluate_vector(arrays[i], points[j]), np.uint64).reshape(1, 2)
            w = O.vec_op("mul", bc[j:j + 1], ac[i:i + 1])
            want = O.vec_op("add", want.reshape(1, 2), O.vec_op("mul", w, e)).reshape(2)
    assert np.array_equal(want, K)
