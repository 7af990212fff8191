"""The oracle's _Whir_Open (or_whir_open_weights, src/Whir.cpp:469-545):
its record is the one-array Whir_Batch_Open record for the same inputs;
total_r is the record's challenges in order (every iteration's sumcheck
challenges, then the closing draws); the final claim is the initial
weights' multilinear value at total_r, checked here as <b, eq(total_r)>
(a different computation from the oracle's evaluate_vector fold); opened
against aggregate_commitments' all-zero-leaf tree, the first iteration's
paths are the zero chain."""
import numpy as np
import pytest

import pyoracle as O


def _parse(rec, n, cw_n, whir_k, first_leaves):
    """Walk a one-array record: (challenges, first iteration's paths)."""
    w = 1
    steps = int(rec[0])
    chal, first_paths = [], None
    leaves = first_leaves
    rt = cw_n // n
    for t in range(steps):
        assert int(rec[w]) == 0
        kappa = int(rec[w + 1]); w += 2
        k = kappa.bit_length() - 1
        w += 6 * k
        for _ in range(k):
            chal.append(rec[w:w + 2].copy()); w += 2
        w += 4                                   # new root
        l = int(rec[w]); w += 1 + l
        w += 2 * l * kappa + 2 * l               # replies, folds
        depth = leaves.bit_length() - 1
        paths = rec[w:w + 4 * l * depth].copy(); w += 4 * l * depth
        if t == 0:
            first_paths = paths.reshape(l, depth, 4)
        w += 2                                   # update_betas challenge
        n //= kappa
        cw_n = n * 2 * rt
        rt *= 2
        leaves = cw_n // 4
    w += 2 * n + 2 + 4                           # final arr, K, root
    for _ in range(n.bit_length() - 1):
        chal.append(rec[w:w + 2].copy()); w += 2
    assert w == rec.shape[0]
    return np.array(chal, dtype=np.uint64), first_paths


@pytest.mark.parametrize("logn,rate,whir_k", [(10, 4, 4), (12, 4, 8), (11, 2, 2), (12, 8, 16)])
def test_whir_open_weights_equals_one_array_batch_open(logn, rate, whir_k):
    n = 1 << logn
    arr, z = O.synth(0xC05A0050, 0, n), O.synth(0xC05A0050, 1, logn)
    b = O.precompute_beta(z)
    K = O.whir_claim(arr, z)
    cw = O.rs_encode(arr, rate)
    nodes = O.merkle_commit(cw, True)
    O.libc().srandom(31)
    w, rec, chal, fin = O.whir_open_weights(arr, b, K, cw, nodes, 0, rate, whir_k)
    assert w > 0
    inp = O.whir_batch_inputs([arr], [z], rate, whir_k)
    O.libc().srandom(31)
    w2, rec2 = O.whir_batch_open(inp)
    assert w2 == 0 and np.array_equal(rec, rec2)
    pc, _ = _parse(rec, n, n * rate, whir_k, n * rate // 4)
    assert np.array_equal(pc, chal)
    # the final claim: <b, eq(total_r)>
    assert np.array_equal(fin, O.inner_product(b, O.precompute_beta(chal)))


@pytest.mark.parametrize("whir_k", [4, 8])
def test_whir_open_weights_on_a_zero_leaf_tree(whir_k):
    n, rate = 1 << 10, 4
    arr, b = O.synth(0xC05A0051, 0, n), O.synth(0xC05A0051, 1, n)
    K = O.inner_product(arr, b)
    cw = O.rs_encode(arr, rate)
    leaves = n * rate // whir_k
    O.libc().srandom(5)
    w, rec, chal, fin = O.whir_open_weights(arr, b, K, cw, None, leaves, rate, whir_k)
    assert w > 0
    pc, paths = _parse(rec, n, n * rate, whir_k, leaves)
    assert np.array_equal(pc, chal)
    # the zero chain: level 0 is 32 zero bytes, level l the self-hash of l-1
    z = [bytes(32)]
    for _ in range(paths.shape[1]):
        z.append(O.blake3_64(z[-1] + z[-1]))
    for q in range(paths.shape[0]):
        for d in range(paths.shape[1]):
            assert paths[q, d].astype('<u8').tobytes() == z[d]
    assert np.array_equal(fin, O.inner_product(b, O.precompute_beta(chal)))


def test_whir_open_weights_wrong_claim_fails():
    n = 1 << 8
    arr, b = O.synth(0xC05A0052, 0, n), O.synth(0xC05A0052, 1, n)
    K = O.inner_product(arr, b)
    K[0] = (int(K[0]) + 1) % O.P
    cw = O.rs_encode(arr, 4)
    O.libc().srandom(1)
    w, *_ = O.whir_open_weights(arr, b, K, cw, O.merkle_commit(cw, True), 0, 4, 4)
    assert w < 0
