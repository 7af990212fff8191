"""NTT / Reed-Solomon encode on the GPU against the oracle (fft, pfft,
Whir_Commit_initial codeword: src/utils.cpp:904-964, src/Whir.cpp:158-163)."""
import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("logn", [1, 2, 3, 5, 8, 10, 11, 12, 13, 16, 17, 20])
def test_ntt_forward_matches_oracle(ctx, logn):
    a = O.synth(0xA0 + logn, 0, 1 << logn)
    print(f"ntt forward logn={logn}", flush=True)
    assert np.array_equal(ctx.ntt(a), O.fft(a))


@pytest.mark.parametrize("logn", [1, 4, 9, 12, 15, 18])
def test_ntt_inverse_matches_oracle(ctx, logn):
    a = O.synth(0xB0 + logn, 1, 1 << logn)
    print(f"ntt inverse logn={logn}", flush=True)
    assert np.array_equal(ctx.ntt(a, inverse=True), O.fft(a, inverse=True))


@pytest.mark.parametrize("logn", [21, 24])
def test_ntt_round_trip_large(ctx, logn):
    a = O.synth(0xC0 + logn, 0, 1 << logn)
    print(f"ntt round trip logn={logn}", flush=True)
    assert np.array_equal(ctx.ntt(ctx.ntt(a), inverse=True), a)


@pytest.mark.parametrize("logn,rate", [(1, 4), (3, 2), (6, 4), (10, 4), (14, 4), (16, 4), (12, 8), (13, 16)])
def test_rs_encode_matches_oracle(ctx, logn, rate):
    a = O.synth(0xD0 + logn, 0, 1 << logn)
    print(f"rs_encode logn={logn} rate={rate}", flush=True)
    assert np.array_equal(ctx.rs_encode(a, rate), O.rs_encode(a, rate))


@pytest.mark.parametrize("logn", [25])
def test_ntt_four_pass_matches_oracle(ctx, logn):
    a = O.synth(0xF0 + logn, 0, 1 << logn)
    print(f"ntt forward logn={logn}", flush=True)
    assert np.array_equal(ctx.ntt(a), O.fft(a))


def test_rs_encode_four_pass_matches_oracle(ctx):
    # 2^22 -> 2^25: three passes, the first fanning out over the padding by 8
    a = O.synth(0xF1, 0, 1 << 22)
    print("rs_encode 2^22 rate 8", flush=True)
    assert np.array_equal(ctx.rs_encode(a, 8), O.rs_encode(a, 8))


@pytest.mark.parametrize("passbits,logn", [(4, 20), (3, 20), (5, 20), (2, 11)])
def test_ntt_many_pass_plans_match_oracle(passbits, logn):
    # Regression: the last (natural-order) pass scatters rows, so it must not
    # run in place; plans with 4-10 passes exercise every digit-reversal path.
    import os
    import cosnark
    os.environ["COSNARK_NTT_PASSBITS"] = str(passbits)
    try:
        c = cosnark.Context(0)
    finally:
        del os.environ["COSNARK_NTT_PASSBITS"]
    try:
        a = O.synth(0x70 + logn, 0, 1 << logn)
        print(f"ntt passbits={passbits} logn={logn}", flush=True)
        assert np.array_equal(c.ntt(a), O.fft(a))
        assert np.array_equal(c.ntt(a, inverse=True), O.fft(a, inverse=True))
        assert np.array_equal(c.rs_encode(a[: (1 << logn) // 4], 4), O.rs_encode(a[: (1 << logn) // 4], 4))
    finally:
        c.close()


@pytest.mark.parametrize("passbits,logn,rate", [(0, 16, 4), (0, 15, 4), (0, 16, 2), (0, 15, 8), (3, 18, 4),
                                                (2, 12, 4), (5, 15, 8)])
def test_rs_encode_fan_out_matches_oracle(passbits, logn, rate):
    """The fan-out first pass (zero padding by 2, 4 or 8 covered by the first
    pass's digit) at sizes below its default workgroup threshold, with the
    threshold lowered, at the default and at small pass radices."""
    import os
    import cosnark
    env = {"COSNARK_NTT_FAN_MIN": "1"}
    if passbits:
        env["COSNARK_NTT_PASSBITS"] = str(passbits)
    os.environ.update(env)
    try:
        c = cosnark.Context(0)
    finally:
        for k in env:
            del os.environ[k]
    try:
        a = O.synth(0x90 + logn + rate, 0, 1 << logn)
        print(f"rs_encode fan-out passbits={passbits} logn={logn} rate={rate}", flush=True)
        assert np.array_equal(c.rs_encode(a, rate), O.rs_encode(a, rate))
    finally:
        c.close()


@pytest.mark.parametrize("split,defer", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("passbits,tile,logn", [(8, 4096, 20), (8, 2048, 18), (7, 4096, 17), (6, 1024, 16),
                                                (4, 512, 13), (5, 4096, 19)])
def test_ntt_exchange_and_twiddle_forms_match_oracle(split, defer, passbits, tile, logn):
    """Both LDS exchanges of a pass -- whole 16-byte elements
    (COSNARK_NTT_SPLIT=0) and real words then imaginary words through a
    half-size buffer (the default where each thread holds at most one item per
    step) -- and both inter-pass twiddle schemes: a running product at each
    pass's store (COSNARK_NTT_DEFER=0) and the factors left to the consuming
    passes (the default for plans of three or more passes of at least 4 bits;
    19 bits at 5 per pass takes four passes).  Forward, inverse, an RS encode
    and a WHIR commit (leaf digests from the last pass, restaged in two halves
    under the split form) under each plan."""
    import os
    import cosnark
    env = {"COSNARK_NTT_SPLIT": str(split), "COSNARK_NTT_DEFER": str(defer)}
    os.environ.update(env)
    try:
        c = cosnark.Context(0)
    finally:
        for k in env:
            del os.environ[k]
    try:
        print(f"ntt exchange split={split} defer={defer} passbits={passbits} tile={tile} logn={logn}", flush=True)
        for lg in (logn, logn + 2):
            c.set_ntt_plan(lg, passbits, tile)
        a = O.synth(0x60 + logn + passbits, split, 1 << logn)
        assert np.array_equal(c.ntt(a), O.fft(a))
        assert np.array_equal(c.ntt(a, inverse=True), O.fft(a, inverse=True))
        q = a[: (1 << logn) // 4]
        assert np.array_equal(c.rs_encode(q, 4), O.rs_encode(q, 4))
        cw, nodes = c.whir_commit(c.upload(a), 1 << logn, 4)
        ref_cw = O.rs_encode(a, 4)
        assert np.array_equal(c.download(cw, ((4 << logn), 2)), ref_cw)
        cnt = cosnark.merkle_node_count(4 << logn)
        assert np.array_equal(c.download(nodes, (cnt, 32), np.uint8), O.merkle_commit(ref_cw, True))
    finally:
        c.close()


@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("logn,rate,passbits,tile", [(17, 4, 6, 1024), (17, 4, 6, 4096), (15, 8, 5, 512),
                                                     (15, 8, 5, 2048)])
def test_rs_encode_fan_out_deferred_matches_oracle(split, logn, rate, passbits, tile):
    """Fan-out first passes in plans of three passes (deferred twiddles):
    with the split exchange and narrow enough tiles the first pass runs in
    block form (one k0 per workgroup, the twists in the per-tile tables):
    2^19 points as (2 + 6) + 6 + 5 bits at tile 1024, 2^18 as (3 + 5) + 5 + 5
    at tile 512; at wider tiles (and without the split exchange) the first
    pass runs its k0 loop inside the workgroup.  The 2^25-point encode of
    test_rs_encode_four_pass_matches_oracle takes the block form at the
    default plan, (3 + 8) + 7 + 7 bits.  The codeword
    and every tree node (leaf digests from the last pass) against the
    oracle."""
    import os
    import cosnark
    env = {"COSNARK_NTT_FAN_MIN": "1", "COSNARK_NTT_SPLIT": str(split)}
    os.environ.update(env)
    try:
        c = cosnark.Context(0)
    finally:
        for k in env:
            del os.environ[k]
    try:
        n = 1 << logn
        lg = logn + rate.bit_length() - 1
        print(f"fan-out deferred split={split} logn={logn} rate={rate} passbits={passbits} tile={tile}", flush=True)
        c.set_ntt_plan(lg, passbits, tile)
        a = O.synth(0xFA + logn + passbits, rate, n)
        ref_cw = O.rs_encode(a, rate)
        assert np.array_equal(c.rs_encode(a, rate), ref_cw)
        cw, nodes = c.whir_commit(c.upload(a), n, rate)
        assert np.array_equal(c.download(cw, (n * rate, 2)), ref_cw)
        cnt = cosnark.merkle_node_count(n * rate)
        assert np.array_equal(c.download(nodes, (cnt, 32), np.uint8), O.merkle_commit(ref_cw, True))
    finally:
        c.close()


@pytest.mark.parametrize("leaf_levels", [1, 3, 64])
@pytest.mark.parametrize("passbits,tile,logn", [(8, 4096, 20), (7, 4096, 17), (6, 4096, 16), (8, 2048, 18),
                                                (8, 4096, 12), (8, 4096, 3)])
def test_leaf_pass_tree_levels_match_oracle(leaf_levels, passbits, tile, logn):
    """The last NTT pass writes the leaf digests; with
    COSNARK_NTT_LEAF_LEVELS=3 tree levels 1 and 2 as well where a half-tile
    holds one two-leaf unit per thread (radix 16 x 16 over 16 columns, 16 x 8
    over 32, 8 x 8 over 64), and with more (the default) every level up to
    the root, each lane climbing from its level-2 node while the node index
    has trailing zeros (16 x 16 columns: one level-2 node per row; 32 and 64
    columns: two and four).  2048-element tiles keep the leaves alone; a
    2^5-point encode is one pass and hashes nothing.  Every codeword element
    and every node of the WHIR commit's tree against the oracle."""
    import os
    import cosnark
    os.environ["COSNARK_NTT_LEAF_LEVELS"] = str(leaf_levels)
    try:
        c = cosnark.Context(0)
    finally:
        del os.environ["COSNARK_NTT_LEAF_LEVELS"]
    try:
        print(f"leaf pass levels={leaf_levels} passbits={passbits} tile={tile} logn={logn}", flush=True)
        c.set_ntt_plan(logn + 2, passbits, tile)
        a = O.synth(0x1EAF + logn + passbits, leaf_levels, 1 << logn)
        cw, nodes = c.whir_commit(c.upload(a), 1 << logn, 4)
        ref_cw = O.rs_encode(a, 4)
        assert np.array_equal(c.download(cw, ((4 << logn), 2)), ref_cw)
        cnt = cosnark.merkle_node_count(4 << logn)
        got = c.download(nodes, (cnt, 32), np.uint8)
        ref = O.merkle_commit(ref_cw, True)
        bad = np.nonzero(np.any(got != ref, axis=1))[0]
        assert bad.size == 0, f"{bad.size} nodes differ, first {bad[:4]}"
    finally:
        c.close()
