"""Batched plain proofs (cosnark_batch_*) against cosnark_prove_plain and the
oracle: every proof byte-identical to the plain path on the same inputs and
transcript, through the hipGraph replay and the plain-launch path, across
replays with new inputs and transcripts (with other device work between
them), at the 64-proof batched workload, and after the NTT autotuner has
changed the radix plans."""
import numpy as np
import pytest

import cosnark
import pyoracle as O

pytestmark = pytest.mark.gpu


def _fill(batch, seed):
    ins = []
    for i, n in enumerate(batch.sizes):
        f, g = O.synth(seed, 2 * i, n), O.synth(seed, 2 * i + 1, n)
        batch.set_inputs(i, f, g)
        ins.append((f, g))
    return ins


def _plain(ctx, f, g, rate, prev=(0, 0)):
    ts = cosnark.Transcript()
    ts.prev[0], ts.prev[1] = prev
    proof, _ = ctx.prove_plain(f, g, rate, ts)
    return proof.tobytes(), ts.value()


def _transcripts(prevs):
    out = []
    for pv in prevs:
        t = cosnark.Transcript()
        t.prev[0], t.prev[1] = pv
        out.append(t)
    return out


def test_batch_matches_plain_mixed_sizes(ctx):
    sizes = [4, 1 << 10, 1 << 12, 1 << 13, 1 << 13, 1 << 16, 1 << 14, 8, 1 << 13]
    print(f"batch mixed sizes {sizes}", flush=True)
    b = ctx.batch(sizes, 4)
    ins = _fill(b, 0xC05A0020)
    proofs, ts, _ = b.prove()
    for i, (f, g) in enumerate(ins):
        p, t = _plain(ctx, f, g, 4)
        assert proofs[i] == p and ts[i].value() == t, f"proof {i} (n={sizes[i]})"
    for i in (0, 1, 2, 7):
        assert proofs[i] == O.plain_proof(ins[i][0], ins[i][1], 4)[0]
    b.close()


@pytest.mark.parametrize("rate", [2, 8])
def test_batch_other_rates(ctx, rate):
    sizes = [1 << 11, 1 << 14, 1 << 11]
    print(f"batch rate {rate}", flush=True)
    b = ctx.batch(sizes, rate)
    ins = _fill(b, 0xC05A0022 + rate)
    proofs, _, _ = b.prove()
    for i, (f, g) in enumerate(ins):
        assert proofs[i] == _plain(ctx, f, g, rate)[0]
    b.close()


def test_graph_replays_with_new_inputs_and_transcripts(ctx):
    sizes = [1 << 12, 1 << 15, 1 << 12, 1 << 17]
    b = ctx.batch(sizes, 4)
    for rep in range(3):
        print(f"batch replay {rep}", flush=True)
        ins = _fill(b, 0xC05A0021 + rep)
        prevs = [tuple(int(v) for v in O.synth(77, rep * 10 + i, 1)[0]) for i in range(len(sizes))]
        proofs, ts, _ = b.prove(_transcripts(prevs))
        for i, (f, g) in enumerate(ins):
            p, t = _plain(ctx, f, g, 4, prevs[i])
            assert proofs[i] == p and ts[i].value() == t, f"replay {rep} proof {i}"
    b.close()


def test_graph_and_plain_launches_agree(ctx):
    sizes = [1 << 13, 1 << 12, 1 << 18, 1 << 12]
    b = ctx.batch(sizes, 4)
    _fill(b, 0xC05A0023)
    with_graph, ts1, _ = b.prove()
    b.set_graph(False)
    without, ts2, _ = b.prove()
    assert with_graph == without
    assert [t.value() for t in ts1] == [t.value() for t in ts2]
    b.close()


def test_batched_workload_64_proofs(ctx):
    sizes = cosnark.cfg4_sizes()
    assert len(sizes) == 64 and min(sizes) >= 1 << 12 and max(sizes) <= 1 << 20
    print(f"batch cfg4: {sum(sizes)} elements", flush=True)
    b = ctx.batch(sizes, 4)
    ins = _fill(b, 0xC05A0004)
    proofs, ts, ms = b.prove()
    proofs2, _, _ = b.prove()
    assert proofs2 == proofs
    for i, (f, g) in enumerate(ins):
        p, t = _plain(ctx, f, g, 4)
        assert proofs[i] == p and ts[i].value() == t, f"proof {i} (n={sizes[i]})"
    smallest = sorted(range(64), key=lambda i: sizes[i])[:2]
    for i in smallest:
        assert proofs[i] == O.plain_proof(ins[i][0], ins[i][1], 4)[0]
    b.close()


def test_autotuned_plans_give_the_same_proofs(ctx):
    sizes = [1 << 12, 1 << 14, 1 << 16, 1 << 16]
    tuned = cosnark.Context(0)
    best = tuned.autotune(14, 18, batch=2)
    assert np.all(best > 0)
    plans = [tuned.ntt_plan(l) for l in range(14, 19)]
    print(f"tuned plans {plans}", flush=True)
    b = tuned.batch(sizes, 4)
    ins = _fill(b, 0xC05A0024)
    proofs, _, _ = b.prove()
    for i, (f, g) in enumerate(ins):
        assert proofs[i] == _plain(ctx, f, g, 4)[0]
    b.close()
    tuned.close()


@pytest.mark.parametrize("block_log", [7, 9])
def test_merkle_schedule_gives_the_same_proofs(ctx, block_log):
    """Batched proofs under a non-default Merkle schedule (shared launches
    of all trees' levels) equal the plain prover's."""
    sizes = [1 << 10, 1 << 12, 1 << 16, 1 << 18]
    c = cosnark.Context(0)
    c.set_merkle_schedule(block_log)
    b = c.batch(sizes, 4)
    ins = _fill(b, 0xC05A0030 + block_log)
    proofs, _, _ = b.prove()
    for i, (f, g) in enumerate(ins):
        assert proofs[i] == _plain(ctx, f, g, 4)[0], f"proof {i} (n={sizes[i]}, b={block_log})"
    b.close()
    c.close()


def test_batch_argument_errors(ctx):
    lib = cosnark.lib()
    import ctypes
    h = ctypes.c_void_p()
    for sizes, rate in (([3], 4), ([1 << 27], 4), ([4], 1), ([16], 3)):
        arr = np.ascontiguousarray(sizes, dtype=np.uint64)
        assert lib.cosnark_batch_create(ctx.h, len(sizes), cosnark._p(arr), rate, ctypes.byref(h)) == cosnark.E_ARG


@pytest.mark.parametrize("levels", [0, 1, 64])
def test_batch_tree_levels_from_the_encode(levels):
    """Each group's last NTT pass writes its proofs' leaf digests and tree
    levels 1 and 2 by default (COSNARK_BATCH_LEAF_LEVELS=3, covered by the
    64-proof commitment test); here the other settings -- no levels (trees
    from the codewords), the leaf digests alone, the whole tree -- on a mixed
    batch: every codeword and tree node against the oracle, and every proof
    equal to the plain prover's."""
    import os
    os.environ["COSNARK_BATCH_LEAF_LEVELS"] = str(levels)
    try:
        c = cosnark.Context(0)
    finally:
        del os.environ["COSNARK_BATCH_LEAF_LEVELS"]
    try:
        sizes = [1 << 12, 1 << 14, 1 << 14, 1 << 16, 1 << 11, 1 << 15]
        print(f"batch tree levels={levels} sizes={sizes}", flush=True)
        b = c.batch(sizes, 4)
        ins = _fill(b, 0xB1E0 + levels)
        proofs, _, _ = b.prove()
        for i, n in enumerate(sizes):
            f, g = ins[i]
            ref_cw = O.rs_encode(f, 4)
            ref_nodes = O.merkle_commit(ref_cw, True)
            cw, nodes = b.commitment(i)
            assert np.array_equal(cw, ref_cw), f"proof {i} codeword"
            bad = np.nonzero(np.any(nodes != ref_nodes, axis=1))[0]
            assert bad.size == 0, f"proof {i} (n={n}): {bad.size} nodes differ, first {bad[:4]}"
            assert proofs[i] == _plain(c, f, g, 4)[0]
        b.close()
    finally:
        c.close()
