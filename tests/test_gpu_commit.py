"""The commitment of the timed provers, every element and every node.

A dup-left root depends on leaf 0 alone (src/merkle_tree.cpp:287-292), so
proof bytes pin only codeword[0..3].  These tests download the whole
codeword and every tree level that the plain prover (cosnark_prove_plain_ex),
the batched prover (cosnark_batch_commitment) and the server
(cosnark_server_submit_ex) leave, and compare them with the oracle's RS
encode and Merkle tree -- at the bench's 2^22 share under the radix plans
the bench picks between, at the 64-proof batched workload, and for a 2^22
request inside a server mix.  The all-cores oracle functions used at the
large sizes are pinned against the single-thread ones in
test_oracle_mt.py."""
import os

import numpy as np
import pytest

import cosnark
import pyoracle as O

pytestmark = pytest.mark.gpu

_cache = {}


def _oracle_commit(seed, stream, n, rate=4):
    key = (seed, stream, n, rate)
    if key not in _cache:
        f = O.synth(seed, stream, n)
        cw = O.rs_encode_mt(f, rate) if n >= 1 << 14 else O.rs_encode(f, rate)
        nodes = O.merkle_commit_mt(cw, True) if n >= 1 << 14 else O.merkle_commit(cw, True)
        _cache.clear()
        _cache[key] = (f, cw, nodes)
    return _cache[key]


def _check_commit(cw, nodes, ref_cw, ref_nodes, what):
    assert cw.shape == ref_cw.shape and nodes.shape == ref_nodes.shape, what
    bad = np.nonzero(np.any(cw != ref_cw, axis=1))[0]
    assert bad.size == 0, f"{what}: {bad.size} codeword elements differ, first {bad[:4]}"
    badn = np.nonzero(np.any(nodes != ref_nodes, axis=1))[0]
    assert badn.size == 0, f"{what}: {badn.size} tree nodes differ, first {badn[:4]}"


@pytest.mark.parametrize("m", [4, 10, 16, 20])
def test_plain_commitment_matches_oracle(ctx, m):
    seed, n = 0xC05A0001, 1 << m
    print(f"plain commitment m={m}", flush=True)
    f, ref_cw, ref_nodes = _oracle_commit(seed, 0, n)
    g = O.synth(seed, 1, n)
    df, dg = ctx.upload(f), ctx.upload(g)
    proof, cw, nodes = ctx.prove_plain_keep(df, dg, n, 4, cosnark.Transcript())
    _check_commit(ctx.download(cw, (4 * n, 2)), ctx.download(nodes, (ref_nodes.shape[0], 32), np.uint8),
                  ref_cw, ref_nodes, f"plain m={m}")
    # the proof is the plain prover's, and its root is the tree's last node
    p2, _ = ctx.prove_plain(None, None, 4, cosnark.Transcript(), dev_inputs=(df, dg, n), stage_times=False)
    assert proof.tobytes() == p2.tobytes()
    assert proof[:32].tobytes() == ref_nodes[-1].tobytes()
    for b in (df, dg, cw, nodes):
        b.free()


@pytest.mark.parametrize("plan", [(0, 0), (8, 4096), (8, 2048), (7, 4096), (6, 4096)])
def test_share_commitment_2_22_every_element_and_node(ctx, plan):
    """The bench's workload: f of the 2^22 share (seed 0xC05A0002), encoded to
    2^24 points with leaf digests from the last NTT pass, tree over 2^22
    leaves; under the context default and the plans the bench compares."""
    seed, n = 0xC05A0002, 1 << 22
    print(f"share commitment plan={plan}", flush=True)
    f, ref_cw, ref_nodes = _oracle_commit(seed, 0, n)
    g = O.synth(seed, 1, n)
    df, dg = ctx.upload(f), ctx.upload(g)
    if plan != (0, 0):
        ctx.set_ntt_plan(24, *plan)
    try:
        proof, cw, nodes = ctx.prove_plain_keep(df, dg, n, 4, cosnark.Transcript())
    finally:
        ctx.set_ntt_plan(24, 0, 0)
    got_cw = ctx.download(cw, (4 * n, 2))
    got_nodes = ctx.download(nodes, (ref_nodes.shape[0], 32), np.uint8)
    _check_commit(got_cw, got_nodes, ref_cw, ref_nodes, f"2^22 share plan {plan}")
    # closed forms besides the oracle: X[0] = sum f
    s = np.array([int(np.sum(f[:, 0].astype(object)) % O.P), int(np.sum(f[:, 1].astype(object)) % O.P)], np.uint64)
    assert np.array_equal(got_cw[0], s)
    assert proof[:32].tobytes() == ref_nodes[-1].tobytes()
    for b in (df, dg, cw, nodes):
        b.free()


def test_batch_commitments_cfg4_every_element_and_node(ctx):
    """All 64 codewords and trees of the batched workload (sizes 2^12..2^20
    from seed 0xC05A0004) after a graph replay, against the oracle."""
    sizes = cosnark.cfg4_sizes()
    print(f"batch commitments cfg4 ({sum(sizes)} elements)", flush=True)
    b = ctx.batch(sizes, 4)
    ins = []
    for i, n in enumerate(sizes):
        f, g = O.synth(0xC05A0004, 2 * i + 10, n), O.synth(0xC05A0004, 2 * i + 11, n)
        b.set_inputs(i, f, g)
        ins.append(f)
    b.prove()            # captures the graph
    proofs, _, _ = b.prove()   # replays it
    for i, n in enumerate(sizes):
        ref_cw = O.rs_encode_mt(ins[i], 4)
        ref_nodes = O.merkle_commit_mt(ref_cw, True)
        cw, nodes = b.commitment(i)
        _check_commit(cw, nodes, ref_cw, ref_nodes, f"batch proof {i} (n={n})")
        assert proofs[i][:32] == ref_nodes[-1].tobytes()
    b.close()


def test_server_mix_with_a_2_22_request_keeps_its_commitment(ctx):
    """A 2^22 request served among smaller ones that arrive around it: its
    proof equals the plain prover's and its codeword and tree equal the
    oracle's; a small request's commitment is checked too."""
    seed = 0xC05A0002
    n_big = 1 << 22
    f, ref_cw, ref_nodes = _oracle_commit(seed, 0, n_big)
    g = O.synth(seed, 1, n_big)
    dfb, dgb = ctx.upload(f), ctx.upload(g)
    small = [1 << 14, 1 << 17, 1 << 20, 1 << 16]
    pool = [(ctx.upload(O.synth(0xC05A0035, 2 * i, n)), ctx.upload(O.synth(0xC05A0035, 2 * i + 1, n)))
            for i, n in enumerate(small)]
    print("server mix with 2^22", flush=True)
    srv = cosnark.Server(12 << 30, rate=4, max_log_n=22, slots=3)
    cwb = ctx.alloc(16 * 4 * n_big)
    ndb = ctx.alloc(32 * cosnark.merkle_node_count(4 * n_big))
    n_s = small[1]
    cws = ctx.alloc(16 * 4 * n_s)
    nds = ctx.alloc(32 * cosnark.merkle_node_count(4 * n_s))
    tickets = [srv.submit(pool[0][0], pool[0][1], small[0])]
    tickets.append(srv.submit_keep(dfb, dgb, n_big, cwb, ndb))
    tickets.append(srv.submit_keep(pool[1][0], pool[1][1], n_s, cws, nds))
    tickets += [srv.submit(pool[i][0], pool[i][1], small[i]) for i in (2, 3)]
    sizes = [small[0], n_big, n_s, small[2], small[3]]
    res = [srv.wait(t, n) for t, n in zip(tickets, sizes)]
    srv.close()
    exp, _ = ctx.prove_plain(None, None, 4, cosnark.Transcript(), dev_inputs=(dfb, dgb, n_big), stage_times=False)
    assert res[1][0] == exp.tobytes()
    _check_commit(ctx.download(cwb, (4 * n_big, 2)), ctx.download(ndb, (ref_nodes.shape[0], 32), np.uint8),
                  ref_cw, ref_nodes, "served 2^22 request")
    fs = O.synth(0xC05A0035, 2, n_s)
    rcw = O.rs_encode_mt(fs, 4)
    rnd = O.merkle_commit_mt(rcw, True)
    _check_commit(ctx.download(cws, (4 * n_s, 2)), ctx.download(nds, (rnd.shape[0], 32), np.uint8), rcw, rnd,
                  "served 2^17 request")
    for i, n in enumerate(sizes):
        if i == 1:
            continue
        j = {0: 0, 2: 1, 3: 2, 4: 3}[i]
        ff = O.synth(0xC05A0035, 2 * j, n)
        gg = O.synth(0xC05A0035, 2 * j + 1, n)
        p, _ = ctx.prove_plain(ff, gg, 4, cosnark.Transcript())
        assert res[i][0] == p.tobytes(), f"request {i}"
    for b in [dfb, dgb, cwb, ndb, cws, nds] + [x for pr in pool for x in pr]:
        b.free()


def test_sumcheck_2_22_at_the_most_workgroups(ctx):
    """The fused first rounds write nine partials per workgroup; at the
    largest round width a context allows (COSNARK_SC_BLOCKS = 2048) the
    2^22 sumcheck must still equal the oracle's (its partial slots hold
    SC_MAX_BLOCKS / 3 such workgroups)."""
    n = 1 << 22
    f, g = O.synth(0xC05A0002, 0, n), O.synth(0xC05A0002, 1, n)
    y = O.inner_product(f, g)
    prev = np.zeros(2, np.uint64)
    rc_o, polys_o, r_o, fin_o = O.product_sumcheck_mt(y, f, g, prev)
    assert rc_o == 0
    old = os.environ.get("COSNARK_SC_BLOCKS")
    os.environ["COSNARK_SC_BLOCKS"] = "2048"
    try:
        c2 = cosnark.Context(0)
    finally:
        if old is None:
            del os.environ["COSNARK_SC_BLOCKS"]
        else:
            os.environ["COSNARK_SC_BLOCKS"] = old
    try:
        print("sumcheck 2^22 at 2048 workgroups", flush=True)
        ts = cosnark.Transcript()
        rc, polys, r, fin = c2.product_sumcheck(f, g, y, ts)
        assert rc == 0
        assert np.array_equal(polys, polys_o) and np.array_equal(r, r_o) and np.array_equal(fin, fin_o)
        # the plain prover takes the same fused rounds
        proof, _ = c2.prove_plain(f, g, 4, cosnark.Transcript())
        assert proof[32:].tobytes() == polys_o.tobytes() + fin_o.tobytes()
    finally:
        c2.close()
