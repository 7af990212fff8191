"""Merkle authentication paths verified on the GPU (cosnark_merkle_verify)
against trees the GPU built, in the reference's dup-left form and the
standard form, and the plain-proof verifier on GPU-made proofs.
Dup-left parents hash their left child twice, so a node's hash goes up only
while it is a left child and the root depends on leaf 0 alone: a changed leaf
reaches the root only at position 0, and a changed sibling at level L only
when pos >> L == 1.  The tests state exactly that for dup-left trees; in
standard trees every change must fail."""
import numpy as np
import pytest

import cosnark
import pyoracle as O

pytestmark = pytest.mark.gpu


def _tree(ctx, n, dup_left, seed):
    e = O.synth(seed, 0, n)
    nodes = O.merkle_commit(e, dup_left)
    d = ctx.upload(e)
    cnt = cosnark.merkle_node_count(n)
    nd = ctx.alloc(32 * cnt)
    ctx._ck(cosnark.lib().cosnark_merkle_commit(ctx.h, d.ptr, n, nd.ptr, int(dup_left)))
    got = ctx.download(nd, (cnt, 32), np.uint8)
    d.free(); nd.free()
    assert np.array_equal(got, nodes)
    return e, nodes


def _paths(nodes, n, pos):
    return np.stack([O.merkle_open(nodes, n, int(p)) for p in pos])


@pytest.mark.parametrize("logn,dup_left", [(12, True), (12, False), (20, True), (3, True), (16, False)])
def test_genuine_paths_accepted(ctx, logn, dup_left):
    print(f"merkle_verify logn={logn} dup_left={dup_left}", flush=True)
    n = 1 << logn
    e, nodes = _tree(ctx, n, dup_left, 0xC05A0060 + logn)
    rng = np.random.default_rng(logn)
    pos = rng.integers(0, n // 4, size=300).astype(np.uint32)
    paths = _paths(nodes, n, pos)
    leaves = np.stack([e[4 * p:4 * p + 4] for p in pos])
    ok, bad = ctx.merkle_verify(leaves, pos, paths, nodes[-1], dup_left)
    assert bad == 0 and ok.all()


@pytest.mark.parametrize("dup_left", [True, False])
def test_changes_rejected(ctx, dup_left):
    n = 1 << 12
    e, nodes = _tree(ctx, n, dup_left, 0xC05A0061)
    pos = np.arange(0, 64, dtype=np.uint32)          # even and odd leaves
    paths = _paths(nodes, n, pos)
    leaves = np.stack([e[4 * p:4 * p + 4] for p in pos])
    # a changed leaf element
    lv = leaves.copy()
    lv[:, 2, 1] ^= np.uint64(1)
    ok, _ = ctx.merkle_verify(lv, pos, paths, nodes[-1], dup_left)
    reaches = (pos == 0) if dup_left else np.ones(len(pos), bool)
    assert not ok[reaches].any() and ok[~reaches].all()
    # a changed sibling at level L
    for L in (0, 1, 4, paths.shape[1] - 1):
        pt = paths.copy()
        pt[:, L, 5] ^= 1
        ok, _ = ctx.merkle_verify(leaves, pos, pt, nodes[-1], dup_left)
        reaches = ((pos >> L) == 1) if dup_left else np.ones(len(pos), bool)
        assert not ok[reaches].any() and ok[~reaches].all(), f"level {L}"
    # the wrong root, and positions out of range
    root = nodes[-1].copy()
    root[0] ^= 1
    ok, bad = ctx.merkle_verify(leaves, pos, paths, root, dup_left)
    assert bad == len(pos) and not ok.any()
    ok, _ = ctx.merkle_verify(leaves, pos + np.uint32(n // 4), paths, nodes[-1], dup_left)
    assert not ok.any()


@pytest.mark.parametrize("logn", [10, 16, 22])
def test_plain_proofs_verify(ctx, logn):
    print(f"verify_plain logn={logn}", flush=True)
    n = 1 << logn
    f, g = O.synth(0xC05A0062, 0, n), O.synth(0xC05A0062, 1, n)
    proof, _ = ctx.prove_plain(f, g, 4, cosnark.Transcript())
    y = O.inner_product(f, g)
    rc, r, bad = cosnark.verify_plain(proof, n, y)
    assert rc == 0 and bad == -1
