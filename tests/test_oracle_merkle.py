"""The oracle's Merkle commit / open (src/merkle_tree.cpp:195-233,
267-299, 320-335).  Pins: leaf digests are BLAKE3 of the raw 64 bytes of
four elements (checked against the reference's compiled BLAKE3 via the
golden digests' function); the reference's parent hashes its left child
twice, so level l node i is the l-fold self-hash chain of leaf 2^l * i;
openings rebuild the root in the left||right mode."""
import numpy as np
import pytest

import pyoracle as O


def h2(d):
    return O.blake3_64(d + d)


@pytest.mark.parametrize("log_elems", [3, 6, 10])
def test_dup_left_closed_form(log_elems):
    n = 1 << log_elems
    e = O.synth(400 + log_elems, 0, n)
    nodes = O.merkle_commit(e, True)
    leaves = n // 4
    # leaves
    for i in range(leaves):
        assert nodes[i].tobytes() == O.blake3_64(e[4 * i:4 * i + 4].tobytes())
    # every level from the chain of its leftmost leaf
    off, size, lvl = leaves, leaves // 2, 1
    while size >= 1:
        for i in range(size):
            d = nodes[(1 << lvl) * i].tobytes()
            for _ in range(lvl):
                d = h2(d)
            assert nodes[off + i].tobytes() == d
        off += size
        size //= 2
        lvl += 1
    assert off == nodes.shape[0]


def test_dup_left_root_depends_only_on_leaf0():
    e = O.synth(410, 0, 256)
    e2 = e.copy()
    e2[4:] = O.synth(411, 0, 252)
    assert O.merkle_commit(e, True)[-1].tobytes() == O.merkle_commit(e2, True)[-1].tobytes()
    assert O.merkle_commit(e, False)[-1].tobytes() != O.merkle_commit(e2, False)[-1].tobytes()


@pytest.mark.parametrize("pos", [0, 1, 6, 63])
def test_open_rebuilds_root_pairs_mode(pos):
    n = 256
    e = O.synth(420, 0, n)
    nodes = O.merkle_commit(e, False)
    path = O.merkle_open(nodes, n, pos)
    d, p = nodes[pos].tobytes(), pos
    for sib in path:
        d = O.blake3_64(d + sib.tobytes()) if p % 2 == 0 else O.blake3_64(sib.tobytes() + d)
        p //= 2
    assert d == nodes[-1].tobytes()
    assert len(path) == 6
