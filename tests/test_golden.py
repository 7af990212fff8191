"""The oracle reproduces the committed oracle fixtures (regression pins)
and the synthetic-input generator is stable."""
import json
import os

import numpy as np

import pyoracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "oracle_vectors.json")


def test_oracle_fixtures():
    g = json.load(open(GOLD))
    s = g["synth_head"]
    assert O.synth(s["seed"], 0, 4).tolist() == s["f"]
    assert O.synth(s["seed"], 1, 4).tolist() == s["g"]
    m = g["merkle_2e10"]
    e = O.synth(m["seed"], 0, 1 << 10)
    nodes = O.merkle_commit(e, True)
    assert nodes[-1].tobytes().hex() == m["root_dup_left"]
    assert O.merkle_commit(e, False)[-1].tobytes().hex() == m["root_pairs"]
    assert [x.tobytes().hex() for x in O.merkle_open(nodes, 1 << 10, 5)] == m["open_pos5"]
    f = g["fft_2e8"]
    a = O.synth(f["seed"], 0, 1 << 8)
    assert O.fft(a)[:8].tolist() == f["forward_head"]
    assert O.fft(a, True)[:8].tolist() == f["inverse_head"]
    h = g["hash_to_field"]
    assert O.hash_to_field(np.zeros(2, np.uint64), np.array(h["elems"], np.uint64)).tolist() == h["out"]


def test_plain_proof_fixture():
    g = json.load(open(GOLD))["plain_proof_cfg1"]
    f, gg = O.synth(g["seed"], 0, g["n"]), O.synth(g["seed"], 1, g["n"])
    proof, y, r = O.plain_proof(f, gg, g["rate"])
    assert proof.hex() == g["proof_hex"]
    assert y.tolist() == g["claim"]
