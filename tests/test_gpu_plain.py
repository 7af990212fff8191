"""The plain single-rank prover (RS-encode f, Merkle commit, product
sumcheck of f*g) on the GPU against the oracle pipeline."""
import numpy as np
import pytest

import cosnark
import pyoracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m", [4, 10, 16, 20])
def test_plain_proof_bytes_match_oracle(ctx, m):
    seed = 0xC05A0001
    n = 1 << m
    f, g = O.synth(seed, 0, n), O.synth(seed, 1, n)
    print(f"plain proof m={m}", flush=True)
    exp, _, _ = O.plain_proof(f, g, 4)
    ts = cosnark.Transcript()
    proof, times = ctx.prove_plain(f, g, 4, ts)
    assert proof.tobytes() == exp


def test_plain_proof_cfg1_matches_fixture(ctx):
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "oracle_vectors.json")))["plain_proof_cfg1"]
    f, gg = O.synth(g["seed"], 0, g["n"]), O.synth(g["seed"], 1, g["n"])
    proof, _ = ctx.prove_plain(f, gg, g["rate"], cosnark.Transcript())
    assert proof.tobytes().hex() == g["proof_hex"]
