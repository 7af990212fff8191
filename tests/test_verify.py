"""The plain-proof verifier (cosnark_verify_plain, host only) on the committed
2^16 proof fixture: it accepts the proof with its claim, reproduces the
fixture's challenges, and rejects a changed claim, round polynomial, final
value or non-canonical element, naming the failing round."""
import json
import os

import numpy as np
import pytest

import cosnark

GOLD = os.path.join(os.path.dirname(__file__), "golden", "oracle_vectors.json")
P = (1 << 61) - 1


def _fixture():
    g = json.load(open(GOLD))["plain_proof_cfg1"]
    return bytearray(bytes.fromhex(g["proof_hex"])), g["n"], g["claim"], g["challenges"]


def _set(proof, off, val):
    proof[off:off + 8] = int(val).to_bytes(8, "little")


def _get(proof, off):
    return int.from_bytes(proof[off:off + 8], "little")


def test_fixture_accepted_and_challenges_reproduced():
    proof, n, y, ch = _fixture()
    rc, r, bad = cosnark.verify_plain(proof, n, y)
    assert rc == 0 and bad == -1
    assert r.tolist() == ch


def test_wrong_claim_fails_round_0():
    proof, n, y, _ = _fixture()
    rc, _, bad = cosnark.verify_plain(proof, n, [(y[0] + 1) % P, y[1]])
    assert rc == cosnark.E_CHECK and bad == 0


@pytest.mark.parametrize("rnd", [0, 5, 15])
def test_changed_constant_term_fails_its_round(rnd):
    proof, n, y, _ = _fixture()
    off = 32 + 48 * rnd + 32            # c of round rnd
    _set(proof, off, (_get(proof, off) + 1) % P)
    rc, _, bad = cosnark.verify_plain(proof, n, y)
    assert rc == cosnark.E_CHECK and bad == rnd


def test_sum_preserving_change_fails_later():
    # a+1, b-1 keeps H(0)+H(1) but moves the challenge and H(r)
    proof, n, y, _ = _fixture()
    off = 32 + 48 * 3
    _set(proof, off, (_get(proof, off) + 1) % P)
    _set(proof, off + 16, (_get(proof, off + 16) - 1) % P)
    rc, _, bad = cosnark.verify_plain(proof, n, y)
    assert rc == cosnark.E_CHECK and bad > 3


def test_changed_final_value_fails_at_the_end():
    proof, n, y, _ = _fixture()
    m = n.bit_length() - 1
    off = 32 + 48 * m
    _set(proof, off, (_get(proof, off) + 1) % P)
    rc, _, bad = cosnark.verify_plain(proof, n, y)
    assert rc == cosnark.E_CHECK and bad == m


def test_non_canonical_element_rejected():
    proof, n, y, _ = _fixture()
    off = 32 + 48 * 2 + 8               # imaginary part of a in round 2, + p (same residue)
    v = _get(proof, off)
    if v + P < (1 << 64):
        _set(proof, off, v + P)
        rc, _, bad = cosnark.verify_plain(proof, n, y)
        assert rc == cosnark.E_CHECK and bad == 2
