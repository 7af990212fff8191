"""The C ABI library loads, exports every function include/cosnark.h
declares, and its host-side functions (no GPU needed) agree with the
oracle."""
import os
import re

import numpy as np

import cosnark
import pyoracle as O

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "cosnark.h")


def declared_functions():
    text = open(HDR).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cosnark_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    L = cosnark.lib()
    names = declared_functions()
    assert len(names) > 15
    for n in names:
        assert hasattr(L, n), n


def test_host_functions_match_oracle():
    assert np.array_equal(cosnark.synth(0xC05A0002, 1, 300), O.synth(0xC05A0002, 1, 300))
    for n in (8, 64, 1 << 20):
        assert cosnark.merkle_node_count(n) == int(O.oracle().or_merkle_node_count(n))
    ts = cosnark.Transcript()
    prev = np.zeros(2, np.uint64)
    e = O.synth(1, 0, 9)
    assert np.array_equal(cosnark.hash_to_field(ts, e), O.hash_to_field(prev, e))
    idx = cosnark.get_indexes(ts, 7, 1000, 33)
    exp = np.zeros(14, np.uint32)
    O.oracle().or_get_indexes(O._p(prev), 7, 1000, 33, O._p(exp))
    assert np.array_equal(idx.reshape(-1), exp)
    assert cosnark.lib().cosnark_plain_proof_bytes(1 << 16) == 32 + 48 * 16 + 32
