"""The benchmark's mirrors of library launch rules (bench.leaf_levels,
bench.merkle_launches), which size the roofline's algorithmic bytes and the
Merkle stage's compression count, checked against plans worked by hand from
ntt_launch_batch and merkle.hip's stage builders."""
import os
import sys

import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
import bench  # noqa: E402


@pytest.mark.parametrize("logn,passbits,tile,want", [
    (24, 8, 4096, 3),   # 8 + 8 + 8: last pass 16 x 16 over 16 columns, R C = 4096
    (24, 8, 2048, 1),   # 8 columns: whole leaves, no four-leaf groups
    (24, 8, 1024, 1),   # 4 columns
    (24, 6, 4096, 3),   # 6 x 4: 8 x 8 over 64 columns
    (19, 7, 4096, 3),   # 7 + 6 + 6... last 6 bits: 8 x 8 over 64 columns
    (21, 7, 4096, 3),   # 7 x 3: 16 x 8 over 32 columns
    (14, 8, 4096, 3),   # two 7-bit passes: 16 x 8 over 32 columns, rows 2^7 apart
    (8, 8, 4096, 0),    # one pass: no leaves from the transform
    (10, 5, 256, 1),    # 5 + 5, 8 columns
])
def test_leaf_levels_rule(monkeypatch, logn, passbits, tile, want):
    monkeypatch.delenv("COSNARK_NTT_SPLIT", raising=False)
    monkeypatch.delenv("COSNARK_NTT_LEAF_LEVELS", raising=False)
    assert bench.leaf_levels(logn, passbits, tile) == want


def test_leaf_levels_follow_the_settings(monkeypatch):
    monkeypatch.setenv("COSNARK_NTT_LEAF_LEVELS", "1")
    assert bench.leaf_levels(24, 8, 4096) == 1
    monkeypatch.setenv("COSNARK_NTT_LEAF_LEVELS", "3")
    monkeypatch.setenv("COSNARK_NTT_SPLIT", "0")
    # whole-element exchange: one step-2 item per thread, no fused levels
    assert bench.leaf_levels(24, 8, 4096) == 1


@pytest.mark.parametrize("leaves,have,want", [
    (1 << 22, 0, 3),    # from the elements: 2^22 -> 2^12 -> 2^2 -> 1
    (1 << 22, 1, 2),    # levels 1..11 (2048 blocks), 12..22
    (1 << 22, 3, 2),    # levels 3..13 (512 blocks), 14..22
    (1 << 10, 3, 1),
    (4, 3, 0),          # level 2 is the root
])
def test_merkle_launches_rule(leaves, have, want):
    assert bench.merkle_launches(leaves, have) == want
