"""The C++ drop-in layer (include/cosnark_dropin.hpp): tests/dropin/test_dropin
calls each reference-named entry point (fft, MT_commit_Blake,
create_tree_blake, open_tree_blake, hash_to_field, get_indexes,
_quadratic_sumcheck, Whir_Commit_initial, Whir_Open, Encode, distributed_MT,
pcs_sumcheck_phase1/2, quadratic_coSumcheck_phase1/2, RS_multifold5) with the
reference's type shapes and checks each against the oracle."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "dropin", "test_dropin")


def test_dropin_program():
    if not os.path.exists(BIN):
        pytest.fail("tests/dropin/test_dropin is not built (run __graft_entry__.build())")
    print("drop-in program", flush=True)
    p = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(p.stdout)
    print(p.stderr)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-2000:]
    assert "PASSED (0 failures)" in p.stdout
