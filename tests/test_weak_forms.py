"""The NTT passes compute on weak values (components in [0, 2^61 + 2^32],
unreduced sums, one fold per radix-4 layer).  A host build of the same
header code is checked against the DFT written out as a sum, with inputs at
the edges of the weak range (no GPU needed)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "shorter-and-faster-post-quantum-secure-scalable-collaborative-snarks_amd", "csrc")


def test_weak_dft_matches_direct_sum(tmp_path):
    exe = tmp_path / "weak_forms"
    r = subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17", "-I", CSRC,
                        "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "hostcheck", "weak_forms.cpp"),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "weak forms: 0 bad" in r.stdout
