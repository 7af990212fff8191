"""The C++ drop-in header compiles with the reference's type shapes and the
test program is built (no GPU needed)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_dropin_header_compiles_standalone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "cosnark_dropin.hpp"\n'
                   "struct F { unsigned long long real, img; };\n"
                   "struct H { unsigned char arr[32]; };\n"
                   "void use(std::vector<F> &a, std::vector<std::vector<H>> &t) {\n"
                   "  cosnark::dropin::fft(a, 4, false);\n"
                   "  cosnark::dropin::MT_commit_Blake(a.data(), t, 16);\n"
                   "  cosnark::dropin::RS_multifold5(a, a, 16);\n"
                   "}\n")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I", os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_dropin_program_builds():
    r = subprocess.run(["make", "-C", os.path.join(HERE, "dropin")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(os.path.join(HERE, "dropin", "test_dropin"))
