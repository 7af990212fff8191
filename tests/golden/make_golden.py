"""Writes the golden fixtures under tests/golden/.

ref_vectors.json    -- outputs of the reference's own compiled code
                       (oracle/_ref: src/fieldElement.cpp, src/polynomial.cpp,
                       src/Blake3_hash.cpp + Blake/ C) on seeded inputs.
oracle_vectors.json -- outputs of the oracle for the functions whose
                       reference source cannot be compiled here (merkle,
                       transcript, fft, sumcheck, the plain proof); these are
                       regression pins, the oracle itself being pinned by the
                       property tests in tests/test_oracle_*.py.

Run from the repo root in a container that has the reference tree:
    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def fe_list(a):
    return [[int(x), int(y)] for x, y in np.asarray(a).reshape(-1, 2)]


def ref_vectors():
    R = O.ref()
    assert R is not None, "oracle/_ref not built (needs the reference tree)"
    p = O._p
    rng = np.random.default_rng(20261203)
    n = 64
    a = rng.integers(0, O.P, size=(n, 2), dtype=np.uint64)
    b = rng.integers(0, O.P, size=(n, 2), dtype=np.uint64)
    # edge values: 0, 1, p-1, 2^31, 2^60
    edge = np.array([[0, 0], [1, 0], [0, 1], [O.P - 1, O.P - 1], [1 << 31, 0], [1 << 60, (1 << 60) + 5]], np.uint64)
    a = np.concatenate([a, edge, edge[::-1]])
    b = np.concatenate([b, edge[::-1], edge])
    out = {}
    for name in ("mul", "add", "sub"):
        o = np.zeros_like(a)
        getattr(R, "ref_" + name)(p(a), p(b), p(o), a.shape[0])
        out[name] = fe_list(o)
    o = np.zeros_like(a)
    R.ref_inv(p(a), p(o), a.shape[0])
    out["inv"] = fe_list(o)
    out["a"], out["b"] = fe_list(a), fe_list(b)
    roots = []
    for logn in range(0, 62):
        r = np.zeros(2, np.uint64)
        R.ref_root_of_unity(logn, p(r))
        roots.append([int(r[0]), int(r[1])])
    out["root_of_unity"] = roots
    msgs = rng.integers(0, 256, size=(16, 64), dtype=np.uint8)
    msgs[0] = np.arange(64, dtype=np.uint8)
    msgs[1] = 0
    dig = []
    for m in msgs:
        d = np.zeros(32, np.uint8)
        R.ref_blake3_hash64(p(np.ascontiguousarray(m)), p(d))
        dig.append(d.tobytes().hex())
    out["blake3_msgs"] = [m.tobytes().hex() for m in msgs]
    out["blake3_digests"] = dig
    rnd = {}
    for seed in (1, 3396, 0xC05A0003):
        O.libc().srandom(seed)
        r = np.zeros((12, 2), np.uint64)
        R.ref_random_elements(p(r), 12)
        rnd[str(seed)] = fe_list(r)
    out["random_elements"] = rnd
    lin = []
    for i in range(8):
        l1 = rng.integers(0, O.P, size=(2, 2), dtype=np.uint64)
        l2 = rng.integers(0, O.P, size=(2, 2), dtype=np.uint64)
        q = np.zeros((3, 2), np.uint64)
        R.ref_linear_mul(p(l1), p(l2), p(q))
        x = rng.integers(0, O.P, size=2, dtype=np.uint64)
        e = np.zeros(2, np.uint64)
        R.ref_quadratic_eval(p(q), p(x), p(e))
        lin.append({"l1": fe_list(l1), "l2": fe_list(l2), "quad": fe_list(q), "x": fe_list(x)[0], "eval": fe_list(e)[0]})
    out["linear_mul"] = lin
    lay = np.zeros(3, np.uint64)
    R.ref_fe_layout(p(lay))
    out["fe_layout"] = [int(v) for v in lay]
    return out


def oracle_vectors():
    out = {}
    seed = 0xC05A0001
    f, g = O.synth(seed, 0, 1 << 16), O.synth(seed, 1, 1 << 16)
    out["synth_head"] = {"seed": seed, "f": fe_list(f[:4]), "g": fe_list(g[:4])}
    proof, y, r = O.plain_proof(f, g, 4)
    out["plain_proof_cfg1"] = {"n": 1 << 16, "rate": 4, "seed": seed, "claim": fe_list(y)[0],
                               "proof_hex": proof.hex(), "challenges": fe_list(r)}
    e = O.synth(0x77, 0, 1 << 10)
    nodes = O.merkle_commit(e, True)
    out["merkle_2e10"] = {"seed": 0x77, "root_dup_left": nodes[-1].tobytes().hex(),
                          "root_pairs": O.merkle_commit(e, False)[-1].tobytes().hex(),
                          "open_pos5": [x.tobytes().hex() for x in O.merkle_open(nodes, 1 << 10, 5)]}
    a = O.synth(0x78, 0, 1 << 8)
    out["fft_2e8"] = {"seed": 0x78, "forward_head": fe_list(O.fft(a)[:8]), "inverse_head": fe_list(O.fft(a, True)[:8])}
    prev = np.zeros(2, np.uint64)
    out["hash_to_field"] = {"elems": fe_list(a[:5]), "out": fe_list(O.hash_to_field(prev, a[:5]))[0]}
    return out


if __name__ == "__main__":
    with open(os.path.join(OUT, "ref_vectors.json"), "w") as fh:
        json.dump(ref_vectors(), fh, indent=0)
    with open(os.path.join(OUT, "oracle_vectors.json"), "w") as fh:
        json.dump(oracle_vectors(), fh, indent=0)
    print("wrote", OUT)
