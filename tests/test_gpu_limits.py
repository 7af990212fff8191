"""The largest sizes the library takes (2^28-point transforms, 2^28-element
trees, a 2^26-element sumcheck), checked through properties that do not
need the oracle to run at that size:
- transform round trips;
- X[0] = sum of the inputs;
- a codeword's coefficients vanish above the message length;
- in the dup-left tree, level l node i is the l-fold self-hash of leaf 2^l * i;
- the sumcheck's round identities hold against a claim computed by an
  independent kernel, and its final values are the multilinear evaluations
  at its challenges."""
import ctypes

import numpy as np
import pytest

import cosnark
import pyoracle as O

pytestmark = pytest.mark.gpu


def _sum_mod_p(col):
    """sum of canonical u64 values mod p, exactly (32-bit halves)."""
    col = np.asarray(col, dtype=np.uint64)
    lo = int(np.sum(col & np.uint64(0xFFFFFFFF), dtype=np.uint64))
    hi = int(np.sum(col >> np.uint64(32), dtype=np.uint64))
    return ((hi << 32) + lo) % O.P


def _d2h_at(ctx, buf, offset, nbytes):
    out = np.zeros(nbytes, np.uint8)
    src = ctypes.c_void_p(buf.ptr.value + offset)
    ctx._ck(cosnark.lib().cosnark_d2h(ctx.h, cosnark._p(out), src, nbytes))
    return out


def test_ntt_round_trip_at_2_28(ctx):
    n = 1 << 28
    print("ntt 2^28", flush=True)
    a = O.synth(0xC05A0050, 0, n)
    d = ctx.upload(a)
    ctx._ck(cosnark.lib().cosnark_ntt(ctx.h, d.ptr, 28, 0))
    x0 = _d2h_at(ctx, d, 0, 16).view(np.uint64)
    assert int(x0[0]) == _sum_mod_p(a[:, 0]) and int(x0[1]) == _sum_mod_p(a[:, 1])
    ctx._ck(cosnark.lib().cosnark_ntt(ctx.h, d.ptr, 28, 1))
    back = ctx.download(d, (n, 2))
    d.free()
    assert np.array_equal(back, a)


def test_rs_encode_2_26_to_2_28(ctx):
    n, rate = 1 << 26, 4
    print("rs_encode 2^26 -> 2^28", flush=True)
    a = O.synth(0xC05A0051, 0, n)
    d = ctx.upload(a)
    cw = ctx.alloc(16 * n * rate)
    ctx._ck(cosnark.lib().cosnark_rs_encode(ctx.h, d.ptr, n, rate, cw.ptr))
    c0 = _d2h_at(ctx, cw, 0, 16).view(np.uint64)
    assert int(c0[0]) == _sum_mod_p(a[:, 0]) and int(c0[1]) == _sum_mod_p(a[:, 1])
    ctx._ck(cosnark.lib().cosnark_ntt(ctx.h, cw.ptr, 28, 1))
    coeffs = ctx.download(cw, (n * rate, 2))
    d.free(); cw.free()
    assert np.array_equal(coeffs[:n], a)
    assert not coeffs[n:].any()


def _h2(dg):
    return O.blake3_64(dg + dg)


def test_merkle_2_28_elements_closed_form(ctx):
    n = 1 << 28
    print("merkle 2^28 elements", flush=True)
    e = O.synth(0xC05A0052, 0, n)
    d = ctx.upload(e)
    cnt = cosnark.merkle_node_count(n)
    nodes = ctx.alloc(32 * cnt)
    ctx._ck(cosnark.lib().cosnark_merkle_commit(ctx.h, d.ptr, n, nodes.ptr, 1))
    leaves = n // 4
    rng = np.random.default_rng(7)
    off, lvl = 0, 0
    while (leaves >> lvl) >= 1:
        size = leaves >> lvl
        for i in sorted({0, size - 1, int(rng.integers(0, size))}):
            leaf = (1 << lvl) * i
            dg = O.blake3_64(e[4 * leaf:4 * leaf + 4].tobytes())
            for _ in range(lvl):
                dg = _h2(dg)
            assert _d2h_at(ctx, nodes, 32 * (off + i), 32).tobytes() == dg, f"level {lvl} node {i}"
        off += size
        lvl += 1
    assert off == cnt
    d.free(); nodes.free()


def test_product_sumcheck_at_2_26(ctx):
    n, m = 1 << 26, 26
    print("sumcheck 2^26", flush=True)
    f, g = O.synth(0xC05A0053, 0, n), O.synth(0xC05A0053, 1, n)
    a, b = ctx.upload(f), ctx.upload(g)
    # the claim from an independent kernel: the d = 2 block sum's even entries
    poly = np.zeros((4, 2), np.uint64)
    ctx._ck(cosnark.lib().cosnark_cosumcheck_phase1(ctx.h, a.ptr, b.ptr, n, 2, cosnark._p(poly)))
    y = np.array([(int(poly[0, 0]) + int(poly[2, 0])) % O.P, (int(poly[0, 1]) + int(poly[2, 1])) % O.P], np.uint64)
    ts = cosnark.Transcript()
    polys = np.zeros((m, 3, 2), np.uint64)
    r = np.zeros((m, 2), np.uint64)
    fin = np.zeros((2, 2), np.uint64)
    rc = cosnark.lib().cosnark_product_sumcheck(ctx.h, a.ptr, b.ptr, n, cosnark._p(y), ctypes.byref(ts), 0,
                                                cosnark._p(polys), cosnark._p(r), cosnark._p(fin))
    a.free(); b.free()
    assert rc == 0
    # the transcript is the fold of the round polynomials
    prev = np.zeros(2, np.uint64)
    for i in range(m):
        assert np.array_equal(O.hash_to_field(prev, polys[i]), r[i])
    assert np.array_equal(O.evaluate_vector(f, r), fin[0])
    assert np.array_equal(O.evaluate_vector(g, r), fin[1])
