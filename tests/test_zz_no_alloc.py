"""cosnark_prove_plain_noalloc (the plain prover with no device allocation,
its temporaries in a fixed per-GPU scratch array) against the allocating
entry point cosnark_prove_plain_ex, called through the C ABI: proof bytes,
transcript, codeword and every tree node equal on two calls in a row, the
second smaller; sizes beyond the scratch are refused on the host."""
import ctypes

import numpy as np
import pytest

import cosnark
import pyoracle as O

pytestmark = pytest.mark.gpu


def _bind():
    L = cosnark.lib()
    vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
    L.cosnark_prove_plain_noalloc.argtypes = [vp, vp, vp, u64, i32, vp, vp, vp, vp]
    L.cosnark_prove_plain_noalloc_max_bytes.restype = u64
    return L


def _run(ctx, fn, df, dg, n, rate, prev):
    L = _bind()
    ts = cosnark.Transcript()
    ts.prev[0], ts.prev[1] = prev
    proof = np.zeros(int(L.cosnark_plain_proof_bytes(n)), np.uint8)
    cw = ctx.alloc(16 * n * rate)
    nodes = ctx.alloc(32 * cosnark.merkle_node_count(n * rate))
    if fn == "noalloc":
        rc = L.cosnark_prove_plain_noalloc(ctx.h, df.ptr, dg.ptr, n, rate, ctypes.byref(ts), cosnark._p(proof),
                                           cw.ptr, nodes.ptr)
    else:
        rc = L.cosnark_prove_plain_ex(ctx.h, df.ptr, dg.ptr, n, rate, ctypes.byref(ts), cosnark._p(proof),
                                      cw.ptr, nodes.ptr, None)
    assert rc == 0, L.cosnark_last_error(ctx.h)
    out = (proof.tobytes(), ts.value(), ctx.download(cw, (n * rate, 2)),
           ctx.download(nodes, (cosnark.merkle_node_count(n * rate), 32), np.uint8))
    cw.free(); nodes.free()
    return out


def test_noalloc_equals_allocating_entry_point_twice_second_smaller():
    with cosnark.Context(0) as ctx:
        for n, rate, seed in ((1 << 20, 4, 0xC05A0041), (1 << 12, 4, 0xC05A0042)):
            print(f"noalloc n={n} rate={rate}", flush=True)
            f, g = O.synth(seed, 0, n), O.synth(seed, 1, n)
            df, dg = ctx.upload(f), ctx.upload(g)
            prev = (int(seed), 7)
            a = _run(ctx, "noalloc", df, dg, n, rate, prev)
            b = _run(ctx, "ex", df, dg, n, rate, prev)
            assert a[0] == b[0] and a[1] == b[1]
            assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
            df.free(); dg.free()


def test_noalloc_other_rates_and_overlap_off():
    """Other blowups, with the stages one after another (the scratch's
    NTT temporary and sumcheck buffers then meet on one stream)."""
    with cosnark.Context(0) as ctx:
        ctx._ck(cosnark.lib().cosnark_set_overlap(ctx.h, 0))
        for n, rate in ((1 << 16, 8), (1 << 14, 2)):
            f, g = O.synth(0xC05A0043, rate, n), O.synth(0xC05A0043, rate + 1, n)
            df, dg = ctx.upload(f), ctx.upload(g)
            a = _run(ctx, "noalloc", df, dg, n, rate, (0, 0))
            assert a[0] == O.plain_proof(f, g, rate)[0]
            df.free(); dg.free()


def test_noalloc_refuses_sizes_beyond_its_scratch():
    L = _bind()
    assert L.cosnark_prove_plain_noalloc_max_bytes() == 96 << 20
    with cosnark.Context(0) as ctx:
        n = 1 << 21                  # 4n + 1.5n elements = 176 MiB > 96 MiB
        buf = ctx.alloc(16)
        ts = cosnark.Transcript()
        proof = np.zeros(int(L.cosnark_plain_proof_bytes(n)), np.uint8)
        rc = L.cosnark_prove_plain_noalloc(ctx.h, buf.ptr, buf.ptr, n, 4, ctypes.byref(ts), cosnark._p(proof),
                                           buf.ptr, buf.ptr)
        assert rc == cosnark.E_RANGE
        assert b"scratch" in L.cosnark_last_error(ctx.h)
        buf.free()
