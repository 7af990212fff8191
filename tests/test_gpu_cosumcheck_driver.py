"""The coSumcheck drivers on the GPU (cosnark_quadratic_cosumcheck,
cosnark_cubic_cosumcheck; src/coSumcheck.cpp:233-448) at N >= 4 ranks:
N threads of this process act as the ranks, each with its own context and
its own shares, exchanging through an in-process all-gather in place of the
reference's MPI gather and broadcast.  Every rank's claims, challenge point,
ip_vec and folded vectors equal the oracle's (all ranks in lockstep), and
the reference's checks pass."""
import threading

import numpy as np
import pytest

import cosnark
import pyoracle as O
from cosc_inputs import cubic_inputs, quadratic_inputs

pytestmark = pytest.mark.gpu


def _run_ranks(N, body):
    ag = cosnark.ThreadAllGather(N)
    out, errs = [None] * N, []

    def worker(m):
        try:
            with cosnark.Context(0) as c:
                out[m] = body(c, m, ag.callback(m))
        except Exception as e:      # reported after the join
            errs.append((m, repr(e)))
            ag.barrier.abort()

    th = [threading.Thread(target=worker, args=(m,)) for m in range(N)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=240)
    assert not any(t.is_alive() for t in th), "a rank did not finish"
    assert not errs, errs
    return out


@pytest.mark.parametrize("N,M,degrees", [(4, 1 << 12, [16, 32, 4]), (8, 1 << 14, [256, 8, 4]),
                                         (4, 1 << 16, [256, 64, 2]), (16, 1 << 11, [1024])])
def test_quadratic_driver_n_ranks(N, M, degrees):
    seed = 0xC05A0080 + N
    p1, p2, R, K, K_r, _, _ = quadratic_inputs(N, M, degrees, seed)
    prev = np.zeros(2, np.uint64)
    rc, claims_o, r_o, ip_o, fails_o = O.quadratic_cosumcheck(list(p1), list(p2), R, K, K_r, degrees, prev)
    assert rc == 0 and fails_o == 0
    print(f"quadratic coSumcheck N={N} M={M} degrees={degrees}", flush=True)

    def body(c, m, cb):
        d1, d2 = c.upload(p1[m]), c.upload(p2[m])
        ts = cosnark.Transcript()
        res = c.quadratic_cosumcheck(d1, d2, M, R[m], K, K_r, degrees, N, m, cb, ts)
        return res + (ts.value(),)

    out = _run_ranks(N, body)
    for m in range(N):
        claims, r, ip, ps, fails, tv = out[m]
        assert fails == 0, (m, fails)
        assert np.array_equal(claims, claims_o) and np.array_equal(r, r_o) and np.array_equal(ip, ip_o), m
        assert tv == (int(prev[0]), int(prev[1]))
        assert (ps > 0) == (m == 0)


@pytest.mark.parametrize("N,logM,degrees", [(4, 13, [16, 512]), (8, 12, [8, 16, 32]), (4, 14, [1024, 4, 4])])
def test_cubic_driver_n_ranks(N, logM, degrees):
    seed = 0xC05A0090 + N + logM
    p1, p2, p3, r1, r2, R, K, K_r, _, _, _ = cubic_inputs(N, logM, degrees, seed)
    prev = np.zeros(2, np.uint64)
    rc, claims_o, r_o, ip_o, fails_o = O.cubic_cosumcheck(list(p1), list(p2), list(p3), r1, r2, R, K, K_r, degrees,
                                                          prev)
    assert rc == 0 and fails_o == 0
    M = 1 << logM
    print(f"cubic coSumcheck N={N} M={M} degrees={degrees}", flush=True)

    def body(c, m, cb):
        d1, d2, d3 = c.upload(p1[m]), c.upload(p2[m]), c.upload(p3[m])
        ts = cosnark.Transcript()
        return c.cubic_cosumcheck(d1, d2, d3, M, r1, r2, R[m], K, K_r, degrees, N, m, cb, ts)

    out = _run_ranks(N, body)
    for m in range(N):
        claims, r, ip, ps, fails = out[m]
        assert fails == 0, (m, fails)
        assert np.array_equal(claims, claims_o) and np.array_equal(r, r_o) and np.array_equal(ip, ip_o), m


def test_driver_rejects_fewer_than_four_ranks(ctx):
    M, degrees = 1 << 8, [8, 4]
    p1, p2, R, K, K_r, _, _ = quadratic_inputs(4, M, degrees, 5)
    d1, d2 = ctx.upload(p1[0]), ctx.upload(p2[0])
    ag = cosnark.ThreadAllGather(1)
    with pytest.raises(cosnark.CosnarkError):
        ctx.quadratic_cosumcheck(d1, d2, M, R[0], K, K_r, degrees, 2, 0, ag.callback(0), cosnark.Transcript())
