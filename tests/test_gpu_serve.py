"""The proof server (cosnark_server_*): every proof is the same alone (the
plain prover, one request at a time) and in the mix, under continuous
batching with requests that arrive while others are on the GPU, under a
budget that admits only a few requests at a time, and with two client
threads submitting at once."""
import threading
import time

import numpy as np
import pytest

import cosnark
import pyoracle as O

pytestmark = pytest.mark.gpu


def _pool(ctx, sizes, seed):
    out = []
    for i, n in enumerate(sizes):
        f, g = O.synth(seed, 2 * i, n), O.synth(seed, 2 * i + 1, n)
        out.append((ctx.upload(f), ctx.upload(g), f, g))
    return out


def _alone(ctx, f, g, prev):
    ts = cosnark.Transcript()
    ts.prev[0], ts.prev[1] = prev
    p, _ = ctx.prove_plain(f, g, 4, ts)
    return p.tobytes(), ts.value()


def _free(pool):
    for a, b, _, _ in pool:
        a.free(); b.free()


def test_server_mix_matches_plain(ctx):
    rng = np.random.default_rng(5)
    sizes = [1 << int(e) for e in rng.integers(2, 17, 24)]
    print(f"server mix {sizes}", flush=True)
    pool = _pool(ctx, sizes, 0xC05A0030)
    prevs = [tuple(int(v) for v in O.synth(0xC05A0031, i, 1)[0]) for i in range(len(sizes))]
    srv = cosnark.Server(1 << 30, rate=4, max_log_n=16, slots=2)
    tickets = []
    for i, n in enumerate(sizes):
        tickets.append(srv.submit(pool[i][0], pool[i][1], n, prevs[i]))
        if i % 5 == 4:
            time.sleep(0.002)          # later requests arrive while earlier ones run
    results = [srv.wait(t, n) for t, n in zip(tickets, sizes)]
    st = srv.stats()
    srv.close()
    print(f"server stats {st}", flush=True)
    assert st["completed"] == len(sizes) and st["passes"] >= 2
    for i in range(len(sizes)):
        p, t = _alone(ctx, pool[i][2], pool[i][3], prevs[i])
        assert results[i][0] == p and results[i][1] == t, f"request {i} (n={sizes[i]})"
        assert results[i][2] > 0
    for i in range(3):
        if sizes[i] <= 1 << 12:
            assert results[i][0] == O.plain_proof(pool[i][2], pool[i][3], 4, np.array(prevs[i], np.uint64))[0]
    _free(pool)


def test_server_small_budget_admits_a_few_at_a_time(ctx):
    n, slots = 1 << 14, 2
    sizes = [n] * 12
    pool = _pool(ctx, sizes, 0xC05A0032)
    # two NTT temporaries of 4n elements, one request's memory, table room,
    # and space for about one more request
    work = 16 * 4 * n + 32 * 2 * (4 * n // 4) + 16 * 3 * n // 2 + (1 << 16)
    budget = slots * 16 * 4 * n + (2 << 20) + 2 * work + (1 << 20)
    print(f"server small budget {budget}", flush=True)
    srv = cosnark.Server(budget, rate=4, max_log_n=14, slots=slots)
    tickets = [srv.submit(pool[i][0], pool[i][1], n) for i in range(len(sizes))]
    results = [srv.wait(t, n) for t in tickets]
    st = srv.stats()
    srv.close()
    print(f"server stats {st}", flush=True)
    assert st["completed"] == len(sizes)
    assert st["max_inflight"] <= 3 and st["passes"] >= 4
    assert st["arena_peak_bytes"] <= st["arena_bytes"]
    for i in range(len(sizes)):
        assert results[i][0] == _alone(ctx, pool[i][2], pool[i][3], (0, 0))[0]
    _free(pool)


def test_server_two_client_threads(ctx):
    sizes = [1 << e for e in (12, 13, 15, 12, 14, 16, 12, 13)]
    pool = _pool(ctx, sizes, 0xC05A0033)
    srv = cosnark.Server(1 << 30, rate=4, max_log_n=16, slots=3)
    results = {}

    def client(idx):
        for i in idx:
            t = srv.submit(pool[i][0], pool[i][1], sizes[i])
            results[i] = srv.wait(t, sizes[i])

    th = [threading.Thread(target=client, args=(list(range(k, len(sizes), 2)),)) for k in (0, 1)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    srv.close()
    assert len(results) == len(sizes)
    for i in range(len(sizes)):
        assert results[i][0] == _alone(ctx, pool[i][2], pool[i][3], (0, 0))[0]
    _free(pool)


def test_server_rejects_oversize_and_bad_sizes(ctx):
    srv = cosnark.Server(1 << 28, rate=4, max_log_n=12, slots=1)
    a = ctx.alloc(16 << 13)
    with pytest.raises(cosnark.CosnarkError) as e:
        srv.submit(a, a, 1 << 13)
    assert e.value.code == cosnark.E_RANGE
    with pytest.raises(cosnark.CosnarkError) as e:
        srv.submit(a, a, 3)
    assert e.value.code == cosnark.E_ARG
    srv.close()
    a.free()
    with pytest.raises(cosnark.CosnarkError):
        cosnark.Server(1 << 20, rate=4, max_log_n=20, slots=2)     # budget below one request
