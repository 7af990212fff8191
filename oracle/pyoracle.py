"""ctypes wrapper of the CPU oracle (liboracle.so) and of oracle/_ref.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Element arrays are numpy
uint64 of shape (n, 2) = (real, img), the reference's element layout.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

P = (1 << 61) - 1


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class _Fe(ctypes.Structure):
    _fields_ = [("real", ctypes.c_uint64), ("img", ctypes.c_uint64)]


def fe(x):
    return _Fe(int(x[0]), int(x[1]))


def unfe(s):
    return np.array([s.real, s.img], dtype=np.uint64)


_o = None
_r = None
_libc = None


def build():
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def oracle():
    global _o
    if _o is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        for f in ("or_add", "or_sub", "or_mul"):
            getattr(L, f).argtypes = [_Fe, _Fe]
            getattr(L, f).restype = _Fe
        L.or_inv.argtypes = [_Fe]
        L.or_inv.restype = _Fe
        L.or_root_of_unity.argtypes = [i32]
        L.or_root_of_unity.restype = _Fe
        for f in ("or_vec_mul", "or_vec_add", "or_vec_sub"):
            getattr(L, f).argtypes = [vp, vp, vp, u64]
        L.or_vec_inv.argtypes = [vp, vp, u64]
        L.or_blake3_hash64.argtypes = [vp, vp]
        L.or_blake3_small.argtypes = [vp, ctypes.c_size_t, vp]
        L.or_blake3_hash64_many.argtypes = [vp, vp, u64]
        L.or_hash_F.argtypes = [_Fe, _Fe]
        L.or_hash_F.restype = _Fe
        L.or_hash_to_field.argtypes = [vp, vp, u64]
        L.or_hash_to_field.restype = _Fe
        L.or_get_indexes.argtypes = [vp, i32, ctypes.c_uint32, ctypes.c_uint32, vp]
        L.or_fft.argtypes = [vp, i32, i32]
        L.or_pfft.argtypes = [vp, i32]
        L.or_rs_encode.argtypes = [vp, u64, i32, vp]
        L.or_merkle_node_count.argtypes = [u64]
        L.or_merkle_node_count.restype = u64
        L.or_merkle_commit.argtypes = [vp, u64, vp, i32]
        L.or_merkle_open.argtypes = [vp, u64, u64, vp]
        L.or_merkle_open.restype = i32
        L.or_precompute_beta.argtypes = [vp, i32, vp]
        L.or_evaluate_vector.argtypes = [vp, u64, vp]
        L.or_evaluate_vector.restype = _Fe
        L.or_compute_lagrange_coeff.argtypes = [_Fe, i32, vp]
        L.or_random_element.restype = _Fe
        L.or_product_sumcheck.argtypes = [_Fe, vp, vp, u64, vp, i32, vp, vp, vp]
        L.or_product_sumcheck.restype = i32
        L.or_synth_fill.argtypes = [u64, u64, vp, u64]
        L.or_whir_query_count.argtypes = [i32]
        L.or_whir_query_count.restype = i32
        L.or_rs_multifold_segment_tensor.argtypes = [vp, i32, vp, i32, u64, u64]
        L.or_rs_multifold_segment_tensor.restype = _Fe
        L.or_whir_open.argtypes = [vp, u64, vp, _Fe, i32, i32, vp, u64]
        L.or_whir_open.restype = ctypes.c_int64
        L.or_copcs_encode.argtypes = [vp, i32, u64, i32, i32, vp]
        L.or_distributed_mt.argtypes = [vp, i32, u64, vp]
        L.or_pcs_sumcheck_phase1.argtypes = [vp, vp, u64, i32, vp]
        L.or_pcs_sumcheck_phase2.argtypes = [vp, vp, u64, i32, _Fe]
        L.or_rs_multifold5.argtypes = [vp, u64, vp, i32, vp]
        L.or_encode_concatenated.argtypes = [vp, u64, i32, i32, i32, vp, vp]
        L.or_rs_multifold_segment.argtypes = [vp, vp, i32, u64, u64]
        L.or_rs_multifold_segment.restype = _Fe
        L.or_row_combine.argtypes = [vp, i32, u64, vp, vp]
        L.or_cosumcheck3_phase1.argtypes = [vp, vp, vp, vp, u64, i32, vp]
        L.or_cosumcheck3_phase2.argtypes = [vp, vp, vp, vp, u64, i32, vp]
        L.or_unpack_secret_share.argtypes = [vp, i32, i32, vp]
        L.or_compute_secret_shares.argtypes = [vp, u64, i32, i32, i32, vp, vp]
        L.or_sumcheck_aggr.argtypes = [vp, i32, i32, vp, vp]
        L.or_whir_rates.argtypes = [vp, i32, i32, i32, vp]
        L.or_whir_aggregate_claims.argtypes = [i32, vp, u64, i32, vp, vp, vp, vp, vp, vp]
        L.or_whir_kappas.argtypes = [vp, i32, i32, vp, vp, i32]
        L.or_whir_kappas.restype = i32
        L.or_whir_batch_open.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, i32, i32, vp, u64]
        L.or_whir_batch_open.restype = ctypes.c_int64
        L.or_fft_mt.argtypes = [vp, i32, i32, i32]
        L.or_rs_encode_mt.argtypes = [vp, u64, i32, vp, i32]
        L.or_merkle_commit_mt.argtypes = [vp, u64, vp, i32, i32]
        L.or_product_sumcheck_mt.argtypes = [_Fe, vp, vp, u64, vp, vp, vp, vp, i32]
        L.or_product_sumcheck_mt.restype = i32
        L.or_plain_proof_mt.argtypes = [vp, vp, u64, i32, vp, vp, vp, i32]
        L.or_plain_proof_mt.restype = i32
        L.or_max_threads.restype = i32
        L.or_whir_set_threads.argtypes = [i32]
        L.or_whir_open_weights.argtypes = [vp, u64, vp, _Fe, vp, u64, vp, u64, i32, i32, vp, u64, vp, vp]
        L.or_whir_open_weights.restype = ctypes.c_int64
        L.or_prepare_data.argtypes = [vp, u64, vp, u64, i32, vp]
        L.or_prepare_data.restype = i32
        L.or_quadratic_cosumcheck.argtypes = [i32, vp, vp, u64, vp, vp, i32, _Fe, _Fe, vp, i32, vp, vp, vp, vp,
                                              ctypes.POINTER(i32)]
        L.or_quadratic_cosumcheck.restype = i32
        L.or_cubic_cosumcheck.argtypes = [i32, vp, vp, vp, u64, vp, i32, vp, i32, vp, vp, i32, _Fe, _Fe, vp, i32, vp,
                                          vp, vp, vp, ctypes.POINTER(i32)]
        L.or_cubic_cosumcheck.restype = i32
        L.or_update_betas_mt.argtypes = [vp, u64, vp, vp, i32, i32]
        _o = L
    return _o


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    """The reference's own compiled field / BLAKE3 / polynomial code."""
    global _r
    if _r is None:
        if not os.path.exists(REF_SO):
            build()
        if not os.path.exists(REF_SO):
            return None
        L = ctypes.CDLL(REF_SO)
        vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        L.ref_fe_layout.argtypes = [vp]
        for f in ("ref_mul", "ref_add", "ref_sub"):
            getattr(L, f).argtypes = [vp, vp, vp, u64]
        L.ref_inv.argtypes = [vp, vp, u64]
        L.ref_pow.argtypes = [vp, u64, u64, vp]
        L.ref_root_of_unity.argtypes = [i32, vp]
        L.ref_from_ll.argtypes = [ctypes.c_longlong, vp]
        L.ref_random_elements.argtypes = [vp, u64]
        L.ref_blake3_hash64.argtypes = [vp, vp]
        L.ref_blake3_any.argtypes = [vp, ctypes.c_size_t, vp]
        L.ref_linear_mul.argtypes = [vp, vp, vp]
        L.ref_quadratic_eval.argtypes = [vp, vp, vp]
        _r = L
    return _r


def libc():
    global _libc
    if _libc is None:
        _libc = ctypes.CDLL("libc.so.6")
        _libc.srandom.argtypes = [ctypes.c_uint]
        _libc.srand.argtypes = [ctypes.c_uint]
    return _libc


# ---- convenience wrappers -------------------------------------------------

def fe_array(n):
    return np.zeros((n, 2), dtype=np.uint64)


def synth(seed, stream, n):
    out = fe_array(n)
    oracle().or_synth_fill(seed, stream, _p(out), n)
    return out


def vec_op(name, a, b=None):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    out = np.zeros_like(a)
    if b is None:
        getattr(oracle(), "or_vec_" + name)(_p(a), _p(out), a.shape[0])
    else:
        b = np.ascontiguousarray(b, dtype=np.uint64)
        getattr(oracle(), "or_vec_" + name)(_p(a), _p(b), _p(out), a.shape[0])
    return out


def root_of_unity(logn):
    return unfe(oracle().or_root_of_unity(logn))


def blake3_64(block):
    b = np.ascontiguousarray(np.frombuffer(bytes(block), dtype=np.uint8))
    out = np.zeros(32, np.uint8)
    oracle().or_blake3_hash64(_p(b), _p(out))
    return out.tobytes()


def blake3_small(data: bytes):
    b = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8).copy()
    out = np.zeros(32, np.uint8)
    oracle().or_blake3_small(_p(b), len(data), _p(out))
    return out.tobytes()


def hash_F(x, y):
    return unfe(oracle().or_hash_F(fe(x), fe(y)))


def hash_to_field(prev, elems):
    """prev: array(2) state, updated in place; returns the new value."""
    st = np.ascontiguousarray(prev, dtype=np.uint64)
    e = np.ascontiguousarray(np.asarray(elems, dtype=np.uint64).reshape(-1, 2))
    r = oracle().or_hash_to_field(_p(st), _p(e), e.shape[0])
    prev[...] = st
    return unfe(r)


def fft(arr, inverse=False):
    a = np.ascontiguousarray(arr, dtype=np.uint64).copy()
    logn = a.shape[0].bit_length() - 1
    oracle().or_fft(_p(a), logn, int(inverse))
    return a


def rs_encode(arr, rate):
    a = np.ascontiguousarray(arr, dtype=np.uint64)
    out = fe_array(a.shape[0] * rate)
    oracle().or_rs_encode(_p(a), a.shape[0], rate, _p(out))
    return out


def merkle_commit(elems, dup_left=True):
    a = np.ascontiguousarray(elems, dtype=np.uint64)
    n = a.shape[0]
    cnt = int(oracle().or_merkle_node_count(n))
    nodes = np.zeros((cnt, 32), np.uint8)
    oracle().or_merkle_commit(_p(a), n, _p(nodes), int(dup_left))
    return nodes


def merkle_open(nodes, n_elems, pos):
    path = np.zeros((64, 32), np.uint8)
    d = oracle().or_merkle_open(_p(np.ascontiguousarray(nodes)), n_elems, pos, _p(path))
    if d < 0:
        raise ValueError("position out of range")
    return path[:d].copy()


def precompute_beta(r):
    r = np.ascontiguousarray(r, dtype=np.uint64)
    m = r.shape[0]
    B = fe_array(1 << m)
    oracle().or_precompute_beta(_p(r), m, _p(B))
    return B


def evaluate_vector(v, r):
    v = np.ascontiguousarray(v, dtype=np.uint64)
    r = np.ascontiguousarray(r, dtype=np.uint64)
    return unfe(oracle().or_evaluate_vector(_p(v), v.shape[0], _p(r)))


def product_sumcheck(y, v1, v2, prev, offset=0):
    """Returns (status, polys (m,3,2), r (m,2), final (2,2)); prev updated."""
    a = np.ascontiguousarray(v1, dtype=np.uint64).copy()
    b = np.ascontiguousarray(v2, dtype=np.uint64).copy()
    n = a.shape[0]
    m = n.bit_length() - 1
    st = np.ascontiguousarray(prev, dtype=np.uint64)
    polys = np.zeros((m, 3, 2), np.uint64)
    r = np.zeros((max(m, 1), 2), np.uint64)
    fin = np.zeros((2, 2), np.uint64)
    rc = oracle().or_product_sumcheck(fe(y), _p(a), _p(b), n, _p(st), offset, _p(polys), _p(r), _p(fin))
    prev[...] = st
    return rc, polys, r[:m], fin


def inner_product(a, b):
    """sum a[i]*b[i] in the field (Python big ints, for claims)."""
    prod = vec_op("mul", a, b)
    re = int(np.sum(prod[:, 0].astype(object)) % P)
    im = int(np.sum(prod[:, 1].astype(object)) % P)
    return np.array([re, im], dtype=np.uint64)


def plain_proof(f, g, rate, prev=None):
    """The plain single-rank prover's proof bytes: Merkle root of the
    rate-blowup codeword of f | per-round (a,b,c) | final v1[0], v2[0]."""
    if prev is None:
        prev = np.zeros(2, np.uint64)
    cw = rs_encode(f, rate)
    nodes = merkle_commit(cw, True)
    y = inner_product(f, g)
    rc, polys, r, fin = product_sumcheck(y, f, g, prev, 0)
    assert rc == 0
    return nodes[-1].tobytes() + polys.tobytes() + fin.tobytes(), y, r


def whir_open(arr, z, K, rate, stop_log=0, cap=1 << 22):
    """Whir_Open record of the oracle (draws from libc random(); seed first)."""
    a = np.ascontiguousarray(arr, dtype=np.uint64)
    zz = np.ascontiguousarray(z, dtype=np.uint64)
    rec = np.zeros(cap, np.uint64)
    w = oracle().or_whir_open(_p(a), a.shape[0], _p(zz), fe(K), rate, stop_log, _p(rec), cap)
    if w < 0:
        return int(w), None
    return 0, rec[:w].copy()


def whir_claim(arr, z):
    return inner_product(arr, precompute_beta(z))


# ---- coPCS rank-local kernels ----------------------------------------------

def copcs_encode(rows, rate, log_rate_offset=0):
    """rows: (nrows, L, 2).  Returns (nrows, rate*L, 2)."""
    r = np.ascontiguousarray(rows, dtype=np.uint64)
    nrows, L = r.shape[0], r.shape[1]
    out = np.zeros((nrows, rate * L, 2), np.uint64)
    oracle().or_copcs_encode(_p(r), nrows, L, rate, log_rate_offset, _p(out))
    return out


def distributed_mt(cw):
    """cw: (nrows <= 32, Lc, 2).  Returns the Base_MT digests (2*Lc-1, 32)."""
    c = np.ascontiguousarray(cw, dtype=np.uint64)
    nrows, Lc = c.shape[0], c.shape[1]
    nodes = np.zeros((2 * Lc - 1, 32), np.uint8)
    oracle().or_distributed_mt(_p(c), nrows, Lc, _p(nodes))
    return nodes


def pcs_sumcheck_phase1(data, B, d):
    a = np.ascontiguousarray(data, dtype=np.uint64)
    b = np.ascontiguousarray(B, dtype=np.uint64)
    poly = fe_array(2 * d)
    oracle().or_pcs_sumcheck_phase1(_p(a), _p(b), a.shape[0], d, _p(poly))
    return poly


def pcs_sumcheck_phase2(data, B, d, r):
    """Returns the folded (M/d) data and B (inputs are not modified)."""
    a = np.ascontiguousarray(data, dtype=np.uint64).copy()
    b = np.ascontiguousarray(B, dtype=np.uint64).copy()
    M = a.shape[0]
    oracle().or_pcs_sumcheck_phase2(_p(a), _p(b), M, d, fe(r))
    return a[:M // d].copy(), b[:M // d].copy()


def rs_multifold5(cw, a):
    c = np.ascontiguousarray(cw, dtype=np.uint64)
    aa = np.ascontiguousarray(a, dtype=np.uint64)
    fd = aa.shape[0]
    out = fe_array(c.shape[0] // fd)
    oracle().or_rs_multifold5(_p(c), c.shape[0], _p(aa), fd, _p(out))
    return out


def lagrange_coeff(r, d):
    L = fe_array(d)
    oracle().or_compute_lagrange_coeff(fe(r), d, _p(L))
    return L


def encode_concatenated(data, rate0, rate1, degree):
    """Returns (cprime (rate0*n, 2), C (S, degree*rate1, 2))."""
    d = np.ascontiguousarray(data, dtype=np.uint64)
    n = d.shape[0]
    m = n * rate0
    S, D = m // degree, degree * rate1
    cp = fe_array(m)
    C = np.zeros((S, D, 2), np.uint64)
    oracle().or_encode_concatenated(_p(d), n, rate0, rate1, degree, _p(cp), _p(C))
    return cp, C


def rs_multifold_segment(segment, L, n, idx):
    seg = np.ascontiguousarray(segment, dtype=np.uint64)
    LL = np.ascontiguousarray(L, dtype=np.uint64)
    return unfe(oracle().or_rs_multifold_segment(_p(seg), _p(LL), LL.shape[0], n, idx))


def row_combine(rows, c):
    r = np.ascontiguousarray(rows, dtype=np.uint64)
    cc = np.ascontiguousarray(c, dtype=np.uint64)
    out = fe_array(r.shape[1])
    oracle().or_row_combine(_p(r), r.shape[0], r.shape[1], _p(cc), _p(out))
    return out


def cosumcheck3_phase1(p1, p2, p3, b1, d):
    """Cubic block sum: 4d elements."""
    a = [np.ascontiguousarray(x, dtype=np.uint64) for x in (p1, p2, p3, b1)]
    H = fe_array(4 * d)
    oracle().or_cosumcheck3_phase1(_p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), a[0].shape[0], d, _p(H))
    return H


def cosumcheck3_phase2(p1, p2, p3, b1, d, L):
    """Returns the four folded (M/d) vectors (inputs are not modified)."""
    a = [np.ascontiguousarray(x, dtype=np.uint64).copy() for x in (p1, p2, p3, b1)]
    LL = np.ascontiguousarray(L, dtype=np.uint64)
    M = a[0].shape[0]
    oracle().or_cosumcheck3_phase2(_p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), M, d, _p(LL))
    return [x[:M // d].copy() for x in a]


def unpack_secret_share(secret, kk):
    """Returns (the transformed shares as the reference leaves them, the kk values)."""
    s = np.ascontiguousarray(secret, dtype=np.uint64).copy()
    v = fe_array(kk)
    oracle().or_unpack_secret_share(_p(s), s.shape[0], kk, _p(v))
    return s, v


def compute_secret_shares(v, N, k, kk, pads=None):
    """v: (M, 2).  Returns shares (N, M/k, 2)."""
    a = np.ascontiguousarray(v, dtype=np.uint64)
    M = a.shape[0]
    sh = np.zeros((N, M // k, 2), np.uint64)
    pp = None if pads is None else np.ascontiguousarray(pads, dtype=np.uint64)
    oracle().or_compute_secret_shares(_p(a), M, N, k, kk, None if pp is None else _p(pp), _p(sh))
    return sh


def sumcheck_aggr(polys, b):
    """polys: (N, len, 2) gathered per rank; b: (>= N/4, 2).  Returns (len, 2)."""
    P = np.ascontiguousarray(polys, dtype=np.uint64)
    N, ln = P.shape[0], P.shape[1]
    bb = np.ascontiguousarray(b, dtype=np.uint64) if b is not None and len(b) else fe_array(1)
    out = fe_array(ln)
    oracle().or_sumcheck_aggr(_p(P), N, ln, _p(bb), _p(out))
    return out


# ---- batched WHIR ---------------------------------------------------------

def whir_rates(sizes, rate, whir_k=4):
    """get_whir_rates: the commitment rate of each array (sizes largest first)."""
    sz = np.ascontiguousarray(sizes, dtype=np.uint64)
    out = np.zeros(len(sizes), np.int32)
    oracle().or_whir_rates(_p(sz), len(sizes), rate, whir_k, _p(out))
    return [int(x) for x in out]


def whir_kappas(sizes, whir_k=4):
    """Whir_Batch_Open's fold schedule: (kappas, aggr)."""
    sz = np.ascontiguousarray(sizes, dtype=np.uint64)
    k = np.zeros(64, np.int32)
    a = np.zeros(64, np.int32)
    nk = oracle().or_whir_kappas(_p(sz), len(sizes), whir_k, _p(k), _p(a), 64)
    return [int(x) for x in k[:nk]], [bool(x) for x in a[:nk]]


def whir_batch_inputs(arrays, zs, rate, whir_k=4):
    """The prover's inputs to Whir_Batch_Open: arrays (largest first), each
    committed at its get_whir_rates rate, its dup-left tree, eq weights
    precompute_beta(z_i) and claim <arr_i, beta_i>."""
    sizes = [int(np.asarray(a).shape[0]) for a in arrays]
    rates = whir_rates(sizes, rate, whir_k)
    cws = [rs_encode(a, r) for a, r in zip(arrays, rates)]
    trees = [merkle_commit(c) for c in cws]
    betas = [precompute_beta(z) for z in zs]
    evals = np.stack([inner_product(a, b) for a, b in zip(arrays, betas)])
    return {"arrays": [np.ascontiguousarray(a, dtype=np.uint64) for a in arrays], "sizes": sizes, "rates": rates,
            "codewords": cws, "trees": trees, "betas": betas, "evals": evals, "rate": rate, "whir_k": whir_k}


def _ptrs(arrs):
    return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def whir_batch_open(inp, cap=1 << 22):
    """Whir_Batch_Open record of the oracle (libc random()/rand(): seed first)."""
    n = len(inp["arrays"])
    sizes = np.ascontiguousarray(inp["sizes"], dtype=np.uint64)
    cws = [np.ascontiguousarray(c) for c in inp["codewords"]]
    cw_sizes = np.ascontiguousarray([c.shape[0] for c in cws], dtype=np.uint64)
    trees = [np.ascontiguousarray(t) for t in inp["trees"]]
    betas = [np.ascontiguousarray(b) for b in inp["betas"]]
    ev = np.ascontiguousarray(inp["evals"], dtype=np.uint64)
    rec = np.zeros(cap, np.uint64)
    w = oracle().or_whir_batch_open(n, _ptrs(inp["arrays"]), _p(sizes), _ptrs(cws), _p(cw_sizes), _ptrs(trees),
                                    _p(ev), _ptrs(betas), inp["rate"], inp["whir_k"], _p(rec), cap)
    if w < 0:
        return int(w), None
    return 0, rec[:w].copy()


def whir_aggregate_claims(arrays, points):
    """aggregate_claims: (arr, b_aggr, K, beta_ch, arr_ch); seed libc first."""
    arrs = [np.ascontiguousarray(a, dtype=np.uint64) for a in arrays]
    n = arrs[0].shape[0]
    pts = np.ascontiguousarray(points, dtype=np.uint64)
    arr, b = fe_array(n), fe_array(n)
    K = fe_array(1)
    bc, ac = fe_array(pts.shape[0]), fe_array(len(arrs))
    oracle().or_whir_aggregate_claims(len(arrs), _ptrs(arrs), n, pts.shape[0], _p(pts), _p(arr), _p(b), _p(K),
                                      _p(bc), _p(ac))
    return arr, b, K[0], bc, ac


# ---- all-cores variants (oracle_mt.c): the same bytes as the functions above

def max_threads():
    return int(oracle().or_max_threads())


def fft_mt(arr, inverse=False, threads=None):
    a = np.ascontiguousarray(arr, dtype=np.uint64).copy()
    oracle().or_fft_mt(_p(a), a.shape[0].bit_length() - 1, int(inverse), threads or max_threads())
    return a


def rs_encode_mt(arr, rate, threads=None):
    a = np.ascontiguousarray(arr, dtype=np.uint64)
    cw = fe_array(a.shape[0] * rate)
    oracle().or_rs_encode_mt(_p(a), a.shape[0], rate, _p(cw), threads or max_threads())
    return cw


def merkle_commit_mt(elems, dup_left=True, threads=None):
    e = np.ascontiguousarray(elems, dtype=np.uint64)
    cnt = int(oracle().or_merkle_node_count(e.shape[0]))
    nodes = np.zeros((cnt, 32), np.uint8)
    oracle().or_merkle_commit_mt(_p(e), e.shape[0], _p(nodes), int(dup_left), threads or max_threads())
    return nodes


def product_sumcheck_mt(y, v1, v2, prev, threads=None):
    """All-Fiat-Shamir product sumcheck on all cores: (status, polys, r, final)."""
    a = np.ascontiguousarray(v1, dtype=np.uint64)
    b = np.ascontiguousarray(v2, dtype=np.uint64)
    n = a.shape[0]
    m = n.bit_length() - 1
    st = np.ascontiguousarray(prev, dtype=np.uint64)
    polys = np.zeros((max(m, 1), 3, 2), np.uint64)
    r = np.zeros((max(m, 1), 2), np.uint64)
    fin = np.zeros((2, 2), np.uint64)
    rc = oracle().or_product_sumcheck_mt(fe(y), _p(a), _p(b), n, _p(st), _p(polys), _p(r), _p(fin),
                                         threads or max_threads())
    prev[...] = st
    return rc, polys[:m], r[:m], fin


def plain_proof_mt(f, g, rate, prev=None, threads=None):
    """plain_proof on all cores: (proof bytes, stage ms [encode, merkle, sumcheck, total])."""
    a = np.ascontiguousarray(f, dtype=np.uint64)
    b = np.ascontiguousarray(g, dtype=np.uint64)
    n = a.shape[0]
    pv = np.zeros(2, np.uint64) if prev is None else np.ascontiguousarray(prev, dtype=np.uint64)
    proof = np.zeros(64 + 48 * (n.bit_length() - 1), np.uint8)
    ms = np.zeros(4, np.float64)
    rc = oracle().or_plain_proof_mt(_p(a), _p(b), n, rate, _p(pv), _p(proof), _p(ms), threads or max_threads())
    if rc:
        raise AssertionError(f"oracle plain proof: round check failed ({rc})")
    if prev is not None:
        prev[...] = pv
    return proof.tobytes(), ms


def whir_threads(t):
    """Thread count of the WHIR restatement's encodes, trees and weight
    updates (1: single-thread; the bytes are the same either way)."""
    oracle().or_whir_set_threads(int(t))


def whir_open_weights(arr, b, K, cw, nodes, zero_leaves, rate, whir_k, cap=1 << 22):
    """_Whir_Open of the oracle: (status or words, record, total_r (log2 n, 2),
    final claim).  nodes None: the first tree is zero_leaves all-zero leaf
    digests.  Seed libc first."""
    a = np.ascontiguousarray(arr, dtype=np.uint64)
    bb = np.ascontiguousarray(b, dtype=np.uint64)
    c = np.ascontiguousarray(cw, dtype=np.uint64)
    nd = None if nodes is None else np.ascontiguousarray(nodes, dtype=np.uint8)
    n = a.shape[0]
    rec = np.zeros(cap, np.uint64)
    chal = np.zeros((max(n.bit_length() - 1, 1), 2), np.uint64)
    fin = _Fe(0, 0)
    w = oracle().or_whir_open_weights(_p(a), n, _p(bb), fe(K), _p(c), c.shape[0], None if nd is None else _p(nd),
                                      zero_leaves, rate, whir_k, _p(rec), cap, _p(chal), ctypes.byref(fin))
    if w < 0:
        return int(w), None, None, None
    return int(w), rec[:w].copy(), chal[:n.bit_length() - 1], unfe(fin)


def _mask_ptrs(R):
    """R[m][i] arrays -> or_fe*** (rank m, iteration i), kept alive by the caller."""
    rows = [(ctypes.c_void_p * len(Rm))(*[a.ctypes.data for a in Rm]) for Rm in R]
    top = (ctypes.c_void_p * len(R))(*[ctypes.cast(r, ctypes.c_void_p).value for r in rows])
    return top, rows


def quadratic_cosumcheck(p1, p2, R, K, K_r, degrees, prev):
    """quadratic_coSumcheck over N ranks in lockstep.  p1, p2: lists of N
    share vectors (copied); R: N lists of masks (copied).  Returns (status,
    claims (5, 2), sumcheck_r (ndeg + 1, 2), ip_vec, failures); prev updated."""
    N = len(p1)
    a = [np.ascontiguousarray(x, dtype=np.uint64).copy() for x in p1]
    b = [np.ascontiguousarray(x, dtype=np.uint64).copy() for x in p2]
    RR = [[np.ascontiguousarray(x, dtype=np.uint64).copy() for x in Rm] for Rm in R]
    top, keep = _mask_ptrs(RR)
    sizes = np.ascontiguousarray([x.shape[0] for x in RR[0]], dtype=np.uint64)
    deg = np.ascontiguousarray(degrees, dtype=np.int32)
    claims = np.zeros((5, 2), np.uint64)
    r = np.zeros((len(degrees) + 1, 2), np.uint64)
    ip = np.zeros((int(sizes.sum()), 2), np.uint64)
    fails = ctypes.c_int(0)
    st = np.ascontiguousarray(prev, dtype=np.uint64)
    rc = oracle().or_quadratic_cosumcheck(N, _ptrs(a), _ptrs(b), a[0].shape[0], top, _p(sizes), len(RR[0]), fe(K),
                                          fe(K_r), _p(deg), len(degrees), _p(st), _p(claims), _p(r), _p(ip),
                                          ctypes.byref(fails))
    prev[...] = st
    return rc, claims, r, ip, fails.value


def cubic_cosumcheck(p1, p2, p3, r1, r2, R, K, K_r, degrees, prev):
    """cubic_coSumcheck over N ranks in lockstep: (status, claims (4, 2),
    sumcheck_r, ip_vec, failures); prev updated."""
    N = len(p1)
    a, b, c = ([np.ascontiguousarray(x, dtype=np.uint64).copy() for x in v] for v in (p1, p2, p3))
    RR = [[np.ascontiguousarray(x, dtype=np.uint64).copy() for x in Rm] for Rm in R]
    top, keep = _mask_ptrs(RR)
    sizes = np.ascontiguousarray([x.shape[0] for x in RR[0]], dtype=np.uint64)
    deg = np.ascontiguousarray(degrees, dtype=np.int32)
    rr1 = np.ascontiguousarray(r1, dtype=np.uint64)
    rr2 = np.ascontiguousarray(r2, dtype=np.uint64)
    claims = np.zeros((4, 2), np.uint64)
    r = np.zeros((len(degrees) + 1, 2), np.uint64)
    ip = np.zeros((int(sizes.sum()), 2), np.uint64)
    fails = ctypes.c_int(0)
    st = np.ascontiguousarray(prev, dtype=np.uint64)
    rc = oracle().or_cubic_cosumcheck(N, _ptrs(a), _ptrs(b), _ptrs(c), a[0].shape[0], _p(rr1), rr1.shape[0],
                                      _p(rr2), rr2.shape[0], top, _p(sizes), len(RR[0]), fe(K), fe(K_r), _p(deg),
                                      len(degrees), _p(st), _p(claims), _p(r), _p(ip), ctypes.byref(fails))
    prev[...] = st
    return rc, claims, r, ip, fails.value


def prepare_data(shares, rnd, q0):
    """prepare_data: (row count or -1, rows (32, n/16, 2))."""
    a = np.ascontiguousarray(shares, dtype=np.uint64)
    r = np.ascontiguousarray(rnd, dtype=np.uint64)
    n = a.shape[0]
    rows = np.zeros((32, n // 16, 2), np.uint64)
    c = oracle().or_prepare_data(_p(a), n, _p(r), r.shape[0], q0, _p(rows))
    return int(c), rows


def pcs_commit(shares, rnd, q0, rate, log_rate_offset):
    """PCS_commit at one rank: (row count, rows, codewords (rows, rate B, 2),
    Base_MT nodes)."""
    c, rows = prepare_data(shares, rnd, q0)
    assert c > 0
    cw = copcs_encode(rows[:c], rate, log_rate_offset)
    return c, rows[:c], cw, distributed_mt(cw)
