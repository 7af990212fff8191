"""ctypes binding of libcosnark.so (the C ABI in include/cosnark.h).

Host arrays are numpy uint64 arrays of shape (n, 2) -- one row per field
element, columns (real, img) -- the byte layout of the reference's
virgo::fieldElement (src/fieldElement.hpp:96-97).  Digests are uint8 (k, 32).

The library is required: importing this module on a machine without the
built libcosnark.so raises, there is no fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcosnark.so")

P = (1 << 61) - 1

E_OK, E_ARG, E_HIP, E_NOMEM, E_CHECK, E_RANGE = 0, -1, -2, -3, -4, -5


class CosnarkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cosnark error {code}: {msg}")
        self.code = code


_lib = None

# cosnark_allgather_fn: (user, mine, elems, all) -> 0 on success
ALLGATHER_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64,
                                ctypes.POINTER(ctypes.c_uint64))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libcosnark.so not built at {LIB_PATH}; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, i32, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32
        L.cosnark_create.argtypes = [i32, ctypes.POINTER(vp)]
        L.cosnark_destroy.argtypes = [vp]
        L.cosnark_destroy.restype = None
        L.cosnark_last_error.argtypes = [vp]
        L.cosnark_last_error.restype = ctypes.c_char_p
        L.cosnark_sync.argtypes = [vp]
        L.cosnark_stream.argtypes = [vp]
        L.cosnark_stream.restype = vp
        L.cosnark_version.restype = ctypes.c_char_p
        L.cosnark_set_overlap.argtypes = [vp, i32]
        L.cosnark_dev_alloc.argtypes = [vp, u64, ctypes.POINTER(vp)]
        L.cosnark_dev_free.argtypes = [vp, vp]
        for f in ("cosnark_h2d", "cosnark_d2h", "cosnark_d2d"):
            getattr(L, f).argtypes = [vp, vp, vp, u64]
        L.cosnark_synth_fill.argtypes = [u64, u64, vp, u64]
        L.cosnark_synth_fill.restype = None
        L.cosnark_ntt.argtypes = [vp, vp, i32, i32]
        L.cosnark_rs_encode.argtypes = [vp, vp, u64, i32, vp]
        L.cosnark_merkle_node_count.argtypes = [u64]
        L.cosnark_merkle_node_count.restype = u64
        L.cosnark_merkle_commit.argtypes = [vp, vp, u64, vp, i32]
        L.cosnark_merkle_tree_from_leaves.argtypes = [vp, u64, vp, i32]
        L.cosnark_merkle_open.argtypes = [vp, vp, u64, u64, vp, ctypes.POINTER(i32)]
        L.cosnark_hash_to_field.argtypes = [vp, vp, u64, vp]
        L.cosnark_get_indexes.argtypes = [vp, i32, u32, u32, vp]
        L.cosnark_product_sumcheck.argtypes = [vp, vp, vp, u64, vp, vp, i32, vp, vp, vp]
        L.cosnark_plain_proof_bytes.argtypes = [u64]
        L.cosnark_plain_proof_bytes.restype = u64
        L.cosnark_prove_plain.argtypes = [vp, vp, vp, u64, i32, vp, vp, vp]
        L.cosnark_prove_plain_ex.argtypes = [vp, vp, vp, u64, i32, vp, vp, vp, vp, vp]
        L.cosnark_whir_commit.argtypes = [vp, vp, u64, i32, vp, vp]
        L.cosnark_whir_claim.argtypes = [vp, vp, u64, vp, vp]
        L.cosnark_whir_open.argtypes = [vp, vp, u64, vp, vp, i32, i32, vp, vp, vp, u64, ctypes.POINTER(u64)]
        L.cosnark_copcs_encode.argtypes = [vp, vp, i32, u64, i32, i32, vp]
        L.cosnark_distributed_mt.argtypes = [vp, vp, i32, u64, vp]
        L.cosnark_encode_concatenated.argtypes = [vp, vp, u64, i32, i32, i32, vp, vp]
        L.cosnark_multifold_segments.argtypes = [vp, vp, u64, vp, i32, vp, i32, vp, vp]
        L.cosnark_row_combine.argtypes = [vp, vp, i32, u64, vp, vp]
        L.cosnark_cosumcheck_phase1.argtypes = [vp, vp, vp, u64, i32, vp]
        L.cosnark_cosumcheck_phase2.argtypes = [vp, vp, vp, u64, i32, vp]
        L.cosnark_cosumcheck_fold.argtypes = [vp, vp, vp, u64, i32, vp]
        L.cosnark_cosumcheck3_phase1.argtypes = [vp, vp, vp, vp, vp, u64, i32, vp]
        L.cosnark_cosumcheck3_fold.argtypes = [vp, vp, vp, vp, vp, u64, i32, vp]
        L.cosnark_cosumcheck3_phase2.argtypes = [vp, vp, vp, vp, vp, u64, i32, vp]
        L.cosnark_share_unpack.argtypes = [vp, i32, i32, vp]
        L.cosnark_whir_rates.argtypes = [vp, i32, i32, i32, vp]
        L.cosnark_precompute_beta.argtypes = [vp, vp, i32, vp]
        L.cosnark_set_ntt_plan.argtypes = [vp, i32, i32, i32]
        L.cosnark_merkle_verify.argtypes = [vp, vp, vp, vp, i32, u64, vp, i32, vp, ctypes.POINTER(u64)]
        L.cosnark_verify_plain.argtypes = [vp, u64, vp, vp, vp, ctypes.POINTER(i32)]
        L.cosnark_whir_aggregate_claims.argtypes = [vp, i32, vp, u64, i32, vp, vp, vp, vp, vp, vp]
        L.cosnark_whir_open_weights.argtypes = [vp, vp, u64, vp, vp, vp, u64, vp, u64, i32, i32, vp, u64,
                                                ctypes.POINTER(u64), vp, vp]
        L.cosnark_quadratic_cosumcheck.argtypes = [vp, vp, vp, u64, vp, vp, i32, vp, vp, vp, i32, i32, i32,
                                                   ALLGATHER_FN, vp, vp, vp, vp, vp,
                                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i32)]
        L.cosnark_cubic_cosumcheck.argtypes = [vp, vp, vp, vp, u64, vp, i32, vp, i32, vp, vp, i32, vp, vp, vp, i32,
                                               i32, i32, ALLGATHER_FN, vp, vp, vp, vp, vp,
                                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i32)]
        L.cosnark_pcs_commit.argtypes = [vp, vp, u64, vp, u64, i32, i32, i32, vp, vp, vp, ctypes.POINTER(i32)]
        L.cosnark_pcs_prepare_rows.argtypes = [u64, u64, i32, ctypes.POINTER(i32), ctypes.POINTER(u64)]
        L.cosnark_field_random.argtypes = [vp]
        L.cosnark_field_random.restype = None
        L.cosnark_blake3_64.argtypes = [vp, vp]
        L.cosnark_blake3_64.restype = None
        L.cosnark_whir_batch_open.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp, vp, i32, i32, vp, u64,
                                              ctypes.POINTER(u64)]
        L.cosnark_sumcheck_aggr.argtypes = [vp, i32, i32, vp, vp]
        L.cosnark_rs_multifold5.argtypes = [vp, vp, u64, vp, i32, vp]
        L.cosnark_batch_create.argtypes = [vp, i32, vp, i32, ctypes.POINTER(vp)]
        L.cosnark_batch_destroy.argtypes = [vp]
        L.cosnark_batch_destroy.restype = None
        L.cosnark_batch_input.argtypes = [vp, i32, ctypes.POINTER(vp), ctypes.POINTER(vp)]
        L.cosnark_batch_proof_offset.argtypes = [vp, i32]
        L.cosnark_batch_proof_offset.restype = u64
        L.cosnark_batch_commitment.argtypes = [vp, i32, ctypes.POINTER(vp), ctypes.POINTER(vp)]
        L.cosnark_batch_set_graph.argtypes = [vp, i32]
        L.cosnark_batch_prove.argtypes = [vp, vp, vp, ctypes.POINTER(ctypes.c_double)]
        L.cosnark_autotune.argtypes = [vp, i32, i32, i32, vp]
        L.cosnark_ntt_plan.argtypes = [vp, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
        L.cosnark_autotune_merkle.argtypes = [vp, vp, i32, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_double)]
        L.cosnark_merkle_schedule.argtypes = [vp, ctypes.POINTER(i32)]
        L.cosnark_set_merkle_schedule.argtypes = [vp, i32]
        L.cosnark_fmul_rate.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.cosnark_blake3_rate.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
        L.cosnark_server_create.argtypes = [i32, u64, i32, i32, i32, ctypes.POINTER(vp)]
        L.cosnark_server_submit.argtypes = [vp, vp, vp, u64, vp, ctypes.POINTER(u64)]
        L.cosnark_server_submit_ex.argtypes = [vp, vp, vp, u64, vp, vp, vp, ctypes.POINTER(u64)]
        L.cosnark_server_poll.argtypes = [vp, u64, ctypes.POINTER(i32)]
        L.cosnark_server_wait.argtypes = [vp, u64, vp, vp, ctypes.POINTER(ctypes.c_double)]
        L.cosnark_server_stats.argtypes = [vp, vp]
        L.cosnark_server_destroy.argtypes = [vp]
        L.cosnark_server_destroy.restype = None
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def fe_array(n):
    return np.zeros((n, 2), dtype=np.uint64)


def synth(seed, stream, n):
    """n uniform GF(p^2) elements of the workload generator."""
    out = fe_array(n)
    lib().cosnark_synth_fill(seed, stream, _p(out), n)
    return out


class Transcript(ctypes.Structure):
    """The reference's running prev_challenge (src/Fiat_Shamir.cpp:5)."""
    _fields_ = [("prev", ctypes.c_uint64 * 2)]

    def value(self):
        return (int(self.prev[0]), int(self.prev[1]))


def hash_to_field(ts: Transcript, elems):
    e = np.ascontiguousarray(np.asarray(elems, dtype=np.uint64).reshape(-1, 2))
    out = np.zeros(2, np.uint64)
    rc = lib().cosnark_hash_to_field(ctypes.byref(ts), _p(e), e.shape[0], _p(out))
    if rc:
        raise CosnarkError(rc, "hash_to_field")
    return out


def get_indexes(ts: Transcript, l, N, M):
    idx = np.zeros((l, 2), np.uint32)
    rc = lib().cosnark_get_indexes(ctypes.byref(ts), l, N, M, _p(idx))
    if rc:
        raise CosnarkError(rc, "get_indexes")
    return idx


def merkle_node_count(n_elems):
    return int(lib().cosnark_merkle_node_count(n_elems))


class DevBuf:
    """A device allocation owned by a Context."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        ctx._ck(lib().cosnark_dev_alloc(ctx.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p
        ctx._bufs.add(self)

    def free(self):
        if self.ptr and self.ctx.h:
            lib().cosnark_dev_free(self.ctx.h, self.ptr)
        self.ptr = None
        self.ctx._bufs.discard(self)

    def __hash__(self):
        return id(self)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    def __init__(self, device=0):
        h = ctypes.c_void_p()
        rc = lib().cosnark_create(device, ctypes.byref(h))
        if rc:
            raise CosnarkError(rc, "cosnark_create failed (no GPU?)")
        self.h = h
        self._bufs = set()

    def close(self):
        if self.h:
            for b in list(self._bufs):
                b.free()
            lib().cosnark_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ck(self, rc):
        if rc:
            raise CosnarkError(rc, lib().cosnark_last_error(self.h).decode())
        return rc

    # ---- memory ----
    def alloc(self, nbytes):
        return DevBuf(self, nbytes)

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        b = DevBuf(self, a.nbytes)
        self._ck(lib().cosnark_h2d(self.h, b.ptr, _p(a), a.nbytes))
        return b

    def download(self, buf, shape, dtype=np.uint64):
        out = np.zeros(shape, dtype=dtype)
        self._ck(lib().cosnark_d2h(self.h, _p(out), buf.ptr, out.nbytes))
        return out

    def sync(self):
        self._ck(lib().cosnark_sync(self.h))

    # ---- kernels ----
    def ntt(self, arr, inverse=False):
        a = np.ascontiguousarray(arr, dtype=np.uint64)
        n = a.shape[0]
        logn = n.bit_length() - 1
        b = self.upload(a)
        self._ck(lib().cosnark_ntt(self.h, b.ptr, logn, int(inverse)))
        return self.download(b, a.shape)

    def rs_encode(self, coeffs, rate):
        a = np.ascontiguousarray(coeffs, dtype=np.uint64)
        n = a.shape[0]
        b = self.upload(a)
        cw = self.alloc(16 * n * rate)
        self._ck(lib().cosnark_rs_encode(self.h, b.ptr, n, rate, cw.ptr))
        return self.download(cw, (n * rate, 2))

    def merkle_commit(self, elems, dup_left=True, keep_device=False):
        a = np.ascontiguousarray(elems, dtype=np.uint64)
        n = a.shape[0]
        b = self.upload(a)
        cnt = merkle_node_count(n)
        nodes = self.alloc(32 * cnt)
        self._ck(lib().cosnark_merkle_commit(self.h, b.ptr, n, nodes.ptr, int(dup_left)))
        host = self.download(nodes, (cnt, 32), np.uint8)
        if keep_device:
            return host, nodes
        return host

    def merkle_open(self, nodes_dev, n_elems, pos):
        path = np.zeros((64, 32), np.uint8)
        depth = ctypes.c_int()
        self._ck(lib().cosnark_merkle_open(self.h, nodes_dev.ptr, n_elems, pos, _p(path), ctypes.byref(depth)))
        return path[: depth.value].copy()

    def product_sumcheck(self, v1, v2, y, ts: Transcript, offset=0):
        a = np.ascontiguousarray(v1, dtype=np.uint64)
        b = np.ascontiguousarray(v2, dtype=np.uint64)
        n = a.shape[0]
        m = n.bit_length() - 1
        d1, d2 = self.upload(a), self.upload(b)
        yy = np.asarray(y, dtype=np.uint64).reshape(2)
        polys = np.zeros((m, 3, 2), np.uint64)
        r = np.zeros((m, 2), np.uint64)
        fin = np.zeros((2, 2), np.uint64)
        rc = lib().cosnark_product_sumcheck(self.h, d1.ptr, d2.ptr, n, _p(yy), ctypes.byref(ts), offset,
                                            _p(polys), _p(r), _p(fin))
        return rc, polys, r, fin

    def prove_plain(self, f, g, rate, ts: Transcript, dev_inputs=None, stage_times=True):
        """Returns (proof bytes, stage times ms [encode, merkle, sumcheck, total]);
        with stage_times=False no stage events are recorded and the times are zero."""
        if dev_inputs is None:
            a = np.ascontiguousarray(f, dtype=np.uint64)
            b = np.ascontiguousarray(g, dtype=np.uint64)
            n = a.shape[0]
            df, dg = self.upload(a), self.upload(b)
        else:
            df, dg, n = dev_inputs
        nb = int(lib().cosnark_plain_proof_bytes(n))
        proof = np.zeros(nb, np.uint8)
        times = np.zeros(4, np.float64)
        self._ck(lib().cosnark_prove_plain(self.h, df.ptr, dg.ptr, n, rate, ctypes.byref(ts), _p(proof),
                                           _p(times) if stage_times else None))
        return proof, times

    def prove_plain_keep(self, df, dg, n, rate, ts: Transcript):
        """cosnark_prove_plain_ex on device inputs: (proof bytes, codeword
        DevBuf (rate*n elements), nodes DevBuf (every tree level))."""
        nb = int(lib().cosnark_plain_proof_bytes(n))
        proof = np.zeros(nb, np.uint8)
        cw = self.alloc(16 * n * rate)
        nodes = self.alloc(32 * merkle_node_count(n * rate))
        self._ck(lib().cosnark_prove_plain_ex(self.h, df.ptr, dg.ptr, n, rate, ctypes.byref(ts), _p(proof),
                                              cw.ptr, nodes.ptr, None))
        return proof, cw, nodes

    def download_ptr(self, ptr, shape, dtype=np.uint64):
        """Download from a raw device address (int)."""
        out = np.zeros(shape, dtype=dtype)
        self._ck(lib().cosnark_d2h(self.h, _p(out), ctypes.c_void_p(ptr), out.nbytes))
        return out

    # ---- WHIR ----
    def whir_commit(self, arr_dev, n, rate):
        """Whir_Commit_initial: (codeword, nodes) device buffers."""
        cw = self.alloc(16 * n * rate)
        nodes = self.alloc(32 * merkle_node_count(n * rate))
        self._ck(lib().cosnark_whir_commit(self.h, arr_dev.ptr, n, rate, cw.ptr, nodes.ptr))
        return cw, nodes

    def whir_claim(self, arr_dev, n, z):
        zz = np.ascontiguousarray(z, dtype=np.uint64)
        K = np.zeros(2, np.uint64)
        self._ck(lib().cosnark_whir_claim(self.h, arr_dev.ptr, n, _p(zz), _p(K)))
        return K

    def whir_open(self, arr_dev, n, z, K, rate, stop_log, cw, nodes, cap=1 << 22):
        """Returns (status, record words)."""
        zz = np.ascontiguousarray(z, dtype=np.uint64)
        KK = np.ascontiguousarray(K, dtype=np.uint64).reshape(2)
        rec = np.zeros(cap, np.uint64)
        words = ctypes.c_uint64()
        rc = lib().cosnark_whir_open(self.h, arr_dev.ptr, n, _p(zz), _p(KK), rate, stop_log, cw.ptr, nodes.ptr,
                                     _p(rec), cap, ctypes.byref(words))
        return rc, rec[: words.value].copy()

    # ---- coPCS rank-local kernels ----------------------------------------

    def copcs_encode(self, rows, rate, log_rate_offset=0):
        """Encode (src/coPCS.cpp:131-147) of rows (nrows, L, 2) -> (nrows, rate*L, 2)."""
        r = np.ascontiguousarray(rows, dtype=np.uint64)
        nrows, L = r.shape[0], r.shape[1]
        d = self.upload(r)
        out = self.alloc(16 * nrows * L * rate)
        self._ck(lib().cosnark_copcs_encode(self.h, d.ptr, nrows, L, rate, log_rate_offset, out.ptr))
        res = self.download(out, (nrows, L * rate, 2))
        d.free(); out.free()
        return res

    def encode_concatenated(self, data, rate0, rate1, degree):
        """encode_concatenated (src/coPCS.cpp:653-677): (cprime, C (S, degree*rate1, 2))."""
        d = np.ascontiguousarray(data, dtype=np.uint64)
        n = d.shape[0]
        m = n * rate0
        S, D = m // degree, degree * rate1
        a = self.upload(d)
        cp, C = self.alloc(16 * m), self.alloc(16 * S * D)
        self._ck(lib().cosnark_encode_concatenated(self.h, a.ptr, n, rate0, rate1, degree, cp.ptr, C.ptr))
        r = self.download(cp, (m, 2)), self.download(C, (S, D, 2))
        a.free(); cp.free(); C.free()
        return r

    def multifold_segments(self, cw, L, idx):
        """delegate_checks' segments and folds: (segments (l, fd, 2), y (l, 2))."""
        c = np.ascontiguousarray(cw, dtype=np.uint64)
        LL = np.ascontiguousarray(L, dtype=np.uint64)
        ii = np.ascontiguousarray(idx, dtype=np.uint32)
        n, fd, l = c.shape[0], LL.shape[0], ii.shape[0]
        d = self.upload(c)
        seg = np.zeros((l, fd, 2), np.uint64)
        y = np.zeros((l, 2), np.uint64)
        self._ck(lib().cosnark_multifold_segments(self.h, d.ptr, n, _p(LL), fd, _p(ii), l, _p(seg), _p(y)))
        d.free()
        return seg, y

    def row_combine(self, rows, c):
        r = np.ascontiguousarray(rows, dtype=np.uint64)
        cc = np.ascontiguousarray(c, dtype=np.uint64)
        d = self.upload(r)
        out = self.alloc(16 * r.shape[1])
        self._ck(lib().cosnark_row_combine(self.h, d.ptr, r.shape[0], r.shape[1], _p(cc), out.ptr))
        res = self.download(out, (r.shape[1], 2))
        d.free(); out.free()
        return res

    def distributed_mt(self, cw):
        """distributed_MT (src/coPCS.cpp:152-183) of cw (nrows, Lc, 2): Base_MT digests (2Lc-1, 32)."""
        c = np.ascontiguousarray(cw, dtype=np.uint64)
        nrows, Lc = c.shape[0], c.shape[1]
        d = self.upload(c)
        nodes = self.alloc(32 * (2 * Lc - 1))
        self._ck(lib().cosnark_distributed_mt(self.h, d.ptr, nrows, Lc, nodes.ptr))
        res = self.download(nodes, (2 * Lc - 1, 32), np.uint8)
        d.free(); nodes.free()
        return res

    def cosumcheck_phase1(self, p1, p2, d, dev=None):
        """Block sumcheck phase 1: the 2d-element round polynomial (2d, 2)."""
        if dev is None:
            a, b = self.upload(p1), self.upload(p2)
            M = np.asarray(p1).shape[0]
        else:
            a, b, M = dev
        out = np.zeros((2 * d, 2), np.uint64)
        self._ck(lib().cosnark_cosumcheck_phase1(self.h, a.ptr, b.ptr, M, d, _p(out)))
        if dev is None:
            a.free(); b.free()
        return out

    def whir_batch_open(self, arrays, codewords, trees, evals, betas, rate, whir_k=4, cap=1 << 22):
        """Whir_Batch_Open (src/Whir.cpp:705-899) over host arrays (largest
        first) with their codewords, trees (node arrays), claims and eq
        weights; uploads them, returns (status, record).  Seed libc first."""
        ups = []
        def up(x):
            b = self.upload(np.ascontiguousarray(x))
            ups.append(b)
            return b.ptr.value
        n = len(arrays)
        pa = (ctypes.c_void_p * n)(*[up(a) for a in arrays])
        pc = (ctypes.c_void_p * n)(*[up(c) for c in codewords])
        pt = (ctypes.c_void_p * n)(*[up(t) for t in trees])
        pb = (ctypes.c_void_p * n)(*[up(b) for b in betas])
        sizes = np.ascontiguousarray([np.asarray(a).shape[0] for a in arrays], dtype=np.uint64)
        cws = np.ascontiguousarray([np.asarray(c).shape[0] for c in codewords], dtype=np.uint64)
        ev = np.ascontiguousarray(evals, dtype=np.uint64)
        rec = np.zeros(cap, np.uint64)
        words = ctypes.c_uint64(0)
        rc = lib().cosnark_whir_batch_open(self.h, n, pa, _p(sizes), pc, _p(cws), pt, _p(ev), pb, rate, whir_k,
                                           _p(rec), cap, ctypes.byref(words))
        for b in ups:
            b.free()
        return rc, rec[:words.value].copy()

    def whir_open_weights(self, arr_dev, n, b_dev, K, cw_dev, cw_n, nodes_dev, zero_leaves, rate, whir_k,
                          cap=1 << 22):
        """_Whir_Open (cosnark_whir_open_weights): (status, record, total_r,
        final claim).  nodes_dev None: the first tree is zero_leaves all-zero
        leaf digests.  Seed libc first."""
        KK = np.ascontiguousarray(K, dtype=np.uint64).reshape(2)
        rec = np.zeros(cap, np.uint64)
        words = ctypes.c_uint64(0)
        m = n.bit_length() - 1
        chal = np.zeros((max(m, 1), 2), np.uint64)
        fin = np.zeros(2, np.uint64)
        rc = lib().cosnark_whir_open_weights(self.h, arr_dev.ptr, n, b_dev.ptr, _p(KK), cw_dev.ptr, cw_n,
                                             None if nodes_dev is None else nodes_dev.ptr, zero_leaves, rate, whir_k,
                                             _p(rec), cap, ctypes.byref(words), _p(chal), _p(fin))
        return rc, rec[:words.value].copy(), chal[:m], fin

    def pcs_commit(self, d_shares, n, d_random, nr, q0, rate, log_rate_offset, download=True):
        """PCS_commit at one rank: (rows count, rows, codewords, Base_MT
        nodes) as host arrays (download=False: the device buffers)."""
        B = n // 16
        rows = self.alloc(16 * 32 * B)
        cws = self.alloc(16 * 32 * B * rate)
        nodes = self.alloc(32 * (2 * B * rate - 1))
        c = ctypes.c_int(0)
        self._ck(lib().cosnark_pcs_commit(self.h, d_shares.ptr, n, d_random.ptr, nr, q0, rate, log_rate_offset,
                                          rows.ptr, cws.ptr, nodes.ptr, ctypes.byref(c)))
        if not download:
            return c.value, rows, cws, nodes
        out = (c.value, self.download(rows, (c.value, B, 2)), self.download(cws, (c.value, B * rate, 2)),
               self.download(nodes, (2 * B * rate - 1, 32), np.uint8))
        rows.free(); cws.free(); nodes.free()
        return out

    def _masks(self, R):
        RR = [np.ascontiguousarray(x, dtype=np.uint64) for x in R]
        ptrs = (ctypes.c_void_p * len(RR))(*[x.ctypes.data for x in RR])
        sizes = np.ascontiguousarray([x.shape[0] for x in RR], dtype=np.uint64)
        return RR, ptrs, sizes

    def quadratic_cosumcheck(self, d_p1, d_p2, M, R, K, K_r, degrees, nranks, rank, gather, ts):
        """quadratic_coSumcheck as rank `rank` of nranks; gather: an
        ALLGATHER_FN.  Returns (claims (5, 2), r (ndeg + 1, 2), ip_vec, ps,
        check-failure mask).  d_p1, d_p2 are folded in place."""
        RR, ptrs, sizes = self._masks(R)
        deg = np.ascontiguousarray(degrees, dtype=np.int32)
        KK = np.ascontiguousarray(K, dtype=np.uint64).reshape(2)
        Kr = np.ascontiguousarray(K_r, dtype=np.uint64).reshape(2)
        claims = np.zeros((5, 2), np.uint64)
        r = np.zeros((len(degrees) + 1, 2), np.uint64)
        ip = np.zeros((int(sizes.sum()), 2), np.uint64)
        ps = ctypes.c_double(0)
        fails = ctypes.c_int(0)
        self._ck(lib().cosnark_quadratic_cosumcheck(self.h, d_p1.ptr, d_p2.ptr, M, ptrs, _p(sizes), len(RR), _p(KK),
                                                    _p(Kr), _p(deg), len(degrees), nranks, rank, gather, None,
                                                    ctypes.byref(ts), _p(claims), _p(r), _p(ip), ctypes.byref(ps),
                                                    ctypes.byref(fails)))
        return claims, r, ip, ps.value, fails.value

    def cubic_cosumcheck(self, d_p1, d_p2, d_p3, M, r1, r2, R, K, K_r, degrees, nranks, rank, gather, ts):
        """cubic_coSumcheck as rank `rank` of nranks: (claims (4, 2), r, ip_vec,
        ps, check-failure mask)."""
        RR, ptrs, sizes = self._masks(R)
        deg = np.ascontiguousarray(degrees, dtype=np.int32)
        rr1 = np.ascontiguousarray(r1, dtype=np.uint64)
        rr2 = np.ascontiguousarray(r2, dtype=np.uint64)
        KK = np.ascontiguousarray(K, dtype=np.uint64).reshape(2)
        Kr = np.ascontiguousarray(K_r, dtype=np.uint64).reshape(2)
        claims = np.zeros((4, 2), np.uint64)
        r = np.zeros((len(degrees) + 1, 2), np.uint64)
        ip = np.zeros((int(sizes.sum()), 2), np.uint64)
        ps = ctypes.c_double(0)
        fails = ctypes.c_int(0)
        self._ck(lib().cosnark_cubic_cosumcheck(self.h, d_p1.ptr, d_p2.ptr, d_p3.ptr, M, _p(rr1), rr1.shape[0],
                                                _p(rr2), rr2.shape[0], ptrs, _p(sizes), len(RR), _p(KK), _p(Kr),
                                                _p(deg), len(degrees), nranks, rank, gather, None, ctypes.byref(ts),
                                                _p(claims), _p(r), _p(ip), ctypes.byref(ps), ctypes.byref(fails)))
        return claims, r, ip, ps.value, fails.value

    def whir_aggregate_claims(self, arrays, points):
        """aggregate_claims (src/Whir.cpp:608-668): (arr, b_aggr, K, beta_ch,
        arr_ch) for host arrays of equal size and claim points (npoints,
        log2 n, 2).  Seed libc first."""
        bufs = [self.upload(np.ascontiguousarray(a)) for a in arrays]
        n = np.asarray(arrays[0]).shape[0]
        pts = np.ascontiguousarray(points, dtype=np.uint64)
        npts = pts.shape[0]
        pa = (ctypes.c_void_p * len(arrays))(*[b.ptr.value for b in bufs])
        da, db = self.alloc(16 * n), self.alloc(16 * n)
        K = np.zeros(2, np.uint64)
        bc = np.zeros((npts, 2), np.uint64)
        ac = np.zeros((len(arrays), 2), np.uint64)
        self._ck(lib().cosnark_whir_aggregate_claims(self.h, len(arrays), pa, n, npts, _p(pts), da.ptr, db.ptr,
                                                      _p(K), _p(bc), _p(ac)))
        arr, b = self.download(da, (n, 2)), self.download(db, (n, 2))
        for x in bufs + [da, db]:
            x.free()
        return arr, b, K, bc, ac

    def merkle_verify(self, leaves, pos, paths, root, dup_left=True):
        """Authentication paths on the GPU: leaves (q, 4, 2), pos (q,), paths
        (q, depth, 32), root (32,) -> per-path 0/1 flags."""
        lv = np.ascontiguousarray(leaves, dtype=np.uint64)
        ps = np.ascontiguousarray(pos, dtype=np.uint32)
        pt = np.ascontiguousarray(paths, dtype=np.uint8)
        q, depth = ps.shape[0], (pt.shape[1] if pt.ndim == 3 else 0)
        bufs = [self.upload(x) for x in (lv, ps, pt if pt.size else np.zeros(32, np.uint8),
                                         np.ascontiguousarray(root, dtype=np.uint8))]
        ok = self.alloc(max(q, 1))
        bad = ctypes.c_uint64(0)
        self._ck(lib().cosnark_merkle_verify(self.h, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, depth, q, bufs[3].ptr,
                                              int(dup_left), ok.ptr, ctypes.byref(bad)))
        flags = self.download(ok, (max(q, 1),), np.uint8)[:q]
        for b in bufs + [ok]:
            b.free()
        return flags, int(bad.value)

    def cosumcheck3_phase1(self, p1, p2, p3, b1, d, dev=None):
        """Cubic block sumcheck phase 1: the 4d-element round polynomial (4d, 2).
        dev: (four device buffers, M) already resident."""
        if dev is None:
            bufs = [self.upload(x) for x in (p1, p2, p3, b1)]
            M = np.asarray(p1).shape[0]
        else:
            bufs, M = dev
        out = np.zeros((4 * d, 2), np.uint64)
        self._ck(lib().cosnark_cosumcheck3_phase1(self.h, *[b.ptr for b in bufs], M, d, _p(out)))
        if dev is None:
            for b in bufs:
                b.free()
        return out

    def cosumcheck3_fold(self, p1, p2, p3, b1, d, L=None, r=None, dev=None):
        """Cubic block sumcheck phase 2 with weights L (or L from r): the four
        folded (M/d, 2) vectors (None when dev is given; folded in place)."""
        if dev is None:
            bufs = [self.upload(x) for x in (p1, p2, p3, b1)]
            M = np.asarray(p1).shape[0]
        else:
            bufs, M = dev
        if L is not None:
            LL = np.ascontiguousarray(L, dtype=np.uint64)
            self._ck(lib().cosnark_cosumcheck3_fold(self.h, *[b.ptr for b in bufs], M, d, _p(LL)))
        else:
            rr = np.ascontiguousarray(np.asarray(r, dtype=np.uint64).reshape(2))
            self._ck(lib().cosnark_cosumcheck3_phase2(self.h, *[b.ptr for b in bufs], M, d, _p(rr)))
        if dev is not None:
            return None
        res = [self.download(b, (M // d, 2)) for b in bufs]
        for b in bufs:
            b.free()
        return res

    def cosumcheck_phase2(self, p1, p2, d, r, dev=None):
        """Block sumcheck phase 2: the folded (M/d, 2) vectors."""
        rr = np.ascontiguousarray(np.asarray(r, dtype=np.uint64).reshape(2))
        if dev is None:
            a, b = self.upload(p1), self.upload(p2)
            M = np.asarray(p1).shape[0]
        else:
            a, b, M = dev
        self._ck(lib().cosnark_cosumcheck_phase2(self.h, a.ptr, b.ptr, M, d, _p(rr)))
        if dev is not None:
            return None
        r1 = self.download(a, (M // d, 2))
        r2 = self.download(b, (M // d, 2))
        a.free(); b.free()
        return r1, r2

    def rs_multifold5(self, cw, a):
        """RS_multifold5 (src/coPCS_utils.cpp:386-437): (n/fd, 2)."""
        c = np.ascontiguousarray(cw, dtype=np.uint64)
        aa = np.ascontiguousarray(a, dtype=np.uint64)
        n, fd = c.shape[0], aa.shape[0]
        d = self.upload(c)
        out = self.alloc(16 * (n // fd))
        self._ck(lib().cosnark_rs_multifold5(self.h, d.ptr, n, _p(aa), fd, out.ptr))
        res = self.download(out, (n // fd, 2))
        d.free(); out.free()
        return res

    # ---- batches and tuning ---------------------------------------------

    def autotune(self, logn_lo, logn_hi, batch=1):
        best = np.zeros(logn_hi - logn_lo + 1, np.float64)
        self._ck(lib().cosnark_autotune(self.h, logn_lo, logn_hi, batch, _p(best)))
        return best

    def fmul_rate(self):
        r = ctypes.c_double(0)
        self._ck(lib().cosnark_fmul_rate(self.h, ctypes.byref(r)))
        return r.value

    def blake3_rate(self):
        r = ctypes.c_double(0)
        self._ck(lib().cosnark_blake3_rate(self.h, ctypes.byref(r)))
        return r.value

    def ntt_plan(self, logn):
        b, t = ctypes.c_int(0), ctypes.c_int(0)
        self._ck(lib().cosnark_ntt_plan(self.h, logn, ctypes.byref(b), ctypes.byref(t)))
        return b.value, t.value

    def set_ntt_plan(self, logn, passbits, tile):
        self._ck(lib().cosnark_set_ntt_plan(self.h, logn, passbits, tile))

    def autotune_merkle(self, leaf_counts):
        """Time the Merkle schedules (2^b first-level nodes per workgroup,
        b in 6..10) on a batch of trees of these leaf counts and keep the
        fastest; returns (b, ms)."""
        arr = np.ascontiguousarray(leaf_counts, dtype=np.uint64)
        b, ms = ctypes.c_int(0), ctypes.c_double(0)
        self._ck(lib().cosnark_autotune_merkle(self.h, _p(arr), len(arr), ctypes.byref(b), ctypes.byref(ms)))
        return b.value, ms.value

    def merkle_schedule(self):
        b = ctypes.c_int(0)
        self._ck(lib().cosnark_merkle_schedule(self.h, ctypes.byref(b)))
        return b.value

    def set_merkle_schedule(self, block_log):
        self._ck(lib().cosnark_set_merkle_schedule(self.h, block_log))

    def batch(self, sizes, rate):
        return Batch(self, sizes, rate)


class Batch:
    """Many plain proofs of uneven size, proven together (cosnark_batch_*)."""

    def __init__(self, ctx, sizes, rate):
        self.ctx = ctx
        self.sizes = [int(x) for x in sizes]
        self.rate = rate
        arr = np.ascontiguousarray(self.sizes, dtype=np.uint64)
        h = ctypes.c_void_p()
        ctx._ck(lib().cosnark_batch_create(ctx.h, len(self.sizes), _p(arr), rate, ctypes.byref(h)))
        self.h = h
        self.offsets = [int(lib().cosnark_batch_proof_offset(h, i)) for i in range(len(self.sizes) + 1)]

    def close(self):
        if self.h:
            lib().cosnark_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def input_ptrs(self, i):
        f, g = ctypes.c_void_p(), ctypes.c_void_p()
        self.ctx._ck(lib().cosnark_batch_input(self.h, i, ctypes.byref(f), ctypes.byref(g)))
        return f.value, g.value

    def set_inputs(self, i, f, g):
        pf, pg = self.input_ptrs(i)
        a = np.ascontiguousarray(f, dtype=np.uint64)
        b = np.ascontiguousarray(g, dtype=np.uint64)
        assert a.shape == (self.sizes[i], 2) and b.shape == a.shape
        self.ctx._ck(lib().cosnark_h2d(self.ctx.h, ctypes.c_void_p(pf), _p(a), a.nbytes))
        self.ctx._ck(lib().cosnark_h2d(self.ctx.h, ctypes.c_void_p(pg), _p(b), b.nbytes))

    def set_graph(self, on):
        self.ctx._ck(lib().cosnark_batch_set_graph(self.h, int(bool(on))))

    def commitment(self, i):
        """Proof i's codeword (rate*n, 2) and tree nodes (count, 32) as the
        last prove left them."""
        c, t = ctypes.c_void_p(), ctypes.c_void_p()
        self.ctx._ck(lib().cosnark_batch_commitment(self.h, i, ctypes.byref(c), ctypes.byref(t)))
        cwn = self.sizes[i] * self.rate
        cw = self.ctx.download_ptr(c.value, (cwn, 2))
        nodes = self.ctx.download_ptr(t.value, (merkle_node_count(cwn), 32), np.uint8)
        return cw, nodes

    def prove(self, transcripts=None):
        """Returns (list of proof bytes, list of transcripts after, device ms)."""
        n = len(self.sizes)
        ts = (Transcript * n)()
        if transcripts is not None:
            for i, t in enumerate(transcripts):
                ts[i].prev[0], ts[i].prev[1] = t.prev[0], t.prev[1]
        out = np.zeros(self.offsets[-1], np.uint8)
        ms = ctypes.c_double(0)
        self.ctx._ck(lib().cosnark_batch_prove(self.h, ctypes.cast(ts, ctypes.c_void_p), _p(out), ctypes.byref(ms)))
        proofs = [out[self.offsets[i]:self.offsets[i + 1]].tobytes() for i in range(n)]
        return proofs, list(ts), ms.value


def cfg4_sizes(count=64, seed=0xC05A0004, lo=12, hi=20):
    """Batched-proving workload: `count` sizes 2^e, e uniform on lo..hi from
    the seeded generator (stream 2, real part modulo the range)."""
    x = synth(seed, 2, count)[:, 0]
    return [1 << (lo + int(v) % (hi - lo + 1)) for v in x]


class Server:
    """Continuous-batching proof server (cosnark_server_*)."""

    def __init__(self, budget_bytes, rate=4, max_log_n=22, slots=3, device=0):
        h = ctypes.c_void_p()
        rc = lib().cosnark_server_create(device, budget_bytes, rate, max_log_n, slots, ctypes.byref(h))
        if rc != 0:
            raise CosnarkError(rc, "server_create failed")
        self.h = h
        self.rate = rate

    def submit(self, d_f, d_g, n, prev=(0, 0)):
        """d_f, d_g: device pointers (int) or DevBuf; returns the ticket."""
        pf = d_f.ptr if hasattr(d_f, "ptr") else ctypes.c_void_p(d_f)
        pg = d_g.ptr if hasattr(d_g, "ptr") else ctypes.c_void_p(d_g)
        pv = np.ascontiguousarray([int(prev[0]), int(prev[1])], dtype=np.uint64)
        t = ctypes.c_uint64(0)
        rc = lib().cosnark_server_submit(self.h, pf, pg, n, _p(pv), ctypes.byref(t))
        if rc != 0:
            raise CosnarkError(rc, "server_submit failed")
        return t.value

    def submit_keep(self, d_f, d_g, n, d_cw, d_nodes, prev=(0, 0)):
        """submit, with the request's codeword and tree written to the
        device buffers d_cw (rate*n elements) and d_nodes (every level)."""
        pv = np.ascontiguousarray([int(prev[0]), int(prev[1])], dtype=np.uint64)
        t = ctypes.c_uint64(0)
        rc = lib().cosnark_server_submit_ex(self.h, d_f.ptr, d_g.ptr, n, _p(pv), d_cw.ptr, d_nodes.ptr,
                                            ctypes.byref(t))
        if rc != 0:
            raise CosnarkError(rc, "server_submit_ex failed")
        return t.value

    def poll(self, ticket):
        d = ctypes.c_int(0)
        rc = lib().cosnark_server_poll(self.h, ticket, ctypes.byref(d))
        if rc != 0:
            raise CosnarkError(rc, "server_poll failed")
        return bool(d.value)

    def wait(self, ticket, n):
        """Returns (proof bytes, transcript after, latency ms)."""
        proof = np.zeros(int(lib().cosnark_plain_proof_bytes(n)), np.uint8)
        pv = np.zeros(2, np.uint64)
        lat = ctypes.c_double(0)
        rc = lib().cosnark_server_wait(self.h, ticket, _p(proof), _p(pv), ctypes.byref(lat))
        if rc != 0:
            raise CosnarkError(rc, "server request failed")
        return proof.tobytes(), (int(pv[0]), int(pv[1])), lat.value

    def stats(self):
        out = np.zeros(6, np.uint64)
        lib().cosnark_server_stats(self.h, _p(out))
        keys = ("passes", "completed", "max_inflight", "arena_peak_bytes", "arena_bytes", "queued")
        return {k: int(v) for k, v in zip(keys, out)}

    def close(self):
        if self.h:
            lib().cosnark_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def share_unpack(shares, kk):
    """unpack_secret_share (src/MPI_utils.cpp:2769-2785) on the host:
    (the shares transformed as the reference leaves them, the kk values)."""
    s = np.ascontiguousarray(shares, dtype=np.uint64).copy()
    out = np.zeros((kk, 2), np.uint64)
    rc = lib().cosnark_share_unpack(_p(s), s.shape[0], kk, _p(out))
    if rc:
        raise CosnarkError(rc, "share_unpack: bad sizes")
    return s, out


def sumcheck_aggr(polys, b):
    """Rank 0's combine of sumcheck_aggr (src/MPI_utils.cpp:2604-2645) over
    the gathered polynomials (N, len, 2) with weights b (N/4, 2)."""
    P = np.ascontiguousarray(polys, dtype=np.uint64)
    N, ln = P.shape[0], P.shape[1]
    bb = np.ascontiguousarray(b if b is not None and len(b) else np.zeros((1, 2)), dtype=np.uint64)
    out = np.zeros((ln, 2), np.uint64)
    rc = lib().cosnark_sumcheck_aggr(_p(P), N, ln, _p(bb), _p(out))
    if rc:
        raise CosnarkError(rc, "sumcheck_aggr: bad sizes")
    return out


def whir_rates(sizes, rate, whir_k=4):
    """get_whir_rates (src/Whir.cpp:18-137): commitment rates, sizes largest first."""
    sz = np.ascontiguousarray(sizes, dtype=np.uint64)
    out = np.zeros(len(sizes), np.int32)
    rc = lib().cosnark_whir_rates(_p(sz), len(sizes), rate, whir_k, _p(out))
    if rc:
        raise CosnarkError(rc, "whir_rates: sizes must be sorted largest first")
    return [int(x) for x in out]


def verify_plain(proof, n, y, ts=None):
    """cosnark_verify_plain: (status, challenges, bad_round)."""
    pr = np.frombuffer(bytes(proof), dtype=np.uint8).copy()
    yy = np.ascontiguousarray(np.asarray(y, dtype=np.uint64).reshape(2))
    t = ts if ts is not None else Transcript()
    m = int(n).bit_length() - 1
    r = np.zeros((m, 2), np.uint64)
    bad = ctypes.c_int(0)
    rc = lib().cosnark_verify_plain(_p(pr), n, _p(yy), ctypes.byref(t), _p(r), ctypes.byref(bad))
    return rc, r, bad.value


class ThreadAllGather:
    """An all-gather among `n` threads of one process, for running n ranks'
    coSumcheck drivers side by side (each thread holds its own Context).
    callback(rank) gives the ALLGATHER_FN for that rank's thread."""

    def __init__(self, n):
        import threading
        self.n = n
        self.barrier = threading.Barrier(n)
        self.parts = [None] * n
        self._keep = []

    def callback(self, rank):
        def fn(user, mine, elems, all_out):
            try:
                self.parts[rank] = np.ctypeslib.as_array(mine, shape=(2 * elems,)).copy()
                self.barrier.wait()
                out = np.ctypeslib.as_array(all_out, shape=(2 * elems * self.n,))
                for m in range(self.n):
                    out[2 * elems * m: 2 * elems * (m + 1)] = self.parts[m]
                self.barrier.wait()          # everyone has read before the next exchange writes
                return 0
            except Exception:
                self.barrier.abort()
                return 1
        cb = ALLGATHER_FN(fn)
        self._keep.append(cb)
        return cb
