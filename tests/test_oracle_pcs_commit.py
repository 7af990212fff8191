"""prepare_data (src/coPCS.cpp:76-128), the layout PCS_commit encodes and
commits: checked against the layout stated as slices -- every share once,
in order, B - q0 per row; q0 random values closing each data row; zeros
after the last share; one random-only row (case 1: B values, or until the
random values run out; case 2: B - (n - used)) -- and against the library's
row count for the same sizes.  At the 2^22 share with q0 = 333 the layout
is 17 data rows and one random row."""
import numpy as np
import pytest

import cosnark
import pyoracle as O


def _expected(shares, rnd, q0):
    n = shares.shape[0]
    B = n // 16
    per = B - q0
    rows = np.zeros((32, B, 2), np.uint64)
    nd = -(-n // per)
    for i in range(nd):
        seg = shares[i * per:(i + 1) * per]
        rows[i, :seg.shape[0]] = seg
        rows[i, per:] = rnd[i * q0:(i + 1) * q0]
    used = nd * q0
    if used > n or n - used > B:
        cnt = min(B, rnd.shape[0] - used)
    else:
        cnt = B - (n - used)
    rows[nd, :cnt] = rnd[used:used + cnt]
    return nd + 1, rows


@pytest.mark.parametrize("logn,q0,nr_extra", [(10, 3, 100), (12, 33, 0), (14, 100, 5000), (16, 333, 7000),
                                              (10, 30, 0)])
def test_prepare_data_layout(logn, q0, nr_extra):
    n = 1 << logn
    B = n // 16
    nd = -(-n // (B - q0))
    nr = nd * q0 + min(B, 64) + nr_extra
    shares, rnd = O.synth(0xC05A00A0, logn, n), O.synth(0xC05A00A1, logn, nr)
    c, rows = O.prepare_data(shares, rnd, q0)
    ce, exp = _expected(shares, rnd, q0)
    assert c == ce and np.array_equal(rows, exp)
    # every share exactly once, in order
    per = B - q0
    data = np.concatenate([rows[i, :per] for i in range(c - 1)])[:n]
    assert np.array_equal(data, shares)
    nrows, bs = _lib_rows(n, nr, q0)
    assert (nrows, bs) == (c, B)


def _lib_rows(n, nr, q0):
    import ctypes
    L = cosnark.lib()
    L.cosnark_pcs_prepare_rows.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64)]
    r, b = ctypes.c_int(0), ctypes.c_uint64(0)
    assert L.cosnark_pcs_prepare_rows(n, nr, q0, ctypes.byref(r), ctypes.byref(b)) == 0
    return r.value, b.value


def test_share_layout_2_22():
    n, q0 = 1 << 22, 333
    nr = 18 * 333 + (1 << 18)
    assert _lib_rows(n, nr, q0) == (18, 1 << 18)


def test_prepare_data_refusals():
    # too few random values, and a row size not above q0
    shares = O.synth(1, 0, 1 << 8)
    c, _ = O.prepare_data(shares, O.synth(1, 1, 10), 3)
    assert c == -1
    c, _ = O.prepare_data(shares, O.synth(1, 1, 1000), 16)
    assert c == -1
    # more than 32 rows of shares: the reference exits
    shares = O.synth(1, 0, 1 << 10)
    c, _ = O.prepare_data(shares, O.synth(1, 1, 1 << 16), 60)
    assert c == -1
    import ctypes
    r, b = ctypes.c_int(0), ctypes.c_uint64(0)
    assert cosnark.lib().cosnark_pcs_prepare_rows(1 << 10, 1 << 16, 60, ctypes.byref(r), ctypes.byref(b)) == cosnark.E_RANGE
