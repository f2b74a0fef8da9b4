"""CPU oracle for the FP64 CSR SpMV path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker.  It wraps
``oracle/liboracle.so`` (plain C, built by ``oracle/Makefile``), a restatement
of the reference's host algorithms; every C function cites the reference
file:line it follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile liboracle.so with gcc (seconds)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    sig = {
        "orc_spmv": (None, [C.c_int, i64p, i32p, f64p, f64p, C.c_double, C.c_double, f64p]),
        "orc_spmv_omp": (None, [C.c_int, i64p, i32p, f64p, f64p, C.c_double, C.c_double, f64p, C.c_int]),
        "orc_spmv_f32val": (None, [C.c_int, i64p, i32p, f32p, f64p, C.c_double, C.c_double, f64p]),
        "orc_spmv_dot2": (None, [C.c_int, i64p, i32p, f64p, f64p, C.c_double, C.c_double, f64p, C.c_void_p]),
        "orc_spmm": (None, [C.c_int, C.c_int, C.c_int, i64p, i32p, f64p, f64p, C.c_double, C.c_double, f64p]),
        "orc_spmv_rows": (None, [C.c_int, C.c_int, i64p, i32p, f64p, f64p, C.c_double, C.c_double, f64p]),
        "orc_get_row_from_index": (C.c_int, [C.c_int, i64p, C.c_int64]),
        "orc_row_of": (C.c_int, [C.c_int, i64p, C.c_int64]),
        "orc_v1_partition": (None, [C.c_int, C.c_int64, i64p, C.c_int, C.c_int, i64p, i64p, i32p, i32p, u8p, u8p]),
        "orc_v2_tasks": (C.c_int, [C.c_int, C.c_int64, i64p, C.c_int64, C.c_int,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "orc_v1_spmv": (C.c_int, [C.c_int, C.c_int64, i64p, i32p, f64p, f64p, C.c_double, C.c_double, f64p, C.c_int]),
        "orc_tile_sigma": (C.c_int, [C.c_int, C.c_int64]),
        "orc_tile_count": (C.c_int64, [C.c_int64, C.c_int]),
        "orc_tile_ptr": (None, [C.c_int, C.c_int64, i64p, C.c_int, C.c_int64, u32p]),
        "orc_tile_flags": (None, [C.c_int, C.c_int64, i64p, C.c_int, C.c_int64, u32p]),
        "orc_tile_transpose_i32": (None, [C.c_int64, C.c_int, C.c_int64, i32p, i32p]),
        "orc_tile_transpose_f64": (None, [C.c_int64, C.c_int, C.c_int64, f64p, f64p]),
        "orc_gen_g": (C.c_int64, [C.c_int, C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]),
        "orc_coo_to_csr_rowptr": (None, [C.c_int, C.c_int64, i32p, i64p]),
        "orc_harness_alpha_beta": (None, [C.c_uint, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "orc_getB": (C.c_double, [C.c_int, C.c_int64]),
        "orc_getFLOP": (C.c_double, [C.c_int64]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _csr(rowptr, col, val):
    return (np.ascontiguousarray(rowptr, np.int64), np.ascontiguousarray(col, np.int32),
            np.ascontiguousarray(val, np.float64))


def spmv(rowptr, col, val, x, alpha=1.0, beta=0.0, y=None):
    """y = alpha*A*x + beta*y by the reference's host loop (main.cu:350-356)."""
    rowptr, col, val = _csr(rowptr, col, val)
    m = len(rowptr) - 1
    out = np.zeros(m) if y is None else np.array(y, dtype=np.float64, copy=True)
    lib().orc_spmv(m, rowptr, col, val, np.ascontiguousarray(x, np.float64), alpha, beta, out)
    return out


def spmv_omp(rowptr, col, val, x, alpha=1.0, beta=0.0, y=None, threads=1):
    """The host loop with rows split statically over `threads` OpenMP threads
    (bitwise equal to spmv)."""
    rowptr, col, val = _csr(rowptr, col, val)
    m = len(rowptr) - 1
    out = np.zeros(m) if y is None else np.array(y, dtype=np.float64, copy=True)
    lib().orc_spmv_omp(m, rowptr, col, val, np.ascontiguousarray(x, np.float64), alpha, beta, out, int(threads))
    return out


def spmv_rows(r0, r1, rowptr, col, val, x, alpha=1.0, beta=0.0, y=None):
    rowptr, col, val = _csr(rowptr, col, val)
    m = len(rowptr) - 1
    out = np.zeros(m) if y is None else np.array(y, dtype=np.float64, copy=True)
    lib().orc_spmv_rows(r0, r1, rowptr, col, val, np.ascontiguousarray(x, np.float64), alpha, beta, out)
    return out


def spmv_f32val(rowptr, col, val32, x, alpha=1.0, beta=0.0, y=None):
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    m = len(rowptr) - 1
    out = np.zeros(m) if y is None else np.array(y, dtype=np.float64, copy=True)
    lib().orc_spmv_f32val(m, rowptr, np.ascontiguousarray(col, np.int32),
                          np.ascontiguousarray(val32, np.float32),
                          np.ascontiguousarray(x, np.float64), alpha, beta, out)
    return out


def spmv_dot2(rowptr, col, val, x, alpha=1.0, beta=0.0, y=None):
    """Compensated reference; returns (y, |alpha|*|A||x|)."""
    rowptr, col, val = _csr(rowptr, col, val)
    m = len(rowptr) - 1
    out = np.zeros(m) if y is None else np.array(y, dtype=np.float64, copy=True)
    mag = np.zeros(m)
    lib().orc_spmv_dot2(m, rowptr, col, val, np.ascontiguousarray(x, np.float64), alpha, beta, out,
                        mag.ctypes.data)
    return out, mag


def spmm(rowptr, col, val, X, alpha=1.0, beta=0.0, Y=None):
    """Column-major X (n, k) given as a Fortran-ordered or (k, n) C array."""
    rowptr, col, val = _csr(rowptr, col, val)
    m = len(rowptr) - 1
    Xc = np.ascontiguousarray(np.asarray(X, np.float64).T)  # (k, n) rows = columns of X
    k, n = Xc.shape
    Yc = np.zeros((k, m)) if Y is None else np.ascontiguousarray(np.asarray(Y, np.float64).T).copy()
    lib().orc_spmm(m, n, k, rowptr, col, val, Xc.ravel(), alpha, beta, Yc.ravel())
    return Yc.T


def get_row_from_index(rowptr, idx):
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    return lib().orc_get_row_from_index(len(rowptr) - 1, rowptr, idx)


def row_of(rowptr, k):
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    return lib().orc_row_of(len(rowptr) - 1, rowptr, k)


def v1_partition(rowptr, P, rule=1):
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    m = len(rowptr) - 1
    nnz = int(rowptr[-1])
    s = np.zeros(P, np.int64); e = np.zeros(P, np.int64)
    sr = np.zeros(P, np.int32); er = np.zeros(P, np.int32)
    sf = np.zeros(P, np.uint8); ef = np.zeros(P, np.uint8)
    lib().orc_v1_partition(m, nnz, rowptr, P, rule, s, e, sr, er, sf, ef)
    return dict(start_idx=s, end_idx=e, start_row=sr, end_row=er, start_flag=sf, end_flag=ef)


def v2_tasks(rowptr, nb, rule=1):
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    m = len(rowptr) - 1
    nnz = int(rowptr[-1])
    T = lib().orc_v2_tasks(m, nnz, rowptr, nb, rule, None, None, None, None, None, None)
    s = np.zeros(T, np.int64); e = np.zeros(T, np.int64)
    sr = np.zeros(T, np.int32); er = np.zeros(T, np.int32)
    sf = np.zeros(T, np.uint8); ef = np.zeros(T, np.uint8)
    lib().orc_v2_tasks(m, nnz, rowptr, nb, rule, s.ctypes.data, e.ctypes.data, sr.ctypes.data,
                       er.ctypes.data, sf.ctypes.data, ef.ctypes.data)
    return dict(start_idx=s, end_idx=e, start_row=sr, end_row=er, start_flag=sf, end_flag=ef)


def v1_spmv(rowptr, col, val, x, alpha, beta, y, P):
    rowptr, col, val = _csr(rowptr, col, val)
    out = np.array(y, dtype=np.float64, copy=True)
    lib().orc_v1_spmv(len(rowptr) - 1, int(rowptr[-1]), rowptr, col, val,
                      np.ascontiguousarray(x, np.float64), alpha, beta, out, P)
    return out


def tile_sigma(m, nnz):
    return lib().orc_tile_sigma(m, nnz)


def tile_count(nnz, sigma):
    return lib().orc_tile_count(nnz, sigma)


def tile_convert(rowptr, col, val, sigma):
    """Host conversion: (tile_ptr[p+1], flags[p*64], col_t, val_t)."""
    rowptr, col, val = _csr(rowptr, col, val)
    m = len(rowptr) - 1
    nnz = int(rowptr[-1])
    p = lib().orc_tile_count(nnz, sigma)
    tp = np.zeros(p + 1, np.uint32)
    fl = np.zeros(max(p, 1) * 64, np.uint32)[: p * 64]
    fl = np.ascontiguousarray(fl)
    lib().orc_tile_ptr(m, nnz, rowptr, sigma, p, tp)
    if p:
        lib().orc_tile_flags(m, nnz, rowptr, sigma, p, fl)
    ct = np.zeros(p * 64 * sigma, np.int32)
    vt = np.zeros(p * 64 * sigma, np.float64)
    if p:
        lib().orc_tile_transpose_i32(nnz, sigma, p, col, ct)
        lib().orc_tile_transpose_f64(nnz, sigma, p, val, vt)
    return tp, fl, ct, vt


def gen_g(n, seed=1):
    """The reference harness's 'g' matrix as COO (dspmv_test.cu:137-199)."""
    L = lib()
    nnz = L.orc_gen_g(n, seed, None, None, None)
    if nnz < 0:
        raise ValueError("g generator needs n % 8 == 0 and n >= 8")
    r = np.zeros(nnz, np.int32); c = np.zeros(nnz, np.int32); v = np.zeros(nnz)
    L.orc_gen_g(n, seed, r.ctypes.data, c.ctypes.data, v.ctypes.data)
    return r, c, v


def coo_to_csr_rowptr(m, coo_row):
    coo_row = np.ascontiguousarray(coo_row, np.int32)
    rp = np.zeros(m + 1, np.int64)
    lib().orc_coo_to_csr_rowptr(m, len(coo_row), coo_row, rp)
    return rp


def harness_alpha_beta(seed, draws_before):
    a = C.c_double(); b = C.c_double()
    lib().orc_harness_alpha_beta(seed, draws_before, C.byref(a), C.byref(b))
    return a.value, b.value


def getB(m, nnz):
    return lib().orc_getB(m, nnz)


def getFLOP(nnz):
    return lib().orc_getFLOP(nnz)
