"""ctypes binding of libspx.so (the gfx950 SpMV library) and libspxgen.so.

The product is the C ABI in include/spx_capi.h; this module only forwards to
it.  Loading fails loudly when libspx.so has not been built: there is no CPU
fallback anywhere in the product path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# SPX_LIBRARY selects another build of libspx.so (the sanitized host build
# in sparse-blas_amd/_asan/ for the CPU tests); the default is the in-tree one
LIB_PATH = os.environ.get("SPX_LIBRARY") or os.path.join(_PKG, "libspx.so")
GEN_PATH = os.path.join(_PKG, "libspxgen.so")

OK, EARG, EMEM, EDEV, EGUARD, ERANGE = 0, -1, -2, -3, -4, -5
MODE_ROWS, MODE_NNZ, MODE_QUEUE = 0, 1, 2
KERNEL_VECTOR, KERNEL_MERGE, KERNEL_TILE, KERNEL_SEG = 1, 2, 3, 4
FLAG_DEVICE_PTRS, FLAG_FUSED_SPMM = 1, 2
PREC_F64, PREC_F32 = 0, 1
FLAG_DEVICE_PTRS = 1


class SpxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"spx error {code}: {msg}")
        self.code = code


class PlanDesc(C.Structure):
    _fields_ = [
        ("m", C.c_int), ("n", C.c_int), ("nnz", C.c_int64),
        ("rowptr64", C.c_void_p), ("rowptr32", C.c_void_p),
        ("col", C.c_void_p), ("val", C.c_void_p),
        ("kernel", C.c_int), ("mode", C.c_int), ("parts", C.c_int),
        ("nb", C.c_int64), ("queues", C.c_int), ("sigma", C.c_int),
        ("lanes", C.c_int), ("items", C.c_int), ("precision", C.c_int),
        ("flags", C.c_int), ("long_row", C.c_int64),
    ]


# name -> (restype, argtypes); every symbol include/spx_capi.h declares
_P, _I, _I64, _D, _LL = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_longlong
CAPI = {
    "spx_plan_desc_init": (None, [_P]),
    "spx_plan_create": (_I, [_P, _P]),
    "spx_plan_destroy": (_I, [_P]),
    "spx_plan_execute": (_I, [_P, _D, _P, _D, _P, _P]),
    "spx_plan_execute_multi": (_I, [_P, _I, _D, _P, _I64, _D, _P, _I64, _P]),
    "spx_plan_info": (_I, [_P, _P, _I]),
    "spx_plan_export": (_I64, [_P, _I, _P, _I64]),
    "spx_plan_time": (_I, [_P, _D, _P, _D, _P, _I, _I, _I64, _P]),
    "spx_plan_time_each": (_I, [_P, _D, _P, _D, _P, _I, _I, _I64, _P, _P, _P]),
    "spx_mm_read": (_I, [C.c_char_p, _P, _P, _P, _P, _P, _P]),
    "spx_mm_free": (None, [_P]),
    "spx_sel_param_count": (_I, []),
    "spx_sel_features": (_I, [_I, _I, _I64, _P, _P, _I64, _P, _P]),
    "spx_mlp_work_floats": (_I64, [_I]),
    "spx_mlp_forward": (_I, [_P, _P, _I, _P, _P, _P]),
    "spx_mlp_train_step": (_I, [_P, _P, _P, _I, C.c_float, C.c_float, _P, _P, _I, _P, _P, _P, _P]),
    "spx_mlp_argmax": (_I, [_P, _I, _P, _P, _P]),
    "spx_dropin_baseline": (_I, [_I, _I, _LL, _P, _P, _P, _P, _P, _P, _P, _I]),
    "spx_dropin_v1": (_I, [_I, _I, _LL, _P, _P, _P, _P, _P, _P, _P, _I, _I]),
    "spx_dropin_v2": (_I, [_I, _I, _LL, _P, _P, _P, _P, _P, _P, _P, _I, _I, _LL, _I]),
    "spx_dropin_csr5_kernel": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "spx_dropin_csr5_kernel_noalloc": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "spx_get_row_from_index": (_I, [_I, _P, _LL]),
    "spx_get_time": (_D, []),
    "spx_get_gpu_availble_mem": (_D, [_I]),
    "spx_set_v1_streams": (_I, [_I]),
    "spx_dropin_last_times": (_I, [_P, _I]),
    "spx_debug_set_free_mem": (None, [_LL]),
    "spx_debug_fail_alloc": (None, [_LL]),
    "spx_malloc": (_P, [_I64]),
    "spx_free": (_I, [_P]),
    "spx_memcpy_h2d": (_I, [_P, _P, _I64]),
    "spx_memcpy_d2h": (_I, [_P, _P, _I64]),
    "spx_memcpy_d2d": (_I, [_P, _P, _I64]),
    "spx_synchronize": (_I, []),
    "spx_device_count": (_I, []),
    "spx_last_error": (C.c_char_p, []),
    "spx_version": (C.c_char_p, []),
    "spx_batch_create": (_I, [_P, _I, _P, _P, _P, _P]),
    "spx_graph_capture": (_I, [_P, _D, _P, _D, _P, _P]),
    "spx_graph_launch": (_I, [_P, _P]),
    "spx_graph_destroy": (_I, [_P]),
    "spx_graph_time": (_I, [_P, _I, _I, _P]),
}


class BatchItem(C.Structure):
    _fields_ = [("m", C.c_int), ("n", C.c_int), ("nnz", C.c_int64),
                ("rowptr", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p)]

# C++-linkage drop-in symbols of include/spmv_kernel.h (Itanium mangling)
CXX_SYMBOLS = {
    "csr5_kernel": "_Z11csr5_kerneliiiPdS_PiS0_S_S_S_",
    "spMV_mgpu_baseline": "_Z18spMV_mgpu_baselineiixPdS_PxPiS_S_S_i",
    "spMV_mgpu_v1": "_Z12spMV_mgpu_v1iixPdS_PxPiS_S_S_ii",
    "spMV_mgpu_v2": "_Z12spMV_mgpu_v2iixPdS_PxPiS_S_S_iixi",
    "get_row_from_index": "_Z18get_row_from_indexiPxx",
    "get_time": "_Z8get_timev",
    "get_gpu_availble_mem": "_Z20get_gpu_availble_memi",
}

_lib = None
_gen = None


def lib():
    """Load libspx.so; raises if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run python sparse-blas_amd/build.py")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in CAPI.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != OK:
        raise SpxError(rc, lib().spx_last_error().decode())
    return rc


def _ptr(a):
    return a.ctypes.data if a is not None else None


# ---- device buffers --------------------------------------------------------------

class DeviceBuffer:
    """Device allocation owned by this object (spx_malloc / spx_free)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = lib().spx_malloc(max(self.nbytes, 1))
        if not self.ptr:
            raise SpxError(EMEM, "spx_malloc failed")

    @classmethod
    def from_array(cls, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        if arr.nbytes:
            check(lib().spx_memcpy_h2d(b.ptr, arr.ctypes.data, arr.nbytes))
        return b

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        if arr.nbytes:
            check(lib().spx_memcpy_h2d(self.ptr, arr.ctypes.data, arr.nbytes))

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().spx_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().spx_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(lib().spx_synchronize())


# ---- plans -----------------------------------------------------------------------------

class Plan:
    """Device-resident SpMV plan (spx_plan_create / execute / destroy)."""

    def __init__(self, rowptr, col, val, n=None, kernel=KERNEL_MERGE, mode=MODE_NNZ, parts=1, nb=0,
                 queues=1, sigma=0, lanes=0, items=0, precision=PREC_F64, long_row=0, flags=0):
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.m = len(rowptr) - 1
        self.n = int(n) if n is not None else (int(self.col.max()) + 1 if len(self.col) else 0)
        self.nnz = int(rowptr[-1]) if self.m >= 0 and len(rowptr) else 0
        d = PlanDesc()
        lib().spx_plan_desc_init(C.byref(d))
        d.m, d.n, d.nnz = self.m, self.n, self.nnz
        d.rowptr64 = rowptr.ctypes.data
        d.col = self.col.ctypes.data if len(self.col) else None
        d.val = self.val.ctypes.data if len(self.val) else None
        d.kernel, d.mode, d.parts, d.nb, d.queues = kernel, mode, parts, nb, queues
        d.sigma, d.lanes, d.items, d.precision = sigma, lanes, items, precision
        d.long_row = long_row
        d.flags = flags
        h = C.c_void_p()
        check(lib().spx_plan_create(C.byref(h), C.byref(d)))
        self.h = h
        self._rowptr = rowptr
        # the host copies are not needed after creation
        self.col = self.val = None

    def execute(self, x_dev, y_dev, alpha=1.0, beta=0.0, stream=None):
        check(lib().spx_plan_execute(self.h, alpha, x_dev.ptr, beta, y_dev.ptr, stream))

    def execute_multi(self, k, X_dev, ldx, Y_dev, ldy, alpha=1.0, beta=0.0, stream=None):
        check(lib().spx_plan_execute_multi(self.h, k, alpha, X_dev.ptr, ldx, beta, Y_dev.ptr, ldy, stream))

    def __call__(self, x, alpha=1.0, beta=0.0, y=None):
        """Host convenience: upload x (and y), execute, synchronize, download y."""
        xd = DeviceBuffer.from_array(np.asarray(x, np.float64))
        yd = DeviceBuffer(self.m * 8)
        if y is not None:
            yd.upload(np.asarray(y, np.float64))
        self.execute(xd, yd, alpha, beta)
        synchronize()
        return yd.download(np.float64, self.m)

    def info(self):
        out = (C.c_int64 * 32)()
        k = lib().spx_plan_info(self.h, out, 32)
        keys = ["m", "n", "nnz", "kernel", "mode", "parts", "sigma", "lanes", "items", "units",
                "carries", "tiles", "bytes", "precision", "queues", "nlong", "long_nnz", "col_blocks",
                "runs", "long_row", "staged_units", "split", "upload_us", "build_us", "bands"]
        return {keys[i]: int(out[i]) for i in range(k)}

    def export(self, what):
        L = lib()
        nbytes = L.spx_plan_export(self.h, what, None, 0)
        if nbytes < 0:
            raise SpxError(EARG, "export failed")
        dt = {0: np.int64, 1: np.uint32, 2: np.uint32, 3: np.int32, 4: np.float64, 5: np.int64}[what]
        out = np.zeros(nbytes // np.dtype(dt).itemsize, dtype=dt)
        if nbytes:
            got = L.spx_plan_export(self.h, what, out.ctypes.data, nbytes)
            if got < 0:
                raise SpxError(EDEV, "export copy failed")
        return out

    def time(self, x_dev, y_dev, alpha=1.0, beta=0.0, warmup=5, reps=20, flush_bytes=0, wall=False):
        """(execute ms, SpMV kernel ms) per execute from HIP events; with wall=True also the
        host wall-clock ms per execute over the same back-to-back run."""
        out = (C.c_double * 3)()
        check(lib().spx_plan_time(self.h, alpha, x_dev.ptr, beta, y_dev.ptr, warmup, reps, flush_bytes, out))
        return (out[0], out[1], out[2]) if wall else (out[0], out[1])

    def time_each(self, x_dev, y_dev, alpha=1.0, beta=0.0, warmup=5, reps=20, flush_bytes=0):
        """(execute ms array, SpMV kernel ms array, wall ms per execute) of `reps`
        back-to-back executes; flush_bytes > 0 reads a buffer of that size
        before each (cold L2 and MALL), < 0 overwrites one."""
        e = np.zeros(reps)
        k = np.zeros(reps)
        w = C.c_double()
        check(lib().spx_plan_time_each(self.h, alpha, x_dev.ptr, beta, y_dev.ptr, warmup, reps, flush_bytes,
                                       e.ctypes.data, k.ctypes.data, C.byref(w)))
        return e, k, w.value

    def time_wall(self, x_dev, y_dev, alpha=1.0, beta=0.0, warmup=5, reps=20):
        """Host wall-clock ms per execute of `reps` back-to-back executes with no
        events between them (synchronized before and after)."""
        w = C.c_double()
        check(lib().spx_plan_time_each(self.h, alpha, x_dev.ptr, beta, y_dev.ptr, warmup, reps, 0,
                                       None, None, C.byref(w)))
        return w.value

    def close(self):
        if getattr(self, "h", None):
            lib().spx_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchPlan(Plan):
    """Many independent SpMVs as one block-diagonal plan (spx_batch_create).
    x_off / y_off give each matrix's slice of the concatenated x and y."""

    def __init__(self, mats, kernel=KERNEL_MERGE, mode=MODE_QUEUE, parts=1, nb=1 << 22, queues=1, sigma=0,
                 lanes=0, items=0, precision=PREC_F64, long_row=0):
        keep = []
        arr = (BatchItem * len(mats))()
        for i, A in enumerate(mats):
            rp = np.ascontiguousarray(A["rowptr"], np.int64)
            c = np.ascontiguousarray(A["col"], np.int32)
            v = np.ascontiguousarray(A["val"], np.float64)
            keep += [rp, c, v]
            arr[i] = BatchItem(A["m"], A["n"], int(rp[-1]), rp.ctypes.data, c.ctypes.data if len(c) else None,
                               v.ctypes.data if len(v) else None)
        d = PlanDesc()
        lib().spx_plan_desc_init(C.byref(d))
        d.kernel, d.mode, d.parts, d.nb, d.queues = kernel, mode, parts, nb, queues
        d.sigma, d.lanes, d.items, d.precision, d.long_row = sigma, lanes, items, precision, long_row
        self.x_off = np.zeros(len(mats) + 1, np.int64)
        self.y_off = np.zeros(len(mats) + 1, np.int64)
        h = C.c_void_p()
        check(lib().spx_batch_create(C.byref(h), len(mats), arr, C.byref(d), self.x_off.ctypes.data,
                                     self.y_off.ctypes.data))
        self.h = h
        self.m, self.n = int(self.y_off[-1]), int(self.x_off[-1])
        self.nnz = sum(int(A["rowptr"][-1]) for A in mats)


class Graph:
    """An instantiated hipGraph of one plan execute (spx_graph_capture)."""

    def __init__(self, plan, x_dev, y_dev, alpha=1.0, beta=0.0):
        h = C.c_void_p()
        check(lib().spx_graph_capture(plan.h, alpha, x_dev.ptr, beta, y_dev.ptr, C.byref(h)))
        self.h = h
        self._keep = (plan, x_dev, y_dev)

    def launch(self, stream=None):
        check(lib().spx_graph_launch(self.h, stream))

    def time(self, warmup=3, reps=20):
        out = (C.c_double * 1)()
        check(lib().spx_graph_time(self.h, warmup, reps, out))
        return out[0]

    def close(self):
        if getattr(self, "h", None):
            lib().spx_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- drop-in entry points --------------------------------------------------------------

def _dropin_args(rowptr, col, val, x, y):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.array(y, dtype=np.float64, copy=True)
    return rowptr, col, val, x, y


def spMV_mgpu_baseline(m, n, nnz, alpha, val, rowptr, col, x, beta, y, ngpu=1):
    """Returns (rc, y_out) -- SB/include/spmv_kernel.h:9-13 through the C ABI."""
    rowptr, col, val, x, y = _dropin_args(rowptr, col, val, x, y)
    a, b = C.c_double(alpha), C.c_double(beta)
    rc = lib().spx_dropin_baseline(m, n, nnz, C.addressof(a), _ptr(val), _ptr(rowptr), _ptr(col), _ptr(x),
                                   C.addressof(b), _ptr(y), ngpu)
    return rc, y


def spMV_mgpu_v1(m, n, nnz, alpha, val, rowptr, col, x, beta, y, ngpu=1, kernel=1):
    rowptr, col, val, x, y = _dropin_args(rowptr, col, val, x, y)
    a, b = C.c_double(alpha), C.c_double(beta)
    rc = lib().spx_dropin_v1(m, n, nnz, C.addressof(a), _ptr(val), _ptr(rowptr), _ptr(col), _ptr(x),
                             C.addressof(b), _ptr(y), ngpu, kernel)
    return rc, y


def spMV_mgpu_v2(m, n, nnz, alpha, val, rowptr, col, x, beta, y, ngpu=1, kernel=1, nb=1, q=1):
    rowptr, col, val, x, y = _dropin_args(rowptr, col, val, x, y)
    a, b = C.c_double(alpha), C.c_double(beta)
    rc = lib().spx_dropin_v2(m, n, nnz, C.addressof(a), _ptr(val), _ptr(rowptr), _ptr(col), _ptr(x),
                             C.addressof(b), _ptr(y), ngpu, kernel, nb, q)
    return rc, y


DROPIN_PHASES = ["plan_ms", "matrix_upload_ms", "xy_upload_ms", "execute_ms", "download_ms", "free_ms", "total_ms"]


def dropin_last_times():
    """Phases (ms of host time) of this thread's last host-pointer drop-in call."""
    out = (C.c_double * 7)()
    k = lib().spx_dropin_last_times(out, 7)
    return {DROPIN_PHASES[i]: float(out[i]) for i in range(k)}


def csr5_kernel(m, n, nnz, alpha, val_dev, rowptr32_dev, col_dev, x_dev, beta, y_dev):
    a, b = C.c_double(alpha), C.c_double(beta)
    return lib().spx_dropin_csr5_kernel(m, n, nnz, C.addressof(a), val_dev.ptr, rowptr32_dev.ptr, col_dev.ptr,
                                        x_dev.ptr, C.addressof(b), y_dev.ptr)


def read_mtx(path):
    """A Matrix Market coordinate file as a CSR dict (rowptr int64, col int32,
    val float64, m, n, nnz) via spx_mm_read; raises SpxError on a bad file."""
    m, n, nnz = C.c_int(), C.c_int(), C.c_int64()
    rp, col, val = C.c_void_p(), C.c_void_p(), C.c_void_p()
    rc = lib().spx_mm_read(os.fsencode(path), C.byref(m), C.byref(n), C.byref(nnz), C.byref(rp), C.byref(col),
                           C.byref(val))
    if rc != 0:
        raise SpxError(rc, f"spx_mm_read({path!r}) failed with {rc}")
    try:
        M, Z = m.value, nnz.value
        rowptr = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int64)), (M + 1,)).copy()
        c = np.ctypeslib.as_array(C.cast(col, C.POINTER(C.c_int32)), (max(Z, 1),))[:Z].copy()
        v = np.ctypeslib.as_array(C.cast(val, C.POINTER(C.c_double)), (max(Z, 1),))[:Z].copy()
    finally:
        for p in (rp, col, val):
            lib().spx_mm_free(p)
    from .workloads import finish
    return finish(rowptr, c, v, m.value, n.value)


def get_row_from_index(rowptr, idx):
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int64)
    return lib().spx_get_row_from_index(len(rowptr) - 1, rowptr.ctypes.data, idx)


def set_v1_streams(parts):
    return lib().spx_set_v1_streams(parts)


# ---- generators (libspxgen.so) -------------------------------------------------------------

def gen():
    global _gen
    if _gen is not None:
        return _gen
    if not os.path.exists(GEN_PATH):
        raise RuntimeError(f"{GEN_PATH} is not built: run python sparse-blas_amd/build.py")
    G = C.CDLL(GEN_PATH)
    G.spx_gen_zipf_mean.restype = C.c_double
    G.spx_gen_zipf_mean.argtypes = [C.c_double, C.c_int]
    G.spx_gen_rowlens.restype = C.c_int64
    G.spx_gen_rowlens.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_void_p, C.c_int, C.c_void_p]
    G.spx_gen_fill.restype = C.c_int
    G.spx_gen_fill.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_void_p, C.c_int, C.c_void_p,
                               C.c_void_p, C.c_void_p]
    G.spx_gen_vector.restype = None
    G.spx_gen_vector.argtypes = [C.c_int64, C.c_uint64, C.c_int, C.c_void_p]
    G.spx_gen_g.restype = C.c_int64
    G.spx_gen_g.argtypes = [C.c_int, C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p]
    _gen = G
    return G
