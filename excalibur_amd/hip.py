"""ctypes binding of libexcalibur_amd.so's C ABI (include/excalibur_hip.h).

Raises ImportError when the library is missing: the package has no CPU path.
"""
import ctypes
import os

import numpy as np

from . import LIB_PATH

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run excalibur_amd.build.build() (or __graft_entry__.build())")

_lib = ctypes.CDLL(LIB_PATH)

_vp = ctypes.c_void_p
_dp = ctypes.c_void_p  # device pointers travel as void*
_I = ctypes.c_int
_L = ctypes.c_long
_D = ctypes.c_double
_S = ctypes.c_size_t
_U = ctypes.c_uint

SIGNATURES = {
    "exc_last_error": (ctypes.c_char_p, []),
    "exc_version": (_I, []),
    "exc_device_count": (_I, [ctypes.POINTER(_I)]),
    "exc_set_device": (_I, [_I]),
    "exc_malloc": (_I, [ctypes.POINTER(_vp), _S]),
    "exc_free": (_I, [_vp]),
    "exc_host_alloc": (_I, [ctypes.POINTER(_vp), _S]),
    "exc_host_free": (_I, [_vp]),
    "exc_memcpy_h2d": (_I, [_vp, _vp, _S, _vp]),
    "exc_memcpy_d2h": (_I, [_vp, _vp, _S, _vp]),
    "exc_memcpy_d2d": (_I, [_vp, _vp, _S, _vp]),
    "exc_memset": (_I, [_vp, _I, _S, _vp]),
    "exc_stream_create": (_I, [ctypes.POINTER(_vp)]),
    "exc_stream_destroy": (_I, [_vp]),
    "exc_stream_sync": (_I, [_vp]),
    "exc_sturm_count": (_I, [_I, _dp, _dp, _dp, _I, _dp, _vp]),
    "exc_bisection_workspace": (_S, [_I, _I]),
    "exc_bisection": (_I, [_I, _dp, _dp, _D, _D, _D, _I, _dp, _dp, _dp, _dp, _dp, _S, _vp]),
    "exc_ldl": (_I, [_I, _dp, _dp, _dp, _dp, _vp]),
    "exc_twisted_workspace": (_S, [_I, _I]),
    "exc_twisted_vectors": (_I, [_I, _I, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _S, _vp]),
    "exc_bidiag_svd_workspace": (_S, [_I, _I]),
    "exc_bidiag_svd": (_I, [_I, _dp, _dp, _dp, _dp, _dp, _I, _dp, _dp, _S, _vp]),
    "exc_givens_half": (_L, [_L]),
    "exc_givens_plan": (_I, [_L, _L, _I, _U, _vp, _vp]),
    "exc_givens_apply": (_I, [_dp, _L, _L, _dp, _dp, _L, _I, _I, _vp]),
    "exc_givens_apply_f32": (_I, [_dp, _L, _L, _dp, _dp, _L, _I, _I, _vp]),
}

EXTRA_SIGNATURES = {}


def _bind(table):
    for name, (res, args) in table.items():
        f = getattr(_lib, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args


_bind(SIGNATURES)


def register(table):
    """Adds signatures for entry points declared in other parts of the ABI."""
    EXTRA_SIGNATURES.update(table)
    _bind(table)


lib = _lib


class HipError(RuntimeError):
    def __init__(self, rc, name):
        super().__init__(f"{name} returned {rc}: {_lib.exc_last_error().decode(errors='replace')}")
        self.rc = rc


def call(name, *args):
    rc = getattr(_lib, name)(*args)
    if rc != 0:
        raise HipError(rc, name)
    return rc


class DeviceArray:
    """A device buffer with a numpy shape/dtype; freed on close()/GC."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = _vp()
        call("exc_malloc", ctypes.byref(p), max(self.nbytes, 1))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.shape, a.dtype)
        if a.nbytes:
            call("exc_memcpy_h2d", d.ptr, a.ctypes.data, a.nbytes, None)
        return d

    def to_host(self):
        out = np.empty(self.shape, dtype=self.dtype)
        if self.nbytes:
            call("exc_memcpy_d2h", out.ctypes.data, self.ptr, self.nbytes, None)
        return out

    def zero(self):
        call("exc_memset", self.ptr, 0, max(self.nbytes, 1), None)
        sync()
        return self

    def close(self):
        if getattr(self, "ptr", None):
            _lib.exc_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sync(stream=None):
    call("exc_stream_sync", stream)


def device_count():
    n = ctypes.c_int(0)
    rc = _lib.exc_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


def workspace(nbytes):
    return DeviceArray((max(int(nbytes), 1),), np.uint8)


# ------------------------------------------------------------------ convenience wrappers
def sturm_count(alpha, beta, mus):
    n = len(alpha)
    da = DeviceArray.from_host(np.asarray(alpha, np.float64))
    db = DeviceArray.from_host(np.asarray(beta if len(beta) else [0.0], np.float64))
    dm = DeviceArray.from_host(np.asarray(mus, np.float64))
    dc = DeviceArray((len(mus),), np.int32)
    call("exc_sturm_count", n, da.ptr, db.ptr, dm.ptr, len(mus), dc.ptr, None)
    sync()
    return dc.to_host()


def bisection(alpha, beta, lo, hi, tol=1e-20, top_k=None):
    n = len(alpha)
    top_k = n if top_k is None else top_k
    da = DeviceArray.from_host(np.asarray(alpha, np.float64))
    db = DeviceArray.from_host(np.asarray(beta if len(beta) else [0.0], np.float64))
    ds = DeviceArray((max(top_k, 1),))
    dlo = DeviceArray((max(top_k, 1),))
    dhi = DeviceArray((max(top_k, 1),))
    dc = DeviceArray((1,), np.int32)
    ws = workspace(_lib.exc_bisection_workspace(n, top_k))
    call("exc_bisection", n, da.ptr, db.ptr, lo, hi, tol, top_k, ds.ptr, dc.ptr, dlo.ptr, dhi.ptr, ws.ptr,
         ws.nbytes, None)
    sync()
    c = int(dc.to_host()[0])
    return ds.to_host()[:c], dlo.to_host()[:c], dhi.to_host()[:c]


def twisted_vectors(alpha, beta, omega):
    n, k = len(alpha), len(omega)
    da = DeviceArray.from_host(np.asarray(alpha, np.float64))
    db = DeviceArray.from_host(np.asarray(beta if len(beta) else [0.0], np.float64))
    do = DeviceArray.from_host(np.asarray(omega, np.float64))
    du = DeviceArray((n, k)); dv = DeviceArray((k, n))
    dt = DeviceArray((2 * max(k, 1),), np.int32)
    ws = workspace(_lib.exc_twisted_workspace(n, k))
    call("exc_twisted_vectors", n, k, do.ptr, da.ptr, db.ptr, du.ptr, dv.ptr, dt.ptr, ws.ptr, ws.nbytes, None)
    sync()
    return du.to_host(), dv.to_host(), dt.to_host()[: 2 * k].reshape(k, 2)


def bidiag_svd(alpha, beta, top_k=None):
    n = len(alpha)
    k = n if top_k is None else top_k
    da = DeviceArray.from_host(np.asarray(alpha, np.float64))
    db = DeviceArray.from_host(np.asarray(beta if len(beta) else [0.0], np.float64))
    du = DeviceArray((n, max(k, 1))); dv = DeviceArray((max(k, 1), n)); ds = DeviceArray((max(k, 1),))
    dt = DeviceArray((2 * max(k, 1),), np.int32)
    ws = workspace(_lib.exc_bidiag_svd_workspace(n, k))
    rc = _lib.exc_bidiag_svd(n, da.ptr, db.ptr, du.ptr, dv.ptr, ds.ptr, k, dt.ptr, ws.ptr, ws.nbytes, None)
    if rc not in (0, -5):
        raise HipError(rc, "exc_bidiag_svd")
    sync()
    return (ds.to_host()[:k], du.to_host()[:, :k], dv.to_host()[:k], dt.to_host()[: 2 * k].reshape(k, 2), rc)


def givens_plan(K, passes, mask, seed):
    half = _lib.exc_givens_half(K)
    pairs = np.zeros((passes, half, 2), np.int32)
    cs = np.zeros((passes, half, 2), np.float64)
    call("exc_givens_plan", K, passes, 1 if mask else 0, seed & 0xFFFFFFFF, pairs.ctypes.data, cs.ctypes.data)
    return pairs, cs


def givens_apply(x, pairs, cs, mask, is_left, f32=False):
    x = np.ascontiguousarray(x, np.float32 if f32 else np.float64)
    m, n = x.shape
    dx = DeviceArray.from_host(x)
    dp = DeviceArray.from_host(pairs)
    dc = DeviceArray.from_host(cs)
    call("exc_givens_apply_f32" if f32 else "exc_givens_apply", dx.ptr, m, n, dp.ptr, dc.ptr, pairs.shape[0],
         1 if mask else 0, 1 if is_left else 0, None)
    sync()
    return dx.to_host()
