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
    "exc_prof_enable": (_I, [_I]),
    "exc_prof_report": (_I, [ctypes.c_char_p, _S]),
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
    "exc_gemm_workspace": (_S, [_L, _L, _L]),
    "exc_dgemm": (_I, [_I, _I, _L, _L, _L, _D, _dp, _L, _dp, _L, _D, _dp, _L, _dp, _S, _vp]),
    "exc_sgemm": (_I, [_I, _I, _L, _L, _L, _D, _dp, _L, _dp, _L, _D, _dp, _L, _dp, _S, _vp]),
    "exc_fact_workspace": (_S, [_I, _L]),
    "exc_potrf_lower": (_I, [_I, _dp, _L, _dp, _dp, _S, _vp]),
    "exc_trsm_left": (_I, [_I, _I, _I, _L, _dp, _L, _dp, _L, _dp, _S, _vp]),
    "exc_trsm_right": (_I, [_I, _I, _I, _L, _dp, _L, _dp, _L, _dp, _S, _vp]),
    "exc_lu_sign": (_I, [_I, _dp, _L, _dp, _I, _dp, _S, _vp]),
    "exc_qr_t_workspace": (_S, [_I, _L, _I]),
    "exc_qr_t": (_I, [_I, _L, _dp, _L, _dp, _dp, _L, _dp, _L, _dp, _dp, _I, _dp, _S, _vp]),
    "exc_wy_workspace": (_S, [_I, _L, _L]),
    "exc_apply_q_right": (_I, [_L, _L, _I, _dp, _L, _dp, _L, _dp, _dp, _S, _vp]),
    "exc_orgqr_t": (_I, [_I, _L, _dp, _L, _dp, _dp, _L, _dp, _S, _vp]),
    "exc_ormqr_house_workspace": (_S, [_I, _L]),
    "exc_ormqr_house_ln": (_I, [_I, _L, _dp, _dp, _dp, _L, _dp, _S, _vp]),
    "exc_ralha_workspace": (_S, [_I, _I]),
    "exc_ralha_step": (_I, [_I, _I, _I, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _vp, _S, _vp]),
    "exc_ralha_step_generic": (_I, [_I, _I, _I, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _vp, _S, _vp]),
    "exc_ralha_loop_workspace": (_S, [_I, _I]),
    "exc_ralha_link_alloc": (_I, [_I, ctypes.POINTER(_vp)]),
    "exc_ralha_link_free": (_I, [_vp]),
    "exc_ralha_loop_fits": (_I, [_I, _I, _I]),
    "exc_ralha_loop_start": (_I, [_I, _I, _dp, _dp, _dp, _dp, _dp, _dp, _vp, _vp, _S, _I, _D, _vp]),
    "exc_ralha_loop_wait": (_I, [_vp, _I, _D, ctypes.POINTER(_vp)]),
    "exc_ralha_loop_post": (_I, [_vp, _I]),
    "exc_ralha_loop_abort": (_I, [_vp]),
    "exc_ralha_loop_finish": (_I, [_vp, _vp]),
    "exc_ralha_loop_grid": (_I, [_I, _I, _I]),
    "exc_ralha_loop_shape": (_I, [_I, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_S)]),
    "exc_ralha_loop_slots": (_I, [_I, _I, _I]),
    "exc_ralha_loops_colocate": (_I, [_I, ctypes.POINTER(_I), ctypes.POINTER(_I), _I]),
    "exc_ralha_box_bytes": (_S, [_I]),
    "exc_ralha_box_clear": (_I, [_dp, _vp]),
    "exc_ralha_loop_start_boxes": (_I, [_I, _I, _dp, _dp, _dp, _dp, _dp, _dp, _vp, _vp, _S, _I, _D, _I, _I,
                                        ctypes.POINTER(_vp), ctypes.POINTER(_I), _vp]),
    "exc_ralha_loop_running": (_I, [_vp]),
    "exc_ipc_handle": (_I, [_dp, _vp]),
    "exc_ipc_open": (_I, [_vp, ctypes.POINTER(_vp)]),
    "exc_ipc_close": (_I, [_dp]),
    "exc_ralha_dot": (_I, [_I, _I, _I, _dp, _dp, _dp, _vp]),
    "exc_ralha_tail_update": (_I, [_I, _I, _I, _dp, _dp, _dp, _dp, _dp, _vp]),
    "exc_ralha_tail_scale": (_I, [_I, _I, _I, _dp, _dp, _dp, _dp, _vp]),
    "exc_lower_to_upper": (_I, [_I, _dp, _dp, _dp, _vp]),
    "exc_sign_normalize": (_I, [_I, _I, _dp, _dp, _dp, _vp]),
    "exc_rot_back": (_I, [_I, _I, _dp, _dp, _vp]),
    "exc_bisection_bound": (_I, [_I, _dp, _dp, _dp, _vp]),
    "exc_copy2d": (_I, [_L, _L, _dp, _L, _dp, _L, _vp]),
    "exc_transpose": (_I, [_L, _L, _dp, _L, _dp, _L, _vp]),
    "exc_fill2d": (_I, [_L, _L, _D, _dp, _L, _vp]),
    "exc_reflector_message_len": (_L, [_L, _L, _L]),
    "exc_pack_reflectors": (_I, [_L, _L, _L, _dp, _L, _dp, _dp, _vp]),
    "exc_unpack_reflectors": (_I, [_L, _L, _L, _dp, _dp, _L, _dp, _vp]),
    "exc_bidiag_svd_noalloc_capacity": (_S, []),
    "exc_bidiag_svd_noalloc": (_I, [_I, _dp, _dp, _dp, _dp, _dp, _I, _dp, _vp]),
    "exc_bidiag_local_workspace": (_S, [_I, _I]),
    "exc_bidiag_local": (_I, [_I, _I, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _vp, _I, _dp, _S, _D, _vp]),
    "exc_pack_lower_f32": (_I, [_I, _vp, _L, _vp, _vp]),
    "exc_pack_lower_f64": (_I, [_I, _dp, _L, _dp, _vp]),
    "exc_unpack_lower_f64": (_I, [_I, _dp, _dp, _L, _vp]),
    "exc_acc_packed": (_I, [_L, _vp, _I, _I, _dp, _vp]),
    "exc_convert_f64_f32": (_I, [_L, _L, _dp, _L, _dp, _L, _vp]),
    "exc_convert_f32_f64": (_I, [_L, _L, _dp, _L, _dp, _L, _vp]),
    "exc_trace": (_I, [_I, _dp, _L, _dp, _vp]),
    "exc_add_diag": (_I, [_I, _dp, _L, _D, _vp]),
    "exc_dev_identity_workspace": (_S, [_I]),
    "exc_dev_identity": (_I, [_I, _dp, _L, _dp, _vp, _S, _vp]),
    "exc_oa_refine_workspace": (_S, [_I]),
    "exc_oa_refine": (_I, [_I, _dp, _L, _I, _dp, _L, _dp, _L, _dp, _I, _D, _D, _dp, _S, _vp]),
    "exc_givens_apply_batched": (_I, [_I, _vp, _L, _L, _vp, _vp, _L, _I, _I, _vp]),
    "exc_block_shift_mask_workspace": (_S, [_L, _L, _L, _I]),
    "exc_block_shift_mask": (_I, [_dp, _L, _L, _L, _I, _I, _U, _vp, _S, _vp]),
    "exc_batched_local_qr_workspace": (_S, [_I, _I, _vp]),
    "exc_batched_local_qr": (_I, [_I, _I, _vp, _vp, _vp, _vp, _vp, _vp, _S, _vp]),
    "exc_gemm_grouped_tuning": (_I, [ctypes.c_char_p, _S]),
    "exc_gk_max_n": (_I, []),
    "exc_gk_bidiagonals": (_I, [_I] + [_vp] * 6 + [_vp]),
    "exc_bidiag_scale": (_I, [_I] + [_vp] * 6 + [_vp]),
    "exc_count_features": (_I, [_I, _I] + [_vp] * 6 + [_vp]),
    "exc_bisect_index": (_I, [_I] + [_vp] * 5 + [_D] +
                         [_vp] * 11 + [_vp]),
    "exc_mlp_param_count": (_L, []),
    "exc_mlp_dims": (_I, [ctypes.POINTER(_I)] * 3),
    "exc_mlp_acts_floats": (_S, [_L]),
    "exc_mlp_backward_scratch_floats": (_S, [_L]),
    "exc_mlp_forward": (_I, [_L] + [_vp] * 5 + [_S, _vp]),
    "exc_mlp_backward": (_I, [_L] + [_vp] * 9 + [_S, _vp]),
    "exc_mlp_adam": (_I, [_vp] * 4 + [ctypes.c_float] * 4 + [_I, _vp]),
    "exc_mlp_to_bf16": (_I, [_vp] * 3),
    "exc_mlp_infer_bf16": (_I, [_L] + [_vp] * 3 + [_vp]),

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


def gemm(A, B, ta=False, tb=False, alpha=1.0, beta=0.0, C=None, f32=False, split=True):
    dt = np.float32 if f32 else np.float64
    A = np.ascontiguousarray(A, dt); B = np.ascontiguousarray(B, dt)
    M = A.shape[1] if ta else A.shape[0]
    K = A.shape[0] if ta else A.shape[1]
    N = B.shape[0] if tb else B.shape[1]
    dA = DeviceArray.from_host(A); dB = DeviceArray.from_host(B)
    dC = DeviceArray.from_host(np.ascontiguousarray(C, dt)) if C is not None else DeviceArray((M, N), dt).zero()
    ws = workspace(_lib.exc_gemm_workspace(M, N, K)) if split else None
    call("exc_sgemm" if f32 else "exc_dgemm", int(ta), int(tb), M, N, K, alpha, dA.ptr, A.shape[1], dB.ptr,
         B.shape[1], beta, dC.ptr, N, ws.ptr if ws else None, ws.nbytes if ws else 0, None)
    sync()
    return dC.to_host()


def potrf_lower(G):
    G = np.ascontiguousarray(G, np.float64)
    n = G.shape[0]
    d = DeviceArray.from_host(G); info = DeviceArray((1,), np.int32).zero()
    ws = workspace(_lib.exc_fact_workspace(n, n))
    call("exc_potrf_lower", n, d.ptr, n, info.ptr, ws.ptr, ws.nbytes, None)
    sync()
    return d.to_host(), int(info.to_host()[0])


def trsm_left(T, X, upper=False, unit=False):
    T = np.ascontiguousarray(T, np.float64); X = np.ascontiguousarray(X, np.float64)
    n, nrhs = X.shape
    dT = DeviceArray.from_host(T); dX = DeviceArray.from_host(X)
    ws = workspace(_lib.exc_fact_workspace(n, nrhs))
    call("exc_trsm_left", int(upper), int(unit), n, nrhs, dT.ptr, n, dX.ptr, nrhs, ws.ptr, ws.nbytes, None)
    sync()
    return dX.to_host()


def trsm_right(T, X, lower_t=False, unit=False):
    T = np.ascontiguousarray(T, np.float64); X = np.ascontiguousarray(X, np.float64)
    rows, n = X.shape
    dT = DeviceArray.from_host(T); dX = DeviceArray.from_host(X)
    ws = workspace(_lib.exc_fact_workspace(n, rows))
    call("exc_trsm_right", int(lower_t), int(unit), n, rows, dT.ptr, n, dX.ptr, n, ws.ptr, ws.nbytes, None)
    sync()
    return dX.to_host()


def lu_sign(A):
    A = np.ascontiguousarray(A, np.float64)
    n = A.shape[0]
    dA = DeviceArray.from_host(A); ds = DeviceArray((n,))
    ws = workspace(_lib.exc_fact_workspace(n, n))
    call("exc_lu_sign", n, dA.ptr, n, ds.ptr, 0, ws.ptr, ws.nbytes, None)
    sync()
    return dA.to_host(), ds.to_host()


def qr_t(X, want_q=True, want_v=False, square_last=True):
    """QR of X^T (X: m x r). Returns dict rt, qt, vt, tau, sgn (numpy)."""
    X = np.ascontiguousarray(X, np.float64)
    m, r = X.shape
    dX = DeviceArray.from_host(X)
    drt = DeviceArray((m, m))
    dqt = DeviceArray((m, r)) if want_q else None
    dvt = DeviceArray((m, r)) if want_v else None
    dtau = DeviceArray((m,)) if want_v else None
    dsg = DeviceArray((m,))
    ws = workspace(_lib.exc_qr_t_workspace(m, r, int(want_q)))
    call("exc_qr_t", m, r, dX.ptr, r, drt.ptr, dqt.ptr if dqt else None, r, dvt.ptr if dvt else None, r,
         dtau.ptr if dtau else None, dsg.ptr, int(square_last), ws.ptr, ws.nbytes, None)
    sync()
    out = dict(rt=drt.to_host(), sgn=dsg.to_host())
    if dqt: out["qt"] = dqt.to_host()
    if dvt: out["vt"] = dvt.to_host(); out["tau"] = dtau.to_host()
    return out


def apply_q_right(Mt, vt, tau):
    Mt = np.ascontiguousarray(Mt, np.float64); vt = np.ascontiguousarray(vt, np.float64)
    rows, L = Mt.shape
    k = vt.shape[0]
    dM = DeviceArray.from_host(Mt); dV = DeviceArray.from_host(vt); dT = DeviceArray.from_host(np.asarray(tau, np.float64))
    ws = workspace(_lib.exc_wy_workspace(k, L, rows))
    call("exc_apply_q_right", rows, L, k, dM.ptr, L, dV.ptr, L, dT.ptr, ws.ptr, ws.nbytes, None)
    sync()
    return dM.to_host()


def orgqr_t(vt, tau):
    vt = np.ascontiguousarray(vt, np.float64)
    k, L = vt.shape
    dV = DeviceArray.from_host(vt); dT = DeviceArray.from_host(np.asarray(tau, np.float64))
    dQ = DeviceArray((k, L))
    ws = workspace(_lib.exc_wy_workspace(k, L, k))
    call("exc_orgqr_t", k, L, dV.ptr, L, dT.ptr, dQ.ptr, L, ws.ptr, ws.nbytes, None)
    sync()
    return dQ.to_host()


def ormqr_house_ln(house, tau, C):
    house = np.ascontiguousarray(house, np.float64); C = np.ascontiguousarray(C, np.float64)
    m, n = C.shape
    dH = DeviceArray.from_host(house); dT = DeviceArray.from_host(np.asarray(tau, np.float64))
    dC = DeviceArray.from_host(C)
    ws = workspace(_lib.exc_ormqr_house_workspace(m, n))
    call("exc_ormqr_house_ln", m, n, dH.ptr, dT.ptr, dC.ptr, n, ws.ptr, ws.nbytes, None)
    sync()
    return dC.to_host()


def ralha_peers(F_list, generic=False, loop=False, cols_per_wave=0, timeout_s=30.0, boxes=False):
    """The overlapped Ralha loop over P simulated peers in this process: GPU kernels per peer,
    the all-reduce done here in each peer's own order (own buffer, then the others by id).
    F_list[p] is peer p's m x c_p share of the rows f_j; U is returned in the same orientation.
    generic selects the streamed (HBM) form of the step kernel instead of the register form;
    loop runs steps -1 .. m-3 as one launch per peer (one stream each, all running at once) with
    this thread all-reducing the messages through the link buffers; with boxes the launches
    exchange the messages through device-memory boxes instead and this thread only waits."""
    P = len(F_list)
    m = F_list[0].shape[0]
    cs = [f.shape[1] for f in F_list]
    dF = [DeviceArray.from_host(np.ascontiguousarray(np.asarray(f, np.float64).T)) for f in F_list]
    dU = [DeviceArray((c, m)).zero() for c in cs]
    da = [DeviceArray((m,)).zero() for _ in range(P)]
    db = [DeviceArray((m,)).zero() for _ in range(P)]
    dh = [DeviceArray((m, m)).zero() for _ in range(P)]
    dt = [DeviceArray((m,)).zero() for _ in range(P)]
    dmsg = [DeviceArray((m + 3,)).zero() for _ in range(P)]
    dg = [DeviceArray((m + 3,)).zero() for _ in range(P)]
    ds = [DeviceArray((4,)).zero() for _ in range(P)]
    ws = [workspace(_lib.exc_ralha_workspace(m, c)) for c in cs]

    def allreduce(srcs, n):
        vals = [a.to_host()[:n].copy() for a in srcs]
        out = []
        for p in range(P):
            acc = vals[p].copy()
            for q in range(P):
                if q != p:
                    acc = acc + vals[q]
            out.append(acc)
        return out

    def put(dst, vec, off=0):
        buf = dst.to_host()
        buf[off:off + vec.size] = vec
        call("exc_memcpy_h2d", dst.ptr, buf.ctypes.data, buf.nbytes, None)

    def step(i, p):
        call("exc_ralha_step_generic" if generic else "exc_ralha_step", i, m, cs[p], dF[p].ptr, dU[p].ptr, dg[p].ptr, dh[p].ptr, dt[p].ptr, da[p].ptr,
             db[p].ptr, dmsg[p].ptr, ws[p].ptr, ws[p].nbytes, None)

    for p in range(P):
        call("exc_ralha_tail_update", -1, m, cs[p], dF[p].ptr, dU[p].ptr, None, db[p].ptr, ds[p].ptr, None)
    sync()
    red = allreduce(ds, 1)
    for p in range(P):
        put(ds[p], red[p], 1)
        call("exc_ralha_tail_scale", 0, m, cs[p], dF[p].ptr, dU[p].ptr, ds[p].ptr + 8, da[p].ptr, None)
        if not loop:
            step(-1, p)
    if loop:
        _ralha_loop_run(m, cs, dF, dU, dh, dt, da, db, cols_per_wave, timeout_s, boxes)
    else:
        for i in range(m - 2):
            sync()
            red = allreduce(dmsg, m + 3 - i)
            for p in range(P):
                put(dg[p], red[p])
                step(i, p)
    sync()
    for i in range(m - 2, m):
        for p in range(P):
            call("exc_ralha_dot", i, m, cs[p], dF[p].ptr, dU[p].ptr, ds[p].ptr, None)
        sync()
        red = allreduce(ds, 1)
        for p in range(P):
            put(ds[p], red[p], 1)
            call("exc_ralha_tail_update", i, m, cs[p], dF[p].ptr, dU[p].ptr, ds[p].ptr + 8, db[p].ptr,
                 ds[p].ptr + 16, None)
        sync()
        red2 = [a.to_host()[2:3].copy() for a in ds]
        for p in range(P):
            acc = red2[p].copy()
            for q in range(P):
                if q != p:
                    acc = acc + red2[q]
            put(ds[p], acc, 3)
            call("exc_ralha_tail_scale", i, m, cs[p], dF[p].ptr, dU[p].ptr, ds[p].ptr + 24, da[p].ptr, None)
        sync()
    return [dict(alpha=da[p].to_host(), beta=db[p].to_host()[: m - 1], house=dh[p].to_host(), tau=dt[p].to_host(),
                 U=np.ascontiguousarray(dU[p].to_host().T)) for p in range(P)]


_LOOP_STREAMS = []


def _loop_streams(P):
    """P streams for concurrent loop launches, created once per process and reused, so their
    hardware queues are assigned while few streams exist (loops on one queue would serialize)."""
    while len(_LOOP_STREAMS) < P:
        st = _vp()
        call("exc_stream_create", ctypes.byref(st))
        _LOOP_STREAMS.append(st.value)
    return _LOOP_STREAMS[:P]


def _ralha_loop_run(m, cs, dF, dU, dh, dt, da, db, cols_per_wave, timeout_s, boxes=False):
    """Steps -1 .. m-3 of every peer as one launch each, on P streams; messages all-reduced here
    (or, with boxes, exchanged by the launches themselves through device-memory boxes)."""
    import time
    P = len(cs)
    sync()
    for c in cs:
        if not _lib.exc_ralha_loop_fits(m, c, cols_per_wave):
            raise HipError(-1, "exc_ralha_loop_fits")
    streams, links, ws = _loop_streams(P), [], []
    if boxes:
        box = [DeviceArray((_lib.exc_ralha_box_bytes(m),), np.uint8) for _ in range(P)]
        for b in box:
            call("exc_ralha_box_clear", b.ptr, None)
        sync()
        box_ptrs = (_vp * P)(*[b.ptr for b in box])
        nwg = (_I * P)(*[_lib.exc_ralha_loop_grid(m, c, cols_per_wave) for c in cs])
    try:
        for p in range(P):
            ln = _vp()
            call("exc_ralha_link_alloc", m, ctypes.byref(ln))
            links.append(ln.value)
            ws.append(workspace(_lib.exc_ralha_loop_workspace(m, cs[p])))
        started = []
        try:
            for p in range(P):
                if boxes:
                    call("exc_ralha_loop_start_boxes", m, cs[p], dF[p].ptr, dU[p].ptr, dh[p].ptr, dt[p].ptr,
                         da[p].ptr, db[p].ptr, links[p], ws[p].ptr, ws[p].nbytes, cols_per_wave, timeout_s, P, p,
                         box_ptrs, nwg, streams[p])
                else:
                    call("exc_ralha_loop_start", m, cs[p], dF[p].ptr, dU[p].ptr, dh[p].ptr, dt[p].ptr, da[p].ptr,
                         db[p].ptr, links[p], ws[p].ptr, ws[p].nbytes, cols_per_wave, timeout_s, streams[p])
                started.append(p)
            t0 = time.time()
            while boxes and any(_lib.exc_ralha_loop_running(st) == 1 for st in streams):
                if time.time() - t0 > timeout_s:
                    raise HipError(-4, "exc_ralha_loop_running")
                time.sleep(2e-4)
            for j in range(0 if not boxes else m - 2, m - 2):
                n = m + 3 - j
                views = []
                for p in range(P):
                    ptr = _vp()
                    call("exc_ralha_loop_wait", links[p], j, timeout_s, ctypes.byref(ptr))
                    views.append(np.ctypeslib.as_array((ctypes.c_double * n).from_address(ptr.value)))
                vals = [v.copy() for v in views]
                for p in range(P):
                    acc = vals[p].copy()
                    for q in range(P):
                        if q != p:
                            acc = acc + vals[q]
                    views[p][:] = acc
                    call("exc_ralha_loop_post", links[p], j)
        except BaseException:
            for p in started:
                _lib.exc_ralha_loop_abort(links[p])
            for p in started:
                _lib.exc_ralha_loop_finish(links[p], streams[p])
            raise
        for p in started:
            call("exc_ralha_loop_finish", links[p], streams[p])
    finally:
        for ln in links:
            _lib.exc_ralha_link_free(ln)


def lower_to_upper(alpha, beta):
    m = len(alpha)
    da = DeviceArray.from_host(np.asarray(alpha, np.float64))
    db = DeviceArray.from_host(np.concatenate([np.asarray(beta, np.float64), [0.0]]))
    dc = DeviceArray((2 * m,)).zero()
    call("exc_lower_to_upper", m, da.ptr, db.ptr, dc.ptr, None)
    sync()
    return da.to_host(), db.to_host()[: m - 1], dc.to_host()[: 2 * (m - 1)]


def sign_normalize(bu, bv):
    bu = np.ascontiguousarray(bu, np.float64); bv = np.ascontiguousarray(bv, np.float64)
    m, k = bu.shape
    du = DeviceArray.from_host(bu); dv = DeviceArray.from_host(bv); df = DeviceArray((max(k, 1),))
    call("exc_sign_normalize", m, k, du.ptr, dv.ptr, df.ptr, None)
    sync()
    return du.to_host(), dv.to_host()


def rot_back(bu, cs):
    bu = np.ascontiguousarray(bu, np.float64)
    m, k = bu.shape
    du = DeviceArray.from_host(bu); dc = DeviceArray.from_host(np.asarray(cs, np.float64))
    call("exc_rot_back", m, k, du.ptr, dc.ptr, None)
    sync()
    return du.to_host()


def oa_refine(L, U, W, iters=1, tau=1e-6, floor_rel=1e-4, lower=True):
    """FP64 Ogita-Aishima refinement of L ~ U diag(s) W^T on the GPU (exc_oa_refine).
    Returns (U, W, sigma)."""
    L = np.ascontiguousarray(L, dtype=np.float64)
    m = L.shape[0]
    dL = DeviceArray.from_host(L)
    dU = DeviceArray.from_host(U)
    dW = DeviceArray.from_host(W)
    ds = DeviceArray((m,)).zero()
    ws = workspace(lib.exc_oa_refine_workspace(m))
    call("exc_oa_refine", m, dL.ptr, m, 1 if lower else 0, dU.ptr, m, dW.ptr, m, ds.ptr, iters, tau, floor_rel,
         ws.ptr, ws.nbytes, None)
    sync()
    return dU.to_host(), dW.to_host(), ds.to_host()


def bidiag_local(F, one_launch=True, timeout_s=30.0):
    """One-sided bidiagonalization of F (m x c, one owner) and lower to upper (exc_bidiag_local).
    Returns dict(alpha, beta, house, tau, U (m x c), cs) in the oracle's layouts."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    m, c = F.shape
    dF = DeviceArray.from_host(np.ascontiguousarray(F.T))  # column-major storage of F
    dU = DeviceArray((c, m)).zero()
    dh = DeviceArray((m, m)).zero()
    dt = DeviceArray((m,)).zero()
    da = DeviceArray((m,)).zero()
    db = DeviceArray((m,)).zero()
    dcs = DeviceArray((2 * m,)).zero()
    ws = workspace(lib.exc_bidiag_local_workspace(m, c))
    link = ctypes.c_void_p()
    if one_launch:
        call("exc_ralha_link_alloc", m, ctypes.byref(link))
    try:
        call("exc_bidiag_local", m, c, dF.ptr, dU.ptr, dh.ptr, dt.ptr, da.ptr, db.ptr, dcs.ptr,
             link if one_launch else None, 1, ws.ptr, ws.nbytes, timeout_s, None)
        sync()
    finally:
        if one_launch:
            lib.exc_ralha_link_free(link)
    return dict(alpha=da.to_host(), beta=db.to_host()[: m - 1], house=dh.to_host(), tau=dt.to_host(),
                U=np.ascontiguousarray(dU.to_host().T), cs=dcs.to_host()[: 2 * (m - 1)])


def block_shift_mask(x, passes, mask, is_left, seed):
    """block_shift_mask (bbtsvd.cpp:533-669) on the GPU, any K (exc_block_shift_mask)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    m, n = x.shape
    dx = DeviceArray.from_host(x)
    ws = workspace(lib.exc_block_shift_mask_workspace(m, n, passes, 1 if is_left else 0) + 1)
    call("exc_block_shift_mask", dx.ptr, m, n, passes, 1 if mask else 0, 1 if is_left else 0, seed & 0xFFFFFFFF,
         ws.ptr, ws.nbytes, None)
    sync()
    return dx.to_host()
