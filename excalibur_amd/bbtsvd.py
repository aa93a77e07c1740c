"""Python mirror of the reference's bbtsvd.hpp functions on the bbtsvd path (bbtsvd.hpp:22-47),
over the host-pointer C ABI of include/excalibur_bbtsvd.h. Same names and argument meaning;
arrays are numpy float64 and results are returned (in-place functions also update their
argument). Every call runs on the GPU; a failure raises HipError (the reference asserts)."""
import ctypes

import numpy as np

from .hip import HipError, register, lib as _lib

_vp = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_S = ctypes.c_size_t
_U = ctypes.c_uint
_ip = ctypes.POINTER(ctypes.c_int)
_sp = ctypes.POINTER(ctypes.c_size_t)

SIGNATURES = {
    "exc_bb_negative_count": (_I, [_I, _vp, _vp, _D, _ip]),
    "exc_bb_bisection_singular_values": (_I, [_I, _vp, _vp, _D, _D, _D, _I, _vp, _ip]),
    "exc_bb_ldl_decomposition": (_I, [_I, _vp, _vp, _vp, _vp]),
    "exc_bb_inverse_iteration": (_I, [_I, _I, _vp, _vp, _vp, _I]),
    "exc_bb_dstqds": (_I, [_I, _vp, _vp, _D, _vp, _vp, _vp]),
    "exc_bb_dpqds": (_I, [_I, _vp, _vp, _D, _vp, _vp, _vp]),
    "exc_bb_one_pair_twisted": (_I, [_I, _I, _D, _vp, _vp, _vp, _vp, _vp, _vp]),
    "exc_bb_two_side_twisted": (_I, [_I, _I, _vp, _vp, _vp, _vp, _vp]),
    "exc_bb_bidiagonal_svd": (_I, [_I, _vp, _vp, _vp, _vp, _vp, _I]),
    "exc_bb_givens_mask": (_I, [_S, _S, _S, _vp, _I, _I, _U]),
    "exc_bb_block_shift_mask": (_I, [_S, _S, _S, _vp, _I, _I, _U]),
    "exc_bb_find_best_convert": (_I, [_S, _sp, _sp, _sp]),
}
register(SIGNATURES)


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _p(a):
    return a.ctypes.data if a.size else None


_EMPTY_PAD = np.zeros(1)  # lives as long as the module, so its pointer never dangles


def _pad(a):
    """A pointer that is valid even for an empty array (n = 1 bidiagonals have no beta)."""
    return _p(a) if a.size else _p(_EMPTY_PAD)


def _call(name, *args):
    rc = getattr(_lib, name)(*args)
    if rc != 0:
        raise HipError(rc, name)


def negative_count(n, x_diagonal, x_upper_diagonal, mu):
    a, b = _f64(x_diagonal), _f64(x_upper_diagonal)
    c = ctypes.c_int(0)
    _call("exc_bb_negative_count", n, _p(a), _pad(b), float(mu), ctypes.byref(c))
    return c.value


def bisection_singular_values(n, x_diagonal, x_upper_diagonal, a, b, tol, top_k):
    """Returns the sigma array (length sigma_count)."""
    d, e = _f64(x_diagonal), _f64(x_upper_diagonal)
    sigma = np.zeros(max(top_k, 1))
    cnt = ctypes.c_int(0)
    _call("exc_bb_bisection_singular_values", n, _p(d), _pad(e), float(a), float(b), float(tol), int(top_k),
          _p(sigma), ctypes.byref(cnt))
    return sigma[: cnt.value]


def ldl_decomposition(n, x_diagonal, x_upper_diagonal):
    a, b = _f64(x_diagonal), _f64(x_upper_diagonal)
    l, d = np.zeros(max(n - 1, 1)), np.zeros(n)
    _call("exc_bb_ldl_decomposition", n, _p(a), _pad(b), _p(l), _p(d))
    return l[: n - 1], d


def inverse_iteration(n, k, l, u, v, ldv):
    """Updates v (float64 array of at least (n-1)*ldv+1 entries) in place and returns it."""
    l, u = _f64(l), _f64(u)
    assert v.dtype == np.float64 and v.flags.c_contiguous and v.size >= (n - 1) * ldv + 1
    _call("exc_bb_inverse_iteration", n, k, _pad(l), _pad(u), _p(v), ldv)
    return v


def dstqds(n, l, d, tau):
    l, d = _f64(l), _f64(d)
    lp, dp, s = np.zeros(max(n - 1, 1)), np.zeros(n), np.zeros(n)
    _call("exc_bb_dstqds", n, _pad(l), _p(d), float(tau), _p(lp), _p(dp), _p(s))
    return lp[: n - 1], dp, s


def dpqds(n, l, d, tau):
    l, d = _f64(l), _f64(d)
    up, rp, p = np.zeros(max(n - 1, 1)), np.zeros(n), np.zeros(n)
    _call("exc_bb_dpqds", n, _pad(l), _p(d), float(tau), _p(up), _p(rp), _p(p))
    return up[: n - 1], rp, p


def one_pair_singular_vector_twisted_factorization(n, omege_count, omega, l, d, x_diagonal, x_upper_diagonal, v, u):
    """v: n contiguous entries, u: n entries at stride omege_count; both updated in place."""
    l, d, a, b = _f64(l), _f64(d), _f64(x_diagonal), _f64(x_upper_diagonal)
    assert v.dtype == np.float64 and u.dtype == np.float64 and u.size >= (n - 1) * omege_count + 1
    _call("exc_bb_one_pair_twisted", n, omege_count, float(omega), _pad(l), _p(d), _p(a), _pad(b), _p(v), _p(u))
    return v, u


def two_side_singular_vectors_twisted_factorization(n, omege_count, omege, x_diagonal, x_upper_diagonal):
    """Returns u (n x omege_count) and v (omege_count x n)."""
    w, a, b = _f64(omege), _f64(x_diagonal), _f64(x_upper_diagonal)
    u, v = np.zeros((n, omege_count)), np.zeros((omege_count, n))
    _call("exc_bb_two_side_twisted", n, omege_count, _pad(w), _p(a), _pad(b), _p(u), _p(v))
    return u, v


def bidiagonal_bisection_twisted_svd(k, alpha, beta, top_k):
    """7-argument form: returns (u (k x top_k), v (top_k x k), sigma (top_k))."""
    a, b = _f64(alpha), _f64(beta)
    u, v, s = np.zeros((k, top_k)), np.zeros((top_k, k)), np.zeros(max(top_k, 1))
    _call("exc_bb_bidiagonal_svd", k, _p(a), _pad(b), _pad(u), _pad(v), _p(s), top_k)
    return u, v, s[:top_k]


def givens_mask_cache_optimized_parallel_with_simd(m, n, mask_times, x, mask, is_left, seed):
    assert x.dtype == np.float64 and x.flags.c_contiguous and x.size == m * n
    _call("exc_bb_givens_mask", m, n, mask_times, _p(x), int(bool(mask)), int(bool(is_left)), seed & 0xFFFFFFFF)
    return x


def block_shift_mask(m, n, mask_times, x, mask, is_left, seed):
    assert x.dtype == np.float64 and x.flags.c_contiguous and x.size == m * n
    _call("exc_bb_block_shift_mask", m, n, mask_times, _p(x), int(bool(mask)), int(bool(is_left)), seed & 0xFFFFFFFF)
    return x


def find_best_convert(k, remain=0):
    """SquareMaskConvert {k, m, n, remain} after find_best_convert: returns (m, n, remain)."""
    mm, nn, extra = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
    _call("exc_bb_find_best_convert", k - remain, ctypes.byref(mm), ctypes.byref(nn), ctypes.byref(extra))
    return mm.value, nn.value, remain + extra.value
