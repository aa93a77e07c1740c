"""ctypes binding of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the checker and by
bench.py's cpu_baseline leg. The product package (excalibur_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_L = ctypes.c_long
_I = ctypes.c_int
_D = ctypes.c_double


def build():
    subprocess.run(["bash", os.path.join(_HERE, "build.sh")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_nrm2": (_D, [_L, _dp, _L]),
            "orc_idamax": (_L, [_L, _dp, _L]),
            "orc_idamin": (_L, [_L, _dp, _L]),
            "orc_drotg": (None, [_dp, _dp, _dp, _dp]),
            "orc_dlarfg": (None, [_L, _dp, _dp, _L, _dp]),
            "orc_dgeqrf": (None, [_L, _L, _dp, _L, _dp]),
            "orc_dorgqr": (None, [_L, _L, _L, _dp, _L, _dp]),
            "orc_dormqr_left": (None, [_I, _L, _L, _L, _dp, _L, _dp, _dp, _L]),
            "orc_negative_count": (_I, [_I, _dp, _dp, _D]),
            "orc_negative_count_batch": (None, [_I, _dp, _dp, _dp, _I, _ip]),
            "orc_bisection": (None, [_I, _dp, _dp, _D, _D, _D, _I, _dp, _ip, _dp, _dp]),
            "orc_ldl": (None, [_I, _dp, _dp, _dp, _dp]),
            "orc_inverse_iteration": (None, [_I, _I, _dp, _dp, _dp, _L]),
            "orc_dstqds": (None, [_I, _dp, _dp, _D, _dp, _dp, _dp]),
            "orc_dpqds": (None, [_I, _dp, _dp, _D, _dp, _dp, _dp]),
            "orc_one_pair": (None, [_I, _L, _D, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip]),
            "orc_two_side": (None, [_I, _I, _dp, _dp, _dp, _dp, _dp, _ip]),
            "orc_bidiag_svd": (_I, [_I, _dp, _dp, _dp, _dp, _dp, _I, _ip]),
            "orc_givens_half": (_L, [_L]),
            "orc_givens_plan": (None, [_L, _L, _I, ctypes.c_uint, _lp, _dp, _dp]),
            "orc_givens_mask": (None, [_L, _L, _L, _dp, _I, _I, ctypes.c_uint]),
            "orc_block_shift_mask": (None, [_L, _L, _L, _dp, _I, _I, ctypes.c_uint]),
            "orc_find_best_convert": (None, [_L, _lp, _lp, _lp]),
            "orc_local_qr": (None, [_L, _L, _dp, _dp, _dp]),
            "orc_global_qr": (None, [_I, _L, _L, _lp, ctypes.POINTER(_dp), _dp, ctypes.POINTER(_dp), _dp]),
            "orc_ralha": (None, [_I, _L, _lp, ctypes.POINTER(_dp), ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                                 ctypes.POINTER(_dp), ctypes.POINTER(_dp), ctypes.POINTER(_dp)]),
            "orc_lower_to_upper": (None, [_L, _dp, _dp, _dp]),
            "orc_sign_normalize": (None, [_L, _L, _dp, _dp]),
            "orc_ormqr_house_ln": (None, [_L, _L, _dp, _dp, _dp]),
            "orc_drot": (None, [_L, _dp, _L, _dp, _L, _D, _D]),
            "orc_gemm": (None, [_I, _I, _L, _L, _L, _dp, _L, _dp, _L, _dp, _L]),
            "orc_run": (ctypes.c_void_p, [_I, _L, _lp, ctypes.POINTER(_dp), ctypes.POINTER(ctypes.c_uint), _I]),
            "orc_run_count": (_I, [ctypes.c_void_p]),
            "orc_run_get": (_L, [ctypes.c_void_p, ctypes.c_char_p, _I, _dp, _L]),
            "orc_run_timer": (_D, [ctypes.c_void_p, ctypes.c_char_p]),
            "orc_run_free": (None, [ctypes.c_void_p]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ------------------------------------------------------------------ elementary pieces
def nrm2(x):
    x = f64(x)
    return lib().orc_nrm2(x.size, _d(x), 1)


def idamax(x):
    x = f64(x)
    return lib().orc_idamax(x.size, _d(x), 1)


def idamin(x):
    x = f64(x)
    return lib().orc_idamin(x.size, _d(x), 1)


def drotg(a, b):
    a_, b_, c_, s_ = (ctypes.c_double(a), ctypes.c_double(b), ctypes.c_double(), ctypes.c_double())
    lib().orc_drotg(ctypes.byref(a_), ctypes.byref(b_), ctypes.byref(c_), ctypes.byref(s_))
    return a_.value, b_.value, c_.value, s_.value


def dlarfg(alpha, x):
    """Returns (beta, v_tail, tau) as LAPACK dlarfg leaves them."""
    x = f64(x).copy()
    a = ctypes.c_double(alpha)
    t = ctypes.c_double()
    lib().orc_dlarfg(x.size + 1, ctypes.byref(a), _d(x), 1, ctypes.byref(t))
    return a.value, x, t.value


def dgeqrf(A):
    """A: (m, n) array. Returns (QR-packed A as a numpy (m, n) array, tau) like LAPACK."""
    m, n = A.shape
    F = np.asfortranarray(A, dtype=np.float64).copy(order="F")
    tau = np.zeros(min(m, n))
    lib().orc_dgeqrf(m, n, F.ctypes.data_as(_dp), m, _d(tau))
    return F, tau


def dorgqr(F, tau, ncols=None):
    m, n = F.shape
    k = tau.size
    ncols = n if ncols is None else ncols
    Q = np.asfortranarray(F[:, :ncols], dtype=np.float64).copy(order="F")
    lib().orc_dorgqr(m, ncols, k, Q.ctypes.data_as(_dp), m, _d(f64(tau)))
    return Q


def dormqr_left(F, tau, C, trans):
    m = C.shape[0]
    n = C.shape[1]
    Af = np.asfortranarray(F, dtype=np.float64)
    Cf = np.asfortranarray(C, dtype=np.float64).copy(order="F")
    lib().orc_dormqr_left(1 if trans else 0, m, n, tau.size, Af.ctypes.data_as(_dp), Af.shape[0],
                          _d(f64(tau)), Cf.ctypes.data_as(_dp), m)
    return Cf


# ------------------------------------------------------------------ bbtsvd.hpp functions
def negative_count(a, b, mu):
    a, b = f64(a), f64(b)
    return lib().orc_negative_count(a.size, _d(a), _d(b if b.size else np.zeros(1)), float(mu))


def negative_count_batch(a, b, mus):
    a, b, mus = f64(a), f64(b), f64(mus)
    out = np.zeros(mus.size, dtype=np.int32)
    lib().orc_negative_count_batch(a.size, _d(a), _d(b if b.size else np.zeros(1)), _d(mus), mus.size,
                                   out.ctypes.data_as(_ip))
    return out


def bisection(a, b, lo, hi, tol=1e-20, top_k=None):
    a, b = f64(a), f64(b)
    n = a.size
    top_k = n if top_k is None else top_k
    sig = np.zeros(max(top_k, 1))
    llo = np.zeros(max(top_k, 1))
    lhi = np.zeros(max(top_k, 1))
    cnt = ctypes.c_int()
    lib().orc_bisection(n, _d(a), _d(b if b.size else np.zeros(1)), lo, hi, tol, top_k, _d(sig),
                        ctypes.byref(cnt), _d(llo), _d(lhi))
    c = cnt.value
    return sig[:c], llo[:c], lhi[:c]


def sigma_upper_bound(a, b):
    na, nb = nrm2(a), nrm2(b)
    return na * na + nb * nb


def ldl(a, b):
    a, b = f64(a), f64(b)
    n = a.size
    l = np.zeros(max(n - 1, 1))
    d = np.zeros(n)
    lib().orc_ldl(n, _d(a), _d(b if b.size else np.zeros(1)), _d(l), _d(d))
    return l[: n - 1], d


def dstqds(l, d, tau):
    l, d = f64(l), f64(d)
    n = d.size
    lp = np.zeros(max(n - 1, 1)); dp = np.zeros(n); s = np.zeros(n + 1)
    lib().orc_dstqds(n, _d(l if l.size else np.zeros(1)), _d(d), tau, _d(lp), _d(dp), _d(s))
    return lp[: n - 1], dp, s[:n]


def dpqds(l, d, tau):
    l, d = f64(l), f64(d)
    n = d.size
    up = np.zeros(max(n - 1, 1)); rp = np.zeros(n + 1); p = np.zeros(n + 1)
    lib().orc_dpqds(n, _d(l if l.size else np.zeros(1)), _d(d), tau, _d(up), _d(rp), _d(p))
    return up[: n - 1], rp[:n], p[:n]


def inverse_iteration(n, k, l, u, v, ldv):
    """bbtsvd.cpp:65-71 on a copy of v (at least (n-1)*ldv+1 entries); returns the copy."""
    v = f64(v).copy()
    l, u = f64(l), f64(u)
    lib().orc_inverse_iteration(n, k, _d(l if l.size else np.zeros(1)), _d(u if u.size else np.zeros(1)), _d(v),
                                ldv)
    return v


def one_pair(n, ldu, omega, l, d, a, b, v, u):
    """bbtsvd.cpp:93-135 on copies of v (n) and u (stride ldu); returns (v, u, twist_v, twist_u)."""
    v, u = f64(v).copy(), f64(u).copy()
    l, d, a, b = f64(l), f64(d), f64(a), f64(b)
    tv, tu = ctypes.c_int(0), ctypes.c_int(0)
    lib().orc_one_pair(n, ldu, omega, _d(l if l.size else np.zeros(1)), _d(d), _d(a), _d(b if b.size else np.zeros(1)),
                       _d(v), _d(u), ctypes.byref(tv), ctypes.byref(tu))
    return v, u, tv.value, tu.value


def two_side(a, b, sigmas):
    a, b, sigmas = f64(a), f64(b), f64(sigmas)
    n, k = a.size, sigmas.size
    u = np.zeros((n, k)); v = np.zeros((k, n))
    tw = np.zeros(2 * max(k, 1), dtype=np.int32)
    lib().orc_two_side(n, k, _d(sigmas), _d(a), _d(b if b.size else np.zeros(1)), _d(u), _d(v),
                       tw.ctypes.data_as(_ip))
    return u, v, tw[: 2 * k].reshape(k, 2)


def bidiag_svd(a, b, top_k=None):
    """7-argument bidiagonal_bisection_twisted_svd: returns (sigma, u (n x k), v (k x n), twists, rc)."""
    a, b = f64(a), f64(b)
    n = a.size
    k = n if top_k is None else top_k
    u = np.zeros((n, k)); v = np.zeros((k, n)); s = np.zeros(max(n, 1))
    tw = np.zeros(2 * max(k, 1), dtype=np.int32)
    rc = lib().orc_bidiag_svd(n, _d(a), _d(b if b.size else np.zeros(1)), _d(u), _d(v), _d(s), k,
                              tw.ctypes.data_as(_ip))
    return s[:k], u, v, tw[: 2 * k].reshape(k, 2), rc


# ------------------------------------------------------------------ Givens
def givens_plan(K, passes, mask, seed):
    half = lib().orc_givens_half(K)
    pairs = np.zeros((passes, half, 2), dtype=np.int64)
    c = np.zeros((passes, half)); s = np.zeros((passes, half))
    lib().orc_givens_plan(K, passes, 1 if mask else 0, seed & 0xFFFFFFFF, pairs.ctypes.data_as(_lp), _d(c), _d(s))
    return pairs, c, s


def givens_mask(x, passes, mask, is_left, seed):
    y = f64(x).copy()
    m, n = y.shape
    lib().orc_givens_mask(m, n, passes, _d(y), 1 if mask else 0, 1 if is_left else 0, seed & 0xFFFFFFFF)
    return y


def block_shift_mask(x, passes, mask, is_left, seed):
    y = f64(x).copy()
    m, n = y.shape
    lib().orc_block_shift_mask(m, n, passes, _d(y), 1 if mask else 0, 1 if is_left else 0, seed & 0xFFFFFFFF)
    return y


def find_best_convert(k):
    m, n, r = ctypes.c_long(), ctypes.c_long(), ctypes.c_long()
    lib().orc_find_best_convert(k, ctypes.byref(m), ctypes.byref(n), ctypes.byref(r))
    return m.value, n.value, r.value


# ------------------------------------------------------------------ pipeline stages
def local_qr(x, want_q=True):
    """x: (m, r) row-major X_i. Returns (R^T (m, m), Q^T (m, r) or None)."""
    x = f64(x)
    m, r = x.shape
    rt = np.zeros((m, m))
    qt = np.zeros((m, r)) if want_q else None
    lib().orc_local_qr(r, m, _d(x), _d(rt), _d(qt) if want_q else None)
    return rt, qt


def global_qr(panels, m):
    """panels: list of (m_part[c], R) row-major arrays. Returns (shrink list (m, m_part), Qg (R, m), tau)."""
    P = len(panels)
    R = panels[0].shape[1]
    mp = np.array([p.shape[0] for p in panels], dtype=np.int64)
    pans = [f64(p).copy() for p in panels]
    shr = [np.zeros((m, int(mp[c]))) for c in range(P)]
    qg = np.zeros(R * m)
    tau = np.zeros(m)
    PA = (_dp * P)(*[_d(p) for p in pans])
    SA = (_dp * P)(*[_d(s) for s in shr])
    lib().orc_global_qr(P, m, R, mp.ctypes.data_as(_lp), PA, _d(qg), SA, _d(tau))
    return shr, qg.reshape(m, R).T.copy(), tau


def ralha(F_list):
    """F_list: per-peer (m, c_p) row-major blocks. Returns per-peer dicts (alpha, beta, house, tau, U)."""
    P = len(F_list)
    m = F_list[0].shape[0]
    c = np.array([f.shape[1] for f in F_list], dtype=np.int64)
    Fs = [f64(f).copy() for f in F_list]
    al = [np.zeros(m) for _ in range(P)]
    be = [np.zeros(m - 1) for _ in range(P)]
    ho = [np.zeros((m, m)) for _ in range(P)]
    ta = [np.zeros(m) for _ in range(P)]
    U = [np.zeros((m, int(c[p]))) for p in range(P)]

    def arr(lst):
        return (_dp * P)(*[_d(a) for a in lst])

    lib().orc_ralha(P, m, c.ctypes.data_as(_lp), arr(Fs), arr(al), arr(be), arr(ho), arr(ta), arr(U))
    return [dict(alpha=al[p], beta=be[p], house=ho[p], tau=ta[p], U=U[p]) for p in range(P)]


def lower_to_upper(alpha, beta):
    a, b = f64(alpha).copy(), f64(beta).copy()
    cs = np.zeros(2 * max(a.size - 1, 1))
    lib().orc_lower_to_upper(a.size, _d(a), _d(b), _d(cs))
    return a, b, cs[: 2 * (a.size - 1)]


def sign_normalize(bu, bv):
    bu, bv = f64(bu).copy(), f64(bv).copy()
    m, cnt = bu.shape
    lib().orc_sign_normalize(m, cnt, _d(bu), _d(bv))
    return bu, bv


def ormqr_house_ln(house, tau, C):
    C = f64(C).copy()
    m, n = C.shape
    lib().orc_ormqr_house_ln(m, n, _d(f64(house)), _d(f64(tau)), _d(C))
    return C


def rot_back(bu, cs):
    """Combine-V back rotation (client.cpp:1127-1128)."""
    bu = f64(bu).copy()
    m, cnt = bu.shape
    for i in range(m - 2, -1, -1):
        x = bu[i]
        y = bu[i + 1]
        xv = x.copy(); yv = y.copy()
        lib().orc_drot(cnt, _d(xv), 1, _d(yv), 1, cs[2 * i], -cs[2 * i + 1])
        bu[i] = xv; bu[i + 1] = yv
    return bu


def gemm(A, B, ta=False, tb=False):
    A, B = f64(A), f64(B)
    M = A.shape[1] if ta else A.shape[0]
    K = A.shape[0] if ta else A.shape[1]
    N = B.shape[0] if tb else B.shape[1]
    C = np.zeros((M, N))
    lib().orc_gemm(int(ta), int(tb), M, N, K, _d(A), A.shape[1], _d(B), B.shape[1], _d(C), N)
    return C


class Run:
    """Whole simulated run() of P peers (short-and-wide branch, mode 1)."""

    def __init__(self, xs, seeds, top_k=-1):
        P = len(xs)
        self.P = P
        self.m = xs[0].shape[0]
        self.xs = [f64(x) for x in xs]
        ln = np.array([x.shape[1] for x in xs], dtype=np.int64)
        sd = (ctypes.c_uint * P)(*[s & 0xFFFFFFFF for s in seeds])
        XA = (_dp * P)(*[_d(x) for x in self.xs])
        self.h = lib().orc_run(P, self.m, ln.ctypes.data_as(_lp), XA, sd, top_k)
        if not self.h:
            raise ValueError("oracle run: configuration outside the short-and-wide branch or bisection failed")
        self.count = lib().orc_run_count(self.h)

    def get(self, name, p=0):
        n = lib().orc_run_get(self.h, name.encode(), p, None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n)
        lib().orc_run_get(self.h, name.encode(), p, _d(out), n)
        return out

    def timer(self, name):
        return lib().orc_run_timer(self.h, name.encode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_run_free(self.h)
            self.h = None
