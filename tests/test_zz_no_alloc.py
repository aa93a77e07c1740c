"""The allocation-free bidiagonal SVD (exc_bidiag_svd_noalloc) through the C ABI: its results
equal exc_bidiag_svd's bit for bit on two calls in a row (the second smaller, so the shared
scratch holds the first call's data), and sizes beyond its fixed scratch are rejected."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run_both(hip, alpha, beta, top_k):
    n = len(alpha)
    H = hip
    da = H.DeviceArray.from_host(alpha)
    db = H.DeviceArray.from_host(beta)
    out = []
    for noalloc in (False, True):
        du = H.DeviceArray((n, top_k)).zero()
        dv = H.DeviceArray((top_k, n)).zero()
        ds = H.DeviceArray((top_k,)).zero()
        dt = H.DeviceArray((2 * top_k,), np.int32).zero()
        if noalloc:
            H.call("exc_bidiag_svd_noalloc", n, da.ptr, db.ptr, du.ptr, dv.ptr, ds.ptr, top_k, dt.ptr, None)
        else:
            ws = H.workspace(H.lib.exc_bidiag_svd_workspace(n, top_k))
            H.call("exc_bidiag_svd", n, da.ptr, db.ptr, du.ptr, dv.ptr, ds.ptr, top_k, dt.ptr, ws.ptr, ws.nbytes,
                   None)
        H.sync()
        out.append((ds.to_host(), du.to_host(), dv.to_host(), dt.to_host()))
    return out


def test_noalloc_equals_workspace_form_two_calls(hip):
    rng = np.random.default_rng(11)
    for n, k in ((1000, 1000), (300, 64)):  # second call smaller than the first
        print(f"[no-alloc] n={n} top_k={k}", flush=True)
        alpha = rng.uniform(0.5, 2.0, n)
        beta = rng.uniform(0.1, 1.0, n - 1)
        ref, got = _run_both(hip, alpha, beta, k)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)


def test_noalloc_rejects_oversize(hip):
    cap = hip.lib.exc_bidiag_svd_noalloc_capacity()
    n = 2048
    assert hip.lib.exc_bidiag_svd_workspace(n, n) > cap
    da = hip.DeviceArray.from_host(np.ones(n))
    db = hip.DeviceArray.from_host(np.ones(n - 1))
    rc = hip.lib.exc_bidiag_svd_noalloc(n, da.ptr, db.ptr, None, None, None, n, None, None)
    assert rc == -3  # EXC_ERR_SPACE, before any device work
    assert b"48 MiB" in hip.lib.exc_last_error()
