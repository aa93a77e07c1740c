"""Batched peers on one device (excalibur_amd/csrc/batched.hip through the C ABI): the local stage
of the reference's run() (local QR, client.cpp:572-589; Givens masks, client.cpp:592-605) for many
peers' blocks at once, and the merge of their masked factors by a global QR.

Blocks the reference reduces by a local QR (local_n / m > 5 in integer division, client.cpp:572)
go through exc_batched_local_qr (FP32-MFMA CholeskyQR pass, FP64 refinement passes, grouped
launches chosen by the autotuner); the others are masked as they are, as the reference does. The
masks are the reference's plans (public left plan shared, a private right plan per block) applied
by batched sweeps, and the masked blocks side by side are factorized by the FP64 QR (exc_qr_t),
which merges them into the R factor of the whole matrix.
"""
import ctypes
import time

import numpy as np

from . import hip

DA = hip.DeviceArray
lib = hip.lib


def _ptrs(arrays):
    return (ctypes.c_void_p * len(arrays))(*[a.ptr for a in arrays])


def _plan_device(K, mask, seed):
    pairs, cs = hip.givens_plan(K, 32, mask, seed)
    return DA.from_host(pairs), DA.from_host(cs)


def mask_batched(blocks, m, cols, public_seed, private_seeds):
    """Left public mask (K = m) then right private mask (K = cols) on same-shape device blocks."""
    nb = len(blocks)
    if nb == 0:
        return
    xs = DA.from_host(np.array([b.ptr for b in blocks], dtype=np.uint64))
    pub = _plan_device(m, True, public_seed)
    pp = DA.from_host(np.array([pub[0].ptr] * nb, dtype=np.uint64))
    pc = DA.from_host(np.array([pub[1].ptr] * nb, dtype=np.uint64))
    hip.call("exc_givens_apply_batched", nb, xs.ptr, m, cols, pp.ptr, pc.ptr, 32, 1, 1, None)
    priv = [_plan_device(cols, True, s) for s in private_seeds]
    qp = DA.from_host(np.array([p[0].ptr for p in priv], dtype=np.uint64))
    qc = DA.from_host(np.array([p[1].ptr for p in priv], dtype=np.uint64))
    hip.call("exc_givens_apply_batched", nb, xs.ptr, m, cols, qp.ptr, qc.ptr, 32, 1, 0, None)
    hip.sync()


def local_stage(xs, public_seed, private_seeds, want_q=False, timings=None):
    """xs: host blocks (m x cols_b). Returns dict(rt, masked, qt, stats, reduced) with device
    arrays; the blocks the reference does not reduce keep their raw columns."""
    m = xs[0].shape[0]
    cols = np.array([x.shape[1] for x in xs], dtype=np.int64)
    reduce_b = (cols // m) > 5
    dx = [DA.from_host(np.ascontiguousarray(x, np.float64)) for x in xs]
    hip.sync()
    t0 = time.perf_counter()
    qi = np.flatnonzero(reduce_b)
    rt = {b: DA((m, m)) for b in qi}
    qt = {b: DA((m, int(cols[b]))) for b in qi} if want_q else {}
    stats = np.zeros(3 * max(1, len(qi)))
    if len(qi):
        qc = (ctypes.c_long * len(qi))(*[int(cols[b]) for b in qi])
        ws = hip.workspace(lib.exc_batched_local_qr_workspace(len(qi), m, qc))
        hx = _ptrs([dx[b] for b in qi])
        hr = _ptrs([rt[b] for b in qi])
        hq = _ptrs([qt[b] for b in qi]) if want_q else None
        hip.call("exc_batched_local_qr", len(qi), m, qc, hx, hr, hq, stats.ctypes.data, ws.ptr, ws.nbytes, None)
        hip.sync()
    t1 = time.perf_counter()
    masked = {}
    for b in qi:
        masked[b] = DA((m, m))
        hip.call("exc_memcpy_d2d", masked[b].ptr, rt[b].ptr, rt[b].nbytes, None)
    for b in np.flatnonzero(~reduce_b):
        masked[b] = dx[b]
    # same-shape groups of blocks share one batched sweep
    for width in sorted(set(int(masked[b].shape[1]) for b in masked)):
        grp = [b for b in sorted(masked) if masked[b].shape[1] == width]
        mask_batched([masked[b] for b in grp], m, width, public_seed, [private_seeds[b] for b in grp])
    t2 = time.perf_counter()
    if timings is not None:
        timings.update({"local_qr_s": t1 - t0, "mask_s": t2 - t1})
    return {"rt": rt, "qt": qt, "masked": masked, "stats": stats.reshape(-1, 3)[:len(qi)], "reduced": reduce_b,
            "inputs": dx}


def merge(masked, m, timings=None):
    """R^T (m x m, lower) of the masked blocks side by side (the global QR of client.cpp:706-794
    for peers on one device), by exc_qr_t in FP64."""
    order = sorted(masked)
    widths = [int(masked[b].shape[1]) for b in order]
    total = int(sum(widths))
    cat = DA((m, total))
    o = 0
    for b, w in zip(order, widths):
        hip.call("exc_copy2d", m, w, masked[b].ptr, w, cat.ptr + 8 * o, total, None)
        o += w
    rt = DA((m, m)).zero()
    ws = hip.workspace(lib.exc_qr_t_workspace(m, total, 0))
    hip.sync()
    t0 = time.perf_counter()
    hip.call("exc_qr_t", m, total, cat.ptr, total, rt.ptr, None, 0, None, 0, None, None, 0, ws.ptr, ws.nbytes, None)
    hip.sync()
    if timings is not None:
        timings["merge_s"] = time.perf_counter() - t0
    return rt, cat


def per_peer_baseline(xs, public_seed, private_seeds):
    """The same local stage one peer after another, as the plain form runs it (exc_qr_t + masks per
    block): seconds."""
    m = xs[0].shape[0]
    dx = [DA.from_host(np.ascontiguousarray(x, np.float64)) for x in xs]
    hip.sync()
    t0 = time.perf_counter()
    for b, d in enumerate(dx):
        c = d.shape[1]
        out = d
        if c // m > 5:
            out = DA((m, m)).zero()
            ws = hip.workspace(lib.exc_qr_t_workspace(m, c, 1))
            hip.call("exc_qr_t", m, c, d.ptr, c, out.ptr, d.ptr, c, None, 0, None, None, 0, ws.ptr, ws.nbytes, None)
        w = out.shape[1]
        for (K, left, seed) in ((m, 1, public_seed), (w, 0, private_seeds[b])):
            pairs, cs = hip.givens_plan(K, 32, True, seed)
            dp, dc = DA.from_host(pairs), DA.from_host(cs)
            hip.call("exc_givens_apply", out.ptr, m, w, dp.ptr, dc.ptr, 32, 1, left, None)
    hip.sync()
    return time.perf_counter() - t0


def tuning():
    buf = ctypes.create_string_buffer(1 << 14)
    lib.exc_gemm_grouped_tuning(buf, len(buf))
    return buf.value.decode()
