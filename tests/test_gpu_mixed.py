"""The mixed-precision form: FP32-MFMA CholeskyQR passes, the FP64 pass that refines Q, the SVD of
the shared m x m factor by one peer, and the FP64 Ogita-Aishima refinement of its triplets.

Bars (stated per test): normwise relative sigma error against the exact singular values of the
generator (or the FP64 oracle's run), ||U^T U - I|| and ||V^T V - I|| in the 2-norm, and
||A^T - U S V^T||_F / ||A||_F, all <= 1e-12. The form is not the reference's algorithm, so no
result is compared bit for bit; the building blocks (the one-owner bidiagonalization, the
refinement step) are compared with the oracle and with a numpy restatement of the published
refinement formulas.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-mixed] {msg}", flush=True)


def spectrum(m, kind):
    if kind == "geometric6":
        return 10.0 ** (-6.0 * np.arange(m) / (m - 1))
    if kind == "clustered8":
        cl = np.sort(1.0 + 1e-8 * np.linspace(-1, 1, 20))[::-1]
        return np.concatenate([cl, np.geomspace(cl[-1] * 0.99, 1e-8, m - 20)])
    raise ValueError(kind)


def blocks(P, rows, m, kind, seed=0):
    rng = np.random.default_rng(seed)
    s = spectrum(m, kind)
    Q1 = np.linalg.qr(rng.standard_normal((P * rows, m)))[0]
    Q2 = np.linalg.qr(rng.standard_normal((m, m)))[0]
    A = (Q1 * s) @ Q2.T
    return [np.ascontiguousarray(A[i * rows:(i + 1) * rows].T) for i in range(P)], s


def measures(res, xs, s_ref):
    p0 = res["peers"][0]
    sg, U = p0["sigma"], p0["FinalU"]
    k = len(sg)
    VT = np.hstack([p["FinalVT"] for p in res["peers"]])
    X = np.hstack(xs)
    G = U.T @ U - np.eye(k)
    H = VT @ VT.T - np.eye(k)
    return {"sigma_rel_err": float(np.max(np.abs(sg - s_ref[:k])) / s_ref[0]),
            "orth_u_2": float(np.linalg.norm(G, 2)), "orth_u_max": float(np.abs(G).max()),
            "orth_v_2": float(np.linalg.norm(H, 2)), "orth_v_max": float(np.abs(H).max()),
            "residual_rel": float(np.linalg.norm(X - (U * sg) @ VT) / np.linalg.norm(X))}


def assert_bars(st, tol=1e-12):
    print(st)
    for key in ("sigma_rel_err", "orth_u_2", "orth_v_2", "residual_rel"):
        assert st[key] <= tol, (key, st[key])


def oa_step_numpy(L, U, W, tau=1e-6, floor_rel=1e-4):
    """One refinement step, restated from Ogita and Aishima (2020) with the same cluster rule."""
    m = L.shape[0]
    R = np.eye(m) - U.T @ U
    S = np.eye(m) - W.T @ W
    T = U.T @ (L @ W)
    st = np.diag(T) / (1 - (np.diag(R) + np.diag(S)) / 2)
    si, sj = st[:, None], st[None, :]
    al = T + sj * R
    be = T.T + sj * S
    clu = np.abs(sj - si) <= tau * np.maximum(np.abs(si), np.abs(sj))
    with np.errstate(divide="ignore", invalid="ignore"):
        F = np.where(clu, R / 2, (sj * al + si * be) / (sj ** 2 - si ** 2))
        G = np.where(clu, S / 2, (si * al + sj * be) / (sj ** 2 - si ** 2))
    np.fill_diagonal(F, np.diag(R) / 2)
    np.fill_diagonal(G, np.diag(S) / 2)
    Un, Wn = U + U @ F, W + W @ G
    np.fill_diagonal(clu, False)
    mem = clu.any(axis=0) & (np.abs(st) > floor_rel * np.abs(st).max())
    Un[:, mem] = (L @ Wn[:, mem]) / st[mem]
    return Un, Wn, st


def test_oa_refine_step_matches_restatement(hip):
    """exc_oa_refine (one step) against the numpy restatement on a perturbed SVD of a lower
    triangular matrix; two steps bring orthogonality and the residual from ~1e-7 to rounding
    level (quadratic convergence)."""
    rng = np.random.default_rng(3)
    m = 300
    s = np.geomspace(1.0, 1e-5, m)
    Q1 = np.linalg.qr(rng.standard_normal((m, m)))[0]
    Q2 = np.linalg.qr(rng.standard_normal((m, m)))[0]
    L = np.linalg.cholesky(((Q1 * s) @ Q2.T) @ ((Q1 * s) @ Q2.T).T + 1e-30 * np.eye(m))
    U, sv, Wt = np.linalg.svd(L)
    W = Wt.T
    E1 = 1e-7 * rng.standard_normal((m, m))
    E2 = 1e-7 * rng.standard_normal((m, m))
    U0, W0 = U + U @ E1, W + W @ E2
    Ug, Wg, sg = hip.oa_refine(L, U0, W0, iters=1)
    Un, Wn, st = oa_step_numpy(L, U0, W0)
    assert np.max(np.abs(sg - st)) <= 1e-13 * st[0]
    assert np.max(np.abs(Ug - Un)) <= 1e-11
    assert np.max(np.abs(Wg - Wn)) <= 1e-11
    # quadratic convergence: ~1e-7 perturbations leave ~(1e-7 / relative gap)^2 after one step,
    # rounding level after the second
    for Q in (Ug, Wg):
        assert np.linalg.norm(Q.T @ Q - np.eye(m), 2) <= 1e-9
    U2, W2, s2 = hip.oa_refine(L, U0, W0, iters=2)
    for Q in (U2, W2):
        assert np.linalg.norm(Q.T @ Q - np.eye(m), 2) <= 1e-13
    assert np.linalg.norm(L - (U2 * s2) @ W2.T) <= 1e-13 * np.linalg.norm(L)


def test_oa_refine_stops_after_one_step_near_orthonormal(hip):
    """Starting vectors within 1e-10 of orthonormal: the first step is the last (its error is
    about the square of the starting one), so asking for two steps returns the one-step result
    bit for bit, at rounding level; 1e-7 starting errors still take both steps (the test above)."""
    rng = np.random.default_rng(5)
    m = 200
    s = np.geomspace(1.0, 1e-5, m)
    Q1 = np.linalg.qr(rng.standard_normal((m, m)))[0]
    Q2 = np.linalg.qr(rng.standard_normal((m, m)))[0]
    L = np.linalg.cholesky(((Q1 * s) @ Q2.T) @ ((Q1 * s) @ Q2.T).T + 1e-30 * np.eye(m))
    U, sv, Wt = np.linalg.svd(L)
    U0 = U + U @ (1e-13 * rng.standard_normal((m, m)))
    W0 = Wt.T + Wt.T @ (1e-13 * rng.standard_normal((m, m)))
    one = hip.oa_refine(L, U0, W0, iters=1)
    two = hip.oa_refine(L, U0, W0, iters=2)
    for a, b in zip(one, two):
        assert np.array_equal(a, b)
    Ug, Wg, sg = two
    for Q in (Ug, Wg):
        assert np.linalg.norm(Q.T @ Q - np.eye(m), 2) <= 1e-13
    assert np.linalg.norm(L - (Ug * sg) @ Wg.T) <= 1e-13 * np.linalg.norm(L)


def test_oa_refine_cluster(hip):
    """A cluster of 8 values within 1e-10 relative: the step orthogonalizes the cluster's vectors
    and keeps the residual at rounding level, where the unclustered formula would divide by ~0."""
    rng = np.random.default_rng(4)
    m = 128
    s = np.concatenate([1.0 + 1e-10 * np.arange(8)[::-1], np.geomspace(0.5, 1e-3, m - 8)])
    Q1 = np.linalg.qr(rng.standard_normal((m, m)))[0]
    Q2 = np.linalg.qr(rng.standard_normal((m, m)))[0]
    L = (Q1 * s) @ Q2.T
    U, Wt = Q1.copy(), Q2.copy()
    U0 = U + U @ (1e-8 * rng.standard_normal((m, m)))
    W0 = Wt + Wt @ (1e-8 * rng.standard_normal((m, m)))
    Ug, Wg, sg = hip.oa_refine(L, U0, W0, iters=2, lower=False)
    assert np.max(np.abs(sg - s)) <= 1e-13
    for Q in (Ug, Wg):
        assert np.linalg.norm(Q.T @ Q - np.eye(m), 2) <= 1e-13
    assert np.linalg.norm(L - (Ug * sg) @ Wg.T) <= 1e-13 * np.linalg.norm(L)


@pytest.mark.parametrize("one_launch", [True, False])
def test_bidiag_local_vs_oracle(hip, oracle, one_launch):
    """One-owner bidiagonalization (one launch, or one launch per step) against the oracle's
    Ralha with one peer followed by lower to upper, at the tolerances of the multi-peer tests."""
    rng = np.random.default_rng(5)
    m = 300
    F = np.linalg.qr(rng.standard_normal((2 * m, m)))[1]  # the R factor the pipeline feeds it
    got = hip.bidiag_local(F, one_launch=one_launch)
    want = oracle.ralha([F])[0]
    a, b, cs = oracle.lower_to_upper(want["alpha"], want["beta"])
    tol = max(1e-11, 1e-13 * m)
    for key, g, w in (("alpha", got["alpha"], a), ("beta", got["beta"], b), ("tau", got["tau"], want["tau"]),
                      ("U", got["U"], want["U"]), ("cs", got["cs"], cs)):
        assert np.max(np.abs(g - w)) <= tol * max(1.0, np.abs(w).max()), key
    for j in range(1, m - 1):
        if want["tau"][j] != 0:
            assert np.max(np.abs(got["house"][j + 1:, j] - want["house"][j + 1:, j])) <= 1e-10


def test_mixed_config1_vs_oracle(hip, oracle):
    """2 peers x 500 x 100, N(0,1): sigma against the FP64 oracle's run (normwise 1e-12), and the
    orthogonality and residual bars."""
    from excalibur_amd.launch import run_peers
    rng = np.random.default_rng(0)
    A = rng.standard_normal((1000, 100))
    xs = [A[:500].T.copy(), A[500:].T.copy()]
    banner("mixed, 2 x 500 x 100")
    res = run_peers(xs, seed=0, form="mixed")
    so = oracle.Run(xs, [0, 0]).get("sigma", 0)
    st = measures(res, xs, so)
    assert_bars(st)
    # every peer holds the same sigma and FinalU
    for p in res["peers"][1:]:
        assert np.array_equal(p["sigma"], res["peers"][0]["sigma"])
        assert np.array_equal(p["FinalU"], res["peers"][0]["FinalU"])


@pytest.mark.parametrize("P,rows,m,kind", [(2, 10_000, 200, "geometric6"), (2, 20_000, 200, "clustered8"),
                                           (3, 7_000, 300, "geometric6")])
def test_mixed_graded_and_clustered(hip, P, rows, m, kind):
    """Spectra of the metric's workloads at reduced size: geometric 1 .. 1e-6 and 20 values in
    1 +- 1e-8 followed by a geometric tail to 1e-8 (where the plain FP64 path loses orthogonality
    inside the cluster). Against the generator's exact singular values."""
    from excalibur_amd.launch import run_peers
    banner(f"mixed, {P} x {rows} x {m}, {kind}")
    xs, s = blocks(P, rows, m, kind)
    res = run_peers(xs, seed=1, form="mixed")
    assert_bars(measures(res, xs, s))


def test_mixed_topk_and_modes(hip):
    """top_k = 32 keeps the leading triplets of the full result; modes 2 and 3 compute one side."""
    from excalibur_amd.launch import run_peers
    xs, s = blocks(2, 4_000, 120, "geometric6")
    full = run_peers(xs, seed=2, form="mixed")
    top = run_peers(xs, seed=2, form="mixed", topk=32)
    for pf, pt in zip(full["peers"], top["peers"]):
        assert np.array_equal(pt["sigma"], pf["sigma"][:32])
        assert np.array_equal(pt["FinalU"], pf["FinalU"][:, :32])
        assert np.max(np.abs(pt["FinalVT"] - pf["FinalVT"][:32])) <= 1e-13
    left = run_peers(xs, seed=2, form="mixed", mode=2)
    right = run_peers(xs, seed=2, form="mixed", mode=3)
    assert "FinalVT" not in left["peers"][0] and "FinalU" not in right["peers"][0]
    assert np.array_equal(left["peers"][0]["FinalU"], full["peers"][0]["FinalU"])
    assert np.array_equal(right["peers"][1]["FinalVT"], full["peers"][1]["FinalVT"])


def test_mixed_config2_full_size(hip):
    """2 peers x 50,000 x 1,000 rows of A = Q1 diag(sigma) Q2^T, sigma_i = 10^(-6i/999): the
    metric's workload, all bars in the 2-norm."""
    from excalibur_amd.launch import run_peers
    banner("mixed, config 2 full size, 2 x 50000 x 1000")
    t0 = time.time()
    xs, s = blocks(2, 50_000, 1_000, "geometric6")
    banner(f"generated in {time.time() - t0:.1f} s")
    res = run_peers(xs, seed=0, form="mixed", timeout=300)
    banner(f"peers done in {time.time() - t0:.1f} s: {[p['result']['total'] for p in res['peers']]}")
    assert res["peers"][0]["result"]["sigma_count"] == 1000
    assert_bars(measures(res, xs, s))


@pytest.mark.parametrize("n", [1, 2, 7, 300, 2001])
def test_pack_unpack_lower(hip, n):
    """The triangle packing the mixed form's Gram exchange uses: exc_pack_lower_f32/f64 give the
    rows of the lower triangle back to back, exc_unpack_lower_f64 the symmetric matrix."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    want = np.concatenate([A[i, :i + 1] for i in range(n)])
    dA = hip.DeviceArray.from_host(A)
    t = n * (n + 1) // 2
    p64 = hip.DeviceArray((t,))
    hip.call("exc_pack_lower_f64", n, dA.ptr, n, p64.ptr, None)
    got = p64.to_host()
    assert np.array_equal(got, want)
    A32 = A.astype(np.float32)
    dA32 = hip.DeviceArray.from_host(A32)
    p32 = hip.DeviceArray((t,), np.float32)
    hip.call("exc_pack_lower_f32", n, dA32.ptr, n, p32.ptr, None)
    assert np.array_equal(p32.to_host(), want.astype(np.float32))
    out = hip.DeviceArray((n, n)).zero()
    hip.call("exc_unpack_lower_f64", n, p64.ptr, out.ptr, n, None)
    L = np.tril(A)
    assert np.array_equal(out.to_host(), L + np.tril(A, -1).T)


@pytest.mark.parametrize("P,n", [(1, 1), (2, 500500), (3, 7), (5, 1000)])
def test_acc_packed_peer_order(hip, P, n):
    """exc_acc_packed called for peers 0..P-1 in turn gives the host loop acc = 0; acc = acc + x_j
    bit for bit, for FP32 terms widened to FP64 and for FP64 terms; a peer's own term read in place
    and the others from staging give the same bits."""
    rng = np.random.default_rng(P * 1000 + n)
    for f32 in (True, False):
        dt = np.float32 if f32 else np.float64
        parts = [(rng.standard_normal(n) * 10.0 ** rng.integers(-8, 8, n)).astype(dt) for _ in range(P)]
        want = np.zeros(n)
        for x in parts:
            want = want + x.astype(np.float64)
        acc = hip.DeviceArray((n,)).zero()
        dparts = [hip.DeviceArray.from_host(x) for x in parts]
        for j, d in enumerate(dparts):
            hip.call("exc_acc_packed", n, d.ptr, int(f32), int(j == 0), acc.ptr, None)
        assert np.array_equal(acc.to_host(), want)
    with pytest.raises(hip.HipError):
        hip.call("exc_acc_packed", -1, None, 0, 1, None, None)


@pytest.mark.parametrize("n", [1, 5, 300, 1500])
def test_trace_and_deviation(hip, n):
    """exc_trace (diagonal sum) and exc_dev_identity (max |A - I| and ||A - I||_F) against numpy."""
    rng = np.random.default_rng(n + 1)
    A = np.eye(n) + 1e-3 * rng.standard_normal((n, n))
    dA = hip.DeviceArray.from_host(A)
    out = hip.DeviceArray((2,)).zero()
    hip.call("exc_trace", n, dA.ptr, n, out.ptr, None)
    assert np.isclose(out.to_host()[0], np.trace(A), rtol=1e-14, atol=0)
    ws = hip.workspace(hip.lib.exc_dev_identity_workspace(n))
    hip.call("exc_dev_identity", n, dA.ptr, n, out.ptr, ws.ptr, ws.nbytes, None)
    got = out.to_host()
    E = A - np.eye(n)
    assert got[0] == np.abs(E).max()
    assert np.isclose(got[1], np.linalg.norm(E), rtol=1e-13, atol=0)
