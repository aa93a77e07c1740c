"""The tall-and-skinny branch of run() (m > global_n; client.cpp:543-569, 627-678, 800-1154) and
the square case m == global_n (short-and-wide stages, tall combine of the right vectors) on the
GPU peers against the oracle's run() of the same files and seeds, and against what the
mathematics fixes at sizes the oracle does not reach."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def problem(m, parts, seed, decay=2.0):
    """X = U diag(s) V^T (m x global_n) by columns. For m > global_n each file holds X_p^T
    (local_n x m row-major, the layout the reference reads in this branch); for m == global_n it
    holds X_p (m x local_n)."""
    rng = np.random.default_rng(seed)
    n = sum(parts)
    U = np.linalg.qr(rng.standard_normal((m, n)))[0]
    V = np.linalg.qr(rng.standard_normal((n, n)))[0]
    s = 10.0 ** (-decay * np.arange(n) / (n - 1))
    X = (U * s) @ V.T
    raw, o = [], 0
    for ln in parts:
        blk = X[:, o:o + ln]
        raw.append(np.ascontiguousarray(blk.T).reshape(m, ln) if m > n else np.ascontiguousarray(blk))
        o += ln
    return X, s, raw


def vectors(res, m, parts):
    p0 = res["peers"][0]
    k = len(p0["sigma"])
    FU = p0["FinalU"]
    U = FU.T if m > sum(parts) else FU  # m x k
    VT = np.hstack([p["FinalVT"] for p in res["peers"]])
    return p0["sigma"], U, VT


@pytest.mark.parametrize("m,parts", [(300, [40, 35]), (301, [50, 50, 27]), (150, [100, 50]), (12000, [40, 40])])
def test_tall_and_square_vs_oracle(hip, oracle, m, parts):
    """sigma within 1e-12 (normwise) of the oracle's, FinalU and FinalVT within 1e-9 of the
    oracle's (both after the reference's sign rule), and X = U diag(sigma) VT to 1e-12.
    300 x [40, 35]: m_part = 150 > global_n, so each peer reduces its rows by a local QR;
    301 x [50, 50, 27]: three peers without the reduction; 150 x [100, 50]: m == global_n;
    12000 x [40, 40]: the public right mask over 12000 lines takes the virtual-matrix mask."""
    from excalibur_amd.launch import run_peers
    print(f"[gpu-tall] m={m} parts={parts}", flush=True)
    X, s, raw = problem(m, parts, seed=m)
    res = run_peers(raw, seed=7, mode=1)
    orc = oracle.Run(raw, [7] * len(parts))
    k = orc.count
    sg, U, VT = vectors(res, m, parts)
    assert len(sg) == k
    so = orc.get("sigma", 0)[:k]
    assert np.max(np.abs(sg - so)) <= 1e-12 * so[0]
    FUo = orc.get("FinalU", 0)
    Uo = FUo.reshape(k, m).T if m > sum(parts) else FUo.reshape(m, k)
    assert np.max(np.abs(U - Uo)) <= 1e-9
    for p, ln in enumerate(parts):
        assert np.max(np.abs(res["peers"][p]["FinalVT"] - orc.get("FinalVT", p).reshape(k, ln))) <= 1e-9
    assert np.linalg.norm(X - (U * sg) @ VT) <= 1e-12 * np.linalg.norm(X)
    assert np.abs(U.T @ U - np.eye(k)).max() <= 1e-12


def test_tall_modes(hip):
    """Modes 2 (left vectors only) and 3 (right only) give the same sigma and vectors as mode 1."""
    from excalibur_amd.launch import run_peers
    m, parts = 400, [60, 40]
    X, s, raw = problem(m, parts, seed=5)
    full = run_peers(raw, seed=3, mode=1)
    left = run_peers(raw, seed=3, mode=2)
    right = run_peers(raw, seed=3, mode=3)
    assert np.array_equal(full["peers"][0]["sigma"], left["peers"][0]["sigma"])
    assert np.array_equal(full["peers"][0]["FinalU"], left["peers"][0]["FinalU"])
    for p in range(2):
        assert np.array_equal(full["peers"][p]["FinalVT"], right["peers"][p]["FinalVT"])


def test_tall_large_properties(hip):
    """20,000 x 500 over two peers (m_part = 10,000 rows reduced by a local QR each; the right
    masks over 20,000 lines take the virtual-matrix mask): sigma against the generator's exact
    values, orthonormal vectors and the reconstruction."""
    from excalibur_amd.launch import run_peers
    m, parts = 20000, [300, 200]
    print(f"[gpu-tall] m={m} parts={parts}", flush=True)
    X, s, raw = problem(m, parts, seed=11)
    res = run_peers(raw, seed=1, mode=1)
    sg, U, VT = vectors(res, m, parts)
    k = len(sg)
    assert k == 500
    assert np.max(np.abs(sg - s)) <= 1e-12 * s[0]
    assert np.linalg.norm(U.T @ U - np.eye(k), 2) <= 1e-12
    assert np.linalg.norm(VT @ VT.T - np.eye(k), 2) <= 1e-12
    assert np.linalg.norm(X - (U * sg) @ VT) <= 1e-12 * np.linalg.norm(X)
