"""Pins the oracle's bidiagonal SVD (bbtsvd.cpp:7-147, 932-968) against mathematics that fixes it:
brute-force eigenvalues, LAPACK through scipy/numpy, factorization identities and closed forms."""
import numpy as np
import pytest

from conftest import random_bidiagonal


def dense(a, b):
    B = np.diag(a)
    if len(b):
        B += np.diag(b, 1)
    return B


def true_sigma(a, b):
    return np.linalg.svd(dense(a, b), compute_uv=False)


@pytest.mark.parametrize("kind", ["gauss", "positive", "graded"])
@pytest.mark.parametrize("n", [1, 2, 7, 30])
def test_negative_count_brute_force(oracle, kind, n):
    """nc(mu) = #{sigma_i < mu}: compared with the eigenvalues of B^T B away from ties."""
    rng = np.random.default_rng(100 + n)
    a, b = random_bidiagonal(rng, n, kind) if n > 1 else (rng.standard_normal(1), np.zeros(0))
    s = np.sort(true_sigma(a, b))
    mus = np.concatenate([s * (1 - 1e-6), s * (1 + 1e-6), [0.0, 1e-300, 10 * s.max() + 1]])
    mus = np.abs(mus)
    got = oracle.negative_count_batch(a, b, mus)
    want = np.array([(s < mu).sum() for mu in mus])
    # exact except where mu lies within the perturbation of a singular value
    gap = np.array([np.min(np.abs(s - mu)) / max(s.max(), 1e-300) for mu in mus])
    ok = gap > 1e-9
    assert np.array_equal(got[ok], want[ok])
    assert got[-3] == 0  # nc(0) = 0
    assert got[-1] == n  # nc(large) = n


def test_negative_count_monotone_and_planted_mistake(oracle):
    rng = np.random.default_rng(1)
    a, b = random_bidiagonal(rng, 40, "gauss")
    mus = np.sort(rng.uniform(0, 6, 400))
    c = oracle.negative_count_batch(a, b, mus)
    assert np.all(np.diff(c) >= 0)
    # a count that ignores the coupling term (b dropped) must disagree with the truth somewhere
    s = true_sigma(a, b)
    want = np.array([(s < mu).sum() for mu in mus])
    wrong = oracle.negative_count_batch(a, np.zeros_like(b), mus)
    assert np.array_equal(c, want) or np.abs(c - want).max() <= 1
    assert not np.array_equal(wrong, want)


@pytest.mark.parametrize("kind", ["gauss", "positive", "graded", "clustered"])
def test_bisection_matches_lapack(oracle, kind):
    rng = np.random.default_rng(7)
    n = 60
    a, b = random_bidiagonal(rng, n, kind)
    ub = oracle.sigma_upper_bound(a, b)
    sig, lo, hi = oracle.bisection(a, b, 0.0, ub, 1e-20, n)
    s_np = true_sigma(a, b)
    if kind == "clustered" or ub < 1.0:
        # the reference's root [0, ||B||_F^2] does not bracket sigma_max when ||B||_F < 1
        assert sig.size <= n
    else:
        assert sig.size == n
        assert np.all(np.diff(sig) <= 0), "descending order"
        assert np.max(np.abs(sig - s_np)) <= 8 * np.finfo(float).eps * s_np[0] * n
        # each leaf is a pair of neighbouring floats (or narrower than tol) around its midpoint
        assert np.all((hi - lo <= 1e-20) | (np.nextafter(lo, np.inf) >= hi) | ((lo + hi) / 2 == lo)
                      | ((lo + hi) / 2 == hi))


def test_bisection_top_k_and_merge_quirk(oracle):
    rng = np.random.default_rng(3)
    a, b = random_bidiagonal(rng, 25, "positive")
    ub = oracle.sigma_upper_bound(a, b)
    full, _, _ = oracle.bisection(a, b, 0.0, ub, 1e-20, 25)
    top, _, _ = oracle.bisection(a, b, 0.0, ub, 1e-20, 5)
    assert np.array_equal(top, full[:5])
    # four equal singular values end in one leaf, recorded once (bbtsvd.cpp:34-37)
    sig, _, _ = oracle.bisection(np.full(4, 2.0), np.zeros(3), 0.0, 16.0, 1e-20, 4)
    assert sig.size == 1 and abs(sig[0] - 2.0) <= 4e-16 * 2
    # root interval that is already a leaf records its midpoint even without singular values
    sig, _, _ = oracle.bisection(np.zeros(3), np.zeros(2), 0.0, 0.0, 1e-20, 3)
    assert sig.size == 1 and sig[0] == 0.0


def test_ldl_identity(oracle):
    rng = np.random.default_rng(5)
    a, b = random_bidiagonal(rng, 50, "positive")
    l, d = oracle.ldl(a, b)
    L = np.eye(50) + np.diag(l, -1)
    B = dense(a, b)
    assert np.allclose(L @ np.diag(d) @ L.T, B.T @ B, rtol=0, atol=1e-13 * np.abs(B.T @ B).max())


def test_qds_identities(oracle):
    """dstqds: L+ D+ L+^T = L D L^T - tau I; dpqds: U+ R+ U+^T = L D L^T - tau I."""
    rng = np.random.default_rng(6)
    n = 40
    a, b = random_bidiagonal(rng, n, "positive")
    l, d = oracle.ldl(a, b)
    L = np.eye(n) + np.diag(l, -1)
    T = L @ np.diag(d) @ L.T
    tau = 0.3
    lp, dp, s = oracle.dstqds(l, d, tau)
    Lp = np.eye(n) + np.diag(lp, -1)
    ref = T - tau * np.eye(n)
    assert np.allclose(Lp @ np.diag(dp) @ Lp.T, ref, atol=1e-11 * np.abs(ref).max())
    up, rp, p = oracle.dpqds(l, d, tau)
    Up = np.eye(n) + np.diag(up, 1)
    assert np.allclose(Up @ np.diag(rp) @ Up.T, ref, atol=1e-11 * np.abs(ref).max())
    # s_i and p_i are the stationary / progressive auxiliaries: d+_i = d_i + s_i, r+_i = ... (closed forms)
    assert np.allclose(dp, d + s, rtol=1e-14, atol=0)


@pytest.mark.parametrize("kind", ["gauss", "positive"])
def test_twisted_vectors_match_svd(oracle, kind):
    rng = np.random.default_rng(11)
    n = 40
    a, b = random_bidiagonal(rng, n, kind)
    B = dense(a, b)
    U, s, Vt = np.linalg.svd(B)
    u, v, tw = oracle.two_side(a, b, s)
    for i in range(n):
        # singular vectors up to sign, for well separated values
        gap = np.min(np.abs(np.delete(s, i) - s[i])) / s[0]
        tol = 1e-13 / max(gap, 1e-12) + 1e-12
        assert abs(abs(v[i] @ Vt[i]) - 1) < tol
        assert abs(abs(u[:, i] @ U[:, i]) - 1) < tol
        # twist index sits where the eigenvector is large (gamma_k ~ delta / v_k^2)
        assert abs(v[i, tw[i, 0]]) >= 0.3 * np.abs(v[i]).max()
        assert abs(u[tw[i, 1], i]) >= 0.3 * np.abs(u[:, i]).max()
        assert np.isclose(np.linalg.norm(v[i]), 1.0, atol=1e-14)


def test_bidiag_svd_sign_rule_and_residual(oracle):
    rng = np.random.default_rng(12)
    n = 50
    a, b = random_bidiagonal(rng, n, "positive")
    B = dense(a, b)
    sig, u, v, tw, rc = oracle.bidiag_svd(a, b)
    assert rc == n
    s_np = np.linalg.svd(B, compute_uv=False)
    assert np.max(np.abs(sig - s_np)) <= 64 * np.finfo(float).eps * s_np[0]
    for i in range(n):
        p = int(np.argmax(np.abs(v[i])))
        btu = (B.T @ u[:, i])[p]
        assert btu * v[i, p] > 0, "v takes the sign of (B^T u)_p"
    assert np.linalg.norm(B @ v.T - u * sig) <= 1e-11 * s_np[0]


def test_bidiag_svd_check_none_quirk(oracle):
    # diagonal B with alpha already descending: sigma_i == alpha_i, vectors become unit vectors
    a = np.array([3.0, 2.0, 1.0, 0.5])
    b = np.zeros(3)
    sig, u, v, tw, rc = oracle.bidiag_svd(a, b)
    assert np.array_equal(sig, a)
    assert np.array_equal(u, np.eye(4)) and np.array_equal(v, np.eye(4))


def test_bidiag_svd_golden(oracle):
    """Committed fixtures (tests/golden/make_golden.py): singular values from LAPACK's SVD of the
    dense bidiagonal."""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "bidiag_lapack.npz"))
    for key in [k for k in z.files if k.startswith("alpha_")]:
        tag = key[len("alpha_"):]
        a, b, s = z[key], z["beta_" + tag], z["sigma_" + tag]
        sig, u, v, tw, rc = oracle.bidiag_svd(a, b)
        assert rc == a.size
        assert np.max(np.abs(sig - s)) <= 32 * np.finfo(float).eps * s[0] * np.sqrt(a.size)
