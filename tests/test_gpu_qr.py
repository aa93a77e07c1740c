"""GPU Householder QR (Cholesky QR + Householder reconstruction) and compact-WY applications
against LAPACK (scipy) and the oracle: R, Q, reflectors and tau in LAPACK's convention."""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-qr] {msg}", flush=True)


def lapack_qr(A):
    qr_, tau, _, info = lapack.dgeqrf(A)
    assert info == 0
    q, _, info = lapack.dorgqr(qr_, tau)
    return qr_, tau, q


def graded_matrix(rng, r, m, cond):
    U, _ = np.linalg.qr(rng.standard_normal((r, m)))
    V, _ = np.linalg.qr(rng.standard_normal((m, m)))
    s = cond ** (-np.arange(m) / max(m - 1, 1))
    return (U * s) @ V.T


@pytest.mark.parametrize("r,m,cond", [(40, 7, 1.0), (300, 100, 1e3), (2000, 257, 1e6), (5000, 1000, 1.0),
                                      (1000, 130, 1e10), (64, 64, 10.0)])
def test_qr_t_matches_lapack(hip, r, m, cond):
    banner(f"qr_t r={r} m={m} cond={cond:g}")
    rng = np.random.default_rng(r + m)
    A = graded_matrix(rng, r, m, cond) if cond != 1.0 else rng.standard_normal((r, m))
    qr_, tau, q = lapack_qr(A)
    R = np.triu(qr_[:m])
    out = hip.qr_t(A.T.copy(), want_q=True, want_v=True, square_last=True)
    nrm = np.linalg.norm(A, 2)
    # R and Q in LAPACK's signs; Q^T Q = I; A = Q R
    assert np.max(np.abs(out["rt"] - R.T)) <= 1e-12 * nrm * np.sqrt(m)
    Qt = out["qt"]
    assert np.max(np.abs(Qt @ Qt.T - np.eye(m))) <= 1e-13 * m
    assert np.linalg.norm(Qt.T @ out["rt"].T - A) <= 1e-14 * nrm * np.sqrt(m * r)
    if cond <= 1e6:
        assert np.max(np.abs(Qt - q.T)) <= 1e-9 * cond
        # reflectors and tau
        V = np.tril(qr_[:, :m], -1) + np.eye(r, m)
        assert np.max(np.abs(out["tau"] - tau)) <= 1e-10 * cond
        assert np.max(np.abs(out["vt"] - V.T)) <= 1e-9 * cond


@pytest.mark.parametrize("L,k,rows", [(40, 7, 9), (2000, 500, 500), (1500, 500, 300)])
def test_apply_q_right_matches_dormqr(hip, L, k, rows):
    banner(f"apply_q_right L={L} k={k} rows={rows}")
    rng = np.random.default_rng(L + k)
    A = rng.standard_normal((L, k))
    qr_, tau, _ = lapack_qr(A)
    M = rng.standard_normal((L, rows))
    ref, _, info = lapack.dormqr("L", "T", qr_, tau, M, lwork=max(1, rows * 64))
    assert info == 0
    vt = (np.tril(qr_[:, :k], -1) + np.eye(L, k)).T.copy()
    got = hip.apply_q_right(M.T.copy(), vt, tau)
    assert np.max(np.abs(got - ref.T)) <= 1e-12 * np.abs(M).max() * np.sqrt(L)


@pytest.mark.parametrize("L,k", [(40, 7), (2000, 1000), (1200, 1000)])
def test_orgqr_t_matches_dorgqr(hip, L, k):
    banner(f"orgqr_t L={L} k={k}")
    rng = np.random.default_rng(L * 3 + k)
    A = rng.standard_normal((L, k))
    qr_, tau, q = lapack_qr(A)
    vt = (np.tril(qr_[:, :k], -1) + np.eye(L, k)).T.copy()
    got = hip.orgqr_t(vt, tau)
    assert np.max(np.abs(got - q.T)) <= 1e-12 * np.sqrt(k)


@pytest.mark.parametrize("m,n", [(6, 4), (300, 300), (1000, 1000)])
def test_ormqr_house_matches_oracle(hip, oracle, m, n):
    banner(f"ormqr_house m={m} n={n}")
    rng = np.random.default_rng(m + n)
    # a Ralha-like reflector matrix: tau_0 = tau_{m-1} = 0, reflector j in column j below row j
    house = np.zeros((m, m))
    tau = np.zeros(m)
    for j in range(1, m - 1):
        x = rng.standard_normal(m - j)
        beta, v, t = oracle.dlarfg(x[0], x[1:])
        house[j + 1:, j] = v
        house[j, j] = 1.0
        tau[j] = t
    house[0, 0] = 1.0
    C = rng.standard_normal((m, n))
    got = hip.ormqr_house_ln(house, tau, C)
    ref = oracle.ormqr_house_ln(house, tau, C)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.sqrt(m)
