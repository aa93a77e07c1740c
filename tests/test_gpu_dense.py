"""GPU checks of the MFMA GEMM and the blocked factorizations against numpy (LAPACK)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-dense] {msg}", flush=True)


@pytest.mark.parametrize("ta", [False, True])
@pytest.mark.parametrize("tb", [False, True])
@pytest.mark.parametrize("shape", [(1, 1, 1), (17, 33, 5), (128, 128, 16), (300, 200, 1000), (130, 7, 5000)])
def test_dgemm(hip, ta, tb, shape):
    M, N, K = shape
    banner(f"dgemm {shape} ta={ta} tb={tb}")
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    C0 = rng.standard_normal((M, N))
    ref = 0.5 * ((A.T if ta else A) @ (B.T if tb else B)) - 2.0 * C0
    got = hip.gemm(A, B, ta, tb, alpha=0.5, beta=-2.0, C=C0)
    scale = np.sqrt(K) * np.abs(A).max() * np.abs(B).max() + np.abs(C0).max()
    assert np.max(np.abs(got - ref)) <= 1e-14 * scale * 8


def test_sgemm(hip):
    banner("sgemm")
    rng = np.random.default_rng(3)
    A = rng.standard_normal((257, 300)); B = rng.standard_normal((300, 129))
    got = hip.gemm(A, B, f32=True)
    ref = A.astype(np.float32).astype(np.float64) @ B.astype(np.float32).astype(np.float64)
    assert np.max(np.abs(got - ref)) <= 1e-5 * np.sqrt(300) * 4


@pytest.mark.parametrize("n", [1, 5, 128, 129, 700])
def test_potrf(hip, n):
    banner(f"potrf n={n}")
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n + 10))
    G = X @ X.T
    L, info = hip.potrf_lower(G)
    assert info == 0
    assert np.allclose(np.triu(L, 1), 0)
    ref = np.linalg.cholesky(G)
    assert np.max(np.abs(L - ref)) <= 1e-12 * np.abs(ref).max()
    # non-positive pivot is reported
    G2 = G.copy(); G2[n // 2, n // 2] = -1.0
    _, info2 = hip.potrf_lower(G2)
    assert info2 > 0


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("unit", [False, True])
@pytest.mark.parametrize("n,nrhs", [(3, 5), (100, 37), (128, 64), (300, 1000), (700, 130), (200, 5000), (64, 4100)])
def test_trsm_left(hip, upper, unit, n, nrhs):
    banner(f"trsm_left n={n} nrhs={nrhs} upper={upper} unit={unit}")
    rng = np.random.default_rng(n * 3 + nrhs)
    T = rng.standard_normal((n, n)) * 0.1 + np.eye(n) * 2
    T = np.triu(T) if upper else np.tril(T)
    Teff = T.copy()
    if unit:
        np.fill_diagonal(Teff, 1.0)
    X = rng.standard_normal((n, nrhs))
    got = hip.trsm_left(T, X, upper, unit)
    ref = np.linalg.solve(Teff, X)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("upper", [False, True])
@pytest.mark.parametrize("unit", [False, True])
@pytest.mark.parametrize("n,nrhs", [(100, 16384), (128, 20000), (300, 16400), (1000, 16384)])
def test_trsm_left_wide(hip, upper, unit, n, nrhs):
    """Wide right-hand sides (the leaves' many-wavefront form): the solution matches LAPACK's, and
    on a graded triangle (an R factor of singular values falling over six decades, the triangles
    the Cholesky QR passes solve with) the solve stays backward stable: ||T X - B|| is at rounding
    level relative to ||T|| ||X|| (relative to ||B|| it can reach u cond(T), for LAPACK too)."""
    import scipy.linalg as sl
    banner(f"trsm_left wide n={n} nrhs={nrhs} upper={upper} unit={unit}")
    rng = np.random.default_rng(n + nrhs + 2 * upper + unit)
    T = rng.standard_normal((n, n)) * 0.1 + np.eye(n) * 2
    T = np.triu(T) if upper else np.tril(T)
    X = rng.standard_normal((n, nrhs))
    got = hip.trsm_left(T, X, upper, unit)
    ref = sl.solve_triangular(T, X, lower=not upper, unit_diagonal=unit)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.abs(ref).max()
    if unit:
        return
    # graded: R factor of a matrix with singular values 1 .. 1e-6
    q1, _ = np.linalg.qr(rng.standard_normal((2 * n, n)))
    q2, _ = np.linalg.qr(rng.standard_normal((n, n)))
    R = np.linalg.qr((q1 * np.geomspace(1.0, 1e-6, n)) @ q2.T, mode="r")
    G = R if upper else np.ascontiguousarray(R.T)
    got = hip.trsm_left(G, X, upper, False)
    # normwise backward error (substitution: 1e-17 and below here; LAPACK's dtrtrs alike)
    back = np.linalg.norm(G @ got - X) / (np.linalg.norm(G) * np.linalg.norm(got))
    assert back <= 1e-15, back


@pytest.mark.parametrize("lower_t", [False, True])
@pytest.mark.parametrize("n,rows", [(3, 5), (128, 64), (300, 1000), (700, 130), (200, 5000), (64, 4100)])
def test_trsm_right(hip, lower_t, n, rows):
    banner(f"trsm_right n={n} rows={rows} lower_t={lower_t}")
    rng = np.random.default_rng(n * 5 + rows)
    T = rng.standard_normal((n, n)) * 0.1 + np.eye(n) * 2
    T = np.tril(T) if lower_t else np.triu(T)
    Op = T.T if lower_t else T
    X = rng.standard_normal((rows, n))
    got = hip.trsm_right(T, X, lower_t)
    ref = np.linalg.solve(Op.T, X.T).T
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("n", [4, 128, 300, 1000])
def test_lu_sign(hip, n):
    banner(f"lu_sign n={n}")
    rng = np.random.default_rng(n)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    LU, s = hip.lu_sign(Q)
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    assert set(np.unique(s)) <= {-1.0, 1.0}
    assert np.max(np.abs(L @ U - (Q - np.diag(s)))) <= 1e-12 * n
    assert np.all(np.abs(np.diag(U)) >= 1.0 - 1e-12), "pivots have magnitude >= 1"
