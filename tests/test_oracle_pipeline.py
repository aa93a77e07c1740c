"""CPU checks that pin the oracle's pipeline stages (local and global QR, Ralha, lower->upper,
the whole simulated run) against LAPACK fixtures and against what the mathematics fixes:
orthogonality, reconstruction, sigma(B) = sigma(A), and LAPACK's reflector sign convention.
Each check is chosen so that a plausible slip (a dropped term, a wrong sign or index) fails it;
the planted-mistake cases show that for the Ralha and run checks."""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

import os

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_elementary_qr_matches_lapack_fixtures(oracle):
    z = np.load(os.path.join(GOLDEN, "qr_lapack.npz"))
    for tag in ("tall40x7", "sq9", "tall100x20"):
        A = z["A_" + tag]
        F, tau = oracle.dgeqrf(A)
        scale = np.abs(z["qr_" + tag]).max()
        assert np.abs(F - z["qr_" + tag]).max() <= 1e-13 * scale
        assert np.abs(tau - z["tau_" + tag]).max() <= 1e-13
        Q = oracle.dorgqr(F, tau)
        assert np.abs(Q - z["q_" + tag]).max() <= 1e-13


@pytest.mark.parametrize("m,r", [(7, 40), (20, 20), (30, 170)])
def test_local_qr_is_lapack_qr_of_xt(oracle, m, r):
    """local_qr_reduction (client.cpp:417-445): R^T and Q^T of dgeqrf/dorgqr of X^T."""
    rng = np.random.default_rng(m * r)
    x = rng.standard_normal((m, r))
    rt, qt = oracle.local_qr(x)
    qr_, tau, _, info = lapack.dgeqrf(x.T.copy())
    assert info == 0
    R = np.triu(qr_[:m, :m])
    q, _, info = lapack.dorgqr(qr_, tau)
    assert np.abs(rt - R.T).max() <= 1e-13 * np.abs(R).max()
    assert np.abs(qt - q[:, :m].T).max() <= 1e-13
    assert np.abs(qt.T @ rt.T - x.T).max() <= 1e-13 * np.abs(x).max()


@pytest.mark.parametrize("P,m,R", [(2, 20, 50), (3, 17, 40)])
def test_global_qr_is_lapack_qr_of_stacked_panels(oracle, P, m, R):
    """The sequential panel protocol (client.cpp:706-794) factors the R x m matrix whose column
    slabs the peers hold; the result must be LAPACK's QR of that matrix."""
    rng = np.random.default_rng(P * 100 + m)
    mp = [m // P + (1 if i < m % P else 0) for i in range(P)]
    Y = rng.standard_normal((R, m))
    pos = np.concatenate([[0], np.cumsum(mp)])
    panels = [Y[:, pos[c]:pos[c + 1]].T.copy() for c in range(P)]
    shr, qg, tau = oracle.global_qr(panels, m)
    qr_, tau_l, _, _ = lapack.dgeqrf(Y.copy())
    Rl = np.triu(qr_[:m, :m])
    q, _, _ = lapack.dorgqr(qr_, tau_l)
    assert np.abs(tau - tau_l).max() <= 1e-13
    for c in range(P):
        assert np.abs(shr[c] - Rl[:, pos[c]:pos[c + 1]]).max() <= 1e-13 * np.abs(Rl).max()
    assert np.abs(qg - q).max() <= 1e-13


def _bidiag_sigma(alpha, beta):
    m = alpha.size
    B = np.diag(alpha) + np.diag(beta, -1)  # lower bidiagonal of the one-sided loop
    return np.linalg.svd(B, compute_uv=False), B


@pytest.mark.parametrize("P,m,cols", [(1, 15, [15]), (2, 24, [12, 12]), (3, 31, [11, 10, 10])])
def test_ralha_bidiagonal_has_the_singular_values_of_f(oracle, P, m, cols):
    """One-sided bidiagonalization (client.cpp:829-910): F H = U B with B lower bidiagonal,
    U orthonormal, H = H_1...H_{m-1} from the stored reflectors; all peers agree on alpha/beta
    up to their summation order."""
    rng = np.random.default_rng(m)
    Rf = np.linalg.qr(rng.standard_normal((2 * m, m)))[1]
    F = [Rf[:, sum(cols[:p]):sum(cols[:p + 1])].copy() for p in range(P)]
    out = oracle.ralha(F)
    s_true = np.linalg.svd(Rf, compute_uv=False)
    for p in range(P):
        s, B = _bidiag_sigma(out[p]["alpha"], out[p]["beta"])
        assert np.abs(s - s_true).max() <= 1e-13 * s_true[0]
        assert np.abs(out[p]["alpha"] - out[0]["alpha"]).max() <= 1e-12 * s_true[0]
    U = np.hstack([out[p]["U"] for p in range(P)])  # m x c_total, rows u_j
    assert np.abs(U @ U.T - np.eye(m)).max() <= 1e-12
    # H from the reflectors (column j of house, implied unit diagonal), then F^T... : U^T B = H^T F^T
    H = np.eye(m)
    house, tau = out[0]["house"], out[0]["tau"]
    for j in range(1, m):
        v = np.zeros(m)
        v[j] = 1.0
        v[j + 1:] = house[j + 1:, j]
        H = H @ (np.eye(m) - tau[j] * np.outer(v, v))
    Ff = np.hstack(F)  # m x c_total with rows f_j
    lhs = H.T @ Ff     # rows: the transformed f's
    assert np.abs(lhs - B @ U).max() <= 1e-12 * s_true[0]
    # planted mistake: dropping the subdiagonal changes the singular values detectably
    s_bad, _ = _bidiag_sigma(out[0]["alpha"], np.zeros(m - 1))
    assert np.abs(s_bad - s_true).max() > 1e-6 * s_true[0]


def test_lower_to_upper_keeps_singular_values(oracle):
    rng = np.random.default_rng(5)
    a, b = rng.uniform(0.5, 2, 30), rng.uniform(0.1, 1, 29)
    s_low = np.linalg.svd(np.diag(a) + np.diag(b, -1), compute_uv=False)
    a2, b2, cs = oracle.lower_to_upper(a, b)
    s_up = np.linalg.svd(np.diag(a2) + np.diag(b2, 1), compute_uv=False)
    assert np.abs(s_low - s_up).max() <= 1e-14 * s_low[0]
    c, s = cs[0::2], cs[1::2]
    assert np.abs(c * c + s * s - 1).max() <= 1e-15


def _check_run(run, xs, sigma_ref, tol_sigma, tol_orth, tol_res):
    P = len(xs)
    m = xs[0].shape[0]
    k = run.count
    X = np.hstack(xs)  # the reference's A^T = [X_1 ... X_P]
    sg = run.get("sigma")[:k]
    U = run.get("FinalU").reshape(m, k)
    VT = np.hstack([run.get("FinalVT", p).reshape(k, xs[p].shape[1]) for p in range(P)])
    assert np.abs(sg - sigma_ref[:k]).max() <= tol_sigma * sigma_ref[0]
    assert np.abs(U.T @ U - np.eye(k)).max() <= tol_orth
    assert np.abs(VT @ VT.T - np.eye(k)).max() <= tol_orth
    assert np.linalg.norm(X - (U * sg) @ VT) <= tol_res * np.linalg.norm(X)
    return U, sg, VT


def test_run_config1_matches_lapack_fixture(oracle):
    """2 peers x 500 x 100 N(0,1) row blocks (seed 0): sigma against the committed LAPACK values."""
    z = np.load(os.path.join(GOLDEN, "svd_lapack.npz"))
    rng = np.random.default_rng(0)
    A = rng.standard_normal((1000, 100))
    xs = [np.ascontiguousarray(A[i * 500:(i + 1) * 500].T) for i in range(2)]
    run = oracle.Run(xs, [0, 0])
    assert run.count == 100
    _check_run(run, xs, z["sigma_c1"], 1e-13, 1e-12, 1e-13)


@pytest.mark.parametrize("P,m,widths,seed", [(2, 20, [60, 45], 3), (3, 16, [100, 20, 40], 9)])
def test_run_small_matches_numpy_svd(oracle, P, m, widths, seed):
    """Local QR fires for the wide blocks (local_n/m > 5), not for the narrow one."""
    rng = np.random.default_rng(seed)
    xs = [rng.standard_normal((m, w)) for w in widths]
    run = oracle.Run(xs, [seed] * P)
    s_ref = np.linalg.svd(np.hstack(xs), compute_uv=False)
    U, sg, VT = _check_run(run, xs, s_ref, 1e-13, 1e-12, 1e-13)
    # planted mistake: a sign slip in one singular vector breaks the reconstruction check
    U_bad = U.copy()
    U_bad[:, 0] *= -1
    X = np.hstack(xs)
    assert np.linalg.norm(X - (U_bad * sg) @ VT) > 1e-3 * np.linalg.norm(X)


def _tall_problem(m, parts, seed, decay=6.0):
    """X = U diag(s) V^T (m x global_n, m >= global_n) split by columns. For m > global_n peer p's
    file holds X_p^T (local_n x m row-major), the layout the reference reads in the
    tall-and-skinny branch; for m == global_n it holds X_p (m x local_n), as in the
    short-and-wide branch whose stages that case takes."""
    rng = np.random.default_rng(seed)
    n = sum(parts)
    U = np.linalg.qr(rng.standard_normal((m, n)))[0]
    V = np.linalg.qr(rng.standard_normal((n, n)))[0]
    s = 10.0 ** (-decay * np.arange(n) / (n - 1))
    X = (U * s) @ V.T
    raw, o = [], 0
    for ln in parts:
        if m > n:
            B = np.ascontiguousarray(X[:, o:o + ln].T)  # local_n x m
            raw.append(B.reshape(m, ln))                  # the same bytes, shaped as the file's m x local_n
        else:
            raw.append(np.ascontiguousarray(X[:, o:o + ln]))
        o += ln
    return X, s, raw


@pytest.mark.parametrize("m,parts,decay", [(300, [40, 35], 2.0), (301, [50, 50, 27], 2.0), (120, [70, 50], 2.0),
                                           (150, [60, 30], 2.0), (300, [40, 35], 6.0), (150, [60, 30], 6.0)])
def test_oracle_run_tall_and_square(oracle, m, parts, decay):
    """The tall-and-skinny branch (m > global_n: masks on the local_n x m view, re-partition by
    columns, local QR of the m_part x global_n blocks when m_part > global_n, bidiagonal of order
    global_n, FinalU k x m) and the square case m == global_n (the short-and-wide stages with the
    tall combine of the right vectors), pinned by what the mathematics fixes: the exact singular
    values, X = FinalU^T diag(sigma) [FinalVT_0 ... FinalVT_{P-1}], and orthonormal vectors. The
    last holds to 1e-12 for a spectrum over two decades; over six, the vectors built from the
    one-sided bidiagonalization's v_i = f_i / alpha_i lose orthogonality to ~1e-6 (the
    reference's algorithm, without reorthogonalization), so only sigma and X are held there."""
    X, s, raw = _tall_problem(m, parts, seed=m + len(parts), decay=decay)
    n = sum(parts)
    run = oracle.Run(raw, [7] * len(parts))
    k = run.count
    assert k == n
    sg = run.get("sigma", 0)[:k]
    assert np.max(np.abs(sg - s)) <= 1e-11 * s[0]
    FU = run.get("FinalU", 0)
    U = FU.reshape(k, m).T if m > n else FU.reshape(m, k)
    VT = np.hstack([run.get("FinalVT", p).reshape(k, ln) for p, ln in enumerate(parts)])
    assert np.linalg.norm(X - (U * sg) @ VT) <= 1e-13 * np.linalg.norm(X)
    if decay <= 2.0:
        assert np.abs(U.T @ U - np.eye(k)).max() <= 1e-12
        assert np.abs(VT @ VT.T - np.eye(k)).max() <= 1e-12
    # every peer holds the gathered FinalU; sigma bits agree for two peers, and within an ulp or
    # two beyond (each peer adds the others' messages to its own first, client.cpp:318-323)
    for p in range(1, len(parts)):
        assert np.array_equal(run.get("FinalU", p), FU)
        if len(parts) == 2:
            assert np.array_equal(run.get("sigma", p), run.get("sigma", 0))
        else:
            assert np.max(np.abs(run.get("sigma", p) - run.get("sigma", 0))) <= 4 * np.finfo(float).eps * s[0]


def test_oracle_run_tall_wrong_unmask_is_caught(oracle):
    """A planted mistake (the peers' files in the short-and-wide layout while m > global_n) gives
    the SVD of another matrix, which the reconstruction check rejects."""
    m, parts = 200, [30, 30]
    X, s, raw = _tall_problem(m, parts, seed=3)
    wrong = [np.ascontiguousarray(X[:, o:o + ln]) for o, ln in ((0, 30), (30, 30))]
    run = oracle.Run(wrong, [7, 7])
    k = run.count
    U = run.get("FinalU", 0).reshape(k, m).T
    VT = np.hstack([run.get("FinalVT", p).reshape(k, 30) for p in range(2)])
    assert np.linalg.norm(X - (U * run.get("sigma", 0)[:k]) @ VT) > 1e-3 * np.linalg.norm(X)
