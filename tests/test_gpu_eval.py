"""The reference's evaluation report (client.cpp:1215-1309 with utils.cpp:15-66) on the GPU
client, against numpy restatements of the same formulas, and the memory-mapped input path
(-f 1, utils.cpp:229-252) against the read path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def problem(m=60, parts=(700, 500), seed=0):
    rng = np.random.default_rng(seed)
    N = sum(parts)
    U = np.linalg.qr(rng.standard_normal((m, m)))[0]
    V = np.linalg.qr(rng.standard_normal((N, m)))[0]
    s = np.geomspace(1.0, 1e-3, m)
    X = (U * s) @ V.T  # m x N = A^T
    xs, o = [], 0
    for r in parts:
        xs.append(np.ascontiguousarray(X[:, o:o + r]))
        o += r
    return xs, U.T.copy(), V.T.copy(), X


def perturbed(ut, vt, seed):
    """A ground truth 1e-4 away from the exact vectors, so that each distance the report computes
    is far above the rounding level of the computed vectors and the comparison is exact to ~1e-9."""
    rng = np.random.default_rng(seed)
    return ut + 1e-4 * rng.standard_normal(ut.shape), vt + 1e-4 * rng.standard_normal(vt.shape)


def test_evaluation_mode1(hip):
    from excalibur_amd.launch import run_peers
    xs, ut0, vt0, X = problem()
    m, k = xs[0].shape[0], xs[0].shape[0]
    ut, vt = perturbed(ut0, vt0, 5)
    res = run_peers(xs, seed=0, mode=1, evaluate=True, ground_truth=(ut, vt))
    p0 = res["peers"][0]
    FU = p0["FinalU"]
    G = FU.T @ FU
    off = G[~np.eye(k, dtype=bool)]
    ev = p0["eval"]
    assert np.isclose(ev["u_orth_mean"], np.abs(off).mean(), rtol=1e-6, atol=1e-18)
    assert np.isclose(ev["u_orth_max"], np.abs(off).max(), rtol=1e-6, atol=1e-18)
    Pgt = ut.T @ ut
    P = FU @ FU.T
    iu = np.triu_indices(m)
    assert np.isclose(ev["projection_distance_u"], np.abs(Pgt - P)[iu].sum() / (m * (m + 1) / 2), rtol=1e-9)
    # the computed subspaces themselves against the exact ones
    assert np.abs(ut0.T @ ut0 - P).max() < 1e-12
    o = 0
    for c, p in enumerate(res["peers"]):
        ln = xs[c].shape[1]
        gv = vt[:, o:o + ln]
        D = p["FinalVT"].T @ p["FinalVT"] - gv.T @ gv
        want = np.abs(D)[np.triu_indices(ln)].sum() / (ln * (ln + 1) / 2)
        assert np.isclose(p["eval"]["projection_distance_vt"], want, rtol=1e-9), c
        gv0 = vt0[:, o:o + ln]
        assert np.abs(p["FinalVT"].T @ p["FinalVT"] - gv0.T @ gv0).max() < 1e-12
        o += ln


def test_evaluation_mode0_and_memmap(hip):
    from excalibur_amd.launch import run_peers
    xs, ut0, vt0, X = problem(seed=1)
    k = xs[0].shape[0]
    ut, vt = perturbed(ut0, vt0, 6)
    res = run_peers(xs, seed=0, mode=0, evaluate=True, memmap=True, ground_truth=(ut, vt))
    plain = run_peers(xs, seed=0, mode=0)
    err, o = 0.0, 0
    for c, (p, q) in enumerate(zip(res["peers"], plain["peers"])):
        assert np.array_equal(p["sigma"], q["sigma"]) and np.array_equal(p["FinalVT"], q["FinalVT"])
        ln = xs[c].shape[1]
        FV = p["FinalVT"]
        err += np.sum((np.abs(FV) - np.abs(vt[:k, o:o + ln])) ** 2)
        R = (p["FinalU"] * p["sigma"]) @ FV - xs[c]
        # a rounding-level residual: its entries depend on the order of each product's sums, so
        # the mean is held to a few percent (a formula mistake is off by orders of magnitude)
        assert np.isclose(p["eval"]["svd_error"], np.abs(R).mean(), rtol=5e-2)
        assert p["eval"]["svd_error"] < 1e-13 * np.abs(xs[c]).max()
        o += ln
    rmse = np.sqrt(err / (k * X.shape[1]))
    for p in res["peers"]:
        assert np.isclose(p["eval"]["singular_vector_error_vt"], rmse, rtol=1e-9)
