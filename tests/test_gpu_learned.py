"""Learned bisection brackets: the training kernels' gradient against the CPU restatement, BF16
inference against the FP64 forward, the generated bidiagonals, the index bisection against the
oracle's bisection (bbtsvd.cpp:19-52) bit for bit, and the trained network's brackets leaving every
singular value unchanged while sparing Sturm counts."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-learned] {msg}", flush=True)


@pytest.fixture(scope="module")
def L(hip):
    from excalibur_amd import learned
    return learned


def test_gradient_matches_oracle(L):
    """Forward loss and the flat gradient of exc_mlp_forward/backward (FP32, MFMA GEMMs) against
    oracle/mlp.py in float64 on the same parameters and batch."""
    from oracle import mlp
    rng = np.random.default_rng(0)
    p = mlp.init(rng)
    p[mlp.N_PARAMS - 2:] = [0.05, -0.2]
    X = rng.uniform(0, 1, (300, mlp.D_IN)).astype(np.float32)
    y = rng.uniform(-1.5, 0, 300).astype(np.float32)
    net = L.BracketNet()
    net.set_params(p.astype(np.float32))
    loss, g = net.grad(X, y)
    lo, go = mlp.grad(p.astype(np.float32).astype(np.float64), X.astype(np.float64), y.astype(np.float64))
    assert abs(loss - lo) <= 1e-4 * abs(lo)
    assert np.linalg.norm(g - go) <= 2e-3 * np.linalg.norm(go)
    # per block, so a wrong small block cannot hide behind a large one
    o = 0
    for name, shape in mlp.SHAPES:
        k = int(np.prod(shape))
        assert np.linalg.norm(g[o:o + k] - go[o:o + k]) <= 5e-3 * max(np.linalg.norm(go[o:o + k]), 1e-6), name
        o += k


def test_bf16_inference_matches_fp64_forward(L):
    from oracle import mlp
    rng = np.random.default_rng(1)
    p = mlp.init(rng)
    X = rng.uniform(0, 1, (1000, mlp.D_IN)).astype(np.float32)
    net = L.BracketNet()
    net.set_params(p.astype(np.float32))
    got = net.infer(X)
    want, _ = mlp.forward(p.astype(np.float32).astype(np.float64), X.astype(np.float64))
    scale = np.abs(want).max()
    assert np.max(np.abs(got - want)) <= 3e-2 * scale


def gk_restated(sig, start, steps):
    """The first steps of Golub-Kahan on diag(sig) in numpy (alpha_j u_j = D v_j - beta_{j-1}
    u_{j-1}, beta_j v_{j+1} = D u_j - alpha_j v_j)."""
    v = start / np.linalg.norm(start)
    u = np.zeros_like(v)
    beta, d, e = 0.0, [], []
    for j in range(steps):
        u = sig * v - beta * u
        alpha = np.linalg.norm(u)
        u /= alpha
        d.append(alpha)
        v = sig * u - alpha * v
        beta = np.linalg.norm(v)
        v /= beta
        e.append(beta)
    return np.array(d), np.array(e)


def test_gk_bidiagonals(L):
    """exc_gk_bidiagonals: the first steps equal the recurrence restated in numpy (before rounding
    errors, amplified by the loss of orthogonality of plain Lanczos, set the sequences apart); the
    bidiagonal's singular values lie in [min sigma, max sigma] (Ritz values), the largest is
    max sigma (it converges first), and the first alpha is ||D v1||. Plain Lanczos keeps no
    orthogonality, so later values are copies of converged ones rather than the prescribed
    spectrum; the labels come from bisection of the bidiagonal itself, so the data stays
    consistent."""
    rng = np.random.default_rng(2)
    for n, sig in ((48, np.linspace(1.0, 0.1, 48)), (3000, np.geomspace(1.0, 1e-6, 3000))):
        start = rng.standard_normal(n)
        bs = L.BidiagSet([n], sig, start)
        d, e = bs.bidiagonal(0)
        rd, re = gk_restated(sig, start, 8)
        assert np.allclose(d[:8], rd, rtol=1e-12, atol=0) and np.allclose(e[:8], re, rtol=1e-12, atol=0)
        assert np.isclose(d[0], np.linalg.norm(sig * start) / np.linalg.norm(start), rtol=1e-14)
        sv = np.linalg.svd(np.diag(d) + np.diag(e, 1), compute_uv=False)
        assert abs(sv[0] - sig.max()) <= 1e-12
        assert sv.min() >= sig.min() - 1e-13 and sv.max() <= sig.max() + 1e-13
        assert np.isclose(bs.scale[0], np.sqrt(np.sum(d * d) + np.sum(e * e)), rtol=1e-14)


def test_bisect_index_matches_oracle_bisection(L, oracle):
    """Every singular value of generated bidiagonals of each family, one lane per index: the leaves
    they land in, one value per leaf, equal the oracle's bisection from the reference's root bit
    for bit (Lanczos copies of a value share a leaf, as in the reference's tree); random brackets
    (valid, wrong and inverted) change nothing."""
    bs = L.make_set(8, (200, 700), seed=3)
    rng = np.random.default_rng(4)
    for b in range(bs.nb):
        n = int(bs.ns[b])
        a, e = bs.bidiagonal(b)
        root = bs.root[b]
        assert root == oracle.sigma_upper_bound(a, e)
        want = oracle.bisection(a, e, 0.0, root, 1e-20, n)[0]
        tb, ti = np.full(n, b), np.arange(n)
        plain = bs.bisect(tb, ti)
        assert np.all(plain["status"] == 0)
        assert np.all(np.diff(plain["sigma"]) <= 0)
        assert np.array_equal(L.leaf_list(plain), want), bs.families[b]
        s = plain["sigma"]
        lo = s * np.exp(rng.uniform(-2, 0.0, n)) * (1 - 1e-9)
        hi = s * np.exp(rng.uniform(0.0, 2, n)) * (1 + 1e-9)
        bad = rng.random(n) < 0.3
        lo[bad], hi[bad] = hi[bad] * 1.5, hi[bad] * 3  # brackets that miss the value
        inv = rng.random(n) < 0.1
        lo[inv], hi[inv] = hi[inv], lo[inv]
        br = bs.bisect(tb, ti, lo, hi)
        assert np.array_equal(br["sigma"], s) and np.array_equal(br["leaf_lo"], plain["leaf_lo"])
        # a bracket above the value bounds it from above (or, past the root, is not used); an
        # inverted one is not used
        assert np.all(np.isin(br["status"][bad & ~inv], (0, 2)))
        assert np.all(br["status"][inv] == 0)
        assert np.all(br["status"][~bad & ~inv] == 1)
        assert br["iters"][br["status"] == 1].mean() < plain["iters"].mean()
        assert br["iters"][br["status"] == 2].mean() < plain["iters"].mean()
        # the enclosing feature shifts as fallback points close the side a missed bracket leaves open
        _, clo, chi = bs.cell(tb, ti)
        fb = bs.bisect(tb, ti, lo, hi, clo, chi)
        assert np.array_equal(fb["sigma"], s) and np.array_equal(fb["leaf_lo"], plain["leaf_lo"])
        assert np.mean(fb["status"] == 1) > np.mean(br["status"] == 1)
        assert fb["iters"].sum() < br["iters"].sum()


def test_learned_brackets_save_counts(L, hip, oracle):
    """Train on generated bidiagonals (n in [512, 4096]) with the library's kernels, then bisect
    every singular value of n = 16,384 bidiagonals of the four families from the BF16 network's
    brackets: identical leaves to the plain bisection (and the leading values to the oracle and
    the production multisection), fewer Sturm counts."""
    banner("training set: 2000 bidiagonals, n in [512, 4096]")
    train = L.make_set(2000, (512, 4096), seed=10)
    X, y = L.training_data(train, per=48, seed=11)
    net = L.BracketNet(seed=0)
    hist = net.train(X, y, steps=1500, batch=4096, lr=2e-3, seed=1, log=banner)
    assert hist["loss"][-1][1] < hist["loss"][0][1]
    held = L.make_set(200, (512, 4096), seed=12)
    k_lo, k_hi = net.calibrate(held, *L.targets(held, per=48, seed=13))
    banner(f"calibrated bracket ends: {k_lo:.2f} and {k_hi:.2f} Laplace scales")
    test = L.make_set(4, (16384, 16384), seed=20)
    saved = []
    for b in range(test.nb):
        r = L.evaluate(net, test, b)
        banner(f"{test.families[b]}: counts {r['counts_plain']} -> {r['counts_bracketed']} "
               f"({100 * r['saved_frac']:.1f}% saved; the enclosing shifts alone "
               f"{100 * r['saved_frac_shifts_only']:.1f}%), bounded on both sides {100 * r['bounded_both']:.1f}%, "
               f"one side {100 * r['bounded_one']:.1f}%, "
               f"{r['ms_plain']:.1f} -> {r['ms_bracketed']:.1f} ms (+{r['ms_inference']:.2f} ms inference)")
        assert r["identical"]
        saved.append(r["saved_frac"])
        a, e = test.bidiagonal(b)
        top = oracle.bisection(a, e, 0.0, test.root[b], 1e-20, 40)[0]
        assert np.array_equal(r["leaves"][:40], top)
        sg = hip.bisection(a, e, 0.0, test.root[b], 1e-20, 16384)[0]
        assert np.array_equal(sg, r["leaves"])
    assert np.mean(saved) > 0.1
