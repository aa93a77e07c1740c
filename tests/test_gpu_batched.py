"""Batched peers: the local QR and masks of many blocks at once (exc_batched_local_qr,
exc_givens_apply_batched, the grouped GEMM and its autotuner) and the merge by the global QR,
against the oracle's per-peer local QR and masks (client.cpp:572-605, bbtsvd.cpp:727-876) and the
QR invariants."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def blocks(nb, m, lo, hi, seed):
    rng = np.random.default_rng(seed)
    cols = rng.integers(lo, hi + 1, nb)
    scale = np.exp(rng.standard_normal(nb))
    return [rng.standard_normal((m, int(c))) * s for c, s in zip(cols, scale)]


def test_batched_factors_of_uneven_blocks(hip):
    """Uneven blocks (K from 257 to 4096, m = 70: partial tiles, split and unsplit grouped
    products chosen by the autotuner): R R^T = X X^T, R lower, Q orthonormal, R Q = X."""
    from excalibur_amd import batched as B
    rng = np.random.default_rng(0)
    xs = [rng.standard_normal((70, c)) for c in (430, 1000, 457, 4096)]
    st = B.local_stage(xs, 5, [1, 2, 3, 4], want_q=True)
    for b, x in enumerate(xs):
        rt = st["rt"][b].to_host()
        assert np.allclose(rt @ rt.T, x @ x.T, rtol=0, atol=1e-12 * np.abs(x @ x.T).max())
        assert np.allclose(np.triu(rt, 1), 0)
        q = st["qt"][b].to_host()
        assert np.linalg.norm(q @ q.T - np.eye(70), 2) <= 1e-13
        assert np.linalg.norm(rt @ q - x) <= 1e-13 * np.linalg.norm(x)
    assert "kchunk" in B.tuning()


def test_batched_local_stage_vs_oracle(hip, oracle):
    """Reduced size: 6 blocks of 48 columns of A with 250..1500 rows (one below the local-QR
    threshold). R factors equal the oracle's LAPACK local QR up to the signs of its rows, masked
    blocks equal the oracle's masks applied to them bit for bit, and the merged R equals the QR of
    the stacked masked blocks up to row signs."""
    from excalibur_amd import batched as B
    m = 48
    xs = blocks(6, m, 300, 1500, seed=1)
    xs[0] = xs[0][:, :200]  # 200 // 48 = 4: not reduced, masked as it is
    pub, priv = 12345, [7, 8, 9, 10, 11, 12]
    st = B.local_stage(xs, pub, priv, want_q=True)
    assert not st["reduced"][0] and st["reduced"][1:].all()
    for b, x in enumerate(xs):
        if st["reduced"][b]:
            rt = st["rt"][b].to_host()
            rto, qto = oracle.local_qr(x)
            assert np.max(np.abs(np.abs(rt) - np.abs(rto))) <= 1e-12 * np.abs(rto).max()
            start = rt
        else:
            start = x
        w = start.shape[1]
        want = oracle.givens_mask(oracle.givens_mask(start, 32, True, True, pub), 32, True, False, priv[b])
        assert np.array_equal(st["masked"][b].to_host(), want), b
    rt, cat = B.merge(st["masked"], m)
    rt = rt.to_host()
    S = cat.to_host()
    ref = np.linalg.qr(S.T, mode="r")
    assert np.max(np.abs(np.abs(rt.T) - np.abs(ref))) <= 1e-11 * np.abs(ref).max()
    assert np.allclose(rt @ rt.T, S @ S.T, rtol=0, atol=1e-12 * np.abs(S @ S.T).max())


def test_batched_ill_conditioned_block(hip):
    """A block with condition number 1e7 needs a second FP64 pass; Q stays orthonormal. The FP64
    passes multiply by explicit inverses of the triangular factors, so the residual R Q - X grows
    with the factor's condition number (about 1e4 after the FP32 pass here): held to 1e-11."""
    from excalibur_amd import batched as B
    rng = np.random.default_rng(3)
    m, c = 64, 2000
    U = np.linalg.qr(rng.standard_normal((m, m)))[0]
    V = np.linalg.qr(rng.standard_normal((c, m)))[0]
    x = (U * np.geomspace(1, 1e-7, m)) @ V.T
    y = rng.standard_normal((m, 3000))
    st = B.local_stage([x, y], 1, [2, 3], want_q=True)
    assert st["stats"][0, 2] >= 2
    for b, z in enumerate((x, y)):
        rt = st["rt"][b].to_host()
        q = st["qt"][b].to_host()
        assert np.linalg.norm(q @ q.T - np.eye(m), 2) <= 1e-12
        assert np.linalg.norm(rt @ q - z) <= 1e-11 * np.linalg.norm(z)


def test_batched_32_peers_properties(hip):
    """32 peers, 512 columns, rows uniform in [2000, 12000], entries N(0,1) times a lognormal(0,1)
    factor per peer: every R factor against its block's Gram matrix, and the merged R against the
    masked blocks; the batched stage against the per-peer loop's time."""
    import time
    from excalibur_amd import batched as B
    m = 512
    xs = blocks(32, m, 2000, 12000, seed=4)
    seeds = list(range(100, 132))
    t = {}
    st = B.local_stage(xs, 99, seeds, timings=t)
    rt, cat = B.merge(st["masked"], m, timings=t)
    base = B.per_peer_baseline(xs, 99, seeds)
    print(f"[gpu-batched] batched local QR {t['local_qr_s'] * 1e3:.1f} ms, masks {t['mask_s'] * 1e3:.1f} ms, "
          f"merge {t['merge_s'] * 1e3:.1f} ms; per-peer loop {base * 1e3:.1f} ms; tuning {B.tuning()}", flush=True)
    for b, x in enumerate(xs):
        if st["reduced"][b]:
            r = st["rt"][b].to_host()
            G = x @ x.T
            assert np.max(np.abs(r @ r.T - G)) <= 1e-12 * np.abs(G).max()
    rt = rt.to_host()
    S = cat.to_host()
    G = S @ S.T
    assert np.max(np.abs(rt @ rt.T - G)) <= 1e-12 * np.abs(G).max()
