"""GPU parity of the host-pointer drop-ins for the reference's bbtsvd.hpp functions
(include/excalibur_bbtsvd.h, include/bbtsvd.hpp) against the oracle: counts, leaves, twist
factorizations, vectors and masked matrices are bit-identical, as for the device-pointer API."""
import numpy as np
import pytest

from conftest import random_bidiagonal

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-compat] {msg}", flush=True)


@pytest.fixture(scope="module")
def bb(hip):
    from excalibur_amd import bbtsvd
    return bbtsvd


@pytest.mark.parametrize("kind", ["gauss", "positive", "graded"])
@pytest.mark.parametrize("n", [1, 2, 33, 250])
def test_bidiagonal_functions_bit_exact(bb, oracle, kind, n):
    banner(f"bidiagonal functions n={n} kind={kind}")
    rng = np.random.default_rng(7 * n + len(kind))
    a, b = random_bidiagonal(rng, n, kind) if n > 1 else (np.array([1.25]), np.zeros(0))
    ub = oracle.sigma_upper_bound(a, b)
    for mu in [0.0, 0.5 * np.sqrt(ub), np.abs(a).max(), ub]:
        assert bb.negative_count(n, a, b, mu) == oracle.negative_count(a, b, mu)
    for top_k in sorted({1, max(1, n // 3), n}):
        got = bb.bisection_singular_values(n, a, b, 0.0, ub, 1e-20, top_k)
        want, _, _ = oracle.bisection(a, b, 0.0, ub, 1e-20, top_k)
        assert np.array_equal(got, want)
    # empty interval and top_k <= 0 record nothing (bbtsvd.cpp:25)
    assert bb.bisection_singular_values(n, a, b, ub, 0.0, 1e-20, n).size == 0
    assert bb.bisection_singular_values(n, a, b, 0.0, ub, 1e-20, 0).size == 0
    l, d = bb.ldl_decomposition(n, a, b)
    lo, do = oracle.ldl(a, b)
    assert np.array_equal(l, lo) and np.array_equal(d, do)
    sig, _, _ = oracle.bisection(a, b, 0.0, ub, 1e-20, n)  # may hold fewer than n leaves
    ns = sig.size
    tau = sig[ns // 2] ** 2
    for got, want in zip(bb.dstqds(n, l, d, tau), oracle.dstqds(lo, do, tau)):
        assert np.array_equal(got, want)
    for got, want in zip(bb.dpqds(n, l, d, tau), oracle.dpqds(lo, do, tau)):
        assert np.array_equal(got, want)
    lp, _, _ = oracle.dstqds(lo, do, tau)
    up, _, _ = oracle.dpqds(lo, do, tau)
    for k in sorted({0, n // 2, n - 1}):
        for ldv in (1, 3):
            v0 = rng.standard_normal((n - 1) * ldv + 1)
            got = bb.inverse_iteration(n, k, lp, up, v0.copy(), ldv)
            want = oracle.inverse_iteration(n, k, lp, up, v0, ldv)
            assert np.array_equal(got, want)
    cnt = n
    u0 = rng.standard_normal((n - 1) * cnt + 1)
    for j in sorted({0, ns - 1}):
        gv, gu = bb.one_pair_singular_vector_twisted_factorization(n, cnt, sig[j], l, d, a, b, np.zeros(n),
                                                                    u0.copy())
        wv, wu, _, _ = oracle.one_pair(n, cnt, sig[j], lo, do, a, b, np.zeros(n), u0)
        assert np.array_equal(gv, wv) and np.array_equal(gu, wu)
    gu, gv = bb.two_side_singular_vectors_twisted_factorization(n, ns, sig, a, b)
    wu, wv, _ = oracle.two_side(a, b, sig)
    assert np.array_equal(gu, wu) and np.array_equal(gv, wv)
    for top_k in sorted({1, ns, n}):
        so, uo, vo, _, rc = oracle.bidiag_svd(a, b, top_k)
        if rc < 0:  # fewer leaves than top_k: the reference's assert (bbtsvd.cpp:938)
            from excalibur_amd.hip import HipError
            with pytest.raises(HipError):
                bb.bidiagonal_bisection_twisted_svd(n, a, b, top_k)
            continue
        u, v, s = bb.bidiagonal_bisection_twisted_svd(n, a, b, top_k)
        assert np.array_equal(s, so) and np.array_equal(u, uo) and np.array_equal(v, vo)


@pytest.mark.parametrize("m,n", [(9, 14), (64, 33), (300, 301)])
@pytest.mark.parametrize("is_left", [True, False])
def test_masks_bit_exact(bb, oracle, m, n, is_left):
    banner(f"masks {m}x{n} left={is_left}")
    rng = np.random.default_rng(m * n)
    x = rng.standard_normal((m, n))
    for mask, seed in [(True, 5), (False, 5), (True, 4294967295)]:
        got = bb.givens_mask_cache_optimized_parallel_with_simd(m, n, 32, x.copy(), mask, is_left, seed)
        want = oracle.givens_mask(x, 32, mask, is_left, seed)
        assert np.array_equal(got, want)
        got = bb.block_shift_mask(m, n, 32, x.copy(), mask, is_left, seed)
        want = oracle.block_shift_mask(x, 32, mask, is_left, seed)
        assert np.array_equal(got, want)
    # mask then unmask with the same seed gives x back (orthogonal rotations)
    y = bb.block_shift_mask(m, n, 32, x.copy(), True, is_left, 11)
    z = bb.block_shift_mask(m, n, 32, y, False, is_left, 11)
    assert np.abs(z - x).max() < 1e-13 * max(1.0, np.abs(x).max())


@pytest.mark.parametrize("mask", [True, False])
def test_virtual_mask_branch_bit_exact(bb, oracle, mask):
    """block_shift_mask above 10000 lines (the virtual-matrix branch, bbtsvd.cpp:548-668) through
    the host-pointer drop-in, bit for bit against the oracle."""
    rng = np.random.default_rng(9)
    x = rng.standard_normal((3, 10001))
    want = oracle.block_shift_mask(x.copy(), 32, mask, False, 123)
    got = x.copy()
    bb.block_shift_mask(3, 10001, 32, got, mask, False, 123)
    assert np.array_equal(got, want)
