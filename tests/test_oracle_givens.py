"""Pins the oracle's Givens masking (bbtsvd.cpp:533-669, 727-876) by what fixes it: the C
library's rand() stream read directly through ctypes, permutation and pairing invariants,
orthogonality (norm and singular values preserved), mask/unmask round trips, and a brute-force
product of the rotation matrices on small cases."""
import ctypes

import numpy as np
import pytest

libc = ctypes.CDLL(None)
libc.rand.restype = ctypes.c_int
libc.srand.argtypes = [ctypes.c_uint]
RAND_MAX = 2147483647


def libc_cs(seed, passes, half):
    """c, s of an even-K plan straight from the C library (no rand() calls for indices)."""
    libc.srand(seed)
    c = np.zeros((passes, half)); s = np.zeros((passes, half))
    for i in range(passes):
        for j in range(half):
            cv = (libc.rand() / RAND_MAX) * 2.0 - 1.0
            sv = (libc.rand() / RAND_MAX) * 2.0 - 1.0
            r = pow(cv * cv + sv * sv, 0.5)
            c[i, j] = cv / r
            s[i, j] = sv / r
    return c, s


@pytest.mark.parametrize("seed", [0, 7, 123456789])
def test_rotation_values_follow_libc_rand(oracle, seed):
    K, passes = 10, 4
    pairs, c, s = oracle.givens_plan(K, passes, True, seed)
    cw, sw = libc_cs(seed, passes, K // 2)
    assert np.array_equal(c, cw) and np.array_equal(s, sw)
    assert np.allclose(c * c + s * s, 1.0, atol=4e-16)


@pytest.mark.parametrize("K", [2, 9, 64, 255])
def test_plan_structure(oracle, K):
    passes = 32
    pairs, c, s = oracle.givens_plan(K, passes, True, 42)
    half = (K + 1) // 2
    assert pairs.shape == (passes, half, 2)
    assert np.all(pairs[..., 0] < pairs[..., 1]), "each pair stored smaller index first"
    for p in range(passes):
        flat = pairs[p].reshape(-1)
        if K % 2 == 0:
            assert np.array_equal(np.sort(flat), np.arange(K)), "a pass is a permutation"
        else:
            # the first K entries (before pairing) are a permutation; one extra index repeats
            vals, counts = np.unique(flat, return_counts=True)
            assert np.array_equal(vals, np.arange(K)) and counts.sum() == K + 1 and counts.max() == 2
            last = pairs[p, -1]
            assert last[0] != last[1]
    # passes use one engine: consecutive passes differ
    if K > 2:
        assert not np.array_equal(pairs[0], pairs[1])
    # unmask plan = passes reversed, each pass's list reversed, values reversed
    up, uc, us = oracle.givens_plan(K, passes, False, 42)
    if K % 2 == 0:
        assert np.array_equal(np.sort(up[0].reshape(-1)), np.arange(K))
    assert np.array_equal(uc, c[::-1, ::-1]) and np.array_equal(us, s[::-1, ::-1])


def rotation_matrix(K, pairs, c, s, mask):
    """Brute-force product of the rotations, in application order, as a K x K matrix G with
    x_new = G x (x a column of the left-masked matrix)."""
    G = np.eye(K)
    for p in range(pairs.shape[0]):
        for j in range(pairs.shape[1]):
            i1, i2 = pairs[p, j]
            R = np.eye(K)
            sp = s[p, j] if mask else -s[p, j]
            R[i1, i1] = c[p, j]; R[i1, i2] = sp
            R[i2, i1] = -sp; R[i2, i2] = c[p, j]
            G = R @ G
    return G


@pytest.mark.parametrize("K", [6, 7])
def test_mask_equals_rotation_product(oracle, K):
    rng = np.random.default_rng(K)
    x = rng.standard_normal((K, 5))
    pairs, c, s = oracle.givens_plan(K, 32, True, 99)
    G = rotation_matrix(K, pairs, c, s, True)
    got = oracle.givens_mask(x, 32, True, True, 99)
    assert np.allclose(got, G @ x, atol=1e-13)
    assert np.allclose(G @ G.T, np.eye(K), atol=1e-13)
    # right mask on rows: x G^T with K = columns
    y = rng.standard_normal((4, K))
    got = oracle.givens_mask(y, 32, True, False, 99)
    assert np.allclose(got, y @ G.T, atol=1e-13)


@pytest.mark.parametrize("shape", [(100, 37), (37, 100), (64, 64), (1, 50), (51, 1)])
@pytest.mark.parametrize("is_left", [True, False])
def test_mask_unmask_round_trip(oracle, shape, is_left):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape)
    y = oracle.block_shift_mask(x, 32, True, is_left, 1234)
    K = shape[0] if is_left else shape[1]
    if K > 1:
        assert not np.allclose(x, y)
    assert np.isclose(np.linalg.norm(y), np.linalg.norm(x), rtol=1e-13)
    z = oracle.block_shift_mask(y, 32, False, is_left, 1234)
    assert np.max(np.abs(z - x)) < 1e-13
    sx = np.linalg.svd(x, compute_uv=False); sy = np.linalg.svd(y, compute_uv=False)
    assert np.allclose(sx, sy, rtol=0, atol=1e-12 * sx[0])


def test_find_best_convert(oracle):
    for k in [10001, 12345, 20000, 99991, 1000000]:
        m, n, r = oracle.find_best_convert(k)
        assert m * n + r == k
        assert m < 10000 and n > 1 and m > n and m / n < 2.0


@pytest.mark.slow
def test_virtual_matrix_mask_round_trip(oracle):
    """K > 10000 goes through the virtual sqrt(K) x sqrt(K) Givens + cyclic shift path."""
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 10007))
    y = oracle.block_shift_mask(x, 32, True, False, 5)
    assert np.isclose(np.linalg.norm(y[0]), np.linalg.norm(x[0]), rtol=1e-12)
    z = oracle.block_shift_mask(y, 32, False, False, 5)
    assert np.max(np.abs(z - x)) < 1e-12
