"""GPU parity of the Givens masking sweep (C ABI) against the oracle: bit-identical results for
left/right, mask/unmask, even/odd K; round trips at full size."""
import numpy as np
import pytest

from givens_restate import givens_numpy

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-givens] {msg}", flush=True)


@pytest.mark.parametrize("shape", [(2, 3), (7, 5), (100, 37), (37, 101), (1000, 1000), (513, 1001)])
@pytest.mark.parametrize("is_left", [True, False])
@pytest.mark.parametrize("mask", [True, False])
def test_givens_bit_exact(hip, oracle, shape, is_left, mask):
    banner(f"givens shape={shape} left={is_left} mask={mask}")
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    x = rng.standard_normal(shape)
    K = shape[0] if is_left else shape[1]
    seed = 1804289383
    pairs, cs = hip.givens_plan(K, 32, mask, seed)
    got = hip.givens_apply(x, pairs, cs, mask, is_left)
    want = oracle.givens_mask(x, 32, mask, is_left, seed)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("shape", [(1000, 1000), (2000, 3000), (9000, 16)])
def test_givens_round_trip_large(hip, shape):
    banner(f"givens round trip {shape}")
    rng = np.random.default_rng(5)
    x = rng.standard_normal(shape)
    for is_left in (True, False):
        K = shape[0] if is_left else shape[1]
        p1, c1 = hip.givens_plan(K, 32, True, 77)
        p0, c0 = hip.givens_plan(K, 32, False, 77)
        y = hip.givens_apply(x, p1, c1, True, is_left)
        z = hip.givens_apply(y, p0, c0, False, is_left)
        assert np.isclose(np.linalg.norm(y), np.linalg.norm(x), rtol=1e-13)
        assert np.max(np.abs(z - x)) < 1e-13


def test_givens_f32_close(hip, oracle):
    banner("givens f32")
    rng = np.random.default_rng(9)
    x = rng.standard_normal((300, 200))
    pairs, cs = hip.givens_plan(300, 32, True, 3)
    got = hip.givens_apply(x, pairs, cs, True, True, f32=True)
    want = oracle.givens_mask(x, 32, True, True, 3)
    assert np.max(np.abs(got - want)) < 2e-5 * np.abs(want).max()


@pytest.mark.parametrize("shape", [(7, 5), (37, 101), (1000, 1000), (513, 1001)])
@pytest.mark.parametrize("is_left", [True, False])
@pytest.mark.parametrize("mask", [True, False])
def test_givens_f32_bit_exact(hip, shape, is_left, mask):
    """exc_givens_apply_f32 (the mixed form's masks) against its float32 restatement, bit for bit,
    for even and odd K."""
    banner(f"givens f32 shape={shape} left={is_left} mask={mask}")
    rng = np.random.default_rng(shape[0] + 3 * shape[1])
    x = rng.standard_normal(shape).astype(np.float32)
    K = shape[0] if is_left else shape[1]
    pairs, cs = hip.givens_plan(K, 32, mask, 99)
    got = hip.givens_apply(x, pairs, cs, mask, is_left, f32=True)
    assert np.array_equal(got, givens_numpy(x, pairs, cs, mask, is_left, np.float32))


@pytest.mark.parametrize("m,n,is_left", [(6, 10007, False), (5, 12100, False), (10501, 4, True)])
def test_virtual_matrix_mask_bit_exact(hip, oracle, m, n, is_left):
    """block_shift_mask above 10000 lines (bbtsvd.cpp:548-668): the virtual vm x vn matrix, the
    cyclic shift and the remainder sweeps, mask and unmask, bit for bit with the oracle."""
    rng = np.random.default_rng(m + n)
    x = rng.standard_normal((m, n))
    got = hip.block_shift_mask(x, 32, True, is_left, 4242)
    want = oracle.block_shift_mask(x, 32, True, is_left, 4242)
    assert np.array_equal(got, want)
    back = hip.block_shift_mask(got, 32, False, is_left, 4242)
    assert np.array_equal(back, oracle.block_shift_mask(want, 32, False, is_left, 4242))
    assert np.max(np.abs(back - x)) <= 1e-12
