"""Pins two conventions of the oracle that the end-to-end checks cannot see, because singular
values and reconstructions are invariant under them:

- drotg (client.cpp:951, the lower-to-upper conversion): c, s against LAPACK's BLAS drotg as
  scipy ships it, over every sign and magnitude case, and r (left in the first argument) as
  c*a + s*b with the sign of the larger-magnitude input.
- the peer sign rule (client.cpp:987-995): the first largest-magnitude entry of every column of
  bu ends positive, and the matching row of bv flips with it.

Planted mistakes (an r that is always positive, a rule keyed on the largest value instead of the
largest magnitude, or on the last index of a tie) are shown to fail these checks.
"""
import itertools

import numpy as np
import pytest
from scipy.linalg.blas import drotg as blas_drotg

MAGS = [0.0, 1e-300, 1e-8, 0.5, 1.0, 3.0, 4.0, 1e8, 1e300]


def cases():
    for x, y in itertools.product(MAGS, MAGS):
        for sx, sy in itertools.product((1.0, -1.0), (1.0, -1.0)):
            yield sx * x, sy * y


def same(x, w):
    """Equal to 4 ulps and of the same sign (a sign slip on a tiny c or s is caught too)."""
    return abs(x - w) <= 4 * np.spacing(max(abs(w), 1e-300)) and np.sign(x) == np.sign(w)


def test_drotg_matches_blas(oracle):
    n = 0
    for a, b in cases():
        r, z, c, s = oracle.drotg(a, b)
        cw, sw = blas_drotg(a, b)
        assert same(c, cw) and same(s, sw), (a, b, c, s, cw, sw)
        n += 1
    assert n == len(MAGS) ** 2 * 4


def test_drotg_r_convention(oracle):
    for a, b in cases():
        r, z, c, s = oracle.drotg(a, b)
        if a == 0.0 and b == 0.0:
            assert (r, c, s) == (0.0, 1.0, 0.0)
            continue
        roe = a if abs(a) > abs(b) else b
        assert np.sign(r) == np.sign(roe)
        assert np.isclose(abs(r), np.hypot(a, b), rtol=1e-15)
        # the rotation maps (a, b) to (r, 0)
        assert np.isclose(c * a + s * b, r, rtol=1e-15, atol=0)
        assert abs(-s * a + c * b) <= 1e-15 * abs(r)


def test_drotg_planted_positive_r_is_caught(oracle):
    """A restatement that always takes r >= 0 gives different c, s whenever the larger input is
    negative; the BLAS comparison above would reject it."""
    caught = 0
    for a, b in cases():
        if a == 0.0 and b == 0.0:
            continue
        r = np.hypot(a, b)
        c_bad, s_bad = a / r, b / r
        cw, sw = blas_drotg(a, b)
        if not (same(c_bad, cw) and same(s_bad, sw)):
            caught += 1
    assert caught >= len(MAGS) ** 2


def sign_rule_holds(bu0, bv0, bu, bv):
    m, cnt = bu.shape
    for i in range(cnt):
        col = bu[:, i]
        p = int(np.argmax(np.abs(col)))  # first index of the largest magnitude
        if col[p] < 0:
            return False
        flipped = not np.array_equal(col, bu0[:, i])
        if flipped:
            if not (np.array_equal(col, -bu0[:, i]) and np.array_equal(bv[i], -bv0[i])):
                return False
        elif not np.array_equal(bv[i], bv0[i]):
            return False
    return True


def test_sign_rule(oracle):
    rng = np.random.default_rng(3)
    m, cnt = 9, 6
    bu0 = rng.standard_normal((m, cnt)); bv0 = rng.standard_normal((cnt, m))
    # column 0: largest magnitude negative, largest value positive
    bu0[:, 0] = [0.1, 0.7, -0.9, 0.2, 0.0, 0.3, -0.1, 0.5, 0.4]
    # column 1: a tie in magnitude, the first is negative
    bu0[:, 1] = [0.2, -0.8, 0.1, 0.8, 0.0, 0.3, 0.1, 0.2, 0.1]
    # column 2: a tie, the first is positive
    bu0[:, 2] = [0.2, 0.8, 0.1, -0.8, 0.0, 0.3, 0.1, 0.2, 0.1]
    bu, bv = oracle.sign_normalize(bu0, bv0)
    assert sign_rule_holds(bu0, bv0, bu, bv)
    assert np.array_equal(bu[:, 0], -bu0[:, 0]) and np.array_equal(bu[:, 1], -bu0[:, 1])
    assert np.array_equal(bu[:, 2], bu0[:, 2])


def test_sign_rule_planted_mistakes_are_caught(oracle):
    bu0 = np.array([[0.1, 0.2], [0.7, -0.8], [-0.9, 0.8]]); bv0 = np.ones((2, 3))
    # keyed on the largest value: column 0 would stay (0.7 is the largest value)
    bad_u = bu0.copy(); bad_v = bv0.copy()
    for i in range(2):
        if bu0[np.argmax(bu0[:, i]), i] < 0:
            bad_u[:, i] *= -1; bad_v[i] *= -1
    assert not sign_rule_holds(bu0, bv0, bad_u, bad_v)
    # keyed on the last index of a tie: column 1 would stay (+0.8 at index 2)
    bad_u = bu0.copy(); bad_v = bv0.copy()
    for i in range(2):
        a = np.abs(bu0[:, i]); p = len(a) - 1 - int(np.argmax(a[::-1]))
        if bu0[p, i] < 0:
            bad_u[:, i] *= -1; bad_v[i] *= -1
    assert not sign_rule_holds(bu0, bv0, bad_u, bad_v)
    # u flipped without v
    bad_u = bu0.copy(); bad_u[:, 0] *= -1; bad_u[:, 1] *= -1
    assert not sign_rule_holds(bu0, bv0, bad_u, bv0)
    good_u, good_v = oracle.sign_normalize(bu0, bv0)
    assert sign_rule_holds(bu0, bv0, good_u, good_v)


@pytest.mark.parametrize("seed", [0, 1])
def test_sign_rule_on_a_bidiagonal_svd(oracle, seed):
    """On the vectors the pipeline feeds it (client.cpp:984-995): the rule holds and B v = s u."""
    rng = np.random.default_rng(seed)
    n = 40
    a, b = rng.uniform(0.5, 2.0, n), rng.uniform(0.1, 1.0, n - 1)
    s, u, v, _, rc = oracle.bidiag_svd(a, b)
    assert rc == n
    bu, bv = oracle.sign_normalize(u, v)
    assert sign_rule_holds(u, v, bu, bv)
    B = np.diag(a) + np.diag(b, 1)
    assert np.abs(B @ bv.T - bu * s).max() <= 1e-12 * s[0]
