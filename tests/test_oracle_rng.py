"""Pins the Givens index sequence of the oracle (bbtsvd.cpp:729-761) to an independent
restatement of the C++ library algorithms the reference calls, written here in Python from
their published definitions:

- std::default_random_engine is std::minstd_rand0 in libstdc++: the Lehmer generator
  x <- 16807 x mod (2^31 - 1), seeded with s mod (2^31 - 1) (0 maps to 1), range [1, 2^31 - 2]
  (ISO C++ [rand.predef]; libstdc++ <bits/random.h>).
- std::shuffle in libstdc++ 11 (<bits/stl_algo.h>): when the generator's range R satisfies
  R / n >= n, it draws swap positions two at a time from one uniform_int_distribution over
  [0, (i+1)(i+2) - 1], splitting the draw into x / (i+2) and x % (i+2); for even n the first
  swap, with element 1, takes one draw over [0, 1]. Otherwise it draws one position per element.
- std::uniform_int_distribution for a generator whose range is neither 2^32 - 1 nor 2^64 - 1
  (<bits/uniform_int_dist.h>): rejection with scaling = R / (range + 1),
  past = (range + 1) * scaling, result = (g() - min) / scaling once g() - min < past.
- glibc rand() for the odd-K extra index and the rotation values, read through ctypes.

The planted-mistake tests show that a different engine or a seed off by one changes the plan,
so the equality checks are not vacuous.
"""
import ctypes

import numpy as np
import pytest

libc = ctypes.CDLL(None)
libc.rand.restype = ctypes.c_int
libc.srand.argtypes = [ctypes.c_uint]
RAND_MAX = 2147483647


class Lehmer:
    """Multiplicative congruential engine, modulus 2^31 - 1 (minstd_rand0 for a = 16807)."""
    M = 2147483647

    def __init__(self, seed, a=16807):
        self.a = a
        x = seed % self.M
        self.x = 1 if x == 0 else x

    min = 1
    max = M - 1

    def __call__(self):
        self.x = (self.a * self.x) % self.M
        return self.x


def uniform_int(g, lo, hi):
    urngrange = g.max - g.min
    urange = hi - lo
    assert urngrange > urange
    uerange = urange + 1
    scaling = urngrange // uerange
    past = uerange * scaling
    while True:
        ret = g() - g.min
        if ret < past:
            break
    return ret // scaling + lo


def shuffle(t, g):
    n = len(t)
    if n == 0:
        return
    urngrange = g.max - g.min
    if urngrange // n >= n:
        i = 1
        if n % 2 == 0:
            j = uniform_int(g, 0, 1)
            t[i], t[j] = t[j], t[i]
            i += 1
        while i != n:
            swap_range = i + 1
            x = uniform_int(g, 0, swap_range * (swap_range + 1) - 1)
            p1, p2 = x // (swap_range + 1), x % (swap_range + 1)
            t[i], t[p1] = t[p1], t[i]
            i += 1
            t[i], t[p2] = t[p2], t[i]
            i += 1
        return
    for i in range(1, n):
        j = uniform_int(g, 0, i)
        t[i], t[j] = t[j], t[i]


def plan(K, passes, mask, seed, engine=Lehmer):
    """The whole plan as bbtsvd.cpp:729-779 draws it: pairs (passes, half, 2), c, s."""
    libc.srand(seed)
    g = engine(seed)
    idx = []
    for _ in range(passes):
        t = list(range(K))
        shuffle(t, g)
        if K % 2 == 1:
            t.append(libc.rand() % K)
            while t[K] == t[K - 1]:
                t[K] = libc.rand() % K
        if not mask:
            t.reverse()
        for j in range(0, len(t), 2):
            if t[j] > t[j + 1]:
                t[j], t[j + 1] = t[j + 1], t[j]
        idx.append(t)
    if not mask:
        idx.reverse()
    half = len(idx[0]) // 2
    c = np.zeros((passes, half)); s = np.zeros((passes, half))
    for i in range(passes):
        for j in range(half):
            cv = (libc.rand() / RAND_MAX) * 2.0 - 1.0
            sv = (libc.rand() / RAND_MAX) * 2.0 - 1.0
            r = pow(cv * cv + sv * sv, 0.5)
            c[i, j] = cv / r
            s[i, j] = sv / r
        if not mask:
            c[i] = c[i][::-1].copy(); s[i] = s[i][::-1].copy()
    if not mask:
        c = c[::-1].copy(); s = s[::-1].copy()
    pairs = np.array(idx, dtype=np.int64).reshape(passes, half, 2)
    return pairs, c, s


def test_lehmer_known_values():
    # minstd_rand0 seeded with 1: the 10000th output is 1043618065 (ISO C++ [rand.predef]).
    g = Lehmer(1)
    for _ in range(9999):
        g()
    assert g() == 1043618065
    # seed 0 is mapped to 1, seeds are taken mod 2^31 - 1
    assert Lehmer(0).x == 1 and Lehmer(2147483647).x == 1 and Lehmer(2147483648).x == 1


def test_shuffle_is_a_permutation_and_uses_both_paths():
    g = Lehmer(5)
    for n in (1, 2, 3, 10, 11, 1000):
        t = list(range(n)); shuffle(t, g)
        assert sorted(t) == list(range(n))
    # n^2 > range of the engine: the one-draw-per-element path
    t = list(range(50000)); shuffle(t, Lehmer(3))
    assert sorted(t) == list(range(50000))


@pytest.mark.parametrize("K", [2, 3, 8, 9, 100, 101, 1000, 1001])
@pytest.mark.parametrize("mask", [True, False])
@pytest.mark.parametrize("seed", [0, 42, 1804289383])
def test_oracle_plan_equals_restatement(oracle, K, mask, seed):
    passes = 32 if K <= 101 else 4
    op, oc, os_ = oracle.givens_plan(K, passes, mask, seed)
    rp, rc, rs = plan(K, passes, mask, seed)
    assert np.array_equal(op, rp), "index sequence differs from minstd_rand0 + libstdc++ shuffle"
    assert np.array_equal(oc, rc) and np.array_equal(os_, rs)


def test_oracle_plan_large_K_one_draw_path(oracle):
    """K = 46341 exceeds sqrt(2^31 - 3), so libstdc++ shuffles with one draw per element."""
    K = 46341
    op, _, _ = oracle.givens_plan(K, 1, True, 7)
    rp, _, _ = plan(K, 1, True, 7)
    assert np.array_equal(op, rp)


@pytest.mark.parametrize("K", [100, 101])
def test_planted_mistakes_change_the_plan(oracle, K):
    op, _, _ = oracle.givens_plan(K, 4, True, 42)
    # minstd_rand (multiplier 48271) instead of minstd_rand0
    wp, _, _ = plan(K, 4, True, 42, engine=lambda s: Lehmer(s, a=48271))
    assert not np.array_equal(op, wp)
    # seed off by one
    wp, _, _ = plan(K, 4, True, 43)
    assert not np.array_equal(op, wp)
    # one draw per element instead of pairs of positions
    g = Lehmer(42)
    t = list(range(K))
    for i in range(1, K):
        j = uniform_int(g, 0, i)
        t[i], t[j] = t[j], t[i]
    first = sorted(tuple(sorted(t[j:j + 2])) for j in range(0, K - K % 2, 2))
    assert first != sorted(tuple(p) for p in op[0][: K // 2].tolist())
