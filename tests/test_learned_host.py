"""Host logic of the learned brackets (excalibur_amd/learned.py) without a GPU: the shifts that
enclose each target by the feature counts, and the interpolated-position feature, on a spectrum
given directly (no bidiagonal), against brute force."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def Lm():
    from excalibur_amd import learned
    return learned


def fake_set(Lm, sig):
    """A BidiagSet-like object for one bidiagonal with singular values sig (descending)."""
    bs = object.__new__(Lm.BidiagSet)
    n = len(sig)
    bs.ns = np.array([n], np.int32)
    bs.nb = 1
    bs.scale = np.array([np.sqrt(np.sum(sig ** 2))])
    bs.root = np.array([bs.scale[0] ** 2])
    shifts = bs.scale[0] * 10.0 ** (-0.5 * np.arange(Lm.K_SHIFTS))
    bs.frac = (np.array([[np.sum(sig < m) for m in shifts]]) / n).astype(np.float32)
    return bs


@pytest.mark.parametrize("kind", ["geometric", "linear", "clustered", "marchenko_pastur"])
def test_cell_encloses_every_value(Lm, kind):
    rng = np.random.default_rng(1)
    sig = Lm.spectrum(kind, 3000, rng)
    bs = fake_set(Lm, sig)
    ti = np.arange(3000)
    tb = np.zeros_like(ti)
    k, lo, hi = bs.cell(tb, ti)
    s = sig[ti]
    assert np.all(lo <= s) and np.all(s < hi)
    # the cell is one shift interval (a factor sqrt(10)) or reaches the root / zero
    inner = (lo > 0) & (hi < bs.root[0])
    assert np.allclose(hi[inner] / lo[inner], np.sqrt(10.0))
    # brute force: the count of values below each end
    n = 3000
    t = n - 1 - ti
    nc = lambda m: np.searchsorted(np.sort(sig), m, side="left")
    assert np.all(nc(lo) <= t) and np.all(t < nc(hi))


def test_interpolated_feature_on_a_geometric_spectrum(Lm):
    """For sigma_i = 10^(-6 i / (n - 1)) the counting function is linear in log(sigma), so in the
    shift intervals the values fill, the interpolated position gives log(sigma / scale) to within
    the spacing of the values; in the two partly filled intervals at the ends of the spectrum it
    stays within the interval (a factor sqrt(10))."""
    n = 4000
    sig = 10.0 ** (-6.0 * np.arange(n) / (n - 1))
    bs = fake_set(Lm, sig)
    ti = np.arange(0, n, 7)
    tb = np.zeros_like(ti)
    X = bs.features(tb, ti)
    est = X[:, Lm.K_SHIFTS + 1].astype(np.float64) * Lm.Y_SCALE  # ln(sigma / scale)
    want = np.log(sig[ti] / bs.scale[0])
    step = 6.0 * np.log(10.0) / (n - 1)
    _, lo, hi = bs.cell(tb, ti)
    filled = (hi <= sig.max()) & (lo >= sig.min())
    assert filled.mean() > 0.6
    assert np.max(np.abs(est - want)[filled]) <= 2 * step + 1e-5
    assert np.max(np.abs(est - want)) <= np.log(10.0) / 2 + 1e-5
    assert np.allclose(X[:, Lm.K_SHIFTS], (n - 1 - ti + 0.5) / n)
    assert np.all(X[:, Lm.K_SHIFTS + 2] == 1.0)
