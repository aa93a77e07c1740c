"""CPU checks of the C ABI: the library loads, exports every symbol include/*.h declares, and its
host-only functions (the Givens plan) match the oracle bit for bit. No GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(inc, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(exc_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_declared_symbols():
    from excalibur_amd import LIB_PATH
    assert os.path.exists(LIB_PATH), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(LIB_PATH)
    names = declared_symbols()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_signatures_cover_header():
    from excalibur_amd import hip
    import excalibur_amd.bbtsvd  # noqa: F401  (registers the host-pointer entry points)
    names = declared_symbols()
    bound = set(hip.SIGNATURES) | set(hip.EXTRA_SIGNATURES)
    missing = [n for n in names if n not in bound]
    assert not missing, f"no ctypes signature for: {missing}"


@pytest.mark.parametrize("K", [2, 3, 8, 101, 1000, 1001])
@pytest.mark.parametrize("mask", [True, False])
def test_givens_plan_matches_oracle(oracle, K, mask):
    from excalibur_amd import hip
    passes = 32
    for seed in (0, 1804289383, 4294967295):
        pairs, cs = hip.givens_plan(K, passes, mask, seed)
        op, oc, os_ = oracle.givens_plan(K, passes, mask, seed)
        assert np.array_equal(pairs.astype(np.int64), op)
        assert np.array_equal(cs[..., 0], oc) and np.array_equal(cs[..., 1], os_)


@pytest.mark.parametrize("k", [4, 10, 9973, 10001, 12345, 99_999_999, 20_000_000, 17 * 9001])
def test_find_best_convert_matches_oracle(oracle, k):
    """Host-only integer search of the K > 10000 mask (bbtsvd.cpp:166-188)."""
    from excalibur_amd import bbtsvd
    assert bbtsvd.find_best_convert(k) == oracle.find_best_convert(k)
    m, n, r = bbtsvd.find_best_convert(k)
    assert m * n + r == k and (k < 10000 or (1 < n < m < 10000 and m / n < 2))


def test_cpp_dropin_header_compiles_and_links(tmp_path):
    """A C++ caller written against the reference's bbtsvd.hpp names compiles with
    include/bbtsvd.hpp and links against the library (not run: no GPU here)."""
    import shutil
    import subprocess
    from excalibur_amd import LIB_PATH
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    src = tmp_path / "caller.cpp"
    src.write_text(r"""
#include "bbtsvd.hpp"
#include <vector>
int main() {
  std::vector<double> a = {3, 2, 1}, b = {0.5, 0.25}, l(2), d(3), s(3), u(9), v(9), x(12);
  int cnt = 0;
  int nc = negative_count(3, a.data(), b.data(), 1.5);
  bisection_singular_values(3, a.data(), b.data(), 0.0, 20.0, 1e-20, 3, s.data(), &cnt);
  ldl_decomposition(3, a.data(), b.data(), l.data(), d.data());
  std::vector<double> lp(2), dp(3), ss(3), up(2), rp(3), p(3);
  dstqds(3, l.data(), d.data(), 1.0, lp.data(), dp.data(), ss.data());
  dpqds(3, l.data(), d.data(), 1.0, up.data(), rp.data(), p.data());
  inverse_iteration(3, 1, lp.data(), up.data(), v.data(), 1);
  one_pair_singular_vector_twisted_factorization(3, 3, s[0], l.data(), d.data(), a.data(), b.data(), v.data(), u.data());
  two_side_singular_vectors_twisted_factorization(3, 3, s.data(), a.data(), b.data(), u.data(), v.data());
  bidiagonal_bisection_twisted_svd(3, a.data(), b.data(), u.data(), v.data(), s.data(), 3);
  givens_mask_cache_optimized_parallel_with_simd(3, 4, 32, x.data(), true, true, 7u);
  block_shift_mask(3, 4, 32, x.data(), false, true, 7u);
  SquareMaskConvert smc = {12345, 0, 0, 0};
  find_best_convert(&smc);
  return nc + cnt + (int)smc.m > 0 ? 0 : 1;
}
""")
    inc = os.path.join(ROOT, "include")
    libdir = os.path.dirname(LIB_PATH)
    out = tmp_path / "caller"
    r = subprocess.run(["g++", "-std=c++17", "-I", inc, str(src), "-L", libdir, "-lexcalibur_amd",
                        f"-Wl,-rpath,{libdir}", "-o", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def test_loop_colocation_decision():
    """All peers' one-launch bidiagonalization loops must be resident together on one device,
    since each waits on the others' messages (host-only decision, no GPU call)."""
    import ctypes
    from excalibur_amd import hip
    lib = hip.lib

    def shape(m, c):
        nwg = ctypes.c_int(); lds = ctypes.c_size_t()
        assert lib.exc_ralha_loop_shape(m, c, 0, ctypes.byref(nwg), ctypes.byref(lds)) == 0
        return nwg.value, lds.value

    def decide(P, nwg, slots, same):
        a = (ctypes.c_int * P)(*nwg); b = (ctypes.c_int * P)(*slots)
        return lib.exc_ralha_loops_colocate(P, a, b, 1 if same else 0)

    # P = 3, m = 2048: c = 683, 683, 682 -> 86 workgroups each of 98 KB LDS, so one per CU:
    # 256 slots on 256 CUs hold each grid alone but not all three (258)
    m = 2048
    cs = [683, 683, 682]
    shapes = [shape(m, c) for c in cs]
    assert all(nw == 86 for nw, _ in shapes)
    assert all(160 * 1024 // 2 < lds <= 160 * 1024 for _, lds in shapes)
    assert decide(3, [86, 86, 86], [256, 256, 256], True) == 0
    assert decide(3, [86, 86, 86], [256, 256, 256], False) == 1
    # config 2: two peers of c = 500 at m = 1000 -> 63 workgroups each, two per CU fit
    nw, lds = shape(1000, 500)
    assert nw == 63 and lds * 2 <= 160 * 1024
    assert decide(2, [63, 63], [512, 512], True) == 1
    # a peer without the loop (0 workgroups) makes every peer take the per-step path
    assert decide(2, [63, 0], [512, 0], True) == 0
    # the fewest slots decide
    assert decide(2, [100, 100], [512, 150], True) == 0


@pytest.mark.parametrize("shape", [(7, 5), (37, 101), (100, 37), (64, 9)])
@pytest.mark.parametrize("is_left", [True, False])
@pytest.mark.parametrize("mask", [True, False])
def test_givens_restatement_matches_oracle(oracle, shape, is_left, mask):
    """The numpy restatement that the GPU test of the FP32 sweep checks against, run in FP64, is
    bit-identical to the oracle's sweep (pass order, the odd pair's place, the signs), so only its
    rounding to float differs when it stands in for the FP32 sweep."""
    from excalibur_amd import hip
    from givens_restate import givens_numpy
    rng = np.random.default_rng(shape[0] * 13 + shape[1])
    x = rng.standard_normal(shape)
    K = shape[0] if is_left else shape[1]
    pairs, cs = hip.givens_plan(K, 32, mask, 99)
    want = oracle.givens_mask(x, 32, mask, is_left, 99)
    assert np.array_equal(givens_numpy(x, pairs, cs, mask, is_left, np.float64), want)
