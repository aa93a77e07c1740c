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
