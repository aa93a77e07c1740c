"""excalibur_amd: MI355X (gfx950) implementation of Excalibur's per-peer bbtsvd matrix pipeline.

The compute lives in libexcalibur_amd.so (hand-written HIP kernels behind a C ABI, declared in
include/excalibur_hip.h). This package only loads that library (see excalibur_amd.hip) and
provides thin Python conveniences used by tests and the benchmark. There is no CPU fallback:
importing the binding raises when the library has not been built.
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libexcalibur_amd.so")
PEER_PATH = os.path.join(HERE, "excalibur_peer")


def build(verbose=False):
    from . import builder as _b
    _b.build(verbose=verbose)
