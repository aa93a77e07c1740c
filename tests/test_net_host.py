"""The host TCP transport (excalibur_amd/csrc/net.cpp) without a GPU: P peer processes of a small
driver (tests/cpp/net_main.cpp) run all_reduce_sum_same (ring above P * 10000 values, in peer order
below) and all_reduce_sum_same_f32; every peer must hold the same bits, and the FP32 form must
equal the FP64 sum in peer order of the widened values. receive_direct must deliver every frame
intact, both when the buffer is posted before the frame arrives and when the frame is queued
first."""
import os
import socket
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "cpp", "_build")


@pytest.fixture(scope="module")
def driver():
    os.makedirs(BUILD, exist_ok=True)
    exe = os.path.join(BUILD, "net_main")
    src = [os.path.join(HERE, "cpp", "net_main.cpp"), os.path.join(ROOT, "excalibur_amd", "csrc", "net.cpp")]
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "excalibur_amd", "csrc"), *src, "-o", exe,
                    "-lpthread"], check=True)
    return exe


def free_port(span):
    """A base port whose span ports are all free now (each peer listens on base + its id)."""
    for _ in range(50):
        with socket.socket() as s:
            s.bind(("127.0.0.1", 0))
            p = s.getsockname()[1]
        if p + span >= 65000:
            continue
        try:
            for q in range(p, p + span):
                with socket.socket() as t:
                    t.bind(("", q))
            return p
        except OSError:
            continue
    raise RuntimeError("no port")


def run_peers(driver, P, n):
    """Runs the P peers; another process can take a probed port before a peer binds it, so a bind
    failure starts the peers again on other ports (at most three times)."""
    for attempt in range(3):
        port = free_port(P)
        procs = [subprocess.Popen([driver, str(c), str(P), str(port), str(n)], stdout=subprocess.PIPE,
                                  stderr=subprocess.PIPE, text=True) for c in range(P)]
        res = [p.communicate(timeout=120) for p in procs]
        if any(p.returncode != 0 and "bind failed" in err for p, (_, err) in zip(procs, res)):
            continue
        for p, (_, err) in zip(procs, res):
            assert p.returncode == 0, err
        return [out.split() for out, _ in res]
    raise RuntimeError("no free port span after three attempts")


@pytest.mark.parametrize("P,n", [(2, 1000), (2, 300_000), (3, 50_000), (3, 7)])
def test_sums_identical_on_every_peer(driver, P, n):
    outs = run_peers(driver, P, n)
    assert all(o[:2] == outs[0][:2] for o in outs)
    assert all(o[2] == "0" for o in outs)
    assert all(o[3] == "0" for o in outs)  # receive_direct delivered every frame intact
