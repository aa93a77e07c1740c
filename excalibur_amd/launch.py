"""Runs the GPU DecentralizedClient as P peer processes on loopback (one excalibur_peer process per
peer, the reference's main.cpp command line), and collects their timers and results.

The peers share the one GPU; they talk over host TCP on 127.0.0.1..127.0.0.P at base_port + cid,
the addressing of client.cpp:238-245.
"""
import json
import os
import socket
import subprocess
import tempfile
import time

import numpy as np

from . import PEER_PATH


def free_port(span):
    """A base port with `span` consecutive free ports."""
    for _ in range(50):
        s = socket.socket()
        s.bind(("127.0.0.1", 0))
        base = s.getsockname()[1]
        s.close()
        if base + span >= 65000:
            continue
        ok = True
        for p in range(base, base + span):
            t = socket.socket()
            try:
                t.bind(("", p))
            except OSError:
                ok = False
            finally:
                t.close()
            if ok is False:
                break
        if ok:
            return base
    raise RuntimeError("no free port range")


def write_peer_files(xs, dpath):
    """Client{i}.mat: raw float64, row-major m x local_n_i (client.cpp:577, generate_data.cpp:184)."""
    os.makedirs(dpath, exist_ok=True)
    for i, x in enumerate(xs):
        np.ascontiguousarray(x, dtype=np.float64).tofile(os.path.join(dpath, f"Client{i}.mat"))


def write_ground_truth(u, vt, dpath):
    """u.mat (k x m, row i = u_i) and vt.mat (k x global_n, row i = v_i), the ground-truth files
    generate_data.cpp writes for evaluation (generate_data.cpp:152-186)."""
    os.makedirs(dpath, exist_ok=True)
    np.ascontiguousarray(u, dtype=np.float64).tofile(os.path.join(dpath, "u.mat"))
    np.ascontiguousarray(vt, dtype=np.float64).tofile(os.path.join(dpath, "vt.mat"))


def run_peers(xs, seed=0, mode=1, topk=-1, opt=3, form="fp64", dump=True, workdir=None, timeout=600,
              recv_timeout=300.0, extra=None, evaluate=False, memmap=False, ground_truth=None):
    """Runs len(xs) peers on the column blocks xs (each m x local_n_i). Returns dict with per-peer
    timers and, when dump, sigma / FinalU / FinalVT arrays. evaluate: the reference's evaluation
    report (-e 1; ground_truth = (u, vt) is written next to the data); memmap: -f 1."""
    if not os.path.exists(PEER_PATH):
        raise RuntimeError(f"{PEER_PATH} is not built")
    P = len(xs)
    m = xs[0].shape[0]
    own = workdir is None
    workdir = workdir or tempfile.mkdtemp(prefix="exc_peers_")
    dpath = os.path.join(workdir, "data")
    logdir = os.path.join(workdir, "logs")
    os.makedirs(logdir, exist_ok=True)
    write_peer_files(xs, dpath)
    if ground_truth is not None:
        write_ground_truth(ground_truth[0], ground_truth[1], dpath)
    def launch_once():
        port = free_port(P + 1)
        procs = []
        for cid in range(P):
            cmd = [PEER_PATH, "-l", "127.0.0.1", "-p", str(port), "-m", str(m), "-n", str(xs[cid].shape[1]),
                   "-d", dpath, "-g", logdir, "-k", str(topk), "-o", str(mode), "-t", str(P), "-c", str(cid),
                   "-e", "1" if evaluate else "0", "-f", "1" if memmap else "0", "-s", str(seed), "-a", str(opt),
                   "--form", form, "--dump", "1" if dump else "0", "--recv-timeout", str(recv_timeout)] + list(extra or [])
            out = open(os.path.join(logdir, f"peer{cid}.out"), "w")
            procs.append((subprocess.Popen(cmd, stdout=out, stderr=subprocess.STDOUT), out))
        t0 = time.time()
        rcs = []
        for p, out in procs:
            left = max(1.0, timeout - (time.time() - t0))
            try:
                rcs.append(p.wait(timeout=left))
            except subprocess.TimeoutExpired:
                for q, _ in procs:
                    if q.poll() is None:
                        q.kill()
                rcs.append(-9)
            out.close()
        return rcs

    # another process can take a probed port before a peer binds it: the peers then start again
    # on other ports (once)
    for attempt in range(2):
        rcs = launch_once()
        if all(rc == 0 for rc in rcs) or attempt == 1:
            break
        logs = [open(os.path.join(logdir, f"peer{cid}.out")).read() for cid in range(P)]
        if not any("bind failed" in text for text in logs):
            break
    res = {"workdir": workdir, "rc": rcs, "peers": []}
    for cid in range(P):
        text = open(os.path.join(logdir, f"peer{cid}.out")).read()
        info = None
        for line in text.splitlines():
            if line.startswith("EXC_RESULT "):
                info = json.loads(line[len("EXC_RESULT "):])
        peer = {"log": text, "result": info}
        for line in text.splitlines():
            if line.startswith("EXC_EVAL "):
                peer["eval"] = json.loads(line[len("EXC_EVAL "):])
        if dump and info is not None:
            k = info["sigma_count"]
            base = os.path.join(logdir, f"Client{cid}_")
            peer["sigma"] = np.fromfile(base + "sigma.bin")
            if os.path.exists(base + "FinalU.bin"):
                # k x m in the tall-and-skinny branch (m > global_n), as the reference leaves it
                tall = m > sum(x.shape[1] for x in xs)
                peer["FinalU"] = np.fromfile(base + "FinalU.bin").reshape((k, m) if tall else (m, k))
            if os.path.exists(base + "FinalVT.bin"):
                peer["FinalVT"] = np.fromfile(base + "FinalVT.bin").reshape(k, xs[cid].shape[1])
            for name in ("alpha", "beta", "bidiag_v"):
                if os.path.exists(base + name + ".bin"):
                    peer[name] = np.fromfile(base + name + ".bin")
        res["peers"].append(peer)
    if any(rc != 0 for rc in rcs):
        logs = "\n".join(p["log"][-2000:] for p in res["peers"])
        raise RuntimeError(f"peer processes failed rc={rcs}\n{logs}")
    if own:
        for cid in range(P):
            try:
                os.remove(os.path.join(dpath, f"Client{cid}.mat"))
            except OSError:
                pass
    return res
