#!/usr/bin/env python3
"""Benchmark of the per-peer bbtsvd pipeline on one MI355X: the mixed-precision form against the
plain FP64 form.

Workload (default): 2 peers, each holding a 50,000 x 1,000 float64 row block of
A = Q1 diag(sigma) Q2^T (100,000 x 1,000), Q1, Q2 random orthogonal (QR of seeded Gaussians),
sigma_i = 10^(-6 i / 999); full SVD (all 1,000 triplets). The peers are two processes of
excalibur_amd/excalibur_peer sharing the GPU and talking over loopback TCP, the reference's
DecentralizedClient::run() workflow; a step is one run() of both peers with the input already
resident in HBM, timed from "All Clients Connected" to "Finish" (max over the peers).

`value` is the seconds per SVD of the mixed form (FP32-MFMA CholeskyQR passes, one FP64 pass,
FP64 bisection + twisted factorization of the shared factor and the FP64 Ogita-Aishima
refinement); the plain FP64 form (the reference's algorithm: local QR, Givens masks, global QR,
per-column all-reduce bidiagonalization, bisection + twisted factorization) is measured beside it
on the same input, and both again on the short-wide 2 x 2,000 x 100,000 workload with a
clustered spectrum. Accuracy against the generator's exact singular values is reported for
every run. Prints one JSON line. --gpus N > 1 exits: the peers share one GPU, and only that case is measured.
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0

CONFIGS = {
    # name: (peers, rows per peer, columns, spectrum)
    "c2": (2, 50_000, 1_000, "geometric6"),
    "c3": (2, 100_000, 2_000, "clustered8"),
    "c1": (2, 500, 100, "gaussian"),
    "small": (2, 6_000, 200, "geometric6"),
    "small3": (2, 10_000, 400, "clustered8"),
    "cpu_sample": (2, 10_000, 1_000, "geometric6"),
}

WORKLOADS = {
    "geometric6": "float64 row blocks of A = Q1 diag(sigma) Q2^T, sigma_i = 10^(-6i/(n-1))",
    "clustered8": "float64 column blocks of a short-wide A = Q1 diag(sigma) Q2^T, sigma = 20 values "
                  "1 + 1e-8 linspace(-1, 1, 20), then a geometric decay to 1e-8",
    "gaussian": "N(0,1) row blocks",
}


def spectrum(n, spec):
    if spec == "geometric6":
        return 10.0 ** (-6.0 * np.arange(n) / (n - 1))
    if spec == "clustered8":
        cl = np.sort(1.0 + 1e-8 * np.linspace(-1, 1, 20))[::-1]
        return np.concatenate([cl, np.geomspace(cl[-1] * 0.99, 1e-8, n - 20)])
    raise ValueError(spec)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_blocks(cfg, seed=0):
    """Row blocks A_i (rows x n) of A, returned in the reference's orientation X_i = A_i^T."""
    P, rows, n, spec = CONFIGS[cfg]
    rng = np.random.default_rng(seed)
    if spec == "gaussian":
        A = rng.standard_normal((P * rows, n))
        sigma = None
    else:
        sigma = spectrum(n, spec)
        Q1 = np.linalg.qr(rng.standard_normal((P * rows, n)))[0]
        Q2 = np.linalg.qr(rng.standard_normal((n, n)))[0]
        A = (Q1 * sigma) @ Q2.T
        del Q1
    xs = [np.ascontiguousarray(A[i * rows:(i + 1) * rows].T) for i in range(P)]
    return xs, sigma, A


def ensure_built():
    from excalibur_amd import LIB_PATH, PEER_PATH
    if not (os.path.exists(LIB_PATH) and os.path.exists(PEER_PATH)):
        log("library not built: building")
        from excalibur_amd import builder
        builder.build()


def accuracy(peers, xs, sigma_true, A):
    """Normwise relative sigma error, ||U^T U - I|| (max entry and Frobenius), ||A - U S V^T||/||A||."""
    k = peers[0]["result"]["sigma_count"]
    sg = peers[0]["sigma"]
    U = peers[0]["FinalU"]         # reference FinalU = right singular vectors of A (m x k)
    VT = np.hstack([p["FinalVT"] for p in peers])  # left singular vectors of A, transposed
    ref = sigma_true if sigma_true is not None else np.linalg.svd(A, compute_uv=False)
    out = {"sigma_rel_err": float(np.max(np.abs(sg - ref[:k])) / ref[0])}
    G = U.T @ U - np.eye(k)
    out["orth_u_max"] = float(np.abs(G).max())
    out["orth_u_fro"] = float(np.linalg.norm(G))
    H = VT @ VT.T - np.eye(k)
    out["orth_v_max"] = float(np.abs(H).max())
    out["orth_v_fro"] = float(np.linalg.norm(H))
    # X = U diag(s) VT with X = A^T
    X = A.T
    R = X - (U * sg) @ VT
    out["residual_rel"] = float(np.linalg.norm(R) / np.linalg.norm(X))
    return out


def pmc_traffic(name="k_gemm_final_v_c2_pmc.md"):
    """HBM bytes per launch of the final V product from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy of known size); None when absent."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def isolated_gemms(n, rows, reps=5):
    """The two largest products of the run, each launched alone on the GPU from this process and
    timed with HIP events through the library's profiler: the final V product (n x rows x n) and
    the local-QR Gram matrix (n x n x rows, lower half, split-K)."""
    import ctypes
    from excalibur_amd import hip
    rng = np.random.default_rng(5)
    A = hip.DeviceArray.from_host(rng.standard_normal((n, n)))
    B = hip.DeviceArray.from_host(rng.standard_normal((n, rows)))
    C = hip.DeviceArray((n, rows)).zero()
    G = hip.DeviceArray((n, n)).zero()
    ws = hip.workspace(hip.lib.exc_gemm_workspace(n, n, rows))
    buf = ctypes.create_string_buffer(1 << 16)
    out = {}
    for name, launch, flops in (
            ("final_v", lambda: hip.call("exc_dgemm", 0, 0, n, rows, n, 1.0, A.ptr, n, B.ptr, rows, 0.0, C.ptr, rows,
                                         None, 0, None), 2.0 * n * rows * n),
            ("gram", lambda: hip.call("exc_dgemm", 0, 1, n, n, rows, 1.0, B.ptr, rows, B.ptr, rows, 0.0, G.ptr, n,
                                      ws.ptr, ws.nbytes, None), 2.0 * n * n * rows)):
        launch()
        hip.sync()
        hip.call("exc_prof_enable", 1)
        for _ in range(reps):
            launch()
        hip.sync()
        hip.call("exc_prof_enable", 0)
        hip.call("exc_prof_report", buf, len(buf))
        rec = json.loads(buf.value.decode())["gemm_f64"]
        ms = rec["ms"] / rec["launches"]
        out[name] = {"avg_launch_us": round(1000.0 * ms, 2), "achieved": round(flops / (ms * 1e-3) / 1e12, 3)}
    return out


def isolated_givens(n=1000, passes=32, reps=6, seed=5):
    """The FP64 Givens mask sweep alone on an n x n matrix (32 passes of disjoint rotations, rows
    then columns), timed with HIP events through the library's profiler, for PMC runs. Algorithmic
    bytes per sweep: the matrix read and written once, 2 * 8 * n * n."""
    import ctypes
    from excalibur_amd import hip
    rng = np.random.default_rng(seed)
    x = hip.DeviceArray.from_host(rng.standard_normal((n, n)))
    pairs, cs = hip.givens_plan(n, passes, True, seed)
    dp = hip.DeviceArray.from_host(pairs)
    dc = hip.DeviceArray.from_host(cs)
    buf = ctypes.create_string_buffer(1 << 16)
    out = {}
    for name, left in (("rows", 1), ("columns", 0)):
        def launch():
            hip.call("exc_givens_apply", x.ptr, n, n, dp.ptr, dc.ptr, passes, 1, left, None)
        launch()
        hip.sync()
        hip.call("exc_prof_enable", 1)
        for _ in range(reps):
            launch()
        hip.sync()
        hip.call("exc_prof_enable", 0)
        hip.call("exc_prof_report", buf, len(buf))
        rec = json.loads(buf.value.decode())["givens_f64"]
        ms = rec["ms"] / rec["launches"]
        out[name] = {"avg_launch_us": round(1000.0 * ms, 2), "gb_s": round(2.0 * 8 * n * n / (ms * 1e-3) / 1e9, 1)}
    return out


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_baseline(sample="cpu_sample"):
    """The oracle (the CPU FP64 restatement of the path) timed on the host cores on a bounded
    sample of the workload: the whole simulated run() of both peers in one process, its
    independent columns and rows split over OMP_NUM_THREADS threads. Measured, not scaled."""
    from oracle import pyoracle as O
    xs, _, _ = make_blocks(sample, seed=1)
    model, nproc = cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(nproc, 16)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    O.lib()
    t0 = time.time()
    run = O.Run(xs, [0] * len(xs))
    dt = time.time() - t0
    P, rows, n, _ = CONFIGS[sample]
    stages = {k: round(run.timer(k), 4) for k in ("time_local_qr", "time_global_qr", "time_bi_diag", "time_bsvd",
                                                   "time_combine_svd_remove_mask")}
    return {"value": round(dt, 4), "unit": "s", "cores": threads, "kind": "port",
            "sample": f"oracle (C++ FP64 restatement of the path, one process simulating both peers, "
                      f"{threads} OpenMP threads) on {P} peers x {rows} x {n} rows of A with the same "
                      f"spectrum as the workload, full SVD; measured, not scaled",
            "nproc": nproc, "cpu_model": model, "stages_s": stages}


def run_variant(xs, form, device_exchange, steps, warmup, dump, workdir):
    """One run_peers call: warmup + steps repeats of run() on the resident input. Returns the step
    times (max over peers), the stage breakdown, the HIP-event kernel records and the peers."""
    from excalibur_amd.launch import run_peers
    repeat = warmup + steps
    res = run_peers(xs, seed=0, mode=1, topk=-1, opt=3, form=form, dump=dump, workdir=workdir, timeout=1200,
                    extra=["--repeat", str(repeat), "--profile", "1", "--device-exchange", str(device_exchange)])
    per_peer = []
    for p in res["peers"]:
        runs, profs = [], []
        for line in p["log"].splitlines():
            if line.startswith("EXC_RESULT "):
                runs.append(json.loads(line[len("EXC_RESULT "):]))
            elif line.startswith("EXC_PROF "):
                profs.append(json.loads(line[len("EXC_PROF "):]))
        per_peer.append((runs, profs))
    timed = range(warmup, repeat)
    step_times = [max(pp[0][r]["total"] for pp in per_peer) for r in timed]
    keys = ["time_local_qr", "time_apply_mask", "time_exchange_low_dim", "time_global_qr", "time_bi_diag",
            "time_ralha_allreduce", "time_bsvd", "time_combine_svd_remove_mask", "communication"]
    keys += [k for k in per_peer[0][0][warmup] if k.startswith("time_refine")]
    stages = {k: float(np.mean([max(pp[0][r].get(k, 0.0) for pp in per_peer) for r in timed])) for k in keys}
    # the copy of the results into the host members happens after "Finish", on the last run only
    d2h = max(pp[0][repeat - 1].get("time_results_d2h", 0.0) for pp in per_peer)
    agg = {}
    for pp in per_peer:
        for r in timed:
            for name, v in pp[1][r].items():
                a = agg.setdefault(name, {"launches": 0, "ms": 0.0, "flops": 0.0, "bytes": 0.0})
                for f in ("launches", "ms", "flops", "bytes"):
                    a[f] += v[f]
    return {"step_times": step_times, "ms_per_step": 1000.0 * float(np.mean(step_times)), "stages": stages,
            "agg": agg, "steps": len(timed), "results_d2h_s": d2h, "res": res}


def rate(rec):
    return rec["flops"] / (rec["ms"] * 1e-3) / 1e12 if rec["ms"] > 0 else 0.0


def summarize(run):
    """Per-step kernel records and the FP64 GEMM roofline entry of one measured variant."""
    agg, nsteps = run["agg"], run["steps"]
    g = agg.get("gemm_f64", {"launches": 0, "ms": 0.0, "flops": 0.0, "bytes": 0.0})
    fv = agg.get("gemm_f64_final_v", {"launches": 0, "ms": 0.0, "flops": 0.0, "bytes": 0.0})
    achieved = rate(g)
    roof = {"bound": "mfma", "kernel": "k_gemm<double> (all FP64 MFMA GEMM launches of both peers)",
            "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None,
            "launches_per_step": g["launches"] / max(1, nsteps),
            "avg_launch_us": round(1000.0 * g["ms"] / max(1, g["launches"]), 2),
            "flops_per_step": g["flops"] / max(1, nsteps),
            "final_v": {"launches_per_step": fv["launches"] / max(1, nsteps),
                        "avg_launch_us": round(1000.0 * fv["ms"] / max(1, fv["launches"]), 2),
                        "achieved": round(rate(fv), 3), "frac": round(rate(fv) / PEAK_FP64_TFLOPS, 4)}}
    kernels = {k: {"launches": v["launches"] / nsteps, "ms": v["ms"] / nsteps, "achieved": rate(v),
                   "gbs": (v["bytes"] / (v["ms"] * 1e-3) / 1e9) if v["ms"] > 0 else 0.0}
               for k, v in agg.items()}
    return roof, kernels


def measure(xs, sigma_true, A, form, device_exchange, steps, warmup, check, workdir=None):
    t0 = time.time()
    run = run_variant(xs, form, device_exchange, steps, warmup, check, workdir)
    log(f"{form} ({'TCP' if device_exchange == 0 else 'device exchange'}): {run['ms_per_step']:.2f} ms per SVD "
        f"({time.time() - t0:.1f} s)")
    roof, kernels = summarize(run)
    out = {"ms_per_step": round(run["ms_per_step"], 3), "step_times_s": [round(v, 6) for v in run["step_times"]],
           "stages_s": run["stages"], "results_d2h_s": round(run["results_d2h_s"], 6), "roofline": roof,
           "kernels": kernels}
    if check:
        t0 = time.time()
        out["accuracy"] = accuracy(run["res"]["peers"], xs, sigma_true, A)
        log(f"accuracy of {form} in {time.time() - t0:.1f} s: {out['accuracy']}")
    return out


def bench_batched(nb=32, m=512, lo=2000, hi=60000, seed=4):
    """32 peers, 512 columns, rows uniform in [2000, 60000], entries N(0,1) times a lognormal(0,1)
    factor per peer: the local QR and masks of all peers at once (batched, FP32 bulk + FP64
    refinement), merged by the global QR, against the same stage one peer after another."""
    from excalibur_amd import batched as B
    t0 = time.time()
    rng = np.random.default_rng(seed)
    cols = rng.integers(lo, hi + 1, nb)
    scale = np.exp(rng.standard_normal(nb))
    xs = [(rng.standard_normal((m, int(c))) * s_) for c, s_ in zip(cols, scale)]
    log(f"batched peers: generated {nb} blocks ({int(cols.sum())} rows) in {time.time() - t0:.1f} s")
    seeds = list(range(100, 100 + nb))
    B.local_stage(xs[:2], 99, seeds[:2])  # warm the kernels
    runs = []
    for _ in range(3):
        t = {}
        st = B.local_stage(xs, 99, seeds, timings=t)
        rt, cat = B.merge(st["masked"], m, timings=t)
        runs.append(t)
    base = min(B.per_peer_baseline(xs, 99, seeds) for _ in range(2))
    best = min(runs, key=lambda t: t["local_qr_s"] + t["mask_s"])
    # accuracy: R R^T against the masked blocks' Gram matrix, and each block's R against X X^T
    S = cat.to_host()
    R = rt.to_host()
    G = S @ S.T
    err_merge = float(np.max(np.abs(R @ R.T - G)) / np.abs(G).max())
    err_blocks = 0.0
    for b in np.flatnonzero(st["reduced"])[:8]:
        r = st["rt"][b].to_host()
        Gb = xs[b] @ xs[b].T
        err_blocks = max(err_blocks, float(np.max(np.abs(r @ r.T - Gb)) / np.abs(Gb).max()))
    flops = float(sum(2.0 * m * m * c for c in cols))  # Gram (half) + the FP64 solve, per pass
    out = {"config": {"workload": f"{nb} peers x {m} columns, rows uniform in [{lo}, {hi}] (sum {int(cols.sum())}), "
                                  "N(0,1) x lognormal(0,1) per peer; local QR + Givens masks of all peers at once, "
                                  "merged by the global QR", "peers": nb, "columns": m},
           "batched_local_s": round(best["local_qr_s"] + best["mask_s"], 6),
           "batched_local_qr_s": round(best["local_qr_s"], 6), "batched_mask_s": round(best["mask_s"], 6),
           "merge_s": round(best["merge_s"], 6), "per_peer_loop_s": round(base, 6),
           "speedup_vs_per_peer": round(base / (best["local_qr_s"] + best["mask_s"]), 3),
           "rel_err_merged_gram": err_merge, "rel_err_block_gram_max": err_blocks,
           "fp64_passes": [int(v) for v in st["stats"][:, 2]], "autotuner": json.loads(B.tuning())}
    log(f"batched peers: {out['batched_local_s'] * 1e3:.1f} ms batched vs {base * 1e3:.1f} ms per peer")
    return out


def bench_learned(n_train=10_000, steps=2000, seed=10, per=32):
    """Learned bisection brackets: 10,000 generated bidiagonals with n in [512, 4096] (four spectrum
    families), the network trained by the library's FP32 kernels, then every singular value of
    n = 16,384 bidiagonals of the four families bisected from its BF16 brackets and plainly."""
    from excalibur_amd import learned as Lm
    t0 = time.time()
    train = Lm.make_set(n_train, (512, 4096), seed=seed)
    X, y = Lm.training_data(train, per=per, seed=seed + 1)
    t_data = time.time() - t0
    net = Lm.BracketNet(seed=0)
    hist = net.train(X, y, steps=steps, batch=4096, lr=2e-3, seed=1)
    held = Lm.make_set(500, (512, 4096), seed=seed + 5)
    k_lo, k_hi = net.calibrate(held, *Lm.targets(held, per=32, seed=seed + 6))
    test = Lm.make_set(4, (16384, 16384), seed=seed + 10)
    fams = {}
    for b in range(test.nb):
        r = Lm.evaluate(net, test, b)
        fams[test.families[b]] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()
                                  if k not in ("sigma", "leaves")}
    saved = float(np.mean([f["saved_frac"] for f in fams.values()]))
    saved_shifts = float(np.mean([f["saved_frac_shifts_only"] for f in fams.values()]))
    out = {"config": {"workload": "bidiagonals with n = 16,384 from Golub-Kahan of matrices with geometric, linear, "
                                  "clustered and Marchenko-Pastur spectra; MLP 32-256-256-2, FP32 training on "
                                  f"{n_train} generated bidiagonals (n in [512, 4096]), BF16 inference"},
           "train_samples": int(X.shape[0]), "data_s": round(t_data, 3), "train_s": round(hist["seconds"], 3),
           "loss_first_last": [hist["loss"][0][1], hist["loss"][-1][1]], "families": fams,
           "bracket_ends_laplace_scales": [round(k_lo, 3), round(k_hi, 3)],
           "sturm_counts_saved_frac": round(saved, 4),
           "sturm_counts_saved_frac_shifts_only": round(saved_shifts, 4),
           "all_identical": all(f["identical"] for f in fams.values())}
    log(f"learned brackets: {100 * saved:.1f}% of Sturm counts saved, identical={out['all_identical']}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--form", default="mixed", choices=["fp64", "mixed"],
                    help="the form whose time is `value`; the other form is measured beside it")
    ap.add_argument("--device-exchange", type=int, default=0, choices=[0, 1],
                    help="1: peers on one device exchange data in device memory instead of host TCP")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the extra runs (the other form, device-memory exchange, the second workload)")
    ap.add_argument("--second-config", default="c3", help="workload measured in both forms beside the first")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 32 batched peers workload")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-bracket workload")
    ap.add_argument("--only", choices=["batched", "learned"], default=None,
                    help="run only that workload and print its JSON (for profiling it alone)")
    ap.add_argument("--learned-steps", type=int, default=2000, help="Adam steps of the bracket network")
    ap.add_argument("--learned-per", type=int, default=32, help="training targets per generated bidiagonal")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--workdir", default=None)
    ap.add_argument("--prepare", default=None,
                    help="only write the peers' Client{i}.mat files into this directory and exit")
    ap.add_argument("--isolated-only", action="store_true",
                    help="only time the largest products of the configuration alone on the GPU (for PMC runs)")
    ap.add_argument("--isolated-givens", action="store_true",
                    help="only time the FP64 Givens mask sweep (1000 x 1000, rows then columns) alone (for PMC runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB mebibytes device to device once (HBM traffic calibration for PMC runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: all peers of the workload share one MI355X; "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    ensure_built()
    if args.calibrate_copy:
        from excalibur_amd import hip
        nbytes = args.calibrate_copy << 20
        src = hip.DeviceArray((nbytes // 8,)).zero()
        dst = hip.DeviceArray((nbytes // 8,))
        hip.call("exc_memcpy_d2d", dst.ptr, src.ptr, nbytes, None)
        hip.sync()
        print(json.dumps({"copy_bytes": nbytes}))
        return
    if args.isolated_only:
        P, rows, n, _ = CONFIGS[args.config]
        print(json.dumps(isolated_gemms(n, rows)))
        return
    if args.isolated_givens:
        print(json.dumps(isolated_givens()))
        return
    if args.only:
        print(json.dumps(bench_batched() if args.only == "batched"
                         else bench_learned(steps=args.learned_steps, per=args.learned_per)))
        return
    from excalibur_amd.launch import write_peer_files

    if args.prepare:
        xs, _, _ = make_blocks(args.config)
        write_peer_files(xs, args.prepare)
        log(f"wrote {len(xs)} peer files ({xs[0].shape[0]} x {xs[0].shape[1]}) to {args.prepare}")
        return

    check = not args.no_accuracy
    t0 = time.time()
    xs, sigma_true, A = make_blocks(args.config)
    log(f"generated {args.config} in {time.time() - t0:.1f} s: {len(xs)} peers x {xs[0].shape[1]} x {xs[0].shape[0]}")
    main_run = measure(xs, sigma_true, A, args.form, args.device_exchange, args.steps, args.warmup, check,
                       args.workdir)
    other = "fp64" if args.form == "mixed" else "mixed"
    forms = {args.form: main_run}
    variants = {}
    P, rows, n, spec = CONFIGS[args.config]
    if not args.no_variants:
        forms[other] = measure(xs, sigma_true, A, other, args.device_exchange, args.steps, args.warmup, check)
        # the plain form with its exchanges in device memory (IPC), a deviation from host TCP
        dx = measure(xs, sigma_true, A, "fp64", 1 - args.device_exchange, args.steps, args.warmup, False)
        variants["fp64_" + ("device_exchange" if args.device_exchange == 0 else "host_tcp")] = {
            k: dx[k] for k in ("ms_per_step", "step_times_s", "stages_s")}
    try:
        main_run["roofline"]["isolated"] = isolated_gemms(n, rows)
    except Exception as e:  # extra information; the runs themselves already finished
        log(f"isolated GEMM timing failed: {e}")
    second = None
    if not args.no_variants and args.second_config and args.second_config != args.config:
        del xs, A
        t0 = time.time()
        xs2, s2, A2 = make_blocks(args.second_config)
        P2, rows2, n2, spec2 = CONFIGS[args.second_config]
        log(f"generated {args.second_config} in {time.time() - t0:.1f} s")
        second = {"config": {"workload": f"{P2} peers x {n2} x {rows2} " + WORKLOADS[spec2] + "; full SVD",
                             "peers": P2, "rows_per_peer": rows2, "columns": n2}}
        for f in ("fp64", "mixed"):
            second[f] = measure(xs2, s2, A2, f, args.device_exchange, max(1, args.steps // 2), 1, check)
        second["speedup_mixed_vs_plain"] = round(second["fp64"]["ms_per_step"] / second["mixed"]["ms_per_step"], 3)
        del xs2, A2
    batched = None
    if not args.no_variants and not args.no_batched:
        try:
            batched = bench_batched()
        except Exception as e:  # reported, not fatal: the headline runs already finished
            log(f"batched peers failed: {e}")
            batched = {"error": str(e)}
    learned = None
    if not args.no_variants and not args.no_learned:
        try:
            learned = bench_learned(steps=args.learned_steps, per=args.learned_per)
        except Exception as e:
            log(f"learned brackets failed: {e}")
            learned = {"error": str(e)}
    cpu = None
    if not args.no_cpu:
        t0 = time.time()
        cpu = cpu_baseline()
        log(f"cpu baseline in {time.time() - t0:.1f} s: {cpu['value']} s")
        if not args.no_variants:
            sx, _, _ = make_blocks("cpu_sample", seed=1)
            gs = run_variant(sx, args.form, args.device_exchange, 2, 1, False, None)
            cpu["gpu_same_sample_s"] = round(gs["ms_per_step"] / 1000.0, 6)
    roofline = dict(main_run["roofline"])
    roofline["traffic_final_v_plain"] = pmc_traffic()
    roofline["note"] = ("in-pipeline: both peer processes share the GPU, so a launch's duration includes the "
                        "other peer's concurrent work; 'isolated' is the same kernel alone")
    line = {
        "metric": "SVD wall-time (s) + % roofline, mixed vs plain FP64, at <=1e-12 rel sigma err",
        "value": round(main_run["ms_per_step"] / 1000.0, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_run["ms_per_step"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "fp64" else "f32+f64",
        "data": "synthetic",
        "form": args.form,
        "transport": "host TCP (loopback)" if args.device_exchange == 0 else "device memory (IPC)",
        "config": {"workload": f"{P} peers x {rows} x {n} " + WORKLOADS[spec] + "; full SVD, Givens masks "
                               "(plain: bisection + twisted, mode 1, opt 3)",
                   "peers": P, "rows_per_peer": rows, "columns": n},
        "step_times_s": main_run["step_times_s"],
        "stages_s": main_run["stages_s"],
        "results_d2h_s": main_run["results_d2h_s"],
        "accuracy": main_run.get("accuracy", {}),
        "roofline": roofline,
        "kernels": main_run["kernels"],
        "forms": {f: {k: v for k, v in r.items() if k != "kernels"} for f, r in forms.items() if f != args.form},
        "variants": variants,
        "second_workload": second,
        "batched_peers": batched,
        "learned_brackets": learned,
        "cpu_baseline": cpu,
        "reference_cpu_measured_s": 4.99,
    }
    if other in forms:
        line["speedup_mixed_vs_plain"] = round(forms["fp64"]["ms_per_step"] / forms["mixed"]["ms_per_step"], 3)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
