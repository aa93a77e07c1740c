#!/usr/bin/env python3
"""Benchmark of the per-peer bbtsvd pipeline on one MI355X.

Workload (default): 2 peers, each holding a 50,000 x 1,000 float64 row block of
A = Q1 diag(sigma) Q2^T (100,000 x 1,000), Q1, Q2 random orthogonal (QR of seeded Gaussians),
sigma_i = 10^(-6 i / 999); full SVD (all 1,000 triplets) with Givens masking, the reference's
mode 1 (bisection + twisted factorization), opt 3. The peers are two processes of
excalibur_amd/excalibur_peer sharing the GPU and talking over loopback TCP, exactly the
reference's DecentralizedClient::run() workflow; a step is one run() of both peers with the input
already resident in HBM, timed from "All Clients Connected" to "Finish" (max over the peers).

Prints one JSON line (metric, value = seconds per SVD, roofline of the dominant kernel, CPU
baseline). --gpus N > 1 exits: this build measures one GPU.
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
    "c1": (2, 500, 100, "gaussian"),
    "small": (2, 6_000, 200, "geometric6"),
    "cpu_sample": (2, 10_000, 1_000, "geometric6"),
}


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
        sigma = 10.0 ** (-6.0 * np.arange(n) / (n - 1))
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--form", default="fp64", choices=["fp64", "mixed"])
    ap.add_argument("--device-exchange", type=int, default=0, choices=[0, 1],
                    help="1: peers on one device exchange data in device memory instead of host TCP")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the extra runs (device-memory exchange, the other form, the CPU sample on the GPU)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--workdir", default=None)
    ap.add_argument("--prepare", default=None,
                    help="only write the peers' Client{i}.mat files into this directory and exit")
    ap.add_argument("--isolated-only", action="store_true",
                    help="only time the largest products of the configuration alone on the GPU (for PMC runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only copy MIB mebibytes device to device once (HBM traffic calibration for PMC runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every GPU box of this build has one MI355X and the peers share it; "
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
    from excalibur_amd.launch import write_peer_files

    if args.prepare:
        xs, _, _ = make_blocks(args.config)
        write_peer_files(xs, args.prepare)
        log(f"wrote {len(xs)} peer files ({xs[0].shape[0]} x {xs[0].shape[1]}) to {args.prepare}")
        return

    t0 = time.time()
    xs, sigma_true, A = make_blocks(args.config)
    log(f"generated {args.config} in {time.time() - t0:.1f} s: {len(xs)} peers x {xs[0].shape[1]} x {xs[0].shape[0]}")
    workdir = args.workdir or tempfile.mkdtemp(prefix="exc_bench_")
    t0 = time.time()
    main_run = run_variant(xs, args.form, args.device_exchange, args.steps, args.warmup, not args.no_accuracy,
                           workdir)
    log(f"peers finished in {time.time() - t0:.1f} s: {main_run['ms_per_step']:.2f} ms per SVD")
    step_times, ms_per_step, stages, agg = (main_run["step_times"], main_run["ms_per_step"], main_run["stages"],
                                            main_run["agg"])
    nsteps = main_run["steps"]

    def rate(rec):
        return rec["flops"] / (rec["ms"] * 1e-3) / 1e12 if rec["ms"] > 0 else 0.0
    gname = "gemm_f64"
    g = agg.get(gname, {"launches": 0, "ms": 0.0, "flops": 0.0, "bytes": 0.0})
    fv = agg.get("gemm_f64_final_v", {"launches": 0, "ms": 0.0, "flops": 0.0, "bytes": 0.0})
    achieved = rate(g)
    roofline = {"bound": "mfma", "kernel": "k_gemm<double> (all FP64 MFMA GEMM launches of both peers)",
                "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None,
                "traffic_final_v": pmc_traffic(),
                "launches_per_step": g["launches"] / max(1, nsteps),
                "avg_launch_us": round(1000.0 * g["ms"] / max(1, g["launches"]), 2),
                "flops_per_step": g["flops"] / max(1, nsteps),
                "final_v": {"launches_per_step": fv["launches"] / max(1, nsteps),
                            "avg_launch_us": round(1000.0 * fv["ms"] / max(1, fv["launches"]), 2),
                            "achieved": round(rate(fv), 3), "frac": round(rate(fv) / PEAK_FP64_TFLOPS, 4)},
                "note": "in-pipeline: both peer processes share the GPU, so a launch's duration includes "
                        "the other peer's concurrent work; 'isolated' is the same kernel alone"}
    kernels = {k: {"launches": v["launches"] / nsteps, "ms": v["ms"] / nsteps,
                   "achieved": (v["flops"] / (v["ms"] * 1e-3) / 1e12) if v["ms"] > 0 else 0.0,
                   "gbs": (v["bytes"] / (v["ms"] * 1e-3) / 1e9) if v["ms"] > 0 else 0.0}
               for k, v in agg.items()}
    acc = {}
    if not args.no_accuracy:
        t0 = time.time()
        acc = accuracy(main_run["res"]["peers"], xs, sigma_true, A)
        log(f"accuracy checked in {time.time() - t0:.1f} s: {acc}")
    P, rows, n, _ = CONFIGS[args.config]
    try:
        roofline["isolated"] = isolated_gemms(n, rows)
    except Exception as e:  # the measurement is extra information; the run itself already finished
        log(f"isolated GEMM timing failed: {e}")
    variants = {}
    if not args.no_variants:
        # the same workload with the exchanges in device memory (IPC), a deviation from host TCP
        t0 = time.time()
        dx = run_variant(xs, args.form, 1 - args.device_exchange, args.steps, args.warmup, False, None)
        variants["device_exchange" if args.device_exchange == 0 else "host_tcp"] = {
            "ms_per_step": round(dx["ms_per_step"], 3), "step_times_s": [round(s, 6) for s in dx["step_times"]],
            "stages_s": dx["stages"]}
        log(f"device-exchange variant in {time.time() - t0:.1f} s: {dx['ms_per_step']:.2f} ms")
    cpu = None
    if not args.no_cpu:
        t0 = time.time()
        cpu = cpu_baseline()
        log(f"cpu baseline in {time.time() - t0:.1f} s: {cpu['value']} s")
        if not args.no_variants:
            sx, _, _ = make_blocks("cpu_sample", seed=1)
            gs = run_variant(sx, args.form, args.device_exchange, 2, 1, False, None)
            cpu["gpu_same_sample_s"] = round(gs["ms_per_step"] / 1000.0, 6)
    line = {
        "metric": "SVD wall-time (s) + % roofline, mixed vs plain FP64, at <=1e-12 rel sigma err",
        "value": round(ms_per_step / 1000.0, 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.form == "fp64" else "f32+f64",
        "data": "synthetic",
        "form": args.form,
        "transport": "host TCP (loopback)" if args.device_exchange == 0 else "device memory (IPC)",
        "config": {"workload": f"{P} peers x {rows} x {n} float64 row blocks of A = Q1 diag(sigma) Q2^T, "
                               f"sigma_i = 10^(-6i/{n - 1}); full SVD, Givens masks, bisection + twisted "
                               f"(mode 1, opt 3)" if args.config != "c1" else
                               f"{P} peers x {rows} x {n} N(0,1) row blocks, full SVD",
                   "peers": P, "rows_per_peer": rows, "columns": n},
        "step_times_s": [round(s, 6) for s in step_times],
        "stages_s": stages,
        "results_d2h_s": round(main_run["results_d2h_s"], 6),
        "accuracy": acc,
        "roofline": roofline,
        "kernels": kernels,
        "variants": variants,
        "cpu_baseline": cpu,
        "reference_cpu_measured_s": 4.99,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
