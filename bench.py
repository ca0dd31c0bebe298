"""HBDIA SpMV benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|ref128] [--no-cpu]

Default workload (c2): FP64 27-point stencil on a 256^3 grid (16.8M rows,
449.5M stencil non-zeros; off-diagonals U[-1,0), diagonal 1 + sum |off|) plus
floor(0.005 * 27 * N) = 2,264,924 random off-stencil entries that land in the
COO fallback; block width 64, threshold 2; x ~ U[-1,1).  One step is one
hbdiaSpMV(A, x, y, execCOOCPU=false, execCOOGPU=true) with A, x and y resident
in HBM.

ref128: the reference's published configuration -- create3DStencil27Point(128,
128, 128, 0.0), block width 32, threshold 8, x[i] = i.

ranklocal: rank 3 of an 8-way row partition of the c2 matrix at its real shape
(2,097,152 rows), halo columns in the COO part, x's halo filled on the host.

Prints one JSON line: GFLOP/s of the whole job, the roofline of the dominant
kernel (algorithmic bytes per launch / average launch time from HIP events
on the kernel's stream, against 8 TB/s), and the CPU baseline (the oracle's
OpenMP restatement of the reference path on the host cores, same matrix, same
x).  Before timing, y is checked bit-for-bit against that CPU result.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# The timed CPU baseline runs with the reference's thread placement; libgomp reads this when the
# native libraries load, which happens after this line.
os.environ.setdefault("OMP_PROC_BIND", "spread")

HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = ROOT / "profiles" / "traffic_hbdia_spmv_stencil256.json"


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--workload", choices=["c2", "ref128", "precision", "multirhs", "batch", "serve", "copy", "convert",
                                          "ranklocal"], default="c2")
    p.add_argument("--count", type=int, default=None, help="override the number of batched matrices (testing only)")
    p.add_argument("--grid", type=int, default=None, help="override the grid edge (testing only)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the full-size parity check")
    p.add_argument("--dtype", choices=["f64", "f32"], default="f64",
                   help="value, x and y type of the c2, ref128 and multirhs workloads (HBDIA<double> or HBDIA<float>)")
    p.add_argument("--no-forms", action="store_true",
                   help="default workload only: skip the forms (ref128, precision, multirhs, batch, serve)")
    return p.parse_args()


class BenchError(Exception):
    """A parity or consistency failure inside a workload; its argument is the JSON-able record."""


def _emit(res) -> int:
    if isinstance(res, int):
        return res
    print(json.dumps(res), flush=True)
    return 0


def main() -> int:
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2
    legs = {"precision": bench_precision, "multirhs": bench_multirhs, "batch": bench_batch, "serve": bench_serve,
            "copy": bench_copy, "convert": bench_convert, "ranklocal": bench_ranklocal}
    try:
        if args.workload in legs:
            return _emit(legs[args.workload](args))
        line = bench_spmv(args)
    except BenchError as e:
        print(json.dumps(e.args[0]), flush=True)
        return 4
    if isinstance(line, int):
        return line
    if args.workload == "c2" and not args.no_forms and args.dtype == "f64" and args.grid is None:
        line["forms"] = bench_forms(args)
    print(json.dumps(line), flush=True)
    return 0


SETTLE_S = 0.25


def _settle(fn, sync, seconds: float = SETTLE_S) -> int:
    """Run `fn` back to back for `seconds` (untimed) before a measurement that follows host-only
    work.  The part lowers its clocks while the GPU idles (the CPU baseline, the oracle check):
    in one process, 20 back-to-back c2 launches issued after 1 s of idle plus 5 warm-ups averaged
    0.6195 ms, and 0.591 ms once the GPU had been busy for a few tens of milliseconds.  The
    settle loop makes the timed figures the steady-state rate of back-to-back SpMVs rather than
    the clock ramp.  Returns the number of calls made."""
    calls = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(8):
            fn()
        sync()
        calls += 8
    return calls


def _reference_method(A, X, Y, gpu_coo: bool, iterations: int = 100) -> dict:
    """The reference's own timing (benchmarking/benchHBDIA.cpp:30-50): per iteration y cleared, device
    synchronized, hbdiaSpMV + device synchronize timed on the host clock; the first iteration dropped and the mean
    of the rest taken, as its plots do (benchmarking/Plotting/plot_hbdia_vs_cusparse_runtime.py:30-32,
    :59-61)."""
    from hbdia_amd.api import spmv_time_synced
    t = spmv_time_synced(A, X, Y, False, gpu_coo, iterations)[1:]
    return {"iterations": iterations, "dropped": 1, "mean_ms": round(float(t.mean()), 5),
            "median_ms": round(float(np.median(t)), 5), "min_ms": round(float(t.min()), 5),
            "max_ms": round(float(t.max()), 5)}


def _cpu_baseline(hb: dict, x: np.ndarray, N: int, flop: float, dt, what: str) -> dict:
    """The reference path on the host cores: the oracle's restatement (bdia_spmv_kernel per row with
    OpenMP, cpu_coo_spmv_kernel, y += coo) compiled with the reference's host flags on this machine,
    OMP_PROC_BIND=spread, 2 warm-ups then the median of 10, on the same matrix and x."""
    import oracle
    try:
        base = oracle.CpuBaseline()
        build = f"{oracle.CpuBaseline.FLAGS} (reference CMakeLists.txt:39), built on this host"
        run = base.hbdia_spmv
        cores = base.max_threads()
    except RuntimeError as e:  # no compiler on this host: the checker's build, said so
        log(f"cpu baseline build failed ({e}); timing the oracle's own build")
        build = "-O3 -march=x86-64-v3 -ffp-contract=off -fopenmp (oracle build; the reference flags failed to build)"
        run = lambda hb_, x_, y_, n_, c_: oracle.hbdia_spmv(hb_, x_, y_, n_, c_)  # noqa: E731
        cores = oracle.max_threads()
    y = np.zeros(N, dt)
    coo_buf = np.zeros(N, dt)
    times = []
    for rep in range(12):
        y[:] = 0.0
        ts = time.perf_counter()
        run(hb, x, y, N, coo_buf)
        te = time.perf_counter()
        if rep >= 2:
            times.append(te - ts)
    t_cpu = float(np.median(times))
    return {"value": round(flop / t_cpu / 1e9, 3), "unit": "GFLOP/s", "cores": cores, "kind": "port",
            "ms_per_spmv": round(t_cpu * 1e3, 3), "build": build,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND", ""),
            "sample": f"{what}: median of 10 SpMVs after 2 warm-ups (oracle restatement of bdia_spmv_kernel + "
                      f"cpu_coo_spmv_kernel + add, OpenMP)"}


def _traffic(kernel_ms: float) -> int | None:
    """HBM bytes per launch of the c2 kernel from the committed rocprofv3 counter runs, used only while
    the kernel source is the one those runs measured (its hash is recorded beside the figure)."""
    import hashlib
    try:
        rec = json.loads(PROFILE_TRAFFIC.read_text())
    except Exception:
        return None
    src = ROOT / "hbdia_amd" / "csrc" / "hip" / "hbdia_hip.hip"
    digest = hashlib.sha256(src.read_bytes()).hexdigest()[:16]
    if rec.get("source_sha256_16") != digest:
        log(f"traffic figure not used: profiled source {rec.get('source_sha256_16')} != current {digest}")
        return None
    return rec.get("hbm_bytes_per_launch")


def bench_spmv(args) -> dict | int:
    """The drop-in product: c2 (the headline) or ref128 (the reference's published configuration)."""
    import hbdia_amd
    from hbdia_amd import HBDIA, Event, HBDIAVector, gen_vector, hbdiaSpMV
    from hbdia_amd._native import check_hip, hip

    if not hbdia_amd.gpu_available():
        print("no HIP device visible", file=sys.stderr)
        return 3
    log(f"device {hbdia_amd.device_name()}, host threads {hbdia_amd.api.capi.hbdia_get_num_threads()}")

    t0 = time.time()
    dt = np.float32 if args.dtype == "f32" else np.float64
    s_v = np.dtype(dt).itemsize
    if args.workload == "c2":
        n = args.grid or 256
        bw, thr, seed = 64, 2, 0xC2
        A = HBDIA.stencil(n, 27, 0.005, seed=seed, dtype=dt)
        log(f"generated {n}^3 stencil + scatter in {time.time() - t0:.1f}s")
        N = n ** 3
        x = gen_vector(N, seed + 1).astype(dt)
        workload = {"workload": f"stencil27_scatter_{n}^3", "grid": n, "rows": N, "scatter": 0.005,
                    "block_width": bw, "threshold": thr, "seed": seed, "coo_exec": "gpu_fused"}
    else:
        n = args.grid or 128
        bw, thr = 32, 8
        A = HBDIA(dtype=dt)
        A.create3DStencil27Point(n, n, n, 0.0)
        log(f"generated {n}^3 reference stencil in {time.time() - t0:.1f}s")
        N = n ** 3
        x = np.arange(N, dtype=dt)
        workload = {"workload": f"reference_stencil27_{n}^3", "grid": n, "rows": N, "noise": 0.0,
                    "block_width": bw, "threshold": thr, "coo_exec": "gpu_fused"}
    t1 = time.time()
    A.convertToHBDIAFormat(bw, thr)
    info = A.info()
    log(f"converted + uploaded in {time.time() - t1:.1f}s: {info['num_values']} stored values, "
        f"{info['num_offsets']} (block, offset) pairs, {info['num_coo']} fallback entries")

    nnz = info["nnz"]
    nnz_coo = info["num_coo"]
    nnz_bdia = nnz - nnz_coo
    flop = 2.0 * nnz
    alg_bytes = nnz_bdia * s_v + nnz_coo * (s_v + 4 + 4) + N * s_v + N * s_v
    stored_bytes = info["stored_bytes"] + 2 * N * s_v

    X = HBDIAVector(x)
    Y = HBDIAVector(np.zeros(N, dt))
    gpu_coo = nnz_coo > 0

    def step():
        hbdiaSpMV(A, X, Y, False, gpu_coo)

    step()
    A.synchronize()
    y_gpu = Y.download()

    cpu_baseline = None
    parity = None
    if not args.no_cpu:
        import oracle  # the checker and the CPU baseline only
        hb = A.getHBDIA()
        hb["num_blocks"], hb["block_width"] = A.getNumBlocks(), bw
        y_cpu = oracle.hbdia_spmv(hb, x, np.zeros(N, dt), N, np.zeros(N, dt))
        diff = y_gpu != y_cpu
        rel = np.abs(y_gpu - y_cpu) / np.maximum(np.abs(y_cpu), 1e-300)
        parity = {"rows": N, "rows_differing": int(diff.sum()), "max_rel_err": float(rel.max()),
                  "bit_exact": bool(not diff.any())}
        log(f"parity vs oracle at full size: {parity}")
        if rel.max() > 1e-12:
            raise BenchError({"error": "parity failure", "parity": parity})
        cpu_baseline = _cpu_baseline(hb, x, N, flop, dt, "full matrix")
        log(f"cpu baseline: {cpu_baseline}")
        del hb

    stream = A.device_ptrs()["bdia_stream"]
    ev0, ev1 = Event(), Event()
    # settle, then the W warm-ups and the K back-to-back steps that `value` and the roofline are
    # measured on; the reference's own method (each call alone after a synchronize, which lets the
    # clocks sag between calls) runs after them
    settle_calls = _settle(step, A.synchronize)
    for _ in range(args.warmup):
        step()
    A.synchronize()
    hbdia_amd.synchronize()
    ts = time.perf_counter()
    check_hip(hip.hbdia_event_record(ev0.handle, stream))
    for _ in range(args.steps):
        step()
    check_hip(hip.hbdia_event_record(ev1.handle, stream))
    A.synchronize()
    hbdia_amd.synchronize()
    te = time.perf_counter()
    wall_ms = (te - ts) * 1e3 / args.steps
    kern_ms = ev0.elapsed_ms(ev1) / args.steps
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    log(f"{args.steps} steps: {wall_ms:.4f} ms/step wall, {kern_ms:.4f} ms/launch on the stream, "
        f"{achieved:.1f} GB/s algorithmic")
    refm = _reference_method(A, X, Y, gpu_coo)
    log(f"reference timing method (synchronized per call, 100 rounds, first dropped): {refm}")

    traffic = None
    if args.workload == "c2" and (args.grid or 256) == 256 and args.dtype == "f64":
        traffic = _traffic(kern_ms)

    value = flop / (wall_ms * 1e-3) / 1e9
    vs = None
    if args.workload == "ref128" and (args.grid or 128) == 128 and args.dtype == "f64":
        # the reference's figure, 0.40 ms on a GH200-class GPU, is its own method's mean
        vs = round(0.40 / refm["mean_ms"], 3)
    return {
        "metric": f"SpMV GFLOP/s ({args.dtype.upper().replace('F', 'FP')} hbdiaSpMV) + % HBM roofline",
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": vs,
        "dtype": args.dtype,
        "data": "synthetic (seeded generator)",
        "config": workload,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu_baseline,
        "kernel_ms": round(kern_ms, 5),
        "settle": {"seconds": SETTLE_S, "calls": settle_calls},
        "reference_method": refm,
        "algorithmic_bytes": int(alg_bytes),
        "stored_bytes": int(stored_bytes),
        "flop": int(flop),
        "nnz_bdia": int(nnz_bdia),
        "nnz_coo": int(nnz_coo),
        "parity": parity,
        "setup_s": round(time.time() - t0, 1),
    }


def bench_forms(args) -> dict:
    """The other halves of the metric, each at its configured workload, run after the headline in
    the default invocation: the reference's configuration with its own timing method, the
    reduced-precision, multi-RHS and batched forms (each with its algorithmic bytes and roofline
    fraction) and the served SpMVs/s with p50/p99 latency.  A leg that fails records its error."""
    import copy
    out = {}

    def leg(name, fn, steps, warmup=3, **over):
        a = copy.copy(args)
        a.steps, a.warmup, a.grid, a.count = steps, warmup, None, None
        for k, v in over.items():
            setattr(a, k, v)
        t = time.time()
        try:
            r = fn(a)
        except BenchError as e:
            r = {"error": e.args[0]}
        except Exception as e:  # keep the headline line whatever a form does
            r = {"error": f"{type(e).__name__}: {e}"}
        log(f"form {name} done in {time.time() - t:.1f}s")
        return r

    r = leg("ref128", bench_spmv, 50, 5, workload="ref128")
    if isinstance(r, dict) and "error" not in r:
        out["ref128"] = {"workload": r["config"]["workload"], "kernel_ms": r["kernel_ms"],
                         "ms_per_step_pipelined": r["ms_per_step"], "reference_method": r["reference_method"],
                         "vs_reference_0.40ms": r["vs_baseline"], "gflops": r["value"],
                         "algorithmic_bytes": r["algorithmic_bytes"], "roofline_frac": r["roofline"]["frac"],
                         "parity": r["parity"], "cpu_baseline": r["cpu_baseline"]}
    else:
        out["ref128"] = r
    r = leg("precision", bench_precision, 20)
    out["precision"] = r.get("forms", r) if isinstance(r, dict) else r
    if isinstance(r, dict) and "config" in r:
        out["precision"] = {"config": r["config"], **r["forms"]}
    r = leg("multirhs", bench_multirhs, 20)
    out["multirhs"] = {"config": r["config"], "by_k": r["by_k"]} if isinstance(r, dict) and "by_k" in r else r
    r = leg("batch", bench_batch, 20)
    out["batch"] = {"config": r["config"], **r["forms"]} if isinstance(r, dict) and "forms" in r else r
    r = leg("serve", bench_serve, 0)
    out["serve"] = {"config": r["config"], "results": r["results"]} if isinstance(r, dict) and "results" in r else r
    return out


def _time_launches(fn, sync, stream, steps: int, warmup: int):
    """Average wall time and stream-event time (ms) of `fn` over `steps` calls."""
    import hbdia_amd
    from hbdia_amd import Event
    from hbdia_amd._native import check_hip, hip
    _settle(fn, sync)
    for _ in range(warmup):
        fn()
    sync()
    ev0, ev1 = Event(), Event()
    ts = time.perf_counter()
    check_hip(hip.hbdia_event_record(ev0.handle, stream))
    for _ in range(steps):
        fn()
    check_hip(hip.hbdia_event_record(ev1.handle, stream))
    sync()
    hbdia_amd.synchronize()
    te = time.perf_counter()
    return (te - ts) * 1e3 / steps, ev0.elapsed_ms(ev1) / steps


def bench_precision(args) -> dict:
    """FP64 / FP32 / BF16 value storage on the 192^3 stencil + scatter matrix (x, y FP64).

    Reports, for each storage form, the time per SpMV, algorithmic GB/s (s_v = 8, 4, 2),
    the rows recomputed in FP64 and the max / RMS relative error per row against the FP64
    result, which is the oracle's bit for bit (checked)."""
    import oracle
    from hbdia_amd import HBDIA, HBDIAMixed, HBDIAVector, gen_vector, hbdiaSpMV
    n = args.grid or 192
    seed = 0xC3
    t0 = time.time()
    A = HBDIA.stencil(n, 27, 0.005, seed=seed)
    A.convertToHBDIAFormat(64, 2)
    info = A.info()
    N = n ** 3
    x = gen_vector(N, seed + 1)
    log(f"{n}^3 matrix ready in {time.time() - t0:.1f}s")
    X, Y = HBDIAVector(x), HBDIAVector(np.zeros(N))
    stream = A.device_ptrs()["bdia_stream"]
    nnz, ncoo = info["nnz"], info["num_coo"]
    flop = 2.0 * nnz

    def bytes_for(sv):
        return (nnz - ncoo) * sv + ncoo * (sv + 8) + 2 * N * 8

    hbdiaSpMV(A, X, Y, False, True)
    A.synchronize()
    y64 = Y.download()
    hb = A.getHBDIA()
    hb["num_blocks"], hb["block_width"] = A.getNumBlocks(), 64
    ycpu = oracle.hbdia_spmv(hb, x, np.zeros(N), N)
    if not np.array_equal(ycpu, y64):
        raise BenchError({"error": "FP64 parity failure"})
    del hb
    forms = {}
    wall, kern = _time_launches(lambda: hbdiaSpMV(A, X, Y, False, True), A.synchronize, stream, args.steps, args.warmup)
    forms["fp64"] = {"ms": round(kern, 5), "gbs": round(bytes_for(8) / (kern * 1e-3) / 1e9, 1),
                     "gflops": round(flop / (kern * 1e-3) / 1e9, 1), "algorithmic_bytes": int(bytes_for(8)),
                     "roofline_frac": round(bytes_for(8) / (kern * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "max_rel_err": 0.0, "rms_rel_err": 0.0, "fp64_rows": N}
    denom = np.maximum(np.abs(y64), 1e-300)
    for prec, sv, ratio in (("fp32", 4, 1e4), ("bf16", 2, 1e4), ("bf16", 2, 64.0), ("bf16", 2, 8.0)):
        M = HBDIAMixed(A, prec, ratio)
        M.spmv(X, Y)
        A.synchronize()
        y = Y.download()
        rel = np.abs(y - y64) / denom
        nfix = M.fp64_rows()
        wall, kern = _time_launches(lambda: M.spmv(X, Y), A.synchronize, stream, args.steps, args.warmup)
        key = f"{prec}_ratio{ratio:g}"
        forms[key] = {"ms": round(kern, 5), "gbs": round(bytes_for(sv) / (kern * 1e-3) / 1e9, 1),
                      "gflops": round(flop / (kern * 1e-3) / 1e9, 1), "algorithmic_bytes": int(bytes_for(sv)),
                      "roofline_frac": round(bytes_for(sv) / (kern * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                      "cancellation_ratio": ratio, "guaranteed_max_rel_err": M.rel_error_bound,
                      "max_rel_err": float(rel.max()), "rms_rel_err": float(np.sqrt(np.mean(rel ** 2))),
                      "p99_rel_err": float(np.quantile(rel, 0.99)), "median_rel_err": float(np.median(rel)),
                      "fp64_rows": nfix, "speedup_vs_fp64": round(forms["fp64"]["ms"] / kern, 3)}
        log(f"{key}: {forms[key]}")
        M.close()
    best = min(forms, key=lambda k: forms[k]["ms"])
    line = {"metric": "SpMV GFLOP/s by value storage precision (x, y FP64) + max/RMS rel. error vs FP64 oracle",
            "value": forms[best]["gflops"], "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": forms[best]["ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": best, "data": "synthetic (seeded generator)",
            "config": {"workload": f"stencil27_scatter_{n}^3_precision", "grid": n, "rows": N, "scatter": 0.005,
                       "block_width": 64, "threshold": 2, "cancellation_ratio": 1e4, "accumulate": "fp32"},
            "forms": forms}
    return line


def bench_multirhs(args) -> dict:
    """k right-hand sides in one launch vs k single-vector launches, 128^3 27-point stencil."""
    from hbdia_amd import HBDIA, DeviceArray, HBDIAVector, gen_vector, hbdiaSpMV, hbdiaSpMVMulti, multi_ld
    n = args.grid or 128
    dt = np.float32 if args.dtype == "f32" else np.float64
    s_v = np.dtype(dt).itemsize
    A = HBDIA.stencil(n, 27, 0.0, seed=0xC4, dtype=dt)
    A.convertToHBDIAFormat(64, 2)
    info = A.info()
    N = n ** 3
    nnz = info["nnz"]
    stream = A.device_ptrs()["bdia_stream"]
    X1, Y1 = HBDIAVector(gen_vector(N, 1).astype(dt)), HBDIAVector(np.zeros(N, dt))
    _, t1 = _time_launches(lambda: hbdiaSpMV(A, X1, Y1, False, True), A.synchronize, stream, args.steps, args.warmup)
    rows = {}
    # column-major X and Y with a leading dimension one 128-byte line past N (N * 8 bytes is a power
    # of two)
    ld = multi_ld(N, dt)
    for k in (1, 4, 8, 16, 32):
        Xh = np.concatenate([gen_vector(N, 10 + j) for j in range(k)]).astype(dt)
        Xp = np.zeros((k, ld), dt)
        Xp[:, :N] = Xh.reshape(k, N)
        X, Y = DeviceArray.from_host(Xp.ravel()), DeviceArray(k * ld, dt, zero=True)
        hbdiaSpMVMulti(A, X, Y, k, ldx=ld, ldy=ld)
        A.synchronize()
        Yc = Y.download().reshape(k, ld)[:, :N].ravel()
        # column k-1 must equal the single-vector result
        Xs = HBDIAVector(Xh[(k - 1) * N:])
        Ys = HBDIAVector(np.zeros(N, dt))
        hbdiaSpMV(A, Xs, Ys, False, True)
        A.synchronize()
        if not np.array_equal(Yc[(k - 1) * N:], Ys.download()):
            raise BenchError({"error": f"multi-RHS column mismatch at k={k}"})
        _, tk = _time_launches(lambda: hbdiaSpMVMulti(A, X, Y, k, ldx=ld, ldy=ld), A.synchronize, stream,
                               args.steps, args.warmup)
        b = nnz * s_v + k * N * 2 * s_v
        # the same columns row-major (row r's k values adjacent): every column checked against the
        # column-major result, which the single-vector check above ties to hbdiaSpMV
        Xr = DeviceArray.from_host(np.ascontiguousarray(Xh.reshape(k, N).T).ravel())
        Yr = DeviceArray(k * N, dt, zero=True)
        hbdiaSpMVMulti(A, Xr, Yr, k, layout="row")
        A.synchronize()
        if not np.array_equal(Yr.download().reshape(N, k).T.ravel(), Yc):
            raise BenchError({"error": f"row-major multi-RHS differs from column-major at k={k}"})
        _, tr = _time_launches(lambda: hbdiaSpMVMulti(A, Xr, Yr, k, layout="row"), A.synchronize, stream,
                               args.steps, args.warmup)
        best = min(tk, tr)
        rows[k] = {"ms": round(best, 5), "layout": "row" if tr <= tk else "col",
                   "gflops": round(2.0 * nnz * k / (best * 1e-3) / 1e9, 1),
                   "gbs": round(b / (best * 1e-3) / 1e9, 1), "algorithmic_bytes": int(b),
                   "roofline_frac": round(b / (best * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                   "col_major_ms": round(tk, 5), "row_major_ms": round(tr, 5),
                   "speedup_vs_k_single": round(k * t1 / best, 3)}
        log(f"k={k}: {rows[k]}")
    line = {"metric": f"multi-RHS SpMV GFLOP/s ({args.dtype.upper().replace('F', 'FP')}), k right-hand sides in one launch", "value": rows[32]["gflops"],
            "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": rows[32]["ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.dtype,
            "data": "synthetic (seeded generator)",
            "config": {"workload": f"stencil27_{n}^3_multirhs", "grid": n, "rows": N, "block_width": 64,
                       "ld": ld, "single_vector_ms": round(t1, 5)},
            "by_k": rows}
    return line


def bench_batch(args) -> dict:
    """2048 independent banded matrices (rows log-uniform in [1e3, 1e5], 3-9 diagonal offsets in
    [-64, 64], 1% off-band entries; diagonally dominant) in one persistent launch replayed from a
    HIP graph, with block width and tile schedule chosen by the autotuner, against the plain form:
    the same matrices one single-matrix launch each.  FP64 and FP32; the FP32 result's max / RMS
    relative error per row is taken against the FP64 result, which is checked bit for bit against
    the per-matrix launches and, on a sample of matrices, against the oracle."""
    import oracle
    from hbdia_amd import HBDIA, DeviceArray, HBDIABatch, gen_vector
    count = args.count or 2048
    seed = 0xC4B
    forms = {}
    y64 = None
    x64 = None
    for dt, name in ((np.float64, "fp64"), (np.float32, "fp32")):
        t0 = time.time()
        ms = HBDIA.banded_batch(count=count, seed=seed, dtype=dt)
        log(f"{name}: {count} matrices generated in {time.time() - t0:.1f}s")
        t1 = time.time()
        B = HBDIABatch.autotune(ms, block_widths=(64, 128, 256), threshold=2, reps=10)
        info = B.info()
        best = min(B.table, key=lambda r: r["ms"])
        worst = max(B.table, key=lambda r: r["ms"])
        log(f"{name}: autotuned in {time.time() - t1:.1f}s over {len(B.table)} candidates: chose block width "
            f"{ms[0].getBlockWidth()}, kind {info['kind']}, tiles_per_step {info['tiles_per_step']}, workgroups "
            f"{info['workgroups']}; fastest first timing {best}, slowest {worst['ms']:.4f} ms")
        N = B.rows
        if x64 is None:
            x64 = gen_vector(N, seed + 1)
        x = x64.astype(dt)
        X = DeviceArray.from_host(x)
        Y = DeviceArray(N, dt, zero=True)
        Y2 = DeviceArray(N, dt, zero=True)
        B.spmv(X, Y)
        B.spmv_per_matrix(X, Y2)
        B.synchronize()
        y = Y.download()
        if not np.array_equal(y, Y2.download()):
            raise BenchError({"error": f"{name}: batch differs from per-matrix launches"})
        off = B.row_offsets
        for i in np.linspace(0, count - 1, 8).astype(int):
            A = ms[i]
            hb = A.getHBDIA()
            hb["num_blocks"], hb["block_width"] = A.getNumBlocks(), A.getBlockWidth()
            n_i = A.getNumRows()
            want = oracle.hbdia_spmv(hb, x[off[i]:off[i + 1]].copy(), np.zeros(n_i, dt), n_i)
            if not np.array_equal(y[off[i]:off[i + 1]], want):
                raise BenchError({"error": f"{name}: matrix {i} differs from the oracle"})
        B.capture_graph(X, Y)
        stream = B.stream_handle
        _, kern = _time_launches(B.launch_graph, B.synchronize, stream, args.steps, args.warmup)
        _, plain = _time_launches(lambda: B.spmv_per_matrix(X, Y2), B.synchronize, stream, args.steps, args.warmup)
        abytes, flop = info["algorithmic_bytes"], info["flops"]
        f = {"ms": round(kern, 5), "gflops": round(flop / (kern * 1e-3) / 1e9, 1),
             "gbs": round(abytes / (kern * 1e-3) / 1e9, 1), "roofline_frac": round(abytes / (kern * 1e-3) / 1e9 /
                                                                                   HBM_PEAK_GBS, 4),
             "per_matrix_launches_ms": round(plain, 5), "speedup_vs_per_matrix": round(plain / kern, 3),
             "rows": N, "nnz_bdia": info["bdia_nnz"], "nnz_coo": info["coo_nnz"], "algorithmic_bytes": abytes,
             "stored_bytes": info["stored_bytes"], "block_width": ms[0].getBlockWidth(),
             "schedule": {"kind": ["interleaved", "contiguous", "xcd"][info["kind"]],
                          "tiles_per_step": info["tiles_per_step"], "workgroups": info["workgroups"]},
             "candidates": len(B.table), "worst_candidate_ms": round(worst["ms"], 5), "parity": "bit-exact"}
        if name == "fp64":
            y64 = y
            forms["fp64"] = f
        else:
            denom = np.maximum(np.abs(y64), 1e-300)
            rel = np.abs(y.astype(np.float64) - y64) / denom
            f.update({"max_rel_err_vs_fp64": float(rel.max()), "rms_rel_err_vs_fp64": float(np.sqrt(np.mean(rel ** 2))),
                      "p99_rel_err_vs_fp64": float(np.quantile(rel, 0.99)),
                      "speedup_vs_fp64_batch": round(forms["fp64"]["ms"] / kern, 3),
                      "speedup_vs_fp64_per_matrix": round(forms["fp64"]["per_matrix_launches_ms"] / kern, 3)})
            forms["fp32"] = f
        log(f"{name}: {f}")
        B.close()
        del ms
    line = {"metric": "batched SpMV GFLOP/s (2048 banded matrices, one persistent launch in a HIP graph)",
            "value": forms["fp64"]["gflops"], "unit": "GFLOP/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": forms["fp64"]["ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded generator)",
            "config": {"workload": f"banded_batch_{count}", "matrices": count, "rows": "log-uniform [1e3, 1e5]",
                       "offsets": "3-9 in [-64, 64]", "offband": 0.01, "threshold": 2, "seed": seed},
            "roofline": {"bound": "hbm", "achieved": forms["fp64"]["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": forms["fp64"]["roofline_frac"], "traffic": None},
            "forms": forms}
    return line


def _serve_pool(count: int, seed: int):
    """64 resident matrices: 7- and 27-point stencils (0.5% scatter) on grids with edges log-uniform in
    [32, 160], and banded matrices (rows log-uniform in [1e3, 1e5], 3-9 offsets in [-64, 64], 1% off-band)."""
    from hbdia_amd import HBDIA
    rng = np.random.default_rng(seed)
    ms, desc = [], []
    for i in range(count):
        kind = i % 3
        if kind < 2:
            edge = int(round(np.exp(rng.uniform(np.log(32), np.log(160)))))
            pts = 7 if kind == 0 else 27
            ms.append(HBDIA.stencil(edge, pts, 0.005, seed=seed + i))
            desc.append(f"{pts}pt-{edge}^3")
        else:
            n = int(np.exp(rng.uniform(np.log(1e3), np.log(1e5))))
            k = int(rng.integers(3, 10))
            ms.append(HBDIA.banded(n, k, 64, 0.01, seed=seed + i))
            desc.append(f"banded-{n}x{k}")
    for m in ms:
        m.convertToHBDIAFormatOnHost(64, 2)
    return ms, desc


def bench_serve(args) -> dict:
    """Serving: a continuous-batching request loop over 64 matrices resident in HBM (see _serve_pool).
    Each request brings a fresh x (two seeded U[-1,1] vectors per matrix, pinned host memory or HBM) and
    gets y = A x.  Poisson arrivals at 30%, 60% and 90% of the server's measured capacity (a burst of
    requests); p50 / p99 latency from submission to y in the client's buffer, and served SpMVs/s.  The
    plain form -- one request at a time through the single-matrix kernel -- runs the same arrivals.
    Results are checked bit for bit against the oracle on one request per matrix first."""
    import oracle
    from hbdia_amd import DeviceArray, HBDIAServer, gen_vector
    from hbdia_amd.device import PinnedArray
    count = args.count or 64
    seed = 0xC5
    t0 = time.time()
    ms, desc = _serve_pool(count, seed)
    rows = [m.getNumRows() for m in ms]
    # algorithmic bytes of one request (uniformly random matrix): values, fallback entries with their
    # indices, x read and y written, FP64
    req_bytes = []
    for m in ms:
        mi = m.info()
        req_bytes.append((mi["nnz"] - mi["num_coo"]) * 8 + mi["num_coo"] * 16 + 2 * mi["rows"] * 8)
    mean_bytes = float(np.mean(req_bytes))
    log(f"{count} matrices ({sum(rows)} rows) generated and converted in {time.time() - t0:.1f}s")
    variants = 2
    xh = [PinnedArray.from_host(gen_vector(n, seed * 1000 + 2 * i + v)) for i, n in enumerate(rows)
          for v in range(variants)]
    yh = [PinnedArray(n, np.float64) for n in rows]
    xd = [DeviceArray.from_host(p.array) for p in xh]
    yd = [DeviceArray(n, np.float64, zero=True) for n in rows]
    servers = {"continuous": HBDIAServer(ms, hbm_budget=0, max_jobs=256, max_in_flight=3),
               "plain": HBDIAServer(ms, batching=False)}
    info = servers["continuous"].info()
    log(f"resident {info['resident_bytes'] / 1e9:.2f} GB, request arena {info['arena_bytes'] / 1e9:.2f} GB")
    # parity: one request per matrix through each server against the oracle
    for name, S in servers.items():
        for i, m in enumerate(ms):
            y = np.zeros(rows[i])
            S.wait(S.submit(i, xh[i * variants].array, y))
            hb = m.getHBDIA()
            hb["num_blocks"], hb["block_width"] = m.getNumBlocks(), m.getBlockWidth()
            want = oracle.hbdia_spmv(hb, xh[i * variants].array.copy(), np.zeros(rows[i]), rows[i])
            if not np.array_equal(y, want):
                raise BenchError({"error": f"{name} server: matrix {i} differs from the oracle"})
    log("parity: every matrix bit-exact against the oracle through both servers")

    def stats(lat, wall, n):
        return {"requests": int(n), "served_per_s": round(n / wall, 1) if wall > 0 else None,
                "p50_us": round(float(np.percentile(lat, 50)), 1), "p99_us": round(float(np.percentile(lat, 99)), 1),
                "mean_us": round(float(lat.mean()), 1)}

    results = {}
    for vec in ("host", "device"):
        dev = vec == "device"
        xs = [a for a in xd] if dev else [p.array for p in xh]
        ys = [a for a in yd] if dev else [p.array for p in yh]
        res = {}
        caps = {}
        for name, S in servers.items():
            nb = 2000 if name == "continuous" else 400
            lat, wall = S.open_loop(0.0, nb, seed + 1, xs, variants, ys, device=dev)
            caps[name] = nb / wall
            res[f"{name}_capacity_per_s"] = round(caps[name], 1)
            res[f"{name}_capacity_roofline_frac"] = round(caps[name] * mean_bytes / 1e9 / HBM_PEAK_GBS, 4)
            log(f"{vec} x: {name} capacity {caps[name]:.0f} SpMVs/s (burst of {nb})")
        for frac in (0.3, 0.6, 0.9):
            rate = frac * caps["continuous"]
            row = {"offered_per_s": round(rate, 1)}
            for name, S in servers.items():
                # about 1.5 s of arrivals, fewer when the server cannot keep up
                n = int(min(4000, max(200, rate * 1.5)))
                if name == "plain" and rate > caps["plain"]:
                    n = int(max(100, min(n, caps["plain"] * 1.5)))
                lat, wall = S.open_loop(rate, n, seed + int(frac * 100), xs, variants, ys, device=dev)
                row[name] = stats(lat, wall, n)
            log(f"{vec} x at {int(frac * 100)}% load ({rate:.0f}/s): {row}")
            res[f"load_{int(frac * 100)}"] = row
        results[vec] = res
    for S in servers.values():
        S.close()
    head = results["host"]["load_90"]["continuous"]
    line = {"metric": "served SpMVs/s + p50/p99 latency (continuous batching, 64 resident matrices, Poisson arrivals)",
            "value": head["served_per_s"], "unit": "SpMVs/s", "n_gpus": 1, "steps": None, "warmup": None,
            "ms_per_step": round(1e3 / head["served_per_s"], 5) if head["served_per_s"] else None,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded generators)",
            "config": {"workload": "serve_pool64", "matrices": count, "rows": int(sum(rows)),
                       "resident_bytes": info["resident_bytes"], "arena_bytes": info["arena_bytes"],
                       "pool": desc, "x_variants_per_matrix": variants, "loads": "30/60/90% of measured capacity",
                       "request_algorithmic_bytes_mean": int(mean_bytes),
                       "error_vs_fp64_oracle": "0 (bit-exact)"},
            "results": results}
    return line


def bench_copy(args) -> dict:
    """Counter calibration: a 1 GiB device-to-device copy (hipMemcpyAsync on the null stream), the
    plain copy whose FETCH_SIZE / WRITE_SIZE readings scale the SpMV kernels' counters to bytes."""
    from hbdia_amd import DeviceArray, Event
    from hbdia_amd._native import check_hip, hip
    n = 1 << 27  # doubles: 1 GiB
    src = DeviceArray(n, np.float64, zero=True)
    dst = DeviceArray(n, np.float64, zero=True)
    for _ in range(args.warmup):
        check_hip(hip.hbdia_memcpy_d2d(dst.ptr, src.ptr, n * 8, None))
    e0, e1 = Event(), Event()
    e0.record()
    for _ in range(args.steps):
        check_hip(hip.hbdia_memcpy_d2d(dst.ptr, src.ptr, n * 8, None))
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / args.steps
    line = {"metric": "device-to-device copy GB/s (read + write; counter calibration)", "value": round(2 * n * 8 / (ms * 1e-3) / 1e9, 1),
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 5),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "zeros",
            "config": {"workload": "copy_1GiB", "bytes_read": n * 8, "bytes_written": n * 8}}
    return line


def bench_convert(args) -> dict | int:
    """COO -> HBDIA conversion of the 256^3 stencil + scatter matrix (453M entries, block width 64,
    threshold 2): the device conversion with the COO already in HBM (hbdia_convert_coo, every result
    array left in HBM), the whole convertToHBDIAFormatOnDevice (COO upload, device conversion, host
    copies of the result; and with hostValues=False, the blocked values left on the device only),
    and the host conversion + upload (convertToHBDIAFormat); the two methods must leave the same
    host arrays.  value = entries per second of the device conversion."""
    import ctypes

    import hbdia_amd
    from hbdia_amd import HBDIA, DeviceArray
    from hbdia_amd import kernels as K
    from hbdia_amd._native import check_hip, hip

    if not hbdia_amd.gpu_available():
        print("no HIP device visible", file=sys.stderr)
        return 3
    n = args.grid or 256
    bw, thr, seed = 64, 2, 0xC2
    t0 = time.time()
    A = HBDIA.stencil(n, 27, 0.005, seed=seed)
    r, c, v = A.getCOO()
    nnz = len(v)
    log(f"generated {n}^3 stencil + scatter ({nnz} entries) in {time.time() - t0:.1f}s")

    R, C, V = DeviceArray.from_host(r), DeviceArray.from_host(c), DeviceArray.from_host(v)
    cargs = K.ConvertArgs(R.ptr, C.ptr, V.ptr, None, nnz, n ** 3, n ** 3, bw, thr, 1_000_000_000, 8, 1, 1)
    dev_s = []
    for k in range(max(1, args.warmup) + max(1, args.steps)):
        res = K.ConvertResult()
        t = time.perf_counter()
        check_hip(hip.hbdia_convert_coo(ctypes.byref(cargs), ctypes.byref(res), None))
        dt = time.perf_counter() - t
        pairs = res.num_pairs
        check_hip(hip.hbdia_convert_release(ctypes.byref(res)))
        if k >= max(1, args.warmup):
            dev_s.append(dt)
    log(f"device conversion: median {np.median(dev_s) * 1e3:.1f} ms, min {min(dev_s) * 1e3:.1f}, "
        f"max {max(dev_s) * 1e3:.1f} ({pairs} pairs)")
    del R, C, V
    check_hip(hip.hbdia_convert_trim())

    method_s, host_s = [], []
    reps = 2
    for k in range(reps):
        A.deleteHBDIAFormat() if A.isHBDIAFormat() else None
        t = time.perf_counter()
        A.convertToHBDIAFormatOnDevice(bw, thr)
        method_s.append(time.perf_counter() - t)
        log(f"convertToHBDIAFormatOnDevice {method_s[-1]:.2f} s")
    dev_state = A.getHBDIA()
    donly_s = []
    for k in range(reps):
        A.deleteHBDIAFormat()
        t = time.perf_counter()
        A.convertToHBDIAFormatOnDevice(bw, thr, hostValues=False)
        donly_s.append(time.perf_counter() - t)
        log(f"convertToHBDIAFormatOnDevice(hostValues=False) {donly_s[-1]:.2f} s")
    for k in range(reps):
        A.deleteHBDIAFormat()
        t = time.perf_counter()
        A.convertToHBDIAFormat(bw, thr)
        host_s.append(time.perf_counter() - t)
        log(f"convertToHBDIAFormat (host + upload) {host_s[-1]:.2f} s")
    host_state = A.getHBDIA()
    same = all(np.array_equal(dev_state[k], host_state[k]) for k in host_state)
    if not same:
        print("device and host conversions differ", file=sys.stderr)
        return 1
    del dev_state, host_state
    # CPU baseline: the oracle's conversion (the reference's algorithm: one std::map per block, one
    # thread) on a bounded sample, the 96^3 matrix of the same family
    cpu = None
    if not args.no_cpu:
        import oracle
        sn = 96
        S = HBDIA.stencil(sn, 27, 0.005, seed=seed)
        sr, sc_, sv = S.getCOO()
        t = time.perf_counter()
        oracle.convert(sr, sc_, sv, sn ** 3, sn ** 3, bw, thr)
        dt = time.perf_counter() - t
        cpu = {"value": round(len(sv) / dt, 1), "unit": "entries/s", "cores": 1, "kind": "port",
               "sample": f"oracle conversion of the {sn}^3 stencil + scatter ({len(sv)} entries) in {dt:.2f} s"}
        log(f"cpu baseline: {cpu['sample']}")
    ms = float(np.median(dev_s)) * 1e3
    line = {"metric": "COO->HBDIA conversion entries/s (device, COO resident in HBM)",
            "value": round(nnz / (ms * 1e-3), 1), "unit": "entries/s", "n_gpus": 1, "steps": len(dev_s),
            "warmup": max(1, args.warmup), "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "int32 keys, f64 values", "data": "synthetic",
            "config": {"workload": f"convert_stencil27_scatter_{n}^3", "entries": nnz, "block_width": bw,
                       "threshold": thr, "pairs": int(pairs)},
            "ms_min": round(min(dev_s) * 1e3, 3), "ms_max": round(max(dev_s) * 1e3, 3),
            "method_s": round(min(method_s), 3), "method_device_only_s": round(min(donly_s), 3),
            "host_convert_upload_s": round(min(host_s), 3),
            "host_threads": hbdia_amd.api.capi.hbdia_get_num_threads(),
            "host_over_method": round(min(host_s) / min(method_s), 2), "identical": same, "cpu_baseline": cpu}
    return line




def bench_ranklocal(args) -> dict | int:
    """One rank of eight of the c2 matrix on one GPU: rows [3q, 4q) (q = N / 8) with global-row
    offsets, the columns owned by other ranks in the COO part, x = [left halo | local | right halo]
    filled from the global x.  Parity: bit-exact against the oracle's rank-local SpMV on the same
    layout; CPU baseline: that oracle SpMV on the host cores."""
    import hbdia_amd
    from hbdia_amd import HBDIA, HBDIAVector, gen_vector, hbdiaSpMV

    if not hbdia_amd.gpu_available():
        print("no HIP device visible", file=sys.stderr)
        return 3
    n = args.grid or 256
    size, rank, bw, thr, seed = 8, 3, 64, 2, 0xC2
    t0 = time.time()
    G = HBDIA.stencil(n, 27, 0.005, seed=seed)
    N = n ** 3
    x = gen_vector(N, seed + 1)
    log(f"generated {n}^3 stencil + scatter in {time.time() - t0:.1f}s")
    t1 = time.time()
    A = HBDIA.createRankLocalMatrix(G, rank, size, bw, thr)
    G.close()
    log(f"rank {rank} of {size}: partitioned, converted, data ranges in {time.time() - t1:.1f}s")
    t1 = time.time()
    A.prepareForGPU()
    lay = A.getLocalLayout()
    info = A.info()
    nl = A.getNumRows()
    r0, L, R = lay["first_row"], lay["halo_left"], lay["halo_right"]
    log(f"rank-local COO built + uploaded in {time.time() - t1:.1f}s: rows {nl} from {r0}, halo {L} + {R}, "
        f"COO part {lay['coo']} ({lay['folded']} from the blocked part), {lay['ranges']} data ranges")
    X = HBDIAVector.withHalo(x[r0:r0 + nl], A)
    X.fillHalo(A, x)
    Y = HBDIAVector(np.zeros(nl))

    def step():
        hbdiaSpMV(A, X, Y, False, True)

    step()
    A.synchronize()
    y_gpu = Y.download()
    nnz = info["nnz"]
    ncoo = lay["coo"]
    nnz_bdia = nnz - ncoo
    flop = 2.0 * nnz
    alg_bytes = nnz_bdia * 8 + ncoo * (8 + 4 + 4) + (nl + L + R) * 8 + nl * 8

    cpu_baseline = None
    parity = None
    if not args.no_cpu:
        import oracle  # the checker and the CPU baseline only
        hb = A.getHBDIA()
        hb["num_blocks"], hb["block_width"] = A.getNumBlocks(), bw
        loc = {"hb": hb, "num_rows": nl, "halo_left": L, "coo_rows": lay["coo_rows"], "coo_cols": lay["coo_cols"],
               "coo_values": lay["coo_values"]}
        h = lay["halo_cols"]
        x_ext = np.concatenate([x[h[:L]], x[r0:r0 + nl], x[h[L:]]])
        y_cpu = oracle.rank_local_spmv(loc, x_ext, np.zeros(nl))
        diff = y_gpu != y_cpu
        rel = np.abs(y_gpu - y_cpu) / np.maximum(np.abs(y_cpu), 1e-300)
        parity = {"rows": nl, "rows_differing": int(diff.sum()), "max_rel_err": float(rel.max()),
                  "bit_exact": bool(not diff.any())}
        log(f"parity vs oracle (rank-local layout): {parity}")
        if rel.max() > 1e-12:
            raise BenchError({"error": "parity failure", "parity": parity})
        times = []
        for rep in range(12):
            ts = time.perf_counter()
            oracle.rank_local_spmv(loc, x_ext, np.zeros(nl))
            te = time.perf_counter()
            if rep >= 2:
                times.append(te - ts)
        t_cpu = float(np.median(times))
        cpu_baseline = {"value": round(flop / t_cpu / 1e9, 3), "unit": "GFLOP/s", "cores": oracle.max_threads(),
                        "kind": "port", "ms_per_spmv": round(t_cpu * 1e3, 3),
                        "sample": f"the rank's full matrix, median of 10 SpMVs after 2 warm-ups "
                                  f"(oracle: bdia_spmv_kernel over the local segment + cpu_coo_spmv_kernel over "
                                  f"the rank-local COO part + add, OpenMP)"}
        log(f"cpu baseline: {cpu_baseline}")

    stream = A.device_ptrs()["bdia_stream"]
    wall_ms, kern_ms = _time_launches(step, A.synchronize, stream, args.steps, args.warmup)
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    value = flop / (wall_ms * 1e-3) / 1e9
    log(f"{args.steps} steps: {wall_ms:.4f} ms/step wall, {kern_ms:.4f} ms/launch, {achieved:.1f} GB/s algorithmic")
    line = {
        "metric": "SpMV GFLOP/s (FP64 hbdiaSpMV) + % HBM roofline",
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"rank_local_{rank}_of_{size}_stencil27_scatter_{n}^3", "grid": n, "rank": rank,
                   "ranks": size, "rows": nl, "first_row": r0, "halo_left": L, "halo_right": R,
                   "block_width": bw, "threshold": thr, "seed": seed, "coo_exec": "gpu_fused"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": cpu_baseline,
        "kernel_ms": round(kern_ms, 5),
        "algorithmic_bytes": int(alg_bytes),
        "flop": int(flop),
        "nnz_bdia": int(nnz_bdia),
        "nnz_coo": int(ncoo),
        "nnz_coo_folded": int(lay["folded"]),
        "parity": parity,
        "setup_s": round(time.time() - t0, 1),
    }
    return line


if __name__ == "__main__":
    sys.exit(main())
