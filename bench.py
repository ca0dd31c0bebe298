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

HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = ROOT / "profiles" / "traffic_hbdia_spmv_stencil256.json"


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--workload", choices=["c2", "ref128"], default="c2")
    p.add_argument("--grid", type=int, default=None, help="override the grid edge (testing only)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the full-size parity check")
    p.add_argument("--cpu-reps", type=int, default=5)
    return p.parse_args()


def main() -> int:
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import hbdia_amd
    from hbdia_amd import HBDIA, Event, HBDIAVector, gen_vector, hbdiaSpMV
    from hbdia_amd._native import check_hip, hip

    if not hbdia_amd.gpu_available():
        print("no HIP device visible", file=sys.stderr)
        return 3
    log(f"device {hbdia_amd.device_name()}, host threads {hbdia_amd.api.capi.hbdia_get_num_threads()}")

    t0 = time.time()
    if args.workload == "c2":
        n = args.grid or 256
        bw, thr, seed = 64, 2, 0xC2
        A = HBDIA.stencil(n, 27, 0.005, seed=seed)
        log(f"generated {n}^3 stencil + scatter in {time.time() - t0:.1f}s")
        N = n ** 3
        x = gen_vector(N, seed + 1)
        workload = {"workload": f"stencil27_scatter_{n}^3", "grid": n, "rows": N, "scatter": 0.005,
                    "block_width": bw, "threshold": thr, "seed": seed, "coo_exec": "gpu_fused"}
    else:
        n = args.grid or 128
        bw, thr = 32, 8
        A = HBDIA()
        A.create3DStencil27Point(n, n, n, 0.0)
        log(f"generated {n}^3 reference stencil in {time.time() - t0:.1f}s")
        N = n ** 3
        x = np.arange(N, dtype=np.float64)
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
    alg_bytes = nnz_bdia * 8 + nnz_coo * (8 + 4 + 4) + N * 8 + N * 8
    stored_bytes = info["stored_bytes"] + 2 * N * 8

    X = HBDIAVector(x)
    Y = HBDIAVector(np.zeros(N))
    use_cpu = False
    gpu_coo = nnz_coo > 0

    def step():
        hbdiaSpMV(A, X, Y, use_cpu, gpu_coo)

    step()
    A.synchronize()
    y_gpu = Y.download()

    cpu_baseline = None
    parity = None
    if not args.no_cpu:
        import oracle  # the checker and the CPU baseline only
        hb = A.getHBDIA()
        hb["num_blocks"], hb["block_width"] = A.getNumBlocks(), bw
        cores = oracle.max_threads()
        y_cpu = np.zeros(N)
        coo_buf = np.zeros(N)
        oracle.hbdia_spmv(hb, x, y_cpu, N, coo_buf)
        diff = y_gpu != y_cpu
        rel = np.abs(y_gpu - y_cpu) / np.maximum(np.abs(y_cpu), 1e-300)
        parity = {"rows": N, "rows_differing": int(diff.sum()), "max_rel_err": float(rel.max()),
                  "bit_exact": bool(not diff.any())}
        log(f"parity vs oracle at full size: {parity}")
        if rel.max() > 1e-12:
            print(json.dumps({"error": "parity failure", "parity": parity}))
            return 4
        times = []
        for rep in range(args.cpu_reps + 1):
            y_cpu[:] = 0.0
            ts = time.perf_counter()
            oracle.hbdia_spmv(hb, x, y_cpu, N, coo_buf)
            te = time.perf_counter()
            if rep > 0:
                times.append(te - ts)
        t_cpu = float(np.median(times))
        cpu_baseline = {"value": round(flop / t_cpu / 1e9, 3), "unit": "GFLOP/s", "cores": cores, "kind": "port",
                        "ms_per_spmv": round(t_cpu * 1e3, 3),
                        "sample": f"full matrix, median of {args.cpu_reps} SpMVs after 1 warm-up "
                                  f"(oracle restatement of bdia_spmv_kernel + cpu_coo_spmv_kernel + add, OpenMP)"}
        log(f"cpu baseline: {cpu_baseline}")
        del hb

    stream = A.device_ptrs()["bdia_stream"]
    ev0, ev1 = Event(), Event()
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

    traffic = None
    if PROFILE_TRAFFIC.exists() and args.workload == "c2" and (args.grid or 256) == 256:
        try:
            traffic = json.loads(PROFILE_TRAFFIC.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    value = flop / (wall_ms * 1e-3) / 1e9
    vs = None
    if args.workload == "ref128" and (args.grid or 128) == 128:
        vs = round(value / (flop / 0.40e-3 / 1e9), 3)  # reference figure: 0.40 ms on a GH200-class GPU
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
        "vs_baseline": vs,
        "dtype": "f64",
        "data": "synthetic (seeded generator)",
        "config": workload,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu_baseline,
        "kernel_ms": round(kern_ms, 5),
        "algorithmic_bytes": int(alg_bytes),
        "stored_bytes": int(stored_bytes),
        "flop": int(flop),
        "nnz_bdia": int(nnz_bdia),
        "nnz_coo": int(nnz_coo),
        "parity": parity,
        "setup_s": round(time.time() - t0, 1),
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
