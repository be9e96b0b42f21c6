#!/usr/bin/env python3
"""Benchmark of the H_eff hot path on one MI355X.

A step is one H_eff matvec (MatVecMPOS::MultAx) of the isolated single-site workload at chi=512:
theta (d=2, 512, 512) complex128 with N(0,1) entries normalised, env blocks (512, 512, 5) Hermitian per
MPO index, a random real MPO (5, 5, 2, 2) with symmetric physical blocks, seed 1. Inputs are resident in
HBM before the timed region. value = algorithmic FLOP of K matvecs / wall time of the K steps
(8 real FLOP per complex multiply-accumulate, MACs = d chiL^2 chiR DL + chiL chiR DL DR d^2 + d chiL chiR^2 DR).

Prints one JSON line (rank 0). Only one GPU is used: --gpus above 1 exits with a message.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "H_eff matvec TFLOP/s @χ512; solves/s + p99, mixed vs FP64 @|ΔE|≤1e-10"
PEAK_FP64_MFMA = 78.6  # TFLOP/s, MI355X dense FP64 matrix
PEAK_FP32_MFMA = 157.3  # TFLOP/s, MI355X dense FP32 matrix


def leg_macs(d, cL, cR, DL, DR):
    """MACs of the three legs in the order the library runs them (theta.L first when chiL >= chiR)."""
    if cL >= cR:
        return [d * cR * cL * cL * DL, d * d * DL * DR * cL * cR, d * cL * cR * cR * DR]
    return [d * cL * cR * cR * DR, d * d * DL * DR * cL * cR, d * cR * cL * cL * DL]


def gpu_matvec_bench(chi, steps, warmup, fp32=False, legs_nrep=50):
    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import heff_flops, random_heff_problem

    d, D = 2, 5
    th, W, L, R = random_heff_problem(d, chi, chi, D, D, seed=1)
    op = G.MatVecMPOS([W], L, R, dtype=np.complex128)
    n = op.rows()
    xdt = np.complex64 if fp32 else np.complex128
    dx = G.DeviceArray.from_numpy(th.reshape(n, order="F").astype(xdt))
    dy = G.DeviceArray((n,), xdt)
    for _ in range(warmup):
        op.MultAx_device(dx, dy, fp32=fp32)
    G._lib.call("dmrg_synchronize")
    t0 = time.perf_counter()
    for _ in range(steps):
        op.MultAx_device(dx, dy, fp32=fp32)
    G._lib.call("dmrg_synchronize")
    dt = time.perf_counter() - t0
    flops = heff_flops(d, chi, chi, D, D, True)
    legs_ms = op.time_legs(dx, dy, legs_nrep, fp32=fp32)
    macs = leg_macs(d, chi, chi, D, D)
    op.close()
    return {"ms_per_step": dt / steps * 1e3, "tflops": flops * steps / dt / 1e12, "legs_ms": legs_ms,
            "legs_tflops": [8.0 * m / (t * 1e-3) / 1e12 for m, t in zip(macs, legs_ms)], "flops": flops}


def cpu_baseline(chi=512, budget_s=12.0, max_reps=200):
    """Time the oracle (numpy restatement of the reference contraction) on the host cores."""
    import numpy as np

    from dmrg_amd.workloads import heff_flops, random_heff_problem
    from oracle import heff as O  # cpu_baseline leg: the oracle is the thing timed here, as the CPU port

    threads, blas = None, "unknown"
    try:
        from threadpoolctl import threadpool_info

        info = [i for i in threadpool_info() if i.get("user_api") == "blas"]
        if info:
            threads = int(info[0].get("num_threads"))
            blas = f"{info[0].get('internal_api')} {info[0].get('version')}"
    except Exception:
        pass
    th, W, L, R = random_heff_problem(2, chi, chi, 5, 5, seed=1)
    O.matrix_vector_product(th, W, L, R)  # warm
    reps, t0 = 0, time.perf_counter()
    while reps < max_reps and (time.perf_counter() - t0) < budget_s:
        O.matrix_vector_product(th, W, L, R)
        reps += 1
    dt = time.perf_counter() - t0
    flops = heff_flops(2, chi, chi, 5, 5, True)
    return {"value": round(flops * reps / dt / 1e12, 5), "unit": "TFLOP/s", "cores": threads or os.cpu_count(),
            "kind": "port",
            "sample": f"{reps} matvecs of the chi={chi} cx64 workload by the numpy oracle ({blas}, {threads} threads), "
                      f"{dt:.1f} s"}


def load_traffic(name: str):
    p = ROOT / "profiles" / f"traffic_{name}.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--chi", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (the per-request "
              f"sharding across GPUs is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import dmrg_amd as G

    dev = G.device_info()
    r64 = gpu_matvec_bench(args.chi, args.steps, args.warmup)
    r32 = gpu_matvec_bench(args.chi, max(args.steps // 2, 1), args.warmup, fp32=True)
    sweep = {}
    for chi in (128, 256):
        if chi != args.chi:
            sweep[str(chi)] = round(gpu_matvec_bench(chi, 200, 10)["tflops"], 3)

    i_dom = max(range(3), key=lambda i: r64["legs_ms"][i])
    macs = leg_macs(2, args.chi, args.chi, 5, 5)
    achieved = r64["legs_tflops"][i_dom]
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_MFMA, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_MFMA, 4), "traffic": load_traffic(f"leg{i_dom + 1}_chi{args.chi}"),
                "kernel": f"gemm leg {i_dom + 1} of 3 (complex FP64, {macs[i_dom]:.3e} MACs per launch)",
                "legs_ms": [round(x, 4) for x in r64["legs_ms"]]}
    out = {
        "metric": METRIC,
        "value": round(r64["tflops"], 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r64["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "cx64",
        "data": "synthetic",
        "config": {"workload": f"isolated single-site H_eff matvec, theta (2,{args.chi},{args.chi}) complex128, "
                               f"env ({args.chi},{args.chi},5) Hermitian blocks, real MPO (5,5,2,2), seed 1",
                   "d": 2, "chi": args.chi, "D": 5, "device": dev["name"], "cu_count": dev["cu_count"]},
        "roofline": roofline,
        "fp32_mfma": {"tflops": round(r32["tflops"], 4), "ms_per_step": round(r32["ms_per_step"], 5),
                      "frac_of_fp32_peak": round(r32["tflops"] / PEAK_FP32_MFMA, 4)},
        "chi_sweep_tflops": sweep,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.chi)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
