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
    legs_ms = op.time_legs(dx, dy, legs_nrep, fp32=fp32)  # [leg1, leg2, leg3 main, leg3 reduce]
    macs = leg_macs(d, chi, chi, D, D)
    op.close()
    kernel_ms = legs_ms[:3]
    return {"ms_per_step": dt / steps * 1e3, "tflops": flops * steps / dt / 1e12, "legs_ms": legs_ms,
            "legs_tflops": [8.0 * m / (t * 1e-3) / 1e12 for m, t in zip(macs, kernel_ms)], "flops": flops}


def multivector_rates(chis=(128, 256), nvecs=(4, 16), steps=100):
    """Matvec rate at small chi when one call applies H_eff to a block of nvec vectors (MultAx's blockSize,
    matvec_mpos.cpp:976-984; the library runs the block as one batched GEMM per leg instead of a loop)."""
    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import heff_flops, random_heff_problem

    out = {}
    for chi in chis:
        th, W, L, R = random_heff_problem(2, chi, chi, 5, 5, seed=1)
        op = G.MatVecMPOS([W], L, R, dtype=np.complex128)
        n = op.rows()
        rng = np.random.default_rng(chi)
        for nv in nvecs:
            X = (rng.standard_normal((n, nv)) + 1j * rng.standard_normal((n, nv))).astype(np.complex128)
            dx = G.DeviceArray.from_numpy(np.asfortranarray(X))
            dy = G.DeviceArray((n, nv), np.complex128)
            for _ in range(5):
                op.MultAx_device(dx, dy, nvec=nv)
            G._lib.call("dmrg_synchronize")
            t0 = time.perf_counter()
            for _ in range(steps):
                op.MultAx_device(dx, dy, nvec=nv)
            G._lib.call("dmrg_synchronize")
            dt = (time.perf_counter() - t0) / steps
            tf = heff_flops(2, chi, chi, 5, 5, True) * nv / dt / 1e12
            out[f"chi{chi}_nvec{nv}"] = {"ms_per_call": round(dt * 1e3, 4), "tflops": round(tf, 3),
                                         "frac_of_fp64_peak": round(tf / PEAK_FP64_MFMA, 4)}
            dx.close()
            dy.close()
        op.close()
    return out


def host_info():
    """nproc, CPU model and the BLAS numpy runs on (with its thread count)."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    threads, blas = None, "unknown"
    try:
        from threadpoolctl import threadpool_info

        info = [i for i in threadpool_info() if i.get("user_api") == "blas"]
        if info:
            threads = int(info[0].get("num_threads"))
            blas = f"{info[0].get('internal_api')} {info[0].get('version')}"
    except Exception:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model, "blas": blas, "blas_threads": threads}


def cpu_baseline(chi=512, budget_s=10.0, max_reps=200):
    """Time the oracle's C restatement of the reference contraction (oracle/heff_c.c: the same three
    contractions, packed AVX2 GEMMs, OpenMP) on the host cores: all cores, then one thread; then the numpy
    restatement (its GEMMs in the host BLAS) for comparison. Each sample is bounded to about budget_s seconds."""
    from dmrg_amd.workloads import heff_flops, random_heff_problem
    from oracle import heff as O  # cpu_baseline leg: the oracle is the thing timed here, as the CPU port
    from oracle import heff_c as OC

    host = host_info()
    th, W, L, R = random_heff_problem(2, chi, chi, 5, 5, seed=1)
    flops = heff_flops(2, chi, chi, 5, 5, True)

    def timed(fn, budget):
        fn()  # warm (and, for the C port, its workspace)
        reps, t0 = 0, time.perf_counter()
        while reps < max_reps and (time.perf_counter() - t0) < budget:
            fn()
            reps += 1
        dt = time.perf_counter() - t0
        return reps, dt, flops * reps / dt / 1e12

    def numpy_port(budget):
        r, dt, tf = timed(lambda: O.matrix_vector_product(th, W, L, R), budget)
        return {"value": round(tf, 5), "unit": "TFLOP/s", "cores": host["blas_threads"] or os.cpu_count(),
                "sample": f"{r} matvecs by the numpy oracle ({host['blas']}) in {dt:.1f} s"}

    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or os.cpu_count() or 1
    if not OC.host_supported():  # no AVX2/FMA: the numpy restatement is the baseline
        out = numpy_port(budget_s)
        out.update(kind="port", host=host, note="host CPU lacks AVX2/FMA: C restatement not run")
        return out
    reps, dt, tf = timed(lambda: OC.matrix_vector_product(th, W, L, R, threads=threads), budget_s)
    r1, d1, t1 = timed(lambda: OC.matrix_vector_product(th, W, L, R, threads=1), 0.8 * budget_s)
    out = {"value": round(tf, 5), "unit": "TFLOP/s", "cores": threads, "kind": "port",
           "sample": f"{reps} matvecs of the chi={chi} cx64 workload by the oracle's C restatement "
                     f"(oracle/heff_c.c, OpenMP, {threads} threads) in {dt:.1f} s",
           "host": host,
           "single_thread": {"value": round(t1, 5), "unit": "TFLOP/s", "cores": 1,
                             "sample": f"{r1} matvecs in {d1:.1f} s on one thread"}}
    out["numpy_blas"] = numpy_port(0.8 * budget_s)
    # The H^2 matvec shape (D^2 = 25 squared MPO and env blocks, seed 2): the C restatement on all cores.
    from dmrg_amd.workloads import squared_problem

    th2, W1, L1, R1 = random_heff_problem(2, chi, chi, 5, 5, seed=2)
    W2, L2, R2 = squared_problem(th2, W1, L1, R1)
    del W1, L1, R1
    f2 = heff_flops(2, chi, chi, 25, 25, True)
    OC.matrix_vector_product(th2, W2, L2, R2, threads=threads)  # warm
    r2, t0 = 0, time.perf_counter()
    while r2 < max_reps and (time.perf_counter() - t0) < 0.5 * budget_s:
        OC.matrix_vector_product(th2, W2, L2, R2, threads=threads)
        r2 += 1
    d2 = time.perf_counter() - t0
    out["h2_matvec"] = {"value": round(f2 * r2 / d2 / 1e12, 5), "unit": "TFLOP/s", "cores": threads,
                        "sample": f"{r2} H^2 matvecs (chi={chi}, D^2=25) by the C restatement in {d2:.1f} s"}
    # Left environment update of the same site (D = 5 H blocks), EnvBase::build_block in C on all cores.
    OC.env_build_block("L", L, th, W, threads=threads)  # warm
    r3, t0 = 0, time.perf_counter()
    while r3 < max_reps and (time.perf_counter() - t0) < 0.3 * budget_s:
        OC.env_build_block("L", L, th, W, threads=threads)
        r3 += 1
    d3 = time.perf_counter() - t0
    out["env_update"] = {"value": round(flops * r3 / d3 / 1e12, 5), "unit": "TFLOP/s", "cores": threads,
                         "sample": f"{r3} left env updates (chi={chi}, D=5) by the C restatement in {d3:.1f} s"}
    return out


def load_traffic(name: str):
    p = ROOT / "profiles" / f"traffic_{name}.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def section_env_update(chi=512, steps=30):
    """Environment updates (EnvBase::build_block) of the chi=512 site, left and right, with the H blocks (D=5) and
    the variance blocks (D^2=25), device-resident; each call writes a preallocated block."""
    import ctypes

    import numpy as np

    import dmrg_amd as G
    from dmrg_amd import _dims
    from dmrg_amd.workloads import heff_flops, random_heff_problem, squared_problem

    th, W, L, R = random_heff_problem(2, chi, chi, 5, 5, seed=2)
    W2, L2, R2 = squared_problem(th, W, L, R)
    dth = G.DeviceArray.from_numpy(th)
    out = {}
    for tag, (Wx, Lx, Rx) in (("D5", (W, L, R)), ("D25", (W2, L2, R2))):
        D = Wx.shape[0]
        dW = G.DeviceArray.from_numpy(Wx)
        for side, other in (("L", Lx), ("R", Rx)):
            dother = G.DeviceArray.from_numpy(other)
            dout = G.DeviceArray((chi, chi, D), np.complex128)
            args = (dth.code, ctypes.c_char(side.encode()), dout.ptr, dother.ptr, _dims(dother.shape), dth.ptr,
                    _dims(dth.shape), dW.ptr, _dims(dW.shape), None)
            for _ in range(3):
                G._lib.call("dmrg_env_build_block", *args)
            G._lib.call("dmrg_synchronize")
            t0 = time.perf_counter()
            for _ in range(steps):
                G._lib.call("dmrg_env_build_block", *args)
            G._lib.call("dmrg_synchronize")
            dt = (time.perf_counter() - t0) / steps
            fl = heff_flops(2, chi, chi, D, D, True)
            out[f"{tag}_{side}"] = {"ms": round(dt * 1e3, 4), "tflops": round(fl / dt / 1e12, 3),
                                    "frac_of_fp64_peak": round(fl / dt / 1e12 / PEAK_FP64_MFMA, 4)}
            dother.close()
            dout.close()
        dW.close()
    dth.close()
    out["workload"] = (f"theta (2,{chi},{chi}) cx64 seed 2, H blocks ({chi},{chi},5) and variance blocks "
                       f"({chi},{chi},25); FLOP = 8 x (d chi^3 D + chi^2 D^2 d^2 + d chi^3 D) per update")
    return out


def section_h2_matvec(chi=512, steps=30):
    """Isolated H^2 matvec (D^2 = 25 squared MPO and squared env blocks), FP64 and FP32 matrix cores."""
    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import heff_flops, random_heff_problem, squared_problem

    th, W, L, R = random_heff_problem(2, chi, chi, 5, 5, seed=2)
    W2, L2, R2 = squared_problem(th, W, L, R)
    op = G.MatVecMPOS([W2], L2, R2, dtype=np.complex128)
    n = op.rows()
    out = {"chi": chi, "D2": 25, "gflop_per_matvec": round(heff_flops(2, chi, chi, 25, 25, True) / 1e9, 2)}
    for fp32 in (False, True):
        dt = np.complex64 if fp32 else np.complex128
        dx = G.DeviceArray.from_numpy(th.reshape(n, order="F").astype(dt))
        dy = G.DeviceArray((n,), dt)
        for _ in range(3):
            op.MultAx_device(dx, dy, fp32=fp32)
        G._lib.call("dmrg_synchronize")
        t0 = time.perf_counter()
        for _ in range(steps):
            op.MultAx_device(dx, dy, fp32=fp32)
        G._lib.call("dmrg_synchronize")
        ms = (time.perf_counter() - t0) / steps * 1e3
        tf = heff_flops(2, chi, chi, 25, 25, True) / (ms * 1e-3) / 1e12
        key = "fp32" if fp32 else "fp64"
        out[key] = {"ms": round(ms, 4), "tflops": round(tf, 3),
                    "frac_of_peak": round(tf / (PEAK_FP32_MFMA if fp32 else PEAK_FP64_MFMA), 4)}
    op.close()
    return out


def schmidt_values(theta):
    """Schmidt values of a single-site theta (d, chiL, chiR) across its right bond: the singular values of the
    (d chiL) x chiR matrix (the split of split.cpp:21-346), by the library's Jacobi SVD."""
    import numpy as np

    import dmrg_amd as G

    d, cL, cR = theta.shape
    M = np.asfortranarray(theta.reshape(d * cL, cR, order="F"))
    S = G.svd_truncate(M, 1e-300, 0, 3)["S"]
    return S / np.linalg.norm(S)


def schmidt_error(a, b):
    """Largest difference of two Schmidt spectra relative to their largest value, and over the values above
    1e-6 of it the largest relative difference."""
    import numpy as np

    k = min(len(a), len(b))
    a, b = a[:k], b[:k]
    big = a > 1e-6 * a[0]
    return {"max_abs_over_s0": float(np.max(np.abs(a - b)) / a[0]),
            "max_rel_above_1e-6": float(np.max(np.abs(a[big] - b[big]) / a[big]))}


def mfma_fraction(f64, f32, seconds):
    """Fraction of the matrix-core roofline reached by a solve: the time its FP64 and FP32 matvec FLOP would take
    at the FP64 and FP32 matrix peaks, divided by the solve's time (FP64 alone: achieved / FP64 peak)."""
    return (f64 / (PEAK_FP64_MFMA * 1e12) + f32 / (PEAK_FP32_MFMA * 1e12)) / seconds


def section_eigs(chi=512, reps=100, tol=1e-10):
    """Ground-state eigensolves of the isolated workload's H_eff (reps cold random starts per mode), FP64 vs mixed
    (FP32 matrix-core phase + FP64 refinement): solves/s, p50/p99 latency over the reps, matrix-core roofline
    fraction, |dE| between the modes, energy variance and Schmidt values of the solutions."""
    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import heff_flops, random_heff_problem

    th, W, L, R = random_heff_problem(2, chi, chi, 5, 5, seed=1)
    op = G.MatVecMPOS([W], L, R, dtype=np.complex128)
    op.eigs(nev=1, tol=tol, seed=999)  # warm-up: first-use code loads and GEMM tuning
    op.eigs(nev=1, tol=tol, seed=999, mixed=True)
    flop = heff_flops(2, chi, chi, 5, 5, True)
    res = {}
    energies, first_vec = {}, {}
    n = op.rows()
    dv, dy = G.DeviceArray((n,), np.complex128), G.DeviceArray((n,), np.complex128)

    def variance(v):
        """Energy variance <H^2> - <H>^2 of the normalised vector v (FP64 matvec), evaluated as |Hv - <H> v|^2,
        which equals it exactly and has no cancellation between two numbers of size <H>^2."""
        v = np.ascontiguousarray(v / np.linalg.norm(v))
        G._lib.call("dmrg_memcpy_htod", dv.ptr, v.ctypes.data_as(__import__("ctypes").c_void_p), v.nbytes)
        op.MultAx_device(dv, dy)
        hv = dy.to_numpy()
        e = np.vdot(v, hv).real
        return float(np.linalg.norm(hv - e * v) ** 2)

    for mode in ("fp64", "mixed"):
        lat, mv64, mv32, es, rel_r, var = [], [], [], [], [], []
        for k in range(reps):
            t0 = time.perf_counter()
            w, v, info = op.eigs(nev=1, tol=tol, mixed=(mode == "mixed"), seed=1000 + k)
            lat.append(time.perf_counter() - t0)
            mv64.append(info["matvecs_fp64"])
            mv32.append(info["matvecs_fp32"])
            es.append(w[0])
            rel_r.append(info["rnorm"][0] / info["anorm"])
            if k < 10:
                var.append(variance(v[:, 0]))
            if k == 0:
                first_vec[mode] = v[:, 0].reshape(th.shape, order="F")
        total = float(np.sum(lat))  # solve time only (the variance checks are outside)
        energies[mode] = np.array(es)
        res[mode] = {"solves_per_s": round(reps / total, 3), "samples": reps,
                     "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
                     "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2), "max_ms": round(max(lat) * 1e3, 2),
                     "mean_matvecs_fp64": float(np.mean(mv64)), "mean_matvecs_fp32": float(np.mean(mv32)),
                     "roofline_frac": round(mfma_fraction(np.mean(mv64) * flop, np.mean(mv32) * flop, total / reps), 4),
                     "max_rel_residual": float(np.max(rel_r)),
                     "median_variance": float(np.median(var)), "max_variance": float(np.max(var))}
    dv.close()
    dy.close()
    op.close()
    dE = float(np.max(np.abs(energies["mixed"] - energies["fp64"])))
    spread = float(np.max(energies["fp64"]) - np.min(energies["fp64"]))
    S64, Smx = schmidt_values(first_vec["fp64"]), schmidt_values(first_vec["mixed"])
    return {"workload": f"isolated H_eff chi={chi} cx64 (seed 1), {reps} cold starts per mode, tol {tol:g}*|H|; "
                        f"roofline_frac = (FP64 FLOP / FP64 peak + FP32 FLOP / FP32 peak) / solve time; variances of "
                        f"the first 10 solutions",
            "fp64": res["fp64"], "mixed": res["mixed"], "max_abs_dE_mixed_vs_fp64": dE,
            "variance_ratio_mixed_vs_fp64": res["mixed"]["median_variance"] / max(res["fp64"]["median_variance"], 1e-300),
            "schmidt_values_mixed_vs_fp64": schmidt_error(S64, Smx),
            "speedup_mixed_vs_fp64": round(res["mixed"]["solves_per_s"] / res["fp64"]["solves_per_s"], 3),
            "max_abs_dE_between_fp64_starts": spread, "E0": float(energies["fp64"][0])}


def section_batched_chi512(nops=8, tol=1e-10):
    """Batched ground states of nops operators of the chi=512 workload (seeds 1..nops): one after another in FP64,
    batched in FP64 (grouped matvec launches, one synchronisation per step for all) and batched in mixed
    precision (FP32 batch, then FP64 batch). Solves/s, latency, roofline fraction, |dE| and variance ratio."""
    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import heff_flops, random_heff_problem

    ops = []
    for i in range(nops):
        th, W, L, R = random_heff_problem(2, 512, 512, 5, 5, seed=1 + i)
        ops.append(G.MatVecMPOS([W], L, R, dtype=np.complex128))
    flop = heff_flops(2, 512, 512, 5, 5, True)
    G.eigs_batched(ops[:2], tol=1e-4, seed=3)  # warm-up: plans of the grouped legs, code loads
    G.eigs_batched(ops[:2], tol=1e-4, seed=3, mixed=True)
    ops[0].eigs(nev=1, tol=1e-4, seed=3)
    out = {"workload": f"{nops} ground states of the chi=512 cx64 workload (seeds 1-{nops}), tol {tol:g}*|H|, random starts"}
    lat, e_seq, v_seq, mv = [], [], [], 0
    t0 = time.perf_counter()
    for op in ops:
        t1 = time.perf_counter()
        w, v, info = op.eigs(nev=1, tol=tol, seed=11)
        lat.append(time.perf_counter() - t1)
        e_seq.append(w[0])
        v_seq.append(v[:, 0])
        mv += info["matvecs_fp64"]
    t_seq = time.perf_counter() - t0
    out["single_fp64"] = {"seconds": round(t_seq, 3), "solves_per_s": round(nops / t_seq, 3),
                          "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 1), "max_ms": round(max(lat) * 1e3, 1),
                          "matvecs_fp64": mv, "roofline_frac": round(mfma_fraction(mv * flop, 0, t_seq), 4)}

    def variance(op, v):
        v = v / np.linalg.norm(v)
        hv = op.MultAx(v)
        return float(np.linalg.norm(hv - np.vdot(v, hv).real * v) ** 2)

    res = {}
    for mode in ("fp64", "mixed"):
        t0 = time.perf_counter()
        r = G.eigs_batched(ops, tol=tol, seed=11, mixed=(mode == "mixed"))
        dt = time.perf_counter() - t0
        m64 = sum(x[2]["matvecs_fp64"] for x in r)
        m32 = sum(x[2]["matvecs_fp32"] for x in r)
        fin = sorted(x[2]["seconds"] for x in r)
        res[mode] = r
        out["batched_" + mode] = {"seconds": round(dt, 3), "solves_per_s": round(nops / dt, 3),
                                  "all_converged": all(x[2]["nconv"] == 1 for x in r),
                                  "matvecs_fp64": m64, "matvecs_fp32": m32,
                                  "roofline_frac": round(mfma_fraction(m64 * flop, m32 * flop, dt), 4),
                                  "max_abs_dE_vs_single": float(max(abs(x[0] - e) for x, e in zip(r, e_seq))),
                                  "median_variance": float(np.median([variance(op, x[1]) for op, x in zip(ops, r)]))}
    var64 = float(np.median([variance(op, v) for op, v in zip(ops, v_seq)]))
    out["single_fp64"]["median_variance"] = var64
    out["max_abs_dE_mixed_vs_fp64"] = float(max(abs(a[0] - b[0]) for a, b in zip(res["mixed"], res["fp64"])))
    out["variance_ratio_mixed_vs_fp64"] = out["batched_mixed"]["median_variance"] / max(out["batched_fp64"]["median_variance"], 1e-300)
    out["speedup_batched_mixed_vs_single_fp64"] = round(t_seq / out["batched_mixed"]["seconds"], 3)
    out["speedup_batched_mixed_vs_batched_fp64"] = round(out["batched_fp64"]["seconds"] / out["batched_mixed"]["seconds"], 3)
    out["schmidt_values_mixed_vs_fp64"] = schmidt_error(schmidt_values(res["fp64"][0][1].reshape(2, 512, 512, order="F")),
                                                        schmidt_values(res["mixed"][0][1].reshape(2, 512, 512, order="F")))
    for op in ops:
        op.close()
    return out


def section_block_vectors(chi=512, reps=50):
    """The eigensolver's HBM-bound block-vector kernels on n = 2 chi^2 complex vectors (the chi=512 problem) with
    8 and 16 basis columns: achieved algorithmic GB/s (bytes the kernel must move / its duration, HIP events) and
    the fraction of the 8 TB/s HBM peak."""
    import ctypes as C

    import dmrg_amd as G

    n = 2 * chi * chi
    out = {"workload": f"n = {n} complex128 vectors, {reps} launches per kernel"}
    for m in (8, 16):
        ms = (C.c_float * 3)()
        by = (C.c_double * 3)()
        G._lib.call("dmrg_block_vector_time", G.DMRG_CX64, n, m, reps, ms, by)
        for i, name in enumerate(("projections_VHx", "update_y_minus_Vc", "ritz_vector_and_residual")):
            gbs = by[i] / (ms[i] * 1e-3) / 1e9
            out[f"{name}_m{m}"] = {"ms": round(ms[i], 4), "bytes": int(by[i]), "GB_per_s": round(gbs, 1),
                                   "frac_of_hbm_peak": round(gbs / 8000.0, 3)}
    return out


def section_sweep(cpu=True):
    """Full fDMRG sweep of ising_tf_rf: L=16, J=1, g=1, h_i ~ U[-1,1] (seed 0), chi_max=64, variance < 1e-12,
    real FP64, on the GPU, with two-site growth (expand_mode 0, the algorithm of the oracle's CPU sweep, which is
    timed beside it on the host cores) and with the reference's default subspace expansion (expand_mode 43)."""
    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import ising_fields
    from oracle import heff as O  # checker: closed-form energies and entropies

    L = 16
    h = ising_fields(L, 1.0, seed=0)
    G.fdmrg_ising_tf_rf(L, h, chi_max=64, expand_mode=0)  # warm-up (code loads, tuning)
    G.fdmrg_ising_tf_rf(L, h, chi_max=64, expand_mode=43)
    t0 = time.perf_counter()
    r = G.fdmrg_ising_tf_rf(L, h, chi_max=64, expand_mode=0)
    gpu_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    rx = G.fdmrg_ising_tf_rf(L, h, chi_max=64, expand_mode=43)
    gpu_x = time.perf_counter() - t0
    e0 = O.ising_free_fermion_ground_energy(1.0, 1.0, h)
    S0 = O.ising_free_fermion_entropies(1.0, 1.0, h)
    out = {"workload": "ising_tf_rf L=16 J=1 g=1 h~U[-1,1] seed 0, chi_max=64, single-site to var<1e-12, FP64 real",
           "gpu_s": round(gpu_s, 3), "energy": r["energy"], "exact_energy": e0, "abs_dE_vs_exact": abs(r["energy"] - e0),
           "variance": r["variance"], "max_entropy_err_vs_exact": float(np.max(np.abs(r["entropies"] - S0))),
           "sweeps": r["sweeps"], "eigensolves": r["eigensolves"], "t_eigs_s": round(r["t_eigs"], 3),
           "t_env_s": round(r["t_env"], 3), "t_svd_s": round(r["t_svd"], 3),
           "subspace_expansion_43": {"gpu_s": round(gpu_x, 3), "abs_dE_vs_exact": abs(rx["energy"] - e0),
                                     "variance": rx["variance"], "sweeps": rx["sweeps"], "expansions": rx["expansions"],
                                     "t_expand_s": round(rx["t_expand"], 3),
                                     "max_entropy_err_vs_exact": float(np.max(np.abs(rx["entropies"] - S0)))}}
    if cpu:
        from oracle import sweep as S  # cpu_baseline leg: the oracle sweep is the thing timed here

        threads = None
        try:
            from threadpoolctl import threadpool_info

            info = [i for i in threadpool_info() if i.get("user_api") == "blas"]
            threads = int(info[0]["num_threads"]) if info else None
        except Exception:
            pass
        t0 = time.perf_counter()
        rc = S.fdmrg_ising(L, h, chi_max=64)
        cpu_s = time.perf_counter() - t0
        out["cpu_reference_sweep"] = {"seconds": round(cpu_s, 2), "cores": threads or os.cpu_count(), "kind": "port",
                                      "abs_dE_vs_exact": abs(rc["energy"] - e0), "gpu_speedup": round(cpu_s / gpu_s, 1)}
    return out


def section_serve(nreq=128, workers=16, load=0.7, seed=4, mixed=True, nreq_batched=32):
    """Serving stream of independent ising_tf_rf realisations: L in {16,20,24,32}, chi_max in {64,128,256},
    h_i ~ U[-w,w] with w in {0.5,2,6}, seeds 100... A first pass with every request queued at t = 0 measures
    the capacity (solves/s at saturation); the second pass offers Poisson arrivals at `load` x that capacity and
    is the one reported (latencies from arrival to result). total_mfma_tflops = algorithmic matrix-core work of
    all requests (matvecs and env updates) / makespan."""
    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import ising_fields

    rng = np.random.default_rng(seed)
    shapes = [(int(rng.choice([16, 20, 24, 32])), int(rng.choice([64, 128, 256])), float(rng.choice([0.5, 2.0, 6.0])))
              for _ in range(nreq)]
    gaps = rng.exponential(1.0, nreq)

    def requests(rate):
        t, out = 0.0, []
        for i, (L, chi, w) in enumerate(shapes):
            if rate:
                t += float(gaps[i]) / rate
            out.append({"L": L, "chi_max": chi, "h_rand": ising_fields(L, w, seed=100 + i), "arrival_s": t, "seed": 100 + i})
        return out

    _, cap = G.serve(requests(None), workers=workers, hbm_budget_bytes=128 << 30, mixed=mixed)
    capacity = cap["solves_per_s"]
    rate = load * capacity
    res, st = G.serve(requests(rate), workers=workers, hbm_budget_bytes=128 << 30, mixed=mixed)
    gflop = sum(r["gflop"] for r in res)
    # Same saturation pass with the requests' FP64 energy solves batched into one eigensolver (the mixed-precision
    # solves of large problems stay with their request).
    res_b, cap_b = G.serve(requests(None)[:nreq_batched], workers=workers, hbm_budget_bytes=128 << 30, batch_eigs=True,
                           mixed=mixed)
    _, cap_s = G.serve(requests(None)[:nreq_batched], workers=workers, hbm_budget_bytes=128 << 30, mixed=mixed)
    batched = {"requests": nreq_batched, "capacity_solves_per_s": round(cap_b["solves_per_s"], 3),
               "capacity_solves_per_s_unbatched": round(cap_s["solves_per_s"], 3),
               "p50_latency_s_at_saturation": round(cap_b["p50_latency_s"], 3),
               "p50_latency_s_at_saturation_unbatched": round(cap_s["p50_latency_s"], 3),
               "batch_steps": cap_b["batch_steps"], "peak_batch": cap_b["peak_batch"],
               "mean_problems_per_step": round(sum(r["matvecs"] for r in res_b) / max(1, cap_b["batch_steps"]), 2),
               "step_ms": round(cap_b["batch_step_s"] / max(1, cap_b["batch_steps"]) * 1e3, 3),
               "rebuild_s": round(cap_b["batch_rebuild_s"], 2),
               "max_abs_dE_vs_unbatched": float(max(abs(a["energy"] - b["energy"]) for a, b in zip(res_b, res)))}
    lat = [r["latency_s"] for r in res]
    return {"workload": f"{nreq} requests, L in {{16,20,24,32}}, chi_max in {{64,128,256}}, w in {{0.5,2,6}}, "
                        f"{workers} workers, two-site growth (expand_mode 0), "
                        f"{'mixed precision (n >= 131072)' if mixed else 'FP64'}; Poisson arrivals "
                        f"at {load:g} x the measured capacity",
            "capacity_solves_per_s": round(capacity, 3), "arrival_rate_per_s": round(rate, 3),
            "solves_per_s": round(st["solves_per_s"], 3), "latency_samples": len(lat),
            "p50_latency_s": round(st["p50_latency_s"], 3), "p99_latency_s": round(st["p99_latency_s"], 3),
            "max_latency_s": round(max(lat), 3), "mean_latency_s": round(st["mean_latency_s"], 3),
            "completed": st["completed"], "failed": st["failed"], "makespan_s": round(st["makespan_s"], 2),
            "total_mfma_tflops": round(gflop / st["makespan_s"] / 1e3, 3), "total_gflop": round(gflop, 1),
            "peak_concurrency": st["peak_concurrency"], "peak_admitted_gb": round(st["peak_admitted_bytes"] / 2**30, 2),
            "max_variance": max(r["variance"] for r in res), "batched_eigensolves": batched}


def serve_stream(nreq=512, workers=16, load=0.7, seed=4, chunk=128, mixed=True):
    """The serving configuration at its full size: nreq ising_tf_rf requests (L in {16,20,24,32}, chi_max in
    {64,128,256}, w in {0.5,2,6}, seeds 100..), drawn as in section_serve (its first 32 are the same), with
    mixed precision for the problems large enough for it. Run as consecutive serve calls of `chunk` requests
    (a progress line after each): first all queued at t = 0 (capacity = requests / summed makespans), then
    Poisson arrivals at `load` x capacity; latencies are pooled over the chunks."""
    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import ising_fields

    rng = np.random.default_rng(seed)
    shapes = [(int(rng.choice([16, 20, 24, 32])), int(rng.choice([64, 128, 256])), float(rng.choice([0.5, 2.0, 6.0])))
              for _ in range(nreq)]
    gaps = rng.exponential(1.0, nreq)

    def run(rate, tag):
        lat, makespan, done, gflop, var = [], 0.0, 0, 0.0, 0.0
        for c0 in range(0, nreq, chunk):
            reqs, t = [], 0.0
            for i in range(c0, min(nreq, c0 + chunk)):
                L, chi, w = shapes[i]
                if rate:
                    t += float(gaps[i]) / rate
                reqs.append({"L": L, "chi_max": chi, "h_rand": ising_fields(L, w, seed=100 + i),
                             "arrival_s": t - (float(gaps[c0]) / rate if rate else 0.0), "seed": 100 + i})
            res, st = G.serve(reqs, workers=workers, hbm_budget_bytes=128 << 30, mixed=mixed)
            makespan += st["makespan_s"]
            done += st["completed"]
            lat += [r["latency_s"] for r in res]
            gflop += sum(r["gflop"] for r in res)
            var = max(var, max(r["variance"] for r in res))
            print(f"[serve {tag}] requests {c0}..{min(nreq, c0 + chunk) - 1}: {st['completed']} done in "
                  f"{st['makespan_s']:.1f} s", file=sys.stderr, flush=True)
        return lat, makespan, done, gflop, var

    _, ms_cap, _, _, _ = run(None, "capacity")
    capacity = nreq / ms_cap
    rate = load * capacity
    lat, ms, done, gflop, var = run(rate, "load")
    return {"serve_stream": f"{nreq} requests, L in {{16,20,24,32}}, chi_max in {{64,128,256}}, w in {{0.5,2,6}}, "
                            f"seeds 100-{100 + nreq - 1}, {workers} workers, {'mixed' if mixed else 'FP64'}, "
                            f"chunks of {chunk}",
            "capacity_solves_per_s": round(capacity, 3), "arrival_rate_per_s": round(rate, 3),
            "solves_per_s": round(nreq / ms, 3), "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
            "p99_latency_s": round(float(np.percentile(lat, 99)), 3), "completed": done,
            "total_gflop": round(gflop, 1), "max_variance": var}


def section_grouped(nops=32, reps=200, seed=6):
    """Serving-sized operators (chi in {32, 48, 64, 96, 128}, d=2, D=5, complex): their matvecs one after
    another (3-4 launches each), as one grouped launch per leg, and as a replayed HIP graph of those launches."""
    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import random_heff_problem

    rng = np.random.default_rng(seed)
    chis = [int(rng.choice([32, 48, 64, 96, 128])) for _ in range(nops)]
    ops, xs, ys = [], [], []
    for i, chi in enumerate(chis):
        th, W, L, R = random_heff_problem(2, chi, chi, 5, 5, seed=100 + i)
        ops.append(G.MatVecMPOS([W], L, R, dtype=np.complex128))
        xs.append(G.DeviceArray.from_numpy(th.reshape(-1, order="F")))
        ys.append(G.DeviceArray((th.size,), np.complex128))
    out = {"workload": f"{nops} operators, chi in {{32,48,64,96,128}} (drawn, seed {seed}), d=2, D=5, cx64, {reps} runs"}

    def timed(fn):
        for _ in range(5):
            fn()
        G._lib.call("dmrg_synchronize")
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        G._lib.call("dmrg_synchronize")
        return (time.perf_counter() - t0) / reps

    def sequential():
        for op, x, y in zip(ops, xs, ys):
            op.MultAx_device(x, y)

    t_seq = timed(sequential)
    flop = None
    for name, graph in (("grouped", False), ("grouped_graph", True)):
        grp = G.MatvecGroup(ops, xs, ys, graph=graph)
        flop = grp.flop
        t = timed(grp.run)
        out[name] = {"ms_per_batch": round(t * 1e3, 4), "matvecs_per_s": round(nops / t, 1), "tflops": round(flop / t / 1e12, 3)}
        grp.close()
    out["sequential"] = {"ms_per_batch": round(t_seq * 1e3, 4), "matvecs_per_s": round(nops / t_seq, 1),
                         "tflops": round(flop / t_seq / 1e12, 3)}
    out["speedup_graph_vs_sequential"] = round(out["sequential"]["ms_per_batch"] / out["grouped_graph"]["ms_per_batch"], 2)
    for a in xs + ys:
        a.close()
    for op in ops:
        op.close()
    return out


def section_batched_eigs(nops=32, tol=1e-10, seed=6):
    """Ground states of serving-sized operators (chi in {32,...,128}, d=2, D=5, complex): one dmrg_eigs after
    another against one dmrg_eigs_batched call over all of them (lockstep Davidson, grouped matvecs)."""
    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import random_heff_problem

    rng = np.random.default_rng(seed)
    chis = [int(rng.choice([32, 48, 64, 96, 128])) for _ in range(nops)]
    ops = []
    for i, chi in enumerate(chis):
        th, W, L, R = random_heff_problem(2, chi, chi, 5, 5, seed=200 + i)
        ops.append(G.MatVecMPOS([W], L, R, dtype=np.complex128))
    # warm-up (GEMM plans, code loads)
    for op in ops[:5]:
        op.eigs(nev=1, tol=tol, seed=3)
    G.eigs_batched(ops[:4], tol=tol, seed=3)
    lat, e_seq, mv_seq = [], [], 0
    t0 = time.perf_counter()
    for op in ops:
        t1 = time.perf_counter()
        w, _, info = op.eigs(nev=1, tol=tol, seed=11)
        lat.append(time.perf_counter() - t1)
        e_seq.append(w[0])
        mv_seq += info["matvecs_fp64"]
    t_seq = time.perf_counter() - t0
    t0 = time.perf_counter()
    out = G.eigs_batched(ops, tol=tol, seed=11)
    t_bat = time.perf_counter() - t0
    e_bat = [o[0] for o in out]
    anorm = max(o[2]["anorm"] for o in out)
    # Continuous batching: Poisson arrivals at 0.7x the batched rate, served by dmrg_serve_eigs (requests join and
    # leave the running batch) and, on the same arrival times, by one dmrg_eigs after another in arrival order.
    rate = 0.7 * nops / t_bat
    arrivals = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / rate, nops))
    arrivals -= arrivals[0]
    res_c, st_c = G.serve_eigs(ops, arrivals=arrivals, tol=tol, seed=11)
    lat_s, t0 = [], time.perf_counter()
    for op, a in zip(ops, arrivals):
        wait = a - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        op.eigs(nev=1, tol=tol, seed=11)
        lat_s.append(time.perf_counter() - t0 - a)
    t_fifo = time.perf_counter() - t0
    e_con = [r[0] for r in res_c]
    for op in ops:
        op.close()
    cont = {"arrival_rate_per_s": round(rate, 1),
            "continuous_batching": {"solves_per_s": round(st_c["solves_per_s"], 2),
                                    "p50_ms": round(st_c["p50_latency_s"] * 1e3, 2),
                                    "p99_ms": round(st_c["p99_latency_s"] * 1e3, 2),
                                    "peak_concurrency": st_c["peak_concurrency"],
                                    "all_converged": st_c["completed"] == nops},
            "one_at_a_time": {"solves_per_s": round(nops / t_fifo, 2),
                              "p50_ms": round(float(np.percentile(lat_s, 50)) * 1e3, 2),
                              "p99_ms": round(float(np.percentile(lat_s, 99)) * 1e3, 2)},
            "max_abs_dE_vs_sequential": float(np.max(np.abs(np.array(e_con) - np.array(e_seq))))}
    return {"workload": f"{nops} ground states, chi in {{32,48,64,96,128}} (drawn, seed {seed}), d=2, D=5, cx64, "
                        f"tol {tol:g}*|H|",
            "poisson_arrivals": cont,
            "sequential": {"seconds": round(t_seq, 3), "solves_per_s": round(nops / t_seq, 2),
                           "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
                           "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2), "matvecs": int(mv_seq)},
            "batched": {"seconds": round(t_bat, 3), "solves_per_s": round(nops / t_bat, 2),
                        "all_converged": all(o[2]["nconv"] == 1 for o in out),
                        "matvecs": int(sum(o[2]["matvecs_fp64"] for o in out)),
                        "steps": int(max(o[2]["iters"] for o in out))},
            "speedup_batched_vs_sequential": round(t_seq / t_bat, 2),
            "max_abs_dE_batched_vs_sequential": float(np.max(np.abs(np.array(e_bat) - np.array(e_seq)))),
            "anorm_max": anorm}


def section_eigs_h2(chi=512, tol=1e-10, sigma_rel=1e-3):
    """Eigensolve on the squared operator: the eigenvalue nearest 0 of the shifted H^2, (H - sigma)^2 with
    sigma = E0 - sigma_rel |H| below the ground state (E0 from an FP64 solve of H), FP64 vs mixed precision;
    one cold start each. The folded solution is checked against H: its energy against E0 and its variance."""
    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import folded_problem, random_heff_problem

    th, W, L, R = random_heff_problem(2, chi, chi, 5, 5, seed=2)
    h = G.MatVecMPOS([W], L, R, dtype=np.complex128)
    w, _, info = h.eigs(nev=1, tol=1e-12, seed=7)
    E0, anorm = float(w[0]), info["anorm"]
    sigma = E0 - sigma_rel * anorm
    Wt, Lt, Rt = folded_problem(th, W, L, R, sigma)
    h2 = G.MatVecMPOS([Wt], Lt, Rt, dtype=np.complex128)
    h2.eigs(nev=1, tol=1e-6, seed=5, max_matvecs=50)  # warm-up: GEMM plans of the D=31 operator
    h2.eigs(nev=1, tol=1e-6, seed=5, max_matvecs=50, mixed=True)
    n = h.rows()
    out = {"workload": f"(H - sigma)^2 of the chi={chi} cx64 H_eff (seed 2): MPO^2 (25) + H (5) + identity (1) "
                       f"channels, D_total={Wt.shape[0]}, sigma = E0 - {sigma_rel:g}|H|, tol {tol:g}*|A|, one cold start "
                       f"per mode", "E0_of_H": E0, "sigma": sigma,
           "gflop_per_matvec": round(8.0 * (2 * chi ** 3 * Wt.shape[0] * 2 + chi * chi * Wt.shape[0] ** 2 * 4) / 1e9, 2)}
    dx, dy = G.DeviceArray((n,), np.complex128), G.DeviceArray((n,), np.complex128)
    for mode in ("fp64", "mixed"):
        t0 = time.perf_counter()
        lam, v, i2 = h2.eigs(nev=1, tol=tol, seed=11, mixed=(mode == "mixed"), max_matvecs=8000)
        dt = time.perf_counter() - t0
        x = np.ascontiguousarray(v[:, 0] / np.linalg.norm(v[:, 0]))
        G._lib.call("dmrg_memcpy_htod", dx.ptr, x.ctypes.data_as(__import__("ctypes").c_void_p), x.nbytes)
        h.MultAx_device(dx, dy)
        hx = dy.to_numpy()
        E = float(np.vdot(x, hx).real)
        out[mode] = {"seconds": round(dt, 3), "solves_per_s": round(1.0 / dt, 3), "converged": i2["nconv"] == 1,
                     "eigenvalue": float(lam[0]), "matvecs_fp64": i2["matvecs_fp64"], "matvecs_fp32": i2["matvecs_fp32"],
                     "rel_residual": i2["rnorm"][0] / i2["anorm"], "energy": E, "abs_dE_vs_E0": abs(E - E0),
                     "variance_H": float(np.linalg.norm(hx - E * x) ** 2)}
    dx.close()
    dy.close()
    h.close()
    h2.close()
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["seconds"] / out["mixed"]["seconds"], 3)
    out["abs_dE_mixed_vs_fp64"] = abs(out["mixed"]["energy"] - out["fp64"]["energy"])
    return out


def section_expansion(chi=512, reps=5):
    """Subspace expansion at chi: the optimally mixed block (8 matvecs: 4 with H over D=5 envs, 4 with the
    variance operator over D^2=25 envs) and the L/R expansion terms (energy block with the MPO, variance block
    with the MPO^2)."""
    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import random_heff_problem

    th, W, L, R = random_heff_problem(2, chi, chi, 5, 5, seed=1)
    _, W2, L2, R2 = random_heff_problem(2, chi, chi, 25, 25, seed=2)
    h1 = G.MatVecMPOS([W], L, R, dtype=np.complex128)
    h2 = G.MatVecMPOS([W2], L2, R2, dtype=np.complex128)
    dth = G.DeviceArray.from_numpy(th)
    G.optimally_mixed_block(h1, h2, dth)  # warm-up (GEMM plans)
    G._lib.call("dmrg_synchronize")
    t0 = time.perf_counter()
    for _ in range(reps):
        m, alpha, val = G.optimally_mixed_block(h1, h2, dth)
        m.close()
    ms_mixed = (time.perf_counter() - t0) / reps * 1e3
    out = {"workload": f"theta (2,{chi},{chi}) cx64, H envs D=5, variance envs D=25, seed 1/2",
           "mixed_block": {"ms": round(ms_mixed, 3), "matvecs": 8, "alphas": [round(a, 6) for a in alpha]}}
    dW, dW2 = G.DeviceArray.from_numpy(W), G.DeviceArray.from_numpy(W2)
    dL, dR, dL2 = G.DeviceArray.from_numpy(L), G.DeviceArray.from_numpy(R), G.DeviceArray.from_numpy(L2)
    for name, side, blk, mpo, D in (("term_L_energy", "L", dL, dW, 5), ("term_R_energy", "R", dR, dW, 5),
                                    ("term_L_variance", "L", dL2, dW2, 25)):
        P = G.env_expansion_term(side, blk, dth, mpo, 0.5)
        P.close()
        G._lib.call("dmrg_synchronize")
        t0 = time.perf_counter()
        for _ in range(reps):
            P = G.env_expansion_term(side, blk, dth, mpo, 0.5)
            P.close()
        G._lib.call("dmrg_synchronize")
        ms = (time.perf_counter() - t0) / reps * 1e3
        flop = 8.0 * (chi * chi * D * 2 * chi + 2 * 2 * D * D * chi * chi)  # block.mps GEMM + MPO leg
        out[name] = {"ms": round(ms, 3), "tflops": round(flop / ms / 1e9, 2)}
    for x in (dth, dW, dW2, dL, dR, dL2):
        x.close()
    h1.close()
    h2.close()
    return out


def section_two_site(chi=1024, D=8, steps=10, tol=1e-10):
    """Two-site workload: theta (4, chi, chi) cx64, env blocks (chi, chi, 8) Hermitian per w, two random
    range-6 site MPOs (8, 8, 2, 2) merged into (8, 8, 4, 4), seed 3. Matvec FP64 vs FP32 matrix cores, one
    FP64 and one mixed ground-state solve from theta, then the truncated SVD of the solution as a
    (2 chi) x (2 chi) matrix down to rank chi: |S[chi:]|_2 and sum S[chi:]^2, for each rocSOLVER routine."""
    import ctypes as C

    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import heff_flops, theta_as_matrix, two_site_problem

    th, mpos, L, R = two_site_problem(chi, D, seed=3)
    op = G.MatVecMPOS(mpos, L, R, dtype=np.complex128)
    del L, R
    n = op.rows()
    flops = heff_flops(4, chi, chi, D, D, True)
    out = {"workload": f"two-site theta (4,{chi},{chi}) cx64, env ({chi},{chi},{D}), merged MPO ({D},{D},4,4), seed 3",
           "gflop_per_matvec": round(flops / 1e9, 1)}
    for fp32 in (False, True):
        dt = np.complex64 if fp32 else np.complex128
        dx = G.DeviceArray.from_numpy(th.reshape(n, order="F").astype(dt))
        dy = G.DeviceArray((n,), dt)
        for _ in range(2):
            op.MultAx_device(dx, dy, fp32=fp32)
        G._lib.call("dmrg_synchronize")
        t0 = time.perf_counter()
        for _ in range(steps):
            op.MultAx_device(dx, dy, fp32=fp32)
        G._lib.call("dmrg_synchronize")
        ms = (time.perf_counter() - t0) / steps * 1e3
        tf = flops / (ms * 1e-3) / 1e12
        out["fp32" if fp32 else "fp64"] = {"ms": round(ms, 3), "tflops": round(tf, 3),
                                           "frac_of_peak": round(tf / (PEAK_FP32_MFMA if fp32 else PEAK_FP64_MFMA), 4)}
        dx.close()
        dy.close()
    sol = {}
    vec = None
    for mode in ("fp64", "mixed"):  # warm-up: first-use allocations of the eigensolver bases
        op.eigs(nev=1, tol=tol, x0=th, mixed=(mode == "mixed"), max_matvecs=8)
    for mode in ("fp64", "mixed"):
        t0 = time.perf_counter()
        w, v, info = op.eigs(nev=1, tol=tol, x0=th, mixed=(mode == "mixed"))
        sol[mode] = {"seconds": round(time.perf_counter() - t0, 3), "E0": float(w[0]),
                     "matvecs_fp64": info["matvecs_fp64"], "matvecs_fp32": info["matvecs_fp32"],
                     "rel_residual": info["rnorm"][0] / info["anorm"]}
        if mode == "fp64":
            vec = v[:, 0]
        else:
            vec_mixed = v[:, 0]
    sol["abs_dE_mixed_vs_fp64"] = abs(sol["mixed"]["E0"] - sol["fp64"]["E0"])
    # Schmidt values of the two solutions across the middle bond (the (2 chi) x (2 chi) split matrix)
    sv = {}
    for mode, x in (("fp64", vec), ("mixed", vec_mixed)):
        S = G.svd_truncate(theta_as_matrix(x.reshape((4, chi, chi), order="F")), 1e-300, 0, 3)["S"]
        sv[mode] = S / np.linalg.norm(S)
    sol["schmidt_values_mixed_vs_fp64"] = schmidt_error(sv["fp64"], sv["mixed"])
    out["eigs"] = sol
    op.close()
    M = theta_as_matrix(vec.reshape((4, chi, chi), order="F"))
    m, k = M.shape[0], min(M.shape)
    dM = G.DeviceArray.from_numpy(M)
    dU, dV = G.DeviceArray((m, k), np.complex128), G.DeviceArray((k, M.shape[1]), np.complex128)
    S = np.empty(k)
    svd = {"matrix": f"{m}x{M.shape[1]} complex, keep rank {chi}"}
    for method, name in ((3, "jacobi"), (0, "gesvd")):
        rank, terr, disc = C.c_int64(), C.c_double(), C.c_double()
        t0 = time.perf_counter()
        G._lib.call("dmrg_svd_truncate", G.DMRG_CX64, m, M.shape[1], dM.ptr, 1e-14, chi, method, dU.ptr,
                    S.ctypes.data_as(C.POINTER(C.c_double)), dV.ptr, C.byref(rank), C.byref(terr), C.byref(disc), None)
        svd[name] = {"seconds": round(time.perf_counter() - t0, 3), "rank": rank.value,
                     "truncation_error": terr.value, "discarded_weight": disc.value, "S0": float(S[0])}
    for a in (dM, dU, dV):
        a.close()
    out["svd"] = svd
    return out


def host_buffer_matvec(chi=512, reps=50):
    """PCIe-inclusive rate: the PRIMME-callback form MultAx(x, y) with host buffers (x copied in, y out)."""
    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import heff_flops, random_heff_problem

    th, W, L, R = random_heff_problem(2, chi, chi, 5, 5, seed=1)
    op = G.MatVecMPOS([W], L, R, dtype=np.complex128)
    x = np.ascontiguousarray(th.reshape(-1, order="F"))
    for _ in range(3):
        op.MultAx(x)
    t0 = time.perf_counter()
    for _ in range(reps):
        op.MultAx(x)
    dt = (time.perf_counter() - t0) / reps
    op.close()
    return {"ms_per_matvec": round(dt * 1e3, 4), "tflops": round(heff_flops(2, chi, chi, 5, 5, True) / dt / 1e12, 3),
            "note": "host x and y (8 MiB each) cross PCIe every call; not the headline value"}


def profile_eigs(mode, chi=512, tol=1e-10):
    """Warm-up plus 3 eigensolves of the chi cx64 workload in one mode; the per-solve matvec counts."""
    import numpy as np

    import dmrg_amd as G
    from dmrg_amd.workloads import random_heff_problem

    th, W, L, R = random_heff_problem(2, chi, chi, 5, 5, seed=1)
    op = G.MatVecMPOS([W], L, R, dtype=np.complex128)
    op.eigs(nev=1, tol=tol, seed=999, mixed=(mode == "mixed"))
    out = []
    for k in range(3):
        t0 = time.perf_counter()
        w, _, info = op.eigs(nev=1, tol=tol, seed=1000 + k, mixed=(mode == "mixed"))
        out.append({"ms": round((time.perf_counter() - t0) * 1e3, 2), "E": float(w[0]), "matvecs_fp64": info["matvecs_fp64"],
                    "matvecs_fp32": info["matvecs_fp32"]})
    op.close()
    return {"profile_eigs": mode, "chi": chi, "solves": out}


def vendor_gemm_rates(chi=512):
    """The vendor BLAS (torch.matmul on the GPU, i.e. hipBLASLt/rocBLAS) on plain column-major GEMMs of the
    matvec legs' sizes at chi and chi/2 (leg 1: M = d chi, N = chi D, K = chi; leg 3: M = d chi, N = chi, K = chi D;
    d = 2, D = 5) and on 4096^3, for comparison with the library's own GEMM legs. FLOP: 8 per complex MAC."""
    import torch

    out = {}
    shapes = [(2 * chi, 5 * chi, chi), (2 * chi, chi, 5 * chi), (chi, 5 * chi // 2, chi // 2), (4096, 4096, 4096)]
    for M, N, K in shapes:
        for dt in (torch.complex128, torch.complex64):
            a = torch.randn(M, K, dtype=dt, device="cuda:0")
            b = torch.randn(K, N, dtype=dt, device="cuda:0")
            for _ in range(5):
                c = a @ b
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = 50 if M < 4096 else 10
            e0.record()
            for _ in range(reps):
                c = a @ b
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
            out[f"{M}x{N}x{K}_{str(dt).split('.')[-1]}"] = {"ms": round(ms, 4), "tflops": round(8 * M * N * K / ms / 1e9, 1)}
            del a, b, c
    return {"vendor_gemm": out, "torch": torch.__version__}


def profile_svd(n, seed=5):
    """Warm-up plus 3 Jacobi SVDs (method 3) of a random n x n complex matrix resident on the device."""
    import numpy as np

    import dmrg_amd as G

    rng = np.random.default_rng(seed)
    A = np.asfortranarray(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    dA = G.DeviceArray.from_numpy(A)
    G.svd_truncate(dA, 1e-14, 0, 3)
    ms = []
    for _ in range(3):
        t0 = time.perf_counter()
        G.svd_truncate(dA, 1e-14, 0, 3)
        ms.append(round((time.perf_counter() - t0) * 1e3, 2))
    dA.close()
    return {"profile_svd": n, "ms": ms}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--chi", type=int, default=512)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quick", action="store_true", help="matvec only (no eigensolve, sweep or serving sections)")
    ap.add_argument("--two-site", action="store_true", help="only the two-site chi=1024 section, then exit")
    ap.add_argument("--profile-matvec", action="store_true", help="only the chi=512 FP64 matvec steps, then exit")
    ap.add_argument("--fp32", action="store_true", help="with --profile-matvec: the FP32 matrix-core matvec instead")
    ap.add_argument("--serve-stream", type=int, metavar="N",
                    help="only the serving stream of N requests (mixed precision, capacity then Poisson pass), then exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="one 256 MiB device-to-device copy (scales the HBM byte counters of a profiler run), then exit")
    ap.add_argument("--profile-eigs", choices=("fp64", "mixed"),
                    help="one warm-up and 3 timed chi=512 eigensolves in this mode (for profiler passes), then exit")
    ap.add_argument("--vendor-gemm", action="store_true",
                    help="only the vendor BLAS (torch.matmul) on GEMMs of the matvec legs' sizes, then exit")
    ap.add_argument("--profile-block-vectors", action="store_true",
                    help="only the eigensolver's block-vector kernels at the chi=512 size (for profiler passes), then exit")
    ap.add_argument("--profile-svd", type=int, metavar="N",
                    help="3 Jacobi SVDs of a random N x N complex matrix (for profiler passes), then exit")
    args = ap.parse_args()
    if args.two_site:
        print(json.dumps(section_two_site()))
        return
    if args.profile_matvec:
        # Only the chi=512 matvec steps (FP64, or FP32 with --fp32) for profiler passes; pair with DMRG_TUNE_CACHE
        # to reuse plans.
        r = gpu_matvec_bench(args.chi, args.steps, args.warmup, fp32=args.fp32, legs_nrep=1)
        print(json.dumps({"profile_matvec_ms": round(r["ms_per_step"], 5), "legs_ms": r["legs_ms"]}))
        return
    if args.profile_eigs:
        print(json.dumps(profile_eigs(args.profile_eigs, args.chi)))
        return
    if args.vendor_gemm:
        print(json.dumps(vendor_gemm_rates(args.chi)))
        return
    if args.profile_block_vectors:
        print(json.dumps(section_block_vectors(args.chi)))
        return
    if args.profile_svd:
        print(json.dumps(profile_svd(args.profile_svd)))
        return
    if args.serve_stream:
        print(json.dumps(serve_stream(args.serve_stream)))
        return
    if args.copy_calibration:
        import dmrg_amd as G

        nbytes = 256 << 20
        a, b2 = G.DeviceArray((nbytes // 8,), "float64"), G.DeviceArray((nbytes // 8,), "float64")
        a.zero()
        G._lib.call("dmrg_synchronize")
        G._lib.call("dmrg_memcpy_dtod", b2.ptr, a.ptr, nbytes)
        G._lib.call("dmrg_synchronize")
        print(json.dumps({"copy_bytes": nbytes}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (requests would be "
              f"spread over GPUs one per device; that layer is not built).", file=sys.stderr)
        sys.exit(2)

    import dmrg_amd as G

    dev = G.device_info()
    r64 = gpu_matvec_bench(args.chi, args.steps, args.warmup)
    r32 = gpu_matvec_bench(args.chi, max(args.steps // 2, 1), args.warmup, fp32=True)
    sweep = {}
    for chi in (128, 256):
        if chi != args.chi:
            sweep[str(chi)] = round(gpu_matvec_bench(chi, 200, 10)["tflops"], 3)

    # Dominant kernel: the longest single kernel launch (leg 1, leg 2, or leg 3's main kernel).
    i_dom = max(range(3), key=lambda i: r64["legs_ms"][i])
    macs = leg_macs(2, args.chi, args.chi, 5, 5)
    achieved = r64["legs_tflops"][i_dom]
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_MFMA, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_MFMA, 4), "traffic": load_traffic(f"leg{i_dom + 1}_chi{args.chi}"),
                "kernel": f"gemm leg {i_dom + 1} of 3 (complex FP64, {macs[i_dom]:.3e} MACs per launch)",
                "kernel_ms": round(r64["legs_ms"][i_dom], 4),
                "legs_ms": {"leg1": round(r64["legs_ms"][0], 4), "leg2_mpo": round(r64["legs_ms"][1], 4),
                            "leg3_main": round(r64["legs_ms"][2], 4), "leg3_splitk_reduce": round(r64["legs_ms"][3], 4)}}
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
        "chi_sweep_multivector": multivector_rates(),
        "host_buffer_matvec": host_buffer_matvec(args.chi),
    }
    if not args.quick:
        out["h2_matvec"] = section_h2_matvec(args.chi)
        out["env_update"] = section_env_update(args.chi)
        out["eigs"] = section_eigs(args.chi)
        out["batched_eigs_chi512"] = section_batched_chi512()
        out["block_vectors"] = section_block_vectors(args.chi)
        out["eigs_h2_folded"] = section_eigs_h2(args.chi)
        out["sweep"] = section_sweep(cpu=not args.no_cpu_baseline)
        out["serve"] = section_serve()
        out["grouped_matvecs"] = section_grouped()
        out["batched_eigs"] = section_batched_eigs()
        out["expansion"] = section_expansion(args.chi)
        out["two_site_chi1024"] = section_two_site()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.chi)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
