"""Batched eigensolves (dmrg_eigs_batched): many operators of uneven shapes solved together in lockstep, against
dense eigenvalues of each operator (oracle matvec columns) and the single-problem solver.

Tolerances: eigenvalues within 1e-10 * |A| of the dense ones; residuals |Hx - lambda x| <= 1e-11 |A| (solver
tol 1e-12); the same eigenvalues as dmrg_eigs within 1e-10 * |A|.
"""
import sys

import numpy as np
import pytest

import dmrg_amd as G
from dmrg_amd.workloads import random_heff_problem
from oracle import heff as O

pytestmark = pytest.mark.gpu

SHAPES = [(2, 8, 8, 3, 3), (2, 12, 10, 5, 5), (2, 6, 15, 4, 2), (4, 6, 7, 3, 3), (2, 14, 14, 5, 5), (2, 16, 9, 2, 5),
          (2, 3, 4, 3, 3), (2, 13, 11, 5, 4)]


def dense(th, W, L, R):
    return O.dense_heff(W, L, R)


def build(cplx, shapes, seed=0):
    probs, ops = [], []
    dt = np.complex128 if cplx else np.float64
    for i, sh in enumerate(shapes):
        th, W, L, R = random_heff_problem(*sh, seed=seed + i, complex_=cplx)
        if not cplx:
            L, R = np.ascontiguousarray(L.real), np.ascontiguousarray(R.real)
        probs.append((th.astype(dt), W.astype(dt), L.astype(dt), R.astype(dt)))
        ops.append(G.MatVecMPOS([W], L, R, dtype=dt))
    return probs, ops


@pytest.mark.parametrize("target", ["SA", "LA"])
@pytest.mark.parametrize("cplx", [False, True])
def test_batched_matches_dense(cplx, target):
    print(f"[case] batched cplx={cplx} target={target}", file=sys.stderr, flush=True)
    probs, ops = build(cplx, SHAPES)
    try:
        out = G.eigs_batched(ops, tol=1e-12, target=target)
        for (th, W, L, R), (lam, vec, info) in zip(probs, out):
            H = dense(th, W, L, R)
            wd = np.linalg.eigvalsh(H)
            ref = wd[0] if target == "SA" else wd[-1]
            anorm = np.abs(wd).max()
            assert info["nconv"] == 1, info
            assert abs(lam - ref) <= 1e-10 * anorm, (th.shape, lam, ref, info)
            assert np.linalg.norm(H @ vec - lam * vec) <= 1e-11 * anorm * max(1.0, np.linalg.norm(vec))
            assert abs(np.linalg.norm(vec) - 1) <= 1e-10
    finally:
        for op in ops:
            op.close()


def test_batched_matches_single_solver_and_start_vectors():
    """Larger problems (n up to 2 x 64 x 64), some with start vectors, converging at different steps so that the
    batch shrinks and its grouped launches are rebuilt."""
    shapes = [(2, 64, 64, 5, 5), (2, 32, 48, 5, 5), (2, 16, 16, 5, 5), (2, 48, 40, 5, 5), (2, 8, 8, 5, 5), (4, 24, 24, 5, 5)]
    probs, ops = build(True, shapes, seed=20)
    try:
        x0s = [None, probs[1][0], None, probs[3][0], None, None]
        out = G.eigs_batched(ops, tol=1e-12, x0s=x0s)
        iters = [info["iters"] for _, _, info in out]
        assert len(set(iters)) > 1  # the problems left the batch at different steps
        for op, (lam, vec, info) in zip(ops, out):
            w, v, i1 = op.eigs(nev=1, tol=1e-12)
            assert info["nconv"] == 1 and i1["nconv"] == 1
            assert abs(lam - w[0]) <= 1e-10 * info["anorm"], (lam, w[0])
    finally:
        for op in ops:
            op.close()


def test_batched_single_problem_and_errors():
    probs, ops = build(False, SHAPES[:1])
    try:
        (lam, vec, info), = G.eigs_batched(ops, tol=1e-12)
        wd = np.linalg.eigvalsh(dense(*probs[0]))
        assert abs(lam - wd[0]) <= 1e-10 * np.abs(wd).max()
    finally:
        for op in ops:
            op.close()
    a, opsa = build(True, SHAPES[:1])
    b, opsb = build(False, SHAPES[1:2])
    try:
        with pytest.raises(G.DmrgError):
            G.eigs_batched(opsa + opsb)  # one dtype per batch
    finally:
        for op in opsa + opsb:
            op.close()


def overlap(a, b):
    return abs(np.vdot(a, b)) / (np.linalg.norm(a) * np.linalg.norm(b))


@pytest.mark.parametrize("cplx", [False, True])
def test_serve_eigs_alone_vs_mix(cplx):
    """Continuous batching: requests arriving over time, joining and leaving a running batch (at most 3 at once),
    give the eigenpairs each request gets when solved alone (eigenvalues within 1e-10 |A|, eigenvectors with
    overlap 1 - 1e-9), and the dense ones (the single-problem solver's for the two larger operators)."""
    print(f"[case] serve_eigs alone-vs-mix cplx={cplx}", file=sys.stderr, flush=True)
    shapes = SHAPES + [(2, 32, 40, 5, 5), (2, 48, 48, 5, 5)]
    probs, ops = build(cplx, shapes, seed=40)
    try:
        alone = [G.eigs_batched([op], tol=1e-12)[0] for op in ops]
        arrivals = [0.004 * i for i in range(len(ops))]
        mix, stats = G.serve_eigs(ops, arrivals=arrivals, max_batch=3, tol=1e-12)
        assert stats["completed"] == len(ops) and stats["failed"] == 0, stats
        assert 1 <= stats["peak_concurrency"] <= 3
        assert stats["p50_latency_s"] <= stats["p99_latency_s"]
        for i, ((th, W, L, R), (la, va, ia), (lm, vm, im)) in enumerate(zip(probs, alone, mix)):
            if i < len(SHAPES):
                wd = np.linalg.eigvalsh(dense(th, W, L, R))
                ref, anorm = wd[0], np.abs(wd).max()
            else:
                w1, _, i1 = ops[i].eigs(nev=1, tol=1e-12)
                ref, anorm = w1[0], i1["anorm"]
            assert im["nconv"] == 1
            assert abs(lm - la) <= 1e-10 * anorm and abs(lm - ref) <= 1e-10 * anorm, (lm, la, ref)
            assert overlap(va, vm) >= 1 - 1e-9
            assert im["start_s"] >= im["arrival_s"] - 1e-9 and im["finish_s"] >= im["start_s"]
            assert abs(im["latency_s"] - (im["finish_s"] - im["arrival_s"])) <= 1e-12
    finally:
        for op in ops:
            op.close()


def test_serve_eigs_budget_admission():
    """The HBM budget bounds the working memory of the running requests: with room for two of the largest
    requests at a time, no more than that is ever admitted, and every request still completes."""
    print("[case] serve_eigs budget", file=sys.stderr, flush=True)
    shapes = [(2, 48, 48, 5, 5)] * 3 + [(2, 16, 16, 5, 5)] * 3
    probs, ops = build(True, shapes, seed=60)
    try:
        out, stats = G.serve_eigs(ops, tol=1e-12)  # all at t = 0, unlimited budget
        fp = [info["footprint_bytes"] for _, _, info in out]
        assert stats["peak_concurrency"] == len(ops) and stats["peak_admitted_bytes"] == sum(fp)
        budget = 2 * max(fp)
        out2, stats2 = G.serve_eigs(ops, hbm_budget_bytes=budget, tol=1e-12)
        assert stats2["completed"] == len(ops)
        assert stats2["peak_admitted_bytes"] <= budget and stats2["peak_concurrency"] < len(ops)
        for st in (stats, stats2):  # throughput counts converged solves, as dmrg_serve
            assert abs(st["solves_per_s"] * st["makespan_s"] - st["completed"]) <= 1e-6 * st["completed"]
        for (l1, v1, _), (l2, v2, _) in zip(out, out2):
            assert abs(l1 - l2) <= 1e-10 * max(1.0, abs(l1)) and overlap(v1, v2) >= 1 - 1e-9
        with pytest.raises(G.DmrgError):
            G.serve_eigs(ops[:1], hbm_budget_bytes=0)
        # One request larger than the whole budget fails the call up front (DMRG_ERR_OOM), even when it would
        # arrive into an empty batch.
        small = min(fp)
        with pytest.raises(G.DmrgError) as ei:
            G.serve_eigs([ops[3], ops[0]], hbm_budget_bytes=small, tol=1e-12)
        assert ei.value.status == 3
    finally:
        for op in ops:
            op.close()


@pytest.mark.parametrize("target", ["SA", "LA"])
@pytest.mark.parametrize("cplx", [False, True])
def test_batched_mixed_matches_dense_and_fp64(cplx, target):
    """Mixed-precision batch (FP32 phase on the FP32 matrix cores, then FP64): the dense eigenvalues within
    1e-10 |A|, residuals at the FP64 tolerance, eigenvectors equal to the FP64 batch's (overlap 1 - 1e-9), and
    both phases' matvecs counted."""
    print(f"[case] batched mixed cplx={cplx} target={target}", file=sys.stderr, flush=True)
    shapes = SHAPES + [(2, 48, 40, 5, 5), (2, 64, 64, 5, 5)]
    probs, ops = build(cplx, shapes, seed=70)
    try:
        mix = G.eigs_batched(ops, tol=1e-12, target=target, mixed=True)
        f64 = G.eigs_batched(ops, tol=1e-12, target=target)
        for (th, W, L, R), (lam, vec, info), (l64, v64, i64) in zip(probs, mix, f64):
            n = vec.shape[0]
            if n <= 1024:
                wd = np.linalg.eigvalsh(dense(th, W, L, R))
                ref, anorm = (wd[0] if target == "SA" else wd[-1]), np.abs(wd).max()
            else:
                ref, anorm = l64, i64["anorm"]
            hv = O.matrix_vector_product(vec.reshape(th.shape, order="F"), W, L, R).reshape(-1, order="F")
            assert np.linalg.norm(hv - lam * vec) <= 1e-11 * anorm
            assert info["nconv"] == 1, info
            assert abs(lam - ref) <= 1e-10 * anorm and abs(lam - l64) <= 1e-10 * anorm, (lam, ref, l64)
            assert overlap(vec, v64) >= 1 - 1e-9
            assert info["matvecs_fp32"] > 0 and info["matvecs_fp64"] > 0, info
    finally:
        for op in ops:
            op.close()


def test_serve_eigs_mixed_alone_vs_mix():
    """Continuous batching in mixed precision: requests arriving over time (at most 3 at once) give the eigenpairs
    they get alone in mixed precision and the FP64 ones (eigenvalues within 1e-10 |A|, overlap 1 - 1e-9)."""
    print("[case] serve_eigs mixed alone-vs-mix", file=sys.stderr, flush=True)
    shapes = SHAPES[:5] + [(2, 32, 40, 5, 5), (2, 48, 48, 5, 5)]
    probs, ops = build(True, shapes, seed=90)
    try:
        alone = [G.eigs_batched([op], tol=1e-12, mixed=True)[0] for op in ops]
        f64 = [G.eigs_batched([op], tol=1e-12)[0] for op in ops]
        arrivals = [0.003 * i for i in range(len(ops))]
        mix, stats = G.serve_eigs(ops, arrivals=arrivals, max_batch=3, tol=1e-12, mixed=True)
        assert stats["completed"] == len(ops) and stats["failed"] == 0, stats
        for (la, va, ia), (lm, vm, im), (l64, v64, i64) in zip(alone, mix, f64):
            anorm = i64["anorm"]
            assert im["nconv"] == 1 and im["matvecs_fp32"] > 0
            assert abs(lm - la) <= 1e-10 * anorm and abs(lm - l64) <= 1e-10 * anorm, (lm, la, l64)
            assert overlap(va, vm) >= 1 - 1e-9 and overlap(vm, v64) >= 1 - 1e-9
    finally:
        for op in ops:
            op.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_batched_large_problems_solved_alone(mixed):
    """A batch with problems of n >= 2^17 (chi = 256): those are solved one after another by the single-problem
    solver, the small ones in the lockstep batch; every eigenvalue equals the single FP64 solver's within
    1e-10 |A|, and in mixed precision the large problems report FP32 matvecs."""
    print(f"[case] batched large alone mixed={mixed}", file=sys.stderr, flush=True)
    shapes = [(2, 256, 256, 5, 5), (2, 16, 16, 5, 5), (2, 256, 256, 5, 5), (2, 24, 20, 5, 5)]
    probs, ops = build(True, shapes, seed=80)
    try:
        out = G.eigs_batched(ops, tol=1e-10, mixed=mixed)
        for op, (lam, vec, info) in zip(ops, out):
            w, _, i1 = op.eigs(nev=1, tol=1e-10)
            assert info["nconv"] == 1 and i1["nconv"] == 1, info
            assert abs(lam - w[0]) <= 1e-10 * i1["anorm"], (lam, w[0])
            if mixed:
                assert info["matvecs_fp32"] > 0, info
            if op.rows() >= 131072 and mixed:
                assert info["matvecs_fp64"] < 30, info  # correction-equation expansions: few FP64 matvecs
    finally:
        for op in ops:
            op.close()
