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
