"""GPU eigensolver (Davidson) parity: reference dense fixtures, dense eigh of small H_eff, the Ising
free-fermion ground energy, and residual checks with the oracle matvec at larger sizes.

Tolerances: eigenvalues within 1e-10 (relative to |A|) of the exact value for FP64 and for the mixed
FP32+FP64 mode; residual |Hx - lambda x| <= tol |A| with tol = 1e-12 unless stated.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import dmrg_amd as G
from dmrg_amd.workloads import random_heff_problem
from oracle import heff as O

pytestmark = pytest.mark.gpu
FIX = Path(__file__).resolve().parent / "golden" / "lapacke_matrices.npz"


def dense_operator(A):
    """A as an H_eff with d=1, chiR=1, D=1: out[a'] = sum_a L[a, a'] theta[a], so L = A^T."""
    n = A.shape[0]
    L = np.asfortranarray(A.T.reshape(n, n, 1))
    return G.MatVecMPOS([np.ones((1, 1, 1, 1))], L, np.ones((1, 1, 1)), dtype=np.float64)


@pytest.mark.parametrize("tag", ["1A", "1B", "2A", "2B"])
def test_reference_fixture_matrices(tag):
    d = np.load(FIX)
    A, w_ref, V_ref = d[f"{tag}_matrix"], d[f"{tag}_eigvals"].ravel(), d[f"{tag}_eigvecs"]
    op = dense_operator(A)
    scale = np.abs(w_ref).max()
    w, v, info = op.eigs(nev=1, target="SA", tol=1e-13)
    assert abs(w[0] - w_ref[0]) <= 1e-10 * scale, (w, w_ref[0], info)
    assert abs(abs(np.vdot(v[:, 0], V_ref[:, 0])) - 1) <= 1e-9
    w, v, info = op.eigs(nev=1, target="LA", tol=1e-13)
    assert abs(w[0] - w_ref[-1]) <= 1e-10 * scale
    nev = 3
    w, v, info = op.eigs(nev=nev, target="SA", tol=1e-13, ncv=min(A.shape[0], 12))
    np.testing.assert_allclose(w, w_ref[:nev], atol=1e-10 * scale)
    sigma = 0.5 * (w_ref[2] + w_ref[3]) + 0.1 * (w_ref[3] - w_ref[2])
    w, v, info = op.eigs(nev=1, target="CLOSEST", shift=sigma, tol=1e-13, ncv=A.shape[0])
    assert abs(w[0] - w_ref[np.argmin(np.abs(w_ref - sigma))]) <= 1e-10 * scale
    op.close()


def dense_from_oracle(th, W, L, R):
    n = th.size
    H = np.empty((n, n), dtype=th.dtype)
    for j in range(n):
        e = np.zeros(n, dtype=th.dtype)
        e[j] = 1
        H[:, j] = O.matrix_vector_product(e.reshape(th.shape, order="F"), W, L, R).reshape(-1, order="F")
    return H


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("precond", [False, True])
def test_random_heff_ground_state(cplx, mixed, precond):
    th, W, L, R = random_heff_problem(2, 20, 24, 5, 5, seed=2, complex_=cplx)
    H = dense_from_oracle(th, W, L, R)
    wd, vd = np.linalg.eigh(H)
    op = G.MatVecMPOS([W], L, R, dtype=th.dtype)
    w, v, info = op.eigs(nev=1, tol=1e-12, mixed=mixed, precond=precond)
    anorm = np.abs(wd).max()
    assert info["nconv"] == 1, info
    assert abs(w[0] - wd[0]) <= 1e-10 * anorm, (w[0], wd[0], info)
    x = v[:, 0]
    assert np.linalg.norm(H @ x - w[0] * x) <= 1e-11 * anorm
    if mixed:
        assert info["matvecs_fp32"] > 0
    op.close()


@pytest.mark.parametrize("L", [6, 10])
def test_ising_full_chain_free_fermion(L):
    rng = np.random.default_rng(L)
    h = rng.uniform(-1, 1, L)
    mpos = O.ising_tf_rf_mpos(L, 1.0, 1.0, h)
    one = np.ones((1, 1, 1))
    op = G.MatVecMPOS(mpos, one, one, dtype=np.float64)
    assert op.rows() == 2 ** L
    w, v, info = op.eigs(nev=1, tol=1e-13)
    e0 = O.ising_free_fermion_ground_energy(1.0, 1.0, h)
    assert abs(w[0] - e0) <= 1e-10 * abs(e0), (w[0], e0, info)
    S = O.dense_entropies(v[:, 0], L)
    np.testing.assert_allclose(S, O.ising_free_fermion_entropies(1.0, 1.0, h), atol=1e-8)
    op.close()


@pytest.mark.parametrize("target", ["SA", "LA"])
@pytest.mark.parametrize("inner_max", [0, 5, 20])
@pytest.mark.parametrize("cplx", [False, True])
def test_mixed_correction_solves(target, inner_max, cplx):
    """Mixed solves whose FP64 phase expands with FP32 correction-equation CG steps (inner_max > 0) or with
    plain residuals (0): same eigenpair as the dense reference at both ends of the spectrum."""
    print(f"[case] mixed target={target} inner_max={inner_max} cplx={cplx}", file=sys.stderr, flush=True)
    th, W, L, R = random_heff_problem(2, 20, 24, 5, 5, seed=4, complex_=cplx)
    H = dense_from_oracle(th, W, L, R)
    wd = np.linalg.eigvalsh(H)
    ref = wd[0] if target == "SA" else wd[-1]
    anorm = np.abs(wd).max()
    op = G.MatVecMPOS([W], L, R, dtype=th.dtype)
    try:
        w, v, info = op.eigs(nev=1, target=target, tol=1e-12, mixed=True, inner_max=inner_max)
        assert info["nconv"] == 1, info
        assert abs(w[0] - ref) <= 1e-10 * anorm, (w[0], ref, info)
        x = v[:, 0]
        assert np.linalg.norm(H @ x - w[0] * x) <= 1e-11 * anorm
    finally:
        op.close()


@pytest.mark.parametrize("mixed,inner_max", [(False, None), (True, 0), (True, 20)])
def test_large_residual_with_oracle(mixed, inner_max):
    """chi=128 complex (n=32768): residual and Rayleigh quotient checked with the oracle matvec."""
    th, W, L, R = random_heff_problem(2, 128, 128, 5, 5, seed=1)
    op = G.MatVecMPOS([W], L, R, dtype=np.complex128)
    w, v, info = op.eigs(nev=1, tol=1e-12, mixed=mixed, inner_max=inner_max)
    x = v[:, 0].reshape(th.shape, order="F")
    hx = O.matrix_vector_product(x, W, L, R)
    lam = np.vdot(x, hx).real / np.vdot(x, x).real
    assert abs(lam - w[0]) <= 1e-10 * info["anorm"]
    assert np.linalg.norm(hx - w[0] * x) <= 1e-11 * info["anorm"], info
    assert info["nconv"] == 1
    op.close()


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7])
@pytest.mark.parametrize("ncv", [0, 2, 3])
def test_tiny_problems_basis_limits(n, ncv):
    """Regression: tiny problems where the restart would fill the basis (n=4 gives a 2-vector basis by the
    size rule) must stay within the basis and still converge."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A = A + A.T
    op = dense_operator(A)
    w, v, info = op.eigs(nev=1, tol=1e-13, ncv=ncv)
    wd = np.linalg.eigvalsh(A)
    assert abs(w[0] - wd[0]) <= 1e-10 * np.abs(wd).max(), (w, wd, info)
    op.close()
