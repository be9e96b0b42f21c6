"""CPU tests that pin the subspace-expansion oracle (oracle/expansion.py).

expansion_term: literal loops over the defining sums with explicit column-major offsets, and an independent
route through the (brute-force pinned) matvec oracle with one-hot environment blocks. optimally_mixed_block:
the MGS basis equals a sign-fixed QR basis, K1/K2 equal dense projections, K3's chosen vector solves the
3 x 3 eigenproblem at its smallest eigenvalue, and the result lies in the Krylov span.
"""
import numpy as np
import pytest

from oracle import expansion as X
from dmrg_amd.workloads import random_heff_problem
from oracle import heff as O


def rand(rng, cplx, *shape):
    a = rng.standard_normal(shape)
    return a + 1j * rng.standard_normal(shape) if cplx else a


SHAPES = [(2, 3, 2, 2, 3), (2, 2, 4, 3, 1), (3, 2, 3, 1, 2), (2, 1, 1, 2, 2), (4, 3, 2, 2, 2)]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("side", ["L", "R"])
def test_expansion_term_matches_loops(side, shape, cplx):
    d, cL, cR, DL, DR = shape
    rng = np.random.default_rng(sum(shape) + (7 if cplx else 0))
    mps, mpo = rand(rng, cplx, d, cL, cR), rand(rng, cplx, DL, DR, d, d)
    block = rand(rng, cplx, cL, cL, DL) if side == "L" else rand(rng, cplx, cR, cR, DR)
    for alpha in (1.0, -0.37):
        P = X.expansion_term(side, block, mps, mpo, alpha)
        ref = X.expansion_term_loops(side, block, mps, mpo, alpha)
        assert P.shape == ref.shape
        np.testing.assert_allclose(P, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    assert not np.any(X.expansion_term(side, block, mps, mpo, 0.0))


@pytest.mark.parametrize("cplx", [False, True])
def test_expansion_term_through_matvec_with_one_hot_envs(cplx):
    """Side L equals H_eff applied with R[b,b',w'] = delta(b,b') delta(w',k); side R's 4-index tensor equals
    H_eff applied with L[a,a',w] = delta(a,a') delta(w,k), placed at s' + d(b' + chiR(a + chiL k))."""
    d, cL, cR, DL, DR = 2, 3, 4, 2, 3
    rng = np.random.default_rng(5)
    mps, mpo = rand(rng, cplx, d, cL, cR), rand(rng, cplx, DL, DR, d, d)
    bL, bR = rand(rng, cplx, cL, cL, DL), rand(rng, cplx, cR, cR, DR)
    PL = X.expansion_term("L", bL, mps, mpo, 1.0)
    for k in range(DR):
        Rk = np.zeros((cR, cR, DR))
        Rk[:, :, k] = np.eye(cR)
        out = O.matrix_vector_product(mps, mpo, bL, Rk)
        np.testing.assert_allclose(PL[:, :, k * cR:(k + 1) * cR], out, rtol=1e-12, atol=1e-12)
    PR = X.expansion_term("R", bR, mps, mpo, 1.0).reshape(-1, order="F")
    for k in range(DL):
        Lk = np.zeros((cL, cL, DL))
        Lk[:, :, k] = np.eye(cL)
        out = O.matrix_vector_product(mps, mpo, Lk, bR)  # out[s', a, b']
        for sp in range(d):
            for bp in range(cR):
                for a in range(cL):
                    assert abs(PR[sp + d * (bp + cR * (a + cL * k))] - out[sp, a, bp]) <= 1e-12 * (1 + abs(out[sp, a, bp]))


def test_modified_gram_schmidt_equals_sign_fixed_qr_and_skips_zero_columns():
    rng = np.random.default_rng(2)
    V = rand(rng, True, 40, 3)
    Q = X.modified_gram_schmidt(V)
    q, r = np.linalg.qr(V)
    ph = np.diag(r) / np.abs(np.diag(r))
    np.testing.assert_allclose(Q, q * ph, atol=1e-12)
    Z = V.copy()
    Z[:, 1] = 0
    Qz = X.modified_gram_schmidt(Z)
    assert not np.any(Qz[:, 1])                                       # kept as it is (zero)
    np.testing.assert_allclose(Qz[:, 0], V[:, 0] / np.linalg.norm(V[:, 0]))
    assert abs(np.vdot(Qz[:, 0], Qz[:, 2])) <= 1e-12                  # projected against column 0 only
    assert abs(np.linalg.norm(Qz[:, 2]) - 1) <= 1e-12


def random_ops(cplx, seed):
    """Two unrelated Hermitian H_eff operators on one shape (H2 need not be H1^2 for the algorithm's steps)."""
    th, W1, L1, R1 = random_heff_problem(2, 4, 5, 3, 3, seed=seed, complex_=cplx)
    _, W2, L2, R2 = random_heff_problem(2, 4, 5, 4, 2, seed=seed + 100, complex_=cplx)
    return th, (W1, L1, R1), (W2, L2, R2)


def lower_view(K):
    return np.tril(K) + np.tril(K, -1).conj().T


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("mode", [1, 2, 3])
def test_optimally_mixed_block_properties(mode, cplx):
    th, (W1, L1, R1), (W2, L2, R2) = random_ops(cplx, seed=10 + mode)
    H1d, H2d = O.dense_heff(W1, L1, R1), O.dense_heff(W2, L2, R2)
    mixed, alpha, val, K1, K2, K3 = X.optimally_mixed_block(th, W1, L1, R1, L2, R2, mode=mode, mpo2=W2)
    # the basis: sign-fixed QR of the Krylov vectors
    psi = th.reshape(-1, order="F")
    q, r = np.linalg.qr(np.stack([psi, H1d @ psi, H2d @ psi], axis=1))
    Q = q * (np.diag(r) / np.abs(np.diag(r)))
    np.testing.assert_allclose(K1, lower_view(Q.conj().T @ H1d @ Q), atol=1e-11 * np.abs(H1d).max())
    np.testing.assert_allclose(K2, lower_view(Q.conj().T @ H2d @ Q), atol=1e-11 * np.abs(H2d).max())
    K3ref = K2 - K1 @ K1
    if not (mode & X.H1):
        K3ref[1, 0] = K3ref[1, 1] = K3ref[2, 1] = 0
    if not (mode & X.H2):
        K3ref[2, 0] = K3ref[2, 1] = K3ref[2, 2] = 0
    np.testing.assert_allclose(K3, K3ref, atol=1e-10 * np.abs(K3ref).max())
    w, Z = np.linalg.eigh(lower_view(K3ref))
    tol = min(max(K3ref[0, 0].real ** 2, np.finfo(float).eps), 1e-10)
    zero_row = any(np.any(np.max(np.abs(K), axis=1) <= tol) for K in (K1, K2, K3ref))
    opt = int(np.argmax(np.abs(Z[0, :]))) if zero_row else 0
    assert abs(val - w[opt]) <= 1e-10 * np.abs(w).max()
    c = np.array(alpha)
    z = Z[:, opt]
    z = z / (z[np.argmax(np.abs(z))] / np.abs(z).max())  # largest component real positive
    np.testing.assert_allclose(c, z.real, atol=1e-9)
    if not cplx:  # real problem: a unit eigenvector of the real symmetric K3
        assert abs(np.linalg.norm(c) - 1) <= 1e-12
    np.testing.assert_allclose(mixed.reshape(-1, order="F"), Q @ c, atol=1e-10)
    # the mixed block lies in span{psi, H1 psi, H2 psi}
    coef, *_ = np.linalg.lstsq(Q, mixed.reshape(-1, order="F"), rcond=None)
    assert np.linalg.norm(Q @ coef - mixed.reshape(-1, order="F")) <= 1e-10


def test_expansion_term_real_cast_branch():
    """EnvBase.cpp:305-307: with real mps and mpo in complex storage the term uses only the real part of a complex
    env block (real arithmetic, then cast to complex); with a complex mps the full complex contraction is used."""
    rng = np.random.default_rng(4)
    d, cL, cR, DL, DR = 2, 5, 4, 3, 2
    block = rng.standard_normal((cL, cL, DL)) + 1j * rng.standard_normal((cL, cL, DL))
    mps = rng.standard_normal((d, cL, cR)).astype(np.complex128)
    mpo = rng.standard_normal((DL, DR, d, d)).astype(np.complex128)
    P = X.expansion_term("L", block, mps, mpo, 0.7)
    assert P.dtype == np.complex128 and not np.any(P.imag)
    assert np.allclose(P, X.expansion_term_loops("L", block.real, mps.real, mpo.real, 0.7), atol=1e-13)
    full = X.expansion_term_loops("L", block, mps, mpo, 0.7)
    assert np.linalg.norm(P - full) > 1e-3 * np.linalg.norm(full)  # the imaginary part of the block is dropped
    mps_c = mps + 1e-3j * rng.standard_normal(mps.shape)
    assert np.allclose(X.expansion_term("L", block, mps_c, mpo, 0.7), X.expansion_term_loops("L", block, mps_c, mpo, 0.7),
                       atol=1e-12)
