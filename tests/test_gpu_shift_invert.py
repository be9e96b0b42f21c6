"""Shift-invert on the GPU: MatVecMPOS::FactorOP / MultOPv (matvec_mpos.cpp:887-975) against numpy solves with the
dense H_eff of the oracle. set_shift adds -sigma I to the MPO's bottom-left block (matvec_mpos.cpp:1433-1472), which
is -sigma times the identity of H_eff when the env blocks hold the identity at the MPO's edge channels (left block
at w = DL - 1, right block at w = 0), as on a chain; the test problems are built that way. Bars: relative residual
of (H - sigma) y = x at most 1e-10 (LU, LDLT, and LLT with H - sigma positive definite); error paths as the
reference's (no factor, NONE, not positive definite)."""
import sys

import numpy as np
import pytest

import dmrg_amd as G
from dmrg_amd.workloads import random_heff_problem
from oracle import heff as O

pytestmark = pytest.mark.gpu


def chain_edges(L, R):
    """Identity blocks at the MPO's edge channels, so that the set_shift corner shifts H_eff by -sigma I."""
    L, R = L.copy(order="F"), R.copy(order="F")
    L[:, :, -1] = np.eye(L.shape[0])
    R[:, :, 0] = np.eye(R.shape[0])
    return L, R


def rel_residual(A, y, x):
    return np.linalg.norm(A @ y - x) / (np.linalg.norm(A, 2) * np.linalg.norm(y))


@pytest.mark.parametrize("fact", ["LU", "LDLT", "LLT"])
@pytest.mark.parametrize("cplx", [False, True])
def test_factor_op_mult_opv(cplx, fact):
    print(f"[case] shift-invert cplx={cplx} fact={fact}", file=sys.stderr, flush=True)
    dt = np.complex128 if cplx else np.float64
    th, W, L, R = random_heff_problem(2, 10, 12, 5, 5, seed=41, complex_=cplx)
    L, R = chain_edges(L, R)
    H = O.dense_heff(W.astype(dt), L.astype(dt), R.astype(dt))
    w = np.linalg.eigvalsh(0.5 * (H + H.conj().T))
    n = H.shape[0]
    sigma = w[0] - 0.5 if fact == "LLT" else w[n // 2] + 0.4 * (w[n // 2 + 1] - w[n // 2])
    op = G.MatVecMPOS([W], L, R, dtype=dt)
    try:
        assert not op.isReadyFactorOp()
        op.set_shift(sigma)
        op.FactorOP(fact)
        assert op.isReadyFactorOp()
        rng = np.random.default_rng(2)
        X = rng.standard_normal((n, 3)) + (1j * rng.standard_normal((n, 3)) if cplx else 0)
        Y = op.MultOPv(X)
        A = H - sigma * np.eye(n)
        for j in range(3):
            assert rel_residual(A, Y[:, j], X[:, j]) <= 1e-10
        assert np.allclose(Y, np.linalg.solve(A, X), rtol=0, atol=1e-8 * np.abs(np.linalg.solve(A, X)).max())
        # MultAx of the same operator applies H - sigma
        assert np.linalg.norm(op.MultAx(Y[:, 0]) - X[:, 0]) <= 1e-9 * np.linalg.norm(X[:, 0])
    finally:
        op.close()


def test_shift_invert_error_paths():
    th, W, L, R = random_heff_problem(2, 8, 8, 3, 3, seed=42, complex_=False)
    L, R = chain_edges(L, R)
    H = O.dense_heff(W, L, R).real
    w = np.linalg.eigvalsh(0.5 * (H + H.T))
    n = H.shape[0]
    op = G.MatVecMPOS([W], L, R, dtype=np.float64)
    try:
        x = np.ones(n)
        with pytest.raises(G.DmrgError):
            op.MultOPv(x)  # no FactorOP yet
        op.FactorOP("NONE")  # ready without a factor, as the reference
        assert op.isReadyFactorOp()
        with pytest.raises(G.DmrgError) as e:
            op.MultOPv(x)
        assert e.value.status == 7
        op.set_shift(0.5 * (w[0] + w[-1]))  # indefinite H - sigma; set_shift clears the factor
        assert not op.isReadyFactorOp()
        with pytest.raises(G.DmrgError) as e:
            op.FactorOP("LLT")
        assert e.value.status == 2 and "positive definite" in str(e.value)
        assert not op.isReadyFactorOp()
        with pytest.raises(G.DmrgError) as e:
            op.FactorOP("QR")
        assert e.value.status == 7
        op.FactorOP("LU")
        y = op.MultOPv(x)
        A = H - 0.5 * (w[0] + w[-1]) * np.eye(n)
        assert rel_residual(A, y, x) <= 1e-10
    finally:
        op.close()
