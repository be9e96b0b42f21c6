"""CPU test of the host dense Hermitian eigensolver (used for Rayleigh-Ritz) against numpy.linalg.eigh."""
import ctypes as C

import numpy as np
import pytest

from dmrg_amd import _lib


def heev(A):
    m = A.shape[0]
    a = np.asfortranarray(A.astype(np.complex128))
    w = np.empty(m)
    Z = np.empty((m, m), dtype=np.complex128, order="F")
    _lib.call("dmrg_host_heev", m, a.ctypes.data_as(C.POINTER(C.c_double)), w.ctypes.data_as(C.POINTER(C.c_double)),
              Z.ctypes.data_as(C.POINTER(C.c_double)))
    return w, Z


@pytest.mark.parametrize("m", [1, 2, 3, 5, 16, 33, 64, 120])
@pytest.mark.parametrize("cplx", [False, True])
def test_heev_matches_eigh(m, cplx):
    rng = np.random.default_rng(m)
    A = rng.standard_normal((m, m)) + (1j * rng.standard_normal((m, m)) if cplx else 0)
    A = A + A.conj().T
    w, Z = heev(A)
    w_ref = np.linalg.eigh(A)[0]
    np.testing.assert_allclose(w, w_ref, atol=1e-12 * np.abs(w_ref).max())
    np.testing.assert_allclose(Z.conj().T @ Z, np.eye(m), atol=1e-12)
    np.testing.assert_allclose(A @ Z, Z * w, atol=1e-11 * np.abs(w_ref).max())


def test_heev_degenerate_and_diagonal():
    A = np.diag([3.0, 1.0, 1.0, -2.0, 0.0])
    w, Z = heev(A)
    np.testing.assert_allclose(w, [-2, 0, 1, 1, 3], atol=1e-15)
    np.testing.assert_allclose(np.abs(Z.conj().T @ Z), np.eye(5), atol=1e-14)
    w, Z = heev(np.zeros((4, 4)))
    np.testing.assert_allclose(w, 0)
    # clustered spectrum
    rng = np.random.default_rng(1)
    Q, _ = np.linalg.qr(rng.standard_normal((20, 20)) + 1j * rng.standard_normal((20, 20)))
    lam = np.concatenate([np.full(10, 1.0), 1.0 + 1e-13 * np.arange(10)])
    A = (Q * lam) @ Q.conj().T
    w, Z = heev(A)
    np.testing.assert_allclose(w, np.sort(lam), atol=1e-13)
    np.testing.assert_allclose(A @ Z, Z * w, atol=1e-13)


def test_heev_wilkinson_and_graded():
    """Wilkinson's W21+ (pairs of eigenvalues equal to ~1e-14, the classic hard case for the shift strategy), a
    block-diagonal matrix whose tridiagonal form splits, and a graded spectrum at m = 256, against numpy."""
    n = 21
    W = np.diag(np.abs(np.arange(n) - 10).astype(float)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1)
    w, Z = heev(W)
    np.testing.assert_allclose(w, np.linalg.eigvalsh(W), atol=1e-13)
    np.testing.assert_allclose(W @ Z, Z * w, atol=1e-12)
    np.testing.assert_allclose(Z.conj().T @ Z, np.eye(n), atol=1e-12)
    B = np.zeros((12, 12), dtype=complex)
    rng = np.random.default_rng(3)
    for s in (slice(0, 5), slice(5, 6), slice(6, 12)):
        k = s.stop - s.start
        X = rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k))
        B[s, s] = X + X.conj().T
    w, Z = heev(B)
    np.testing.assert_allclose(w, np.linalg.eigvalsh(B), atol=1e-12 * np.abs(w).max())
    np.testing.assert_allclose(B @ Z, Z * w, atol=1e-11 * np.abs(w).max())
    m = 256
    Q, _ = np.linalg.qr(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))
    lam = np.sign(rng.standard_normal(m)) * np.logspace(0, -10, m)
    A = (Q * lam) @ Q.conj().T
    A = 0.5 * (A + A.conj().T)
    w, Z = heev(A)
    np.testing.assert_allclose(w, np.linalg.eigvalsh(A), atol=1e-13)
    np.testing.assert_allclose(A @ Z, Z * w, atol=1e-12)
    np.testing.assert_allclose(Z.conj().T @ Z, np.eye(m), atol=1e-12)
