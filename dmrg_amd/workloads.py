"""Seeded synthetic inputs for the H_eff hot path (host-side numpy, no compute of the path).

random_heff_problem builds the isolated-matvec workload: theta (d, chiL, chiR) with N(0,1)
entries normalised; env blocks whose [:, :, w] slices are Hermitian (A + A^dagger)/2 with
A ~ N(0,1); a real MPO whose [wL, wR, :, :] blocks are symmetric N(0,1). All arrays are
column-major (Fortran order), the reference's layout.
"""
from __future__ import annotations

import numpy as np


def _herm_blocks(rng, chi: int, D: int, complex_: bool) -> np.ndarray:
    env = np.empty((chi, chi, D), dtype=np.complex128 if complex_ else np.float64, order="F")
    for w in range(D):
        A = rng.standard_normal((chi, chi))
        if complex_:
            A = A + 1j * rng.standard_normal((chi, chi))
        env[:, :, w] = (A + A.conj().T) / 2
    return env


def random_mpo(rng, DL: int, DR: int, d: int) -> np.ndarray:
    W = np.empty((DL, DR, d, d), dtype=np.float64, order="F")
    for i in range(DL):
        for j in range(DR):
            A = rng.standard_normal((d, d))
            W[i, j] = (A + A.T) / 2
    return W


def random_heff_problem(d: int, chiL: int, chiR: int, DL: int, DR: int, seed: int, complex_: bool = True):
    """Returns (theta, mpo, envL, envR) for one H_eff matvec."""
    rng = np.random.default_rng(seed)
    theta = rng.standard_normal((d, chiL, chiR))
    if complex_:
        theta = theta + 1j * rng.standard_normal((d, chiL, chiR))
    theta = np.asfortranarray(theta / np.linalg.norm(theta))
    envL = _herm_blocks(rng, chiL, DL, complex_)
    envR = _herm_blocks(rng, chiR, DR, complex_)
    mpo = random_mpo(rng, DL, DR, d)
    if complex_:
        mpo = mpo.astype(np.complex128)
    return theta, np.asfortranarray(mpo), envL, envR


def heff_macs(d: int, chiL: int, chiR: int, DL: int, DR: int) -> float:
    """Multiply-accumulates of one matvec: d chiL^2 chiR DL + chiL chiR DL DR d^2 + d chiL chiR^2 DR."""
    return float(d) * chiL * chiL * chiR * DL + float(chiL) * chiR * DL * DR * d * d + float(d) * chiL * chiR * chiR * DR


def heff_flops(d: int, chiL: int, chiR: int, DL: int, DR: int, complex_: bool) -> float:
    return heff_macs(d, chiL, chiR, DL, DR) * (8.0 if complex_ else 2.0)


def ising_fields(L: int, w: float, seed: int) -> np.ndarray:
    """Random fields h_i ~ Uniform[-w, w] of the ising_tf_rf model (h_mean = 0, h_wdth = 2w,
    IsingRandomField.cpp:143-144), drawn from numpy's PCG64 with the given seed and passed explicitly."""
    return np.random.default_rng(seed).uniform(-w, w, L)


_PZ = np.array([[1.0, 0.0], [0.0, -1.0]])
_PX = np.array([[0.0, 1.0], [1.0, 0.0]])


def random_range_mpo_site(rng, D: int = 8, xi: float = 1.0) -> np.ndarray:
    """One real site MPO (D, D, 2, 2) of a random l-bit-style chain (cf. LBit.cpp:147-260):
    H = sum_i (h_i Z_i + g_i X_i) + sum_i sum_{r=1}^{D-2} J_{i,r} Z_i Z_{i+r},  J_{i,r} ~ N(0,1) e^{-r/xi},
    h_i, g_i ~ N(0,1). Finite-state layout with rows = left bond, columns = right bond, row D-1 starting
    strings and column 0 ending them (the convention of the ising MPO, IsingRandomField.cpp:107-136):
    W[0,0] = W[D-1,D-1] = I, W[D-1,0] = h Z + g X, W[D-1,r] = J_r Z, W[r,r-1] = I (r >= 2), W[1,0] = Z.
    Every physical block is real symmetric."""
    W = np.zeros((D, D, 2, 2), dtype=np.float64, order="F")
    I2 = np.eye(2)
    W[0, 0] = I2
    W[D - 1, D - 1] = I2
    W[D - 1, 0] = rng.standard_normal() * _PZ + rng.standard_normal() * _PX
    for r in range(1, D - 1):
        W[D - 1, r] = rng.standard_normal() * np.exp(-r / xi) * _PZ
        if r >= 2:
            W[r, r - 1] = I2
    W[1, 0] = _PZ
    return W


def two_site_problem(chi: int, D: int, seed: int):
    """Two-site workload: theta (4, chi, chi) cx64 N(0,1) normalised, env blocks (chi, chi, D) Hermitian per w,
    and the two site MPOs (D, D, 2, 2) of random_range_mpo_site (merged by the operator into (D, D, 4, 4))."""
    rng = np.random.default_rng(seed)
    theta = rng.standard_normal((4, chi, chi)) + 1j * rng.standard_normal((4, chi, chi))
    theta = np.asfortranarray(theta / np.linalg.norm(theta))
    envL = _herm_blocks(rng, chi, D, True)
    envR = _herm_blocks(rng, chi, D, True)
    mpos = [random_range_mpo_site(rng, D), random_range_mpo_site(rng, D)]
    return theta, mpos, envL, envR


def theta_as_matrix(theta: np.ndarray) -> np.ndarray:
    """Two-site theta (d1*d2, chiL, chiR) with d1 fastest (contraction_tblis.cpp:478-505) as the split matrix
    M[(s1, a), (s2, b)] of shape (d1*chiL, d2*chiR) (split.cpp:21-346), for d1 = d2."""
    d = int(round(np.sqrt(theta.shape[0])))
    chiL, chiR = theta.shape[1], theta.shape[2]
    t = theta.reshape((d, d, chiL, chiR), order="F").transpose(0, 2, 1, 3)  # (s1, a, s2, b)
    return np.asfortranarray(t.reshape((d * chiL, d * chiR), order="F"))
