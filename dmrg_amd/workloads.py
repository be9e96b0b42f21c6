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
