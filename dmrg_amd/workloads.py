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


def lbit_site_couplings(rng, xi: float = 1.0) -> dict:
    """Random couplings of one site of the l-bit-style chain (cf. LBit.cpp:147-260), amplitudes N(0,1) e^{-r/xi}
    with r the range of the term: 1-body h Z + g X (r = 0), 2-body J2[r-1] Z_i Z_{i+r} for r = 1, 2, 3, and
    3-body K1 Z_i Z_{i+1} Z_{i+2} (r = 2) and K2 Z_i Z_{i+2} Z_{i+3} (r = 3), the terms that start at this site."""
    h, g = rng.standard_normal(), rng.standard_normal()
    J2 = [rng.standard_normal() * np.exp(-r / xi) for r in (1, 2, 3)]
    K1 = rng.standard_normal() * np.exp(-2 / xi)
    K2 = rng.standard_normal() * np.exp(-3 / xi)
    return {"h": h, "g": g, "J2": J2, "K1": K1, "K2": K2}


def lbit_site_mpo(c: dict) -> np.ndarray:
    """The real site MPO (8, 8, 2, 2) of lbit_site_couplings c. Finite-state layout with rows = left bond,
    columns = right bond, row 7 starting strings and column 0 ending them (the convention of the ising MPO,
    IsingRandomField.cpp:107-136): W[0,0] = W[7,7] = I; 1-body W[7,0] = h Z + g X; 2-body W[7,r] = J2_r Z,
    W[r,r-1] = I (r = 2, 3), W[1,0] = Z; 3-body W[7,5] = K1 Z, W[7,6] = K2 Z, W[6,5] = I, W[5,4] = Z, W[4,0] = Z.
    Lower triangular; every physical block is real symmetric."""
    W = np.zeros((8, 8, 2, 2), dtype=np.float64, order="F")
    I2 = np.eye(2)
    W[0, 0] = W[7, 7] = I2
    W[7, 0] = c["h"] * _PZ + c["g"] * _PX
    for r in (1, 2, 3):
        W[7, r] = c["J2"][r - 1] * _PZ
    W[2, 1] = W[3, 2] = I2
    W[1, 0] = _PZ
    W[7, 5] = c["K1"] * _PZ
    W[7, 6] = c["K2"] * _PZ
    W[6, 5] = I2
    W[5, 4] = _PZ
    W[4, 0] = _PZ
    return W


def random_range_mpo_site(rng, D: int = 8, xi: float = 1.0) -> np.ndarray:
    """One real site MPO (8, 8, 2, 2) of the l-bit-style chain with random 1- to 3-body couplings
    (lbit_site_couplings, lbit_site_mpo)."""
    if D != 8:
        raise ValueError("the l-bit-style site MPO has bond dimension 8")
    return lbit_site_mpo(lbit_site_couplings(rng, xi))


def lbit_dense_hamiltonian(couplings: list) -> np.ndarray:
    """Dense Hamiltonian of an open chain with the given per-site couplings, summed term by term from Kronecker
    products of Pauli matrices (site 0 is the fastest index, the layout of the multisite merge): the check of
    lbit_site_mpo. Terms that would reach past the chain end are dropped, as the MPO's edges drop them."""
    L = len(couplings)

    def op(ops: dict) -> np.ndarray:
        out = np.ones((1, 1))
        for i in range(L):
            out = np.kron(ops.get(i, np.eye(2)), out)  # site 0 fastest
        return out

    H = np.zeros((2 ** L, 2 ** L))
    for i, c in enumerate(couplings):
        H += op({i: c["h"] * _PZ + c["g"] * _PX})
        for r in (1, 2, 3):
            if i + r < L:
                H += c["J2"][r - 1] * op({i: _PZ, i + r: _PZ})
        if i + 2 < L:
            H += c["K1"] * op({i: _PZ, i + 1: _PZ, i + 2: _PZ})
        if i + 3 < L:
            H += c["K2"] * op({i: _PZ, i + 2: _PZ, i + 3: _PZ})
    return H


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


def squared_problem(th, W, L, R):
    """Exact H^2 of an isolated workload: MPO^2 of MpoSite.cpp:57-69 and squared env blocks
    L2[:, :, w1 + D w2] = L[:, :, w1] L[:, :, w2] (R likewise), so that H2 theta = H (H theta)."""
    D = W.shape[0]
    t = np.tensordot(W, W.conj(), axes=([3], [2])).transpose(0, 3, 1, 4, 2, 5)
    W2 = np.asfortranarray(t.reshape(D * D, D * D, W.shape[2], W.shape[3], order="F"))
    L2 = np.empty((L.shape[0], L.shape[0], D * D), dtype=L.dtype, order="F")
    R2 = np.empty((R.shape[0], R.shape[0], D * D), dtype=R.dtype, order="F")
    for a in range(D):
        for b in range(D):
            L2[:, :, a + D * b] = L[:, :, a] @ L[:, :, b]
            R2[:, :, a + D * b] = R[:, :, a] @ R[:, :, b]
    return W2, L2, R2


def folded_problem(th, W, L, R, sigma):
    """(H - sigma)^2 = H^2 - 2 sigma H + sigma^2 of an isolated workload as one operator: a block-diagonal MPO of
    the MPO^2 (D^2 channels), -2 sigma W (D channels) and sigma^2 I (one channel), with env blocks stacked the
    same way (squared blocks, the blocks, the identity). D_total = D^2 + D + 1."""
    W2, L2, R2 = squared_problem(th, W, L, R)
    D, DD, chi_l, chi_r = W.shape[0], W2.shape[0], L.shape[0], R.shape[0]
    Dt = DD + D + 1
    Wt = np.zeros((Dt, Dt, W.shape[2], W.shape[3]), dtype=np.complex128, order="F")
    Wt[:DD, :DD] = W2
    Wt[DD:DD + D, DD:DD + D] = -2.0 * sigma * W
    Wt[Dt - 1, Dt - 1] = sigma * sigma * np.eye(W.shape[2])

    def stack(B, B2, chi):
        out = np.zeros((chi, chi, Dt), dtype=np.complex128, order="F")
        out[:, :, :DD], out[:, :, DD:DD + D], out[:, :, Dt - 1] = B2, B, np.eye(chi)
        return out

    return Wt, stack(L, L2, chi_l), stack(R, R2, chi_r)
