"""Oracle (test infrastructure only): CPU finite-DMRG sweep for ising_tf_rf, the restatement the GPU sweep is
checked against and the CPU baseline that bench.py times on the host cores.

Follows fdmrg::run_algorithm / update_state (fdmrg.cpp:170-376) with the same structure as the GPU driver:
two-site sweeps that grow the bond dimension (truncation rule svd.cpp:257-280), then single-site sweeps to an
energy variance below var_tol, the variance measured with MPO^2 built from energy-shifted MPOs
(MpoSite.cpp:57-69, set_energy_shift_mpo). Eigenproblems of size n <= 4096 are solved densely (LAPACK, as the
reference's eig_max_size branch, settings.h:249); larger ones with ARPACK (scipy eigsh), which is also one of the
reference's eigensolver backends (solver_arpack.cpp). Contractions are oracle.heff's.

Pinned by the free-fermion closed form (tests/test_oracle_sweep.py).
"""
from __future__ import annotations

import time

import numpy as np
import scipy.sparse.linalg as sla

from . import heff as O


def _mpo_squared_edges(L, J, g, h, e_shift):
    out = []
    for i in range(L):
        W = O.ising_tf_rf_mpo_bulk(0.0 if i == L - 1 else J, g + float(h[i]), e_shift)
        W2 = O.mpo_squared(W)
        if i == 0:
            W2 = W2[8:9]          # row (2, 2) of the squared left edge
        if i == L - 1:
            W2 = W2[:, 0:1]       # column (0, 0) of the squared right edge
        out.append(np.asfortranarray(W2.real.copy()))
    return out


def _ground_state(W, envL, envR, x0, tol):
    d, cL, cR = W.shape[2], envL.shape[0], envR.shape[0]
    n = d * cL * cR
    shape = (d, cL, cR)
    if n <= 4096:
        H = np.einsum("aAw,wWsS,bBW->SABsab", envL, W, envR).reshape(n, n, order="F")
        H = 0.5 * (H + H.T)
        w, v = np.linalg.eigh(H)
        return w[0], v[:, 0].reshape(shape, order="F")
    op = sla.LinearOperator((n, n), matvec=lambda x: O.matrix_vector_product(
        x.reshape(shape, order="F"), W, envL, envR).reshape(-1, order="F"), dtype=np.float64)
    w, v = sla.eigsh(op, k=1, which="SA", v0=x0.reshape(-1, order="F"), tol=tol, ncv=min(n - 1, 32))
    return w[0], v[:, 0].reshape(shape, order="F")


def _rank(S, lim, rank_max):
    """svd.cpp:257-280."""
    terr = np.array([np.linalg.norm(S[k:]) for k in range(len(S) + 1)])
    r = int(np.count_nonzero(terr >= lim))
    r = min(r, min(len(S), rank_max) if rank_max > 0 else len(S))
    return max(r, min(len(S), 1)), terr[max(r, min(len(S), 1))]


def _entropy(S):
    p = S ** 2
    p = p[p > 0]
    return float(-np.sum(p * np.log(p)))


def fdmrg_ising(L, h, J=1.0, g=1.0, chi_max=64, var_tol=1e-12, trunc_lim=1e-14, two_site_sweeps=10, max_sweeps=20,
                eigs_tol=1e-12, seed=0):
    t0 = time.perf_counter()
    d = 2
    W = [np.asfortranarray(w.real.copy()) for w in O.ising_tf_rf_mpos(L, J, g, h)]
    rng = np.random.default_rng(seed)
    M = []
    for _ in range(L):
        v = rng.uniform(-1, 1, d)
        v[0] += 1.5
        M.append((v / np.linalg.norm(v)).reshape(d, 1, 1))
    EL = [None] * (L + 1)
    ER = [None] * (L + 1)
    EL[0] = np.ones((1, 1, 1))
    ER[L - 1] = np.ones((1, 1, 1))
    for i in range(L - 1, 0, -1):
        ER[i - 1] = O.env_build_block("R", ER[i], M[i], W[i])
    ent = np.zeros(L + 1)

    def two_site(i, right):
        a, b = M[i].shape[1], M[i + 1].shape[2]
        th = np.einsum("sak,tkb->stab", M[i], M[i + 1]).reshape(d * d, a, b, order="F")
        Wm = O.merge_mpos([W[i], W[i + 1]]).real
        E, th = _ground_state(Wm, EL[i], ER[i + 1], th, eigs_tol)
        X = th.reshape(d, d, a, b, order="F").transpose(0, 2, 1, 3).reshape(d * a, d * b, order="F")
        U, S, Vt = np.linalg.svd(X, full_matrices=False)
        r, _ = _rank(S, trunc_lim, chi_max)
        S = S[:r] / np.linalg.norm(S[:r])
        U, Vt = U[:, :r], Vt[:r]
        if right:
            M[i] = U.reshape(d, a, r, order="F")
            M[i + 1] = (S[:, None] * Vt).reshape(r, d, b, order="F").transpose(1, 0, 2)
            EL[i + 1] = O.env_build_block("L", EL[i], M[i], W[i])
        else:
            M[i] = (U * S).reshape(d, a, r, order="F")
            M[i + 1] = Vt.reshape(r, d, b, order="F").transpose(1, 0, 2)
            ER[i] = O.env_build_block("R", ER[i + 1], M[i + 1], W[i + 1])
        ent[i + 1] = _entropy(S)
        return E

    E, Eprev = 0.0, np.inf
    sweeps = 0
    for _ in range(two_site_sweeps):
        for i in range(L - 1):
            E = two_site(i, True)
        for i in range(L - 2, -1, -1):
            E = two_site(i, False)
        sweeps += 1
        if abs(E - Eprev) <= 1e-12 * max(1.0, abs(E)):
            break
        Eprev = E

    # single-site sweeps with variance from the shifted H^2
    es = E / L
    W2 = _mpo_squared_edges(L, J, g, h, es)
    EL2 = [None] * (L + 1)
    ER2 = [None] * (L + 1)
    EL2[0] = np.ones((1, 1, 1))
    ER2[L - 1] = np.ones((1, 1, 1))
    for i in range(L - 1, 0, -1):
        ER2[i - 1] = O.env_build_block("R", ER2[i], M[i], W2[i])

    def one_site(c, right):
        E, th = _ground_state(W[c], EL[c], ER[c], M[c], eigs_tol)
        e2 = O.expectation_value(th, W2[c], EL2[c], ER2[c])
        var = e2 - (E - es * L) ** 2
        dd, a, b = th.shape
        if right and c + 1 < L:
            U, S, Vt = np.linalg.svd(th.reshape(d * a, b, order="F"), full_matrices=False)
            k = len(S)
            M[c] = U.reshape(d, a, k, order="F")
            M[c + 1] = np.einsum("kb,sbe->ske", S[:, None] * Vt, M[c + 1])
            EL[c + 1] = O.env_build_block("L", EL[c], M[c], W[c])
            EL2[c + 1] = O.env_build_block("L", EL2[c], M[c], W2[c])
            ent[c + 1] = _entropy(S)
        elif not right and c > 0:
            X = th.transpose(1, 0, 2).reshape(a, d * b, order="F")
            U, S, Vt = np.linalg.svd(X, full_matrices=False)
            k = len(S)
            M[c] = Vt.reshape(k, d, b, order="F").transpose(1, 0, 2)
            M[c - 1] = np.einsum("sza,ak->szk", M[c - 1], U * S)
            ER[c - 1] = O.env_build_block("R", ER[c], M[c], W[c])
            ER2[c - 1] = O.env_build_block("R", ER2[c], M[c], W2[c])
            ent[c] = _entropy(S)
        else:
            M[c] = th
        return E, var

    var = np.inf
    for _ in range(max_sweeps):
        for c in range(L):
            E, var = one_site(c, True)
        for c in range(L - 1, -1, -1):
            E, var = one_site(c, False)
        sweeps += 1
        if var < var_tol:
            break
    return {"energy": float(E), "variance": float(var), "entropies": ent.copy(), "sweeps": sweeps,
            "chi": [1] + [M[i].shape[2] for i in range(L)], "seconds": time.perf_counter() - t0}
