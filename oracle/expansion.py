"""Oracle (test infrastructure only): CPU restatement of the reference's subspace-expansion kernels.

Only tests/ may import this module, and only as the checker. The product path (dmrg_amd) never calls it.

  expansion_term            - EnvBase::get_expansion_term(mps, mpo, alpha) (EnvBase.cpp:303-361): the
                              enrichment block P built from one env block, the site tensor and the MPO.
                              Pinned by literal index loops over the defining sums (expansion_term_loops)
                              and, on the L side, by P being the H_eff-like action on a product with an
                              identity right block.
  modified_gram_schmidt     - env.cpp:74-93 (and linalg::matrix::modified_gram_schmidt, matrix.cpp:24-44).
  optimally_mixed_block     - env.cpp:96-239 (get_optimally_mixed_block_1site): the 3x3 "variance matrix"
                              Rayleigh-Ritz over span{psi, H psi, H^2 psi}. Pinned by properties that the
                              mathematics fixes: the result lies in the span, its K3 Rayleigh quotient is
                              the smallest eigenvalue of K3, K3 is the variance matrix when the H^2 operator
                              is the square of H, and an exact eigenstate gives alpha_mps = +-1.

Layout notes. Tensors are Eigen column-major; numpy arrays here hold the same index order and are
flattened with order="F" wherever the reference reshapes. Eigen's contract() returns the free indices of
the first operand, then of the second, as numpy.tensordot does.

Sign convention. The reference takes the real part of Eigen's eigenvector of K3; Eigen fixes neither the
sign nor, for complex K3, the phase. Here (and on the GPU) the eigenvector is first rotated so that its
largest-magnitude component is real and positive. Parity of the mixed block and of the three alphas is
therefore up to one overall sign for real problems; the complex-phase choice is parity unpinned.
"""
from __future__ import annotations

import numpy as np

from oracle.heff import matrix_vector_product

H1 = 1  # EnvExpandMode::H1 (source/config/enums.h:224-234)
H2 = 2  # EnvExpandMode::H2


def expansion_term(side: str, block, mps, mpo, alpha: float):
    """P = alpha * (block . mps . mpo), EnvBase.cpp:303-361.

    side 'L': block (chiL, chiL, DL); P[s', a', (b, wR)] with dims (d, chiL, chiR * DR).
    side 'R': block (chiR, chiR, DR); the reference's contraction gives R2[b', a, wL, s'], shuffles it to
    [s', b', a, wL] and reshapes that column-major buffer to (d, chiL * DL, chiR) without reordering.
    """
    if np.iscomplexobj(mps) or np.iscomplexobj(mpo) or np.iscomplexobj(block):
        # EnvBase.cpp:305-307: complex storage whose mps and mpo are real is contracted in real arithmetic with the
        # real part of the (complex) env block, and cast back to complex.
        if not np.any(np.imag(mps)) and not np.any(np.imag(mpo)):
            return expansion_term(side, np.real(block), np.real(mps), np.real(mpo), alpha).astype(np.complex128)
    d, cL, cR = mps.shape
    if side == "L":
        if block.shape[0] != cL:
            raise ValueError("block dimension 0 must equal mps chiL")
        shape = (d, cL, cR * mpo.shape[1])
        if alpha == 0:
            return np.zeros(shape, dtype=np.result_type(block, mps, mpo))
        t = np.tensordot(block, mps, axes=([0], [1]))        # (a', w, s, b)        idx({0},{1})
        t = np.tensordot(t, mpo, axes=([1, 2], [0, 2]))      # (a', b, wR, s')      idx({1,2},{0,2})
        t = t.transpose(3, 0, 1, 2)                          # shuffle {3,0,1,2}: (s', a', b, wR)
    elif side == "R":
        if block.shape[0] != cR:
            raise ValueError("block dimension 0 must equal mps chiR")
        shape = (d, cL * mpo.shape[0], cR)
        if alpha == 0:
            return np.zeros(shape, dtype=np.result_type(block, mps, mpo))
        t = np.tensordot(block, mps, axes=([0], [2]))        # (b', w, s, a)        idx({0},{2})
        t = np.tensordot(t, mpo, axes=([1, 2], [1, 2]))      # (b', a, wL, s')      idx({1,2},{1,2})
        t = t.transpose(3, 0, 1, 2)                          # shuffle {3,0,1,2}: (s', b', a, wL)
    else:
        raise ValueError("side must be 'L' or 'R'")
    P = t.reshape(-1, order="F").reshape(shape, order="F")  # column-major reshape
    return P if alpha == 1.0 else P * alpha


def expansion_term_loops(side: str, block, mps, mpo, alpha: float):
    """The same block by explicit loops over the defining sums and explicit column-major offsets (test pin)."""
    d, cL, cR = mps.shape
    DL, DR = mpo.shape[0], mpo.shape[1]
    dt = np.result_type(block, mps, mpo)
    if side == "L":
        flat = np.zeros(d * cL * cR * DR, dtype=dt)
        for sp in range(d):
            for ap in range(cL):
                for b in range(cR):
                    for wr in range(DR):
                        acc = 0
                        for a in range(cL):
                            for w in range(DL):
                                for s in range(d):
                                    acc += block[a, ap, w] * mps[s, a, b] * mpo[w, wr, s, sp]
                        flat[sp + d * (ap + cL * (b + cR * wr))] = acc
        return alpha * flat.reshape((d, cL, cR * DR), order="F")
    flat = np.zeros(d * cL * DL * cR, dtype=dt)
    for sp in range(d):
        for bp in range(cR):
            for a in range(cL):
                for wl in range(DL):
                    acc = 0
                    for b in range(cR):
                        for w in range(DR):
                            for s in range(d):
                                acc += block[b, bp, w] * mps[s, a, b] * mpo[wl, w, s, sp]
                    flat[sp + d * (bp + cR * (a + cL * wl))] = acc
    return alpha * flat.reshape((d, cL * DL, cR), order="F")


def modified_gram_schmidt(V):
    """env.cpp:74-93: columns orthonormalised in order; a column of norm < eps is kept as it is and later
    columns are not projected against it."""
    V = np.array(V, copy=True)
    Q = np.zeros_like(V)
    for i in range(V.shape[1]):
        Q[:, i] = V[:, i]
        rii = np.linalg.norm(Q[:, i])
        if abs(rii) < np.finfo(np.float64).eps:
            continue
        Q[:, i] /= rii
        for j in range(i + 1, V.shape[1]):
            rij = np.vdot(Q[:, i], V[:, j])  # Eigen dot: conjugate-linear in the first argument
            V[:, j] -= Q[:, i] * rij
    return Q


def _lower_hermitian(K):
    """Eigen's selfAdjointView<Lower>(): the lower triangle mirrored."""
    L = np.tril(K)
    return L + np.tril(K, -1).conj().T


def phase_fixed_real(v):
    """Rotate v so that its largest-magnitude component is real and positive, then take the real part."""
    k = int(np.argmax(np.abs(v)))
    ph = v[k] / abs(v[k]) if abs(v[k]) > 0 else 1.0
    return np.real(v / ph)


def optimally_mixed_block(theta, mpo, envL1, envR1, envL2, envR2, mode: int = H1 | H2, mpo2=None):
    """get_optimally_mixed_block_1site (env.cpp:96-239). H1 uses (mpo, envL1, envR1), H2 uses
    (mpo2 or mpo, envL2, envR2): the reference builds both operators on model.get_mpo(sites) with the
    energy and the variance environments. Returns (mixed theta, (alpha_mps, alpha_h1v, alpha_h2v), optVal,
    K1, K2, K3 as used)."""
    mpo2 = mpo if mpo2 is None else mpo2
    shape = theta.shape

    def h1(x):
        return matrix_vector_product(x.reshape(shape, order="F"), mpo, envL1, envR1).reshape(-1, order="F")

    def h2(x):
        return matrix_vector_product(x.reshape(shape, order="F"), mpo2, envL2, envR2).reshape(-1, order="F")

    v0 = theta.reshape(-1, order="F")
    V = np.stack([v0, h1(v0), h2(v0)], axis=1)
    V = modified_gram_schmidt(V)
    H1V = np.stack([h1(V[:, i]) for i in range(3)], axis=1)
    H2V = np.stack([h2(V[:, i]) for i in range(3)], axis=1)
    K1 = _lower_hermitian(V.conj().T @ H1V)
    K2 = _lower_hermitian(V.conj().T @ H2V)
    K3 = K2 - K1 @ K1
    if not (mode & H1) and K3[2, 2] != 0.0:
        K3[1, 0] = K3[1, 1] = K3[2, 1] = 0.0
    if not (mode & H2) and K3[1, 1] != 0.0:
        K3[2, 0] = K3[2, 1] = K3[2, 2] = 0.0
    evals, evecs = np.linalg.eigh(_lower_hermitian(K3))  # ascending, like Eigen's SelfAdjointEigenSolver
    tol = min(max(K3[0, 0].real ** 2, np.finfo(np.float64).eps), 1e-10)
    zero_rows = max(int(np.sum(np.max(np.abs(K), axis=1) <= tol)) for K in (K1, K2, K3))
    opt = 0
    if zero_rows > 0:
        opt = int(np.argmax(np.abs(evecs[0, :])))  # the column that changes the current state the least
    col = phase_fixed_real(evecs[:, opt])
    mixed = (V @ col).reshape(shape, order="F")
    return mixed, tuple(float(c) for c in col), float(evals[opt]), K1, K2, K3
