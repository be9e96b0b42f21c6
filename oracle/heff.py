"""Oracle (test infrastructure only): CPU restatement of the reference's H_eff hot path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this
module, and only as the checker. The product path (dmrg_amd) never calls it.

Every function restates a reference routine with the same index order (Eigen column-major
tensors; numpy tensordot keeps Eigen's contraction output order: free indices of the first
operand, then of the second). Parity pins for each function are in the tests:
  matrix_vector_product   - literal 6-index brute force (dense_heff), both branches agree,
                            Hermiticity, D=1/identity-env special case, full-chain H.
  env_build_block          - normalised-MPS identity blocks, env-contracted energy = full-chain
                            energy of a dense state vector.
  expectation_value        - brute force sum, = <theta|H theta>.
  merge_mpos / mpo_squared - Kronecker identities against independently built dense H, H^2.
  ising_tf_rf_mpos         - dense Hamiltonian built from Pauli Kronecker products.
"""
from __future__ import annotations

import numpy as np

# Pauli matrices (qm/spin/half.cpp:10-33).
SX = np.array([[0, 1], [1, 0]], dtype=np.complex128)
SY = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
SZ = np.array([[1, 0], [0, -1]], dtype=np.complex128)
ID2 = np.eye(2, dtype=np.complex128)


def matrix_vector_product(mps, mpo, envL, envR, order: str | None = None):
    """H_eff . mps  (contraction_tblis.cpp:213-292, Eigen branch :278-290).

    order 'L' contracts theta.envL first (taken when chiL >= chiR), 'R' theta.envR first."""
    if mps.shape[1] != envL.shape[0] or mps.shape[2] != envR.shape[0] or mps.shape[0] != mpo.shape[2] \
            or envL.shape[2] != mpo.shape[0] or envR.shape[2] != mpo.shape[1]:
        raise ValueError("Dimension mismatch")  # contraction_tblis.cpp:230-234
    if order is None:
        order = "L" if mps.shape[1] >= mps.shape[2] else "R"
    if order == "L":
        t = np.tensordot(mps, envL, axes=([1], [0]))        # (s, b, a', w)     idx({1},{0})
        t = np.tensordot(t, mpo, axes=([3, 0], [0, 2]))      # (b, a', w', s')   idx({3,0},{0,2})
        t = np.tensordot(t, envR, axes=([0, 2], [0, 2]))     # (a', s', b')      idx({0,2},{0,2})
        return np.ascontiguousarray(t.transpose(1, 0, 2))  # shuffle {1,0,2}
    t = np.tensordot(mps, envR, axes=([2], [0]))            # (s, a, b', w')    idx({2},{0})
    t = np.tensordot(t, mpo, axes=([3, 0], [1, 2]))          # (a, b', w, s')    idx({3,0},{1,2})
    t = np.tensordot(t, envL, axes=([0, 2], [0, 2]))         # (b', s', a')      idx({0,2},{0,2})
    return np.ascontiguousarray(t.transpose(1, 2, 0))      # shuffle {1,2,0}


def env_build_block(side: str, other, mps, mpo):
    """EnvBase::build_block (EnvBase.cpp:67-151)."""
    if side == "L":
        if mps.shape[0] != mpo.shape[2] or mps.shape[1] != other.shape[0] or mpo.shape[0] != other.shape[2]:
            raise ValueError("dimension mismatch")
        bmm = np.tensordot(other, mps.conj(), axes=([1], [1]))   # (a, w, s', b')
        bmm = np.tensordot(bmm, mpo, axes=([1, 2], [0, 3]))      # (a, b', w', s)
        return np.tensordot(mps, bmm, axes=([0, 1], [3, 0]))     # (b, b', w')
    if side == "R":
        if mps.shape[0] != mpo.shape[2] or mps.shape[2] != other.shape[0] or mpo.shape[1] != other.shape[2]:
            raise ValueError("dimension mismatch")
        bmm = np.tensordot(other, mps.conj(), axes=([1], [2]))   # (b, w', s', a')
        bmm = np.tensordot(bmm, mpo, axes=([1, 2], [1, 3]))      # (b, a', w, s)
        return np.tensordot(mps, bmm, axes=([0, 2], [3, 0]))     # (a, a', w)
    raise ValueError("side must be L or R")


def expectation_value(mps, mpo, envL, envR):
    """contraction_tblis.cpp:64-69: envL.mps.mpo.conj(mps).envR."""
    t = np.tensordot(envL, mps, axes=([0], [1]))             # (a', w, s, b)
    t = np.tensordot(t, mpo, axes=([2, 1], [2, 0]))          # (a', b, w', s')
    t = np.tensordot(t, mps.conj(), axes=([3, 0], [0, 1]))   # (b, w', b')
    return np.tensordot(t, envR, axes=([0, 2, 1], [0, 1, 2])).item()


def dense_heff(mpo, envL, envR):
    """H[(s',a',b'),(s,a,b)] = sum_{w,w'} L[a,a',w] W[w,w',s,s'] R[b,b',w'] by explicit loops
    (contract_mpo_env, TensorsFinite.cpp:119-142). Row/col index flattened with s fastest."""
    d, cL, cR = mpo.shape[2], envL.shape[0], envR.shape[0]
    DL, DR = mpo.shape[0], mpo.shape[1]
    n = d * cL * cR
    H = np.zeros((n, n), dtype=np.result_type(mpo, envL, envR))
    for b in range(cR):
        for a in range(cL):
            for s in range(d):
                col = s + d * (a + cL * b)
                for bp in range(cR):
                    for ap in range(cL):
                        for sp in range(d):
                            acc = 0
                            for w in range(DL):
                                for wp in range(DR):
                                    acc += envL[a, ap, w] * mpo[w, wp, s, sp] * envR[b, bp, wp]
                            H[sp + d * (ap + cL * bp), col] = acc
    return H


def merge_mpos(mpos):
    """Pre-contract a chain of MPOs into one (DL, DR, prod d, prod d), first site fastest
    (matvec_mpos.cpp:104-119: contract {1},{0}; shuffle {0,3,1,4,2,5})."""
    cur = mpos[0]
    for nxt in mpos[1:]:
        t = np.tensordot(cur, nxt, axes=([1], [0]))              # (wL, s1, s1', wR, s2, s2')
        t = t.transpose(0, 3, 1, 4, 2, 5)                          # (wL, wR, s1, s2, s1', s2')
        dl, dr = t.shape[0], t.shape[1]
        dd = t.shape[2] * t.shape[3]
        cur = t.reshape(dl, dr, dd, dd, order="F")
    return cur


def mpo_squared(mpo):
    """W2[(w1,w2),(w1',w2'),s,s''] = sum_s' W[w1,w1',s,s'] conj(W)[w2,w2',s',s''] (MpoSite.cpp:57-69)."""
    t = np.tensordot(mpo, mpo.conj(), axes=([3], [2]))           # (w1, w1', s, w2, w2', s'')
    t = t.transpose(0, 3, 1, 4, 2, 5)                              # (w1, w2, w1', w2', s, s'')
    D0, D1 = mpo.shape[0] ** 2, mpo.shape[1] ** 2
    return t.reshape(D0, D1, mpo.shape[2], mpo.shape[3], order="F")


def ising_tf_rf_mpo_bulk(J: float, g_i: float, e_shift: float = 0.0):
    """IsingRandomField::get_mpo (IsingRandomField.cpp:107-136): rows wL, cols wR; W[i,j,:,:] = O."""
    W = np.zeros((3, 3, 2, 2), dtype=np.complex128)
    W[0, 0] = ID2
    W[1, 0] = SZ
    W[2, 0] = -g_i * SX - e_shift * ID2
    W[2, 1] = -J * SZ
    W[2, 2] = ID2
    return W


def ising_tf_rf_mpos(L: int, J: float, g: float, h_rand, e_shift: float = 0.0):
    """Site MPOs with edges applied (MpoSite.cpp:20-26, 450-578): site 0 keeps row D-1, site L-1 keeps col 0.
    Field on site i is g + h_rand[i] on sigma^x (IsingRandomField.cpp:30); J on the last site is 0
    (IsingRandomField::set_averages). Returns list of (DL, DR, 2, 2) cx64 arrays."""
    mpos = []
    for i in range(L):
        Ji = 0.0 if i == L - 1 else J
        W = ising_tf_rf_mpo_bulk(Ji, g + float(h_rand[i]), e_shift)
        if i == 0:
            W = W[2:3, :, :, :]
        if i == L - 1:
            W = W[:, 0:1, :, :]
        mpos.append(np.asfortranarray(W))
    return mpos


def ising_dense(L: int, J: float, g: float, h_rand):
    """H = -sum J sz_i sz_{i+1} - sum (g+h_i) sx_i from Kronecker products, basis index s_0 fastest."""
    def op_at(O, i):
        out = np.ones((1, 1), dtype=np.complex128)
        for j in reversed(range(L)):  # site L-1 is the slowest index -> leftmost factor
            out = np.kron(out, O if j == i else ID2)
        return out
    n = 2 ** L
    H = np.zeros((n, n), dtype=np.complex128)
    for i in range(L - 1):
        H -= J * op_at(SZ, i) @ op_at(SZ, i + 1)
    for i in range(L):
        H -= (g + float(h_rand[i])) * op_at(SX, i)
    return H


def chain_dense_from_mpos(mpos):
    """Dense operator of an MPO chain in the matvec convention: out(s') = sum_s W[.., s, s'] theta(s),
    i.e. the matrix M[s', s] of merge_mpos(...)[0, 0].T (edges are 1-dimensional)."""
    Wm = merge_mpos(mpos)
    assert Wm.shape[0] == 1 and Wm.shape[1] == 1
    return Wm[0, 0].T


# ---------------------------------------------------------------------------------- free fermions

def ising_free_fermion_energies(J: float, g: float, h_rand):
    """Single-particle energies eps_k of H = -sum J_i sz_i sz_{i+1} - sum g_i sx_i: the singular values of the
    L x L bidiagonal B with B_ii = g_i and B_{i,i+1} = J_i. Ground energy E0 = -sum eps_k."""
    L = len(h_rand)
    B = np.zeros((L, L))
    for i in range(L):
        B[i, i] = g + float(h_rand[i])
        if i + 1 < L:
            B[i, i + 1] = J
    return np.linalg.svd(B, compute_uv=False)


def ising_free_fermion_ground_energy(J: float, g: float, h_rand) -> float:
    return -float(np.sum(ising_free_fermion_energies(J, g, h_rand)))


def _majorana_matrix(J: float, g: float, h_rand):
    """H = (i/4) sum_mn M_mn gamma_m gamma_n with sx_i = i g_{2i} g_{2i+1}, sz_i sz_{i+1} = i g_{2i+1} g_{2i+2}."""
    L = len(h_rand)
    M = np.zeros((2 * L, 2 * L))
    for i in range(L):
        gi = g + float(h_rand[i])
        M[2 * i, 2 * i + 1] = -2 * gi
        M[2 * i + 1, 2 * i] = 2 * gi
        if i + 1 < L:
            M[2 * i + 1, 2 * i + 2] = -2 * J
            M[2 * i + 2, 2 * i + 1] = 2 * J
    return M


def ising_free_fermion_entropies(J: float, g: float, h_rand):
    """Ground-state entanglement entropies S(l) of the first l sites, l = 0..L (Peschel):
    Gamma = <i g_m g_n> = -polar(M); eigenvalues of Gamma restricted to the first 2l Majoranas are +-i nu_k;
    S = sum_k h((1 + nu_k)/2), h(x) = -x ln x - (1-x) ln(1-x). Natural log as measure.cpp:136."""
    L = len(h_rand)
    M = _majorana_matrix(J, g, h_rand)
    U, s, Vt = np.linalg.svd(M)
    Gam = -(U @ Vt)
    out = [0.0]
    for l in range(1, L + 1):
        G = Gam[: 2 * l, : 2 * l]
        nu = np.abs(np.linalg.eigvals(G).imag)
        nu = np.sort(nu)[::2]  # each nu appears as a +- pair
        S = 0.0
        for v in nu:
            for x in ((1 + v) / 2, (1 - v) / 2):
                if x > 1e-300:
                    S -= x * np.log(x)
        out.append(float(S))
    out[-1] = 0.0 if abs(out[-1]) < 1e-10 else out[-1]
    return np.array(out)


def dense_entropies(psi: np.ndarray, L: int, d: int = 2):
    """Entropies S(l), l=0..L of a dense state vector with site 0 the fastest index."""
    out = [0.0]
    for l in range(1, L):
        m = psi.reshape(d ** l, d ** (L - l), order="F")
        s = np.linalg.svd(m, compute_uv=False)
        p = s ** 2
        p = p[p > 1e-300]
        out.append(float(-np.sum(p * np.log(p))))
    out.append(0.0)
    return np.array(out)


def entanglement_entropy(schmidt) -> float:
    """measure::entanglement_entropy (measure.cpp:134-138): -sum s^2 ln s^2."""
    p = np.abs(np.asarray(schmidt)) ** 2
    p = p[p > 0]
    return float(-np.sum(p * np.log(p)))


def matrix_vector_product_chain(mps, mpos, envL, envR):
    """H_eff . mps for a list of MPOs applied one at a time (contraction_tblis.cpp:307-420, the path MatVecMPOS
    takes for more than 5 MPOs, matvec_mpos.cpp:120-123). mpos unshuffled, each (DL_i, DR_i, d_i, d_i).

    t carries the indices (s_cur, s_rest, b, a', s_done, w): theta.envL first (:358), then MPO i contracts s_cur
    and w (idx {0,5},{0,2} of the {2,3,0,1}-shuffled MPO, :398) and the result is reshaped so that the next site's
    input comes first and this site's output joins s_done as its slowest part (:392-399); finally .envR over
    (b, w') and shuffle {1,0,2} (:417-419)."""
    ds = [m.shape[2] for m in mpos]
    d = int(np.prod(ds))
    if mps.shape[0] != d or mps.shape[1] != envL.shape[0] or mps.shape[2] != envR.shape[0] \
            or envL.shape[2] != mpos[0].shape[0] or envR.shape[2] != mpos[-1].shape[1]:
        raise ValueError("Dimension mismatch")
    cL, cR = envL.shape[0], envR.shape[0]
    t = np.tensordot(mps, envL, axes=([1], [0]))                                 # (s, b, a', w)
    t = t.reshape(ds[0], d // ds[0], cR, cL, 1, envL.shape[2], order="F")          # (s_0, rest, b, a', 1, w)
    for i, W in enumerate(mpos):
        shf = W.transpose(2, 3, 0, 1)                                              # (s, s', wL, wR)
        t = np.tensordot(t, shf, axes=([0, 5], [0, 2]))                            # (rest, b, a', done, s', wR)
        rest, done = t.shape[0], t.shape[3]
        nxt = ds[i + 1] if i + 1 < len(ds) else 1
        t = t.reshape(nxt, rest // nxt, cR, cL, done * W.shape[3], W.shape[1], order="F")
    t = t.reshape(cR, cL, d, mpos[-1].shape[1], order="F")                         # (b, a', s', w')
    t = np.tensordot(t, envR, axes=([0, 3], [0, 2]))                               # (a', s', b')
    return np.ascontiguousarray(t.transpose(1, 0, 2))


def shift_mpos(mpos, sigma_old, sigma_new):
    """MatVecMPOS::set_shift (matvec_mpos.cpp:1437-1471): (sigma_old - sigma_new)/N times the identity added to the
    bottom-left block W[DL-1, 0] of each of the N MPOs the operator holds."""
    out = []
    for W in mpos:
        W = np.array(W, dtype=np.complex128, copy=True)
        W[W.shape[0] - 1, 0] += (sigma_old - sigma_new) / len(mpos) * np.eye(W.shape[2])
        out.append(W)
    return out
