"""CPU tests that pin the oracle (oracle/heff.py) against brute force, closed forms and invariants.

Each check is chosen so that a plausible slip (a transposed operand, a dropped conj, a wrong
index order or sign) in the oracle fails it.
"""
import numpy as np
import pytest

from oracle import heff as O
from dmrg_amd.workloads import random_heff_problem


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def random_general_problem(rng, d, cL, cR, DL, DR, complex_=True):
    """Non-Hermitian, fully general inputs (no symmetry that could hide a transposition)."""
    f = (lambda *s: rand_c(rng, *s)) if complex_ else (lambda *s: rng.standard_normal(s))
    return f(d, cL, cR), f(DL, DR, d, d), f(cL, cL, DL), f(cR, cR, DR)


@pytest.mark.parametrize("complex_", [False, True])
@pytest.mark.parametrize("shape", [(1, 1, 1, 1, 1), (2, 3, 2, 2, 3), (2, 2, 4, 3, 1), (4, 3, 3, 2, 2), (3, 2, 5, 1, 2)])
def test_matvec_brute_force(shape, complex_):
    d, cL, cR, DL, DR = shape
    rng = np.random.default_rng(11)
    th, W, L, R = random_general_problem(rng, d, cL, cR, DL, DR, complex_)
    H = O.dense_heff(W, L, R)
    ref = (H @ th.reshape(-1, order="F")).reshape(th.shape, order="F")
    for order in ("L", "R"):
        out = O.matrix_vector_product(th, W, L, R, order=order)
        np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_matvec_literal_sum():
    """Element-wise 6-index sum, independent of dense_heff's flattening."""
    rng = np.random.default_rng(3)
    d, cL, cR, DL, DR = 2, 3, 2, 2, 3
    th, W, L, R = random_general_problem(rng, d, cL, cR, DL, DR)
    out = O.matrix_vector_product(th, W, L, R)
    sp, ap, bp = 1, 2, 1
    acc = 0
    for s in range(d):
        for a in range(cL):
            for b in range(cR):
                for w in range(DL):
                    for wp in range(DR):
                        acc += L[a, ap, w] * W[w, wp, s, sp] * R[b, bp, wp] * th[s, a, b]
    assert abs(out[sp, ap, bp] - acc) < 1e-12 * max(1, abs(acc))


def test_matvec_orders_agree_large():
    th, W, L, R = random_heff_problem(2, 40, 40, 5, 5, seed=1)
    a = O.matrix_vector_product(th, W, L, R, order="L")
    b = O.matrix_vector_product(th, W, L, R, order="R")
    assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(a)


def test_matvec_hermitian():
    th, W, L, R = random_heff_problem(2, 12, 9, 5, 5, seed=1)
    rng = np.random.default_rng(5)
    x = rand_c(rng, *th.shape)
    hx = O.matrix_vector_product(x, W, L, R)
    hy = O.matrix_vector_product(th, W, L, R)
    assert abs(np.vdot(x, hy) - np.vdot(hx, th)) < 1e-12 * np.linalg.norm(hx) * np.linalg.norm(hy)


def test_matvec_identity_env_applies_transpose():
    """D=1, L=R=I: out(s') = sum_s W[0,0,s,s'] theta(s), i.e. W^T acting on s (Appendix A convention)."""
    rng = np.random.default_rng(2)
    d, cL, cR = 3, 4, 5
    W = rand_c(rng, 1, 1, d, d)
    th = rand_c(rng, d, cL, cR)
    L = np.eye(cL)[:, :, None].astype(complex)
    R = np.eye(cR)[:, :, None].astype(complex)
    out = O.matrix_vector_product(th, W, L, R)
    ref = np.einsum("ts,tab->sab", W[0, 0], th)
    np.testing.assert_allclose(out, ref, atol=1e-13)


def test_expectation_value_brute_force():
    rng = np.random.default_rng(4)
    th, W, L, R = random_general_problem(rng, 2, 3, 4, 2, 3)
    ev = O.expectation_value(th, W, L, R)
    H = O.dense_heff(W, L, R)
    v = th.reshape(-1, order="F")
    assert abs(ev - np.vdot(v, H @ v)) < 1e-12 * abs(ev)


@pytest.mark.parametrize("L", [1, 2, 3, 5, 6])
def test_ising_mpo_matches_kronecker(L):
    rng = np.random.default_rng(L)
    h = rng.uniform(-1, 1, L)
    mpos = O.ising_tf_rf_mpos(L, J=0.7, g=1.1, h_rand=h)
    Hm = O.chain_dense_from_mpos(mpos)
    Hk = O.ising_dense(L, 0.7, 1.1, h)
    np.testing.assert_allclose(Hm, Hk, atol=1e-13)


@pytest.mark.parametrize("L", [2, 3, 4])
def test_mpo_squared_is_h_squared(L):
    rng = np.random.default_rng(7 + L)
    h = rng.uniform(-1, 1, L)
    mpos = O.ising_tf_rf_mpos(L, J=1.0, g=1.0, h_rand=h)
    # Edge-free squares: square the bulk MPOs and apply kron(edge, edge) afterwards (MpoSite.cpp:34-40).
    sq = []
    for i in range(L):
        W2 = O.mpo_squared(O.ising_tf_rf_mpo_bulk(0.0 if i == L - 1 else 1.0, 1.0 + h[i]))
        if i == 0:
            W2 = W2[8:9]          # kron of row-2 edges: index 2 + 3*2
        if i == L - 1:
            W2 = W2[:, 0:1]       # kron of col-0 edges
        sq.append(W2)
    H2 = O.chain_dense_from_mpos(sq)
    H = O.ising_dense(L, 1.0, 1.0, h)
    np.testing.assert_allclose(H2, H @ H, atol=1e-12)


def test_merge_mpos_equals_sequential():
    rng = np.random.default_rng(9)
    A, B = rand_c(rng, 2, 3, 2, 2), rand_c(rng, 3, 4, 3, 3)
    M = O.merge_mpos([A, B])
    assert M.shape == (2, 4, 6, 6)
    for wl in range(2):
        for wr in range(4):
            ref = sum(np.kron(B[w, wr], A[wl, w]) for w in range(3))  # s1 fastest -> B leftmost factor
            np.testing.assert_allclose(M[wl, wr], ref, atol=1e-13)


# ------------------------------------------------------------------------------ environments

def random_mps(rng, L, d, chi, complex_=True):
    """Left-normalised (A) MPS tensors (d, chiL, chiR) with bond dims min(chi, d^l, d^(L-l))."""
    dims = [min(chi, d ** l, d ** (L - l)) for l in range(L + 1)]
    sites = []
    for i in range(L):
        a, b = dims[i], dims[i + 1]
        M = rng.standard_normal((d * a, b))
        if complex_:
            M = M + 1j * rng.standard_normal((d * a, b))
        Q, _ = np.linalg.qr(M)
        sites.append(np.asfortranarray(Q[:, :b].reshape(d, a, b, order="F")))
    return sites


def mps_to_dense(sites):
    psi = np.ones((1, 1), dtype=complex)
    for A in sites:
        d, a, b = A.shape
        new = np.einsum("pa,sab->psb", psi, A)  # prev-flat fastest, then s
        psi = new.reshape(-1, b, order="F")
    return psi[:, 0]


def test_env_left_identity_and_energy():
    rng = np.random.default_rng(21)
    Lc = 6
    h = rng.uniform(-1, 1, Lc)
    mpos = O.ising_tf_rf_mpos(Lc, 1.0, 1.0, h)
    sites = random_mps(rng, Lc, 2, 4)
    env = np.ones((1, 1, 1), dtype=complex)
    for i in range(Lc):
        env = O.env_build_block("L", env, sites[i], mpos[i])
        if 0 < i < Lc - 1:
            np.testing.assert_allclose(env[:, :, 2], np.eye(env.shape[0]), atol=1e-12)
    psi = mps_to_dense(sites)
    e_dense = np.vdot(psi, O.ising_dense(Lc, 1.0, 1.0, h) @ psi) / np.vdot(psi, psi)
    assert env.shape == (1, 1, 1)
    assert abs(env[0, 0, 0] / np.vdot(psi, psi) - e_dense) < 1e-12


def test_env_right_identity_and_energy():
    rng = np.random.default_rng(22)
    Lc = 5
    h = rng.uniform(-1, 1, Lc)
    mpos = O.ising_tf_rf_mpos(Lc, 0.8, 1.3, h)
    # right-normalised B sites: reverse-transpose a left-normalised chain
    A = random_mps(rng, Lc, 2, 4)
    B = [np.asfortranarray(a.transpose(0, 2, 1)) for a in reversed(A)]
    env = np.ones((1, 1, 1), dtype=complex)
    for i in reversed(range(Lc)):
        env = O.env_build_block("R", env, B[i], mpos[i])
        if 0 < i < Lc:
            if env.shape[2] == 3:
                np.testing.assert_allclose(env[:, :, 0], np.eye(env.shape[0]), atol=1e-12)
    psi = mps_to_dense(B)
    e_dense = np.vdot(psi, O.ising_dense(Lc, 0.8, 1.3, h) @ psi) / np.vdot(psi, psi)
    assert abs(env[0, 0, 0] / np.vdot(psi, psi) - e_dense) < 1e-12


def test_heff_is_projected_hamiltonian():
    """With L env from A sites and R env from B sites, H_eff = P^dagger H P (full-chain consistency)."""
    rng = np.random.default_rng(23)
    Lc, c = 5, 2
    h = rng.uniform(-1, 1, Lc)
    mpos = O.ising_tf_rf_mpos(Lc, 1.0, 0.9, h)
    A = random_mps(rng, Lc, 2, 4)
    Bs = [np.asfortranarray(a.transpose(0, 2, 1)) for a in reversed(random_mps(rng, Lc, 2, 4))]
    envL = np.ones((1, 1, 1), dtype=complex)
    for i in range(c):
        envL = O.env_build_block("L", envL, A[i], mpos[i])
    envR = np.ones((1, 1, 1), dtype=complex)
    for i in reversed(range(c + 1, Lc)):
        envR = O.env_build_block("R", envR, Bs[i], mpos[i])
    d, chiL, chiR = 2, A[c - 1].shape[2], Bs[c + 1].shape[1]
    th = rand_c(rng, d, chiL, chiR)
    out = O.matrix_vector_product(th, mpos[c], envL, envR)
    # P maps theta -> dense psi
    def embed(t):
        return mps_to_dense(A[:c] + [t] + Bs[c + 1:])
    Hfull = O.ising_dense(Lc, 1.0, 0.9, h)
    # <e_j| H_eff |th> = <embed(e_j)| H |embed(th)>; H_eff applies the transpose convention on real MPOs,
    # which for this real-symmetric Hamiltonian equals the projection itself.
    psi = embed(th)
    n = th.size
    proj = np.zeros(n, dtype=complex)
    for j in range(n):
        e = np.zeros(n, dtype=complex)
        e[j] = 1
        proj[j] = np.vdot(embed(e.reshape(th.shape, order="F")), Hfull @ psi)
    np.testing.assert_allclose(out.reshape(-1, order="F"), proj, atol=1e-12)


# ------------------------------------------------------------------------------ free fermions

@pytest.mark.parametrize("L,J,g,seed", [(4, 1.0, 1.0, 0), (7, 1.0, 1.0, 1), (9, 0.6, 1.4, 2), (10, 1.0, 0.5, 3)])
def test_free_fermion_ground_state(L, J, g, seed):
    rng = np.random.default_rng(seed)
    h = rng.uniform(-1, 1, L)
    H = O.ising_dense(L, J, g, h)
    w, v = np.linalg.eigh(H)
    assert abs(w[0] - O.ising_free_fermion_ground_energy(J, g, h)) < 1e-11
    S_ff = O.ising_free_fermion_entropies(J, g, h)
    S_ed = O.dense_entropies(v[:, 0], L)
    np.testing.assert_allclose(S_ff, S_ed, atol=1e-9)


def test_entanglement_entropy_definition():
    s = np.array([np.sqrt(0.5), np.sqrt(0.5)])
    assert abs(O.entanglement_entropy(s) - np.log(2)) < 1e-15
    assert O.entanglement_entropy(np.array([1.0])) == 0.0


def _random_chain(rng, ds, Ds, complex_=True):
    f = (lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)) if complex_ else (lambda *s: rng.standard_normal(s))
    return [f(Ds[i], Ds[i + 1], ds[i], ds[i]) for i in range(len(ds))]


@pytest.mark.parametrize("ds,Ds", [([2] * 6, [3, 4, 2, 5, 3, 4, 2]), ([2, 3, 1, 2, 2, 1, 2], [1, 3, 2, 2, 4, 1, 3, 2])])
def test_chain_matvec_equals_merged(ds, Ds):
    """The MPO-chain path (one MPO at a time) is the same operator as the merged MPO (6 and 7 MPOs, uneven
    physical and bond dims), and for tiny shapes it equals the literal dense H_eff of the merged MPO."""
    rng = np.random.default_rng(11)
    mpos = _random_chain(rng, ds, Ds)
    d, cL, cR = int(np.prod(ds)), 3, 4
    th = rng.standard_normal((d, cL, cR)) + 1j * rng.standard_normal((d, cL, cR))
    L = rng.standard_normal((cL, cL, Ds[0])) + 1j * rng.standard_normal((cL, cL, Ds[0]))
    R = rng.standard_normal((cR, cR, Ds[-1])) + 1j * rng.standard_normal((cR, cR, Ds[-1]))
    W = O.merge_mpos(mpos)
    out = O.matrix_vector_product_chain(th, mpos, L, R)
    ref = O.matrix_vector_product(th, W, L, R)
    assert np.linalg.norm(out - ref) <= 1e-12 * np.linalg.norm(ref)
    H = O.dense_heff(W, L, R)
    assert np.linalg.norm(out.reshape(-1, order="F") - H @ th.reshape(-1, order="F")) <= 1e-12 * np.linalg.norm(ref)


def test_chain_shift_per_mpo():
    """set_shift on a list of N MPOs adds (sigma_old - sigma_new)/N on each MPO's bottom-left block. For
    lower-triangular MPOs (the ising chain) the full-chain Hamiltonian moves by -sigma I; for a list that is not
    lower-triangular the per-MPO shift and a shift of the merged MPO give different operators."""
    L, sigma = 6, 0.37
    h = np.random.default_rng(2).uniform(-1, 1, L)
    mpos = O.ising_tf_rf_mpos(L, 1.0, 1.0, h)
    H = O.chain_dense_from_mpos(mpos)
    Hs = O.chain_dense_from_mpos(O.shift_mpos(mpos, 0.0, sigma))
    assert np.linalg.norm(H - O.ising_dense(L, 1.0, 1.0, h)) <= 1e-12 * np.linalg.norm(H)
    assert np.linalg.norm(Hs - (H - sigma * np.eye(H.shape[0]))) <= 1e-12 * np.linalg.norm(H)
    rng = np.random.default_rng(5)
    gen = _random_chain(rng, [2] * 6, [3] * 7)
    merged_shift = O.merge_mpos(gen).copy()
    merged_shift[2, 0] -= sigma * np.eye(merged_shift.shape[2])
    per_mpo = O.merge_mpos(O.shift_mpos(gen, 0.0, sigma))
    assert np.linalg.norm(per_mpo - merged_shift) > 1e-3 * np.linalg.norm(merged_shift)


def test_lbit_site_mpo_matches_kronecker_sum():
    """The l-bit-style site MPO of the two-site workload (1-body h Z + g X, 2-body Z Z at range 1-3, 3-body Z Z Z
    at ranges 2 and 3) reproduces, on a 7-site chain with the ising edges (row 7 on the left, column 0 on the right),
    the Hamiltonian summed term by term from Pauli Kronecker products."""
    from dmrg_amd.workloads import lbit_dense_hamiltonian, lbit_site_couplings, lbit_site_mpo

    rng = np.random.default_rng(3)
    cs = [lbit_site_couplings(rng) for _ in range(7)]
    mpos = [lbit_site_mpo(c) for c in cs]
    mpos[0] = mpos[0][7:8]
    mpos[-1] = mpos[-1][:, 0:1]
    H = O.chain_dense_from_mpos(mpos)
    ref = lbit_dense_hamiltonian(cs)
    assert np.linalg.norm(H - ref) <= 1e-12 * np.linalg.norm(ref)
    assert all(abs(c["K1"]) > 0 and abs(c["K2"]) > 0 for c in cs)  # the 3-body terms are present
