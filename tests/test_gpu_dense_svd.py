"""GPU parity of the dense small-problem path, the chain-end env blocks and the truncated SVD.

Tolerances: explicit H_eff relative 2-norm error <= 1e-12 against the oracle's literal loops; dense
eigenvalues <= 1e-12 * |A| against numpy (LAPACK) and the reference's LAPACKE fixtures; singular values
<= 1e-12 * S[0] against numpy, orthonormality and reconstruction <= 1e-11, for rocSOLVER's gesvd and the
one-sided Jacobi kernels (the LDS-resident path for n <= 32 and the launch-per-round path with one to four
wavefronts per pair); edge blocks exact.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

import dmrg_amd as G
from dmrg_amd import DmrgError
from dmrg_amd.workloads import random_heff_problem
from oracle import heff as O
from oracle import sweep as OS

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def general(rng, d, cL, cR, DL, DR, cplx):
    f = (lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)) if cplx else (lambda *s: rng.standard_normal(s))
    return [np.asfortranarray(x) for x in (f(d, cL, cR), f(DL, DR, d, d), f(cL, cL, DL), f(cR, cR, DR))]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", [(1, 1, 1, 1, 1), (2, 3, 4, 3, 2), (2, 5, 5, 3, 3), (4, 3, 5, 2, 4), (2, 7, 6, 5, 5)])
def test_heff_dense_matches_literal_loops(shape, cplx):
    print(f"[case] heff_dense shape={shape} cplx={cplx}", file=sys.stderr, flush=True)
    rng = np.random.default_rng(11)
    th, W, L, R = general(rng, *shape, cplx)
    H = G.heff_dense(W, L, R)
    Href = O.dense_heff(W, L, R)
    assert rel(H, Href) <= 1e-12
    assert rel(G.heff_dense(W, L, R, transpose=True), Href.T) <= 1e-12
    # H x is the matvec
    assert rel(H @ th.ravel(order="F"), O.matrix_vector_product(th, W, L, R).ravel(order="F")) <= 1e-12


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_operator_get_matrix_and_dense_ground_state(dtype):
    print(f"[case] get_matrix dtype={np.dtype(dtype).name}", file=sys.stderr, flush=True)
    cplx = dtype == np.complex128
    th, W, L, R = random_heff_problem(2, 12, 10, 3, 3, seed=5, complex_=cplx)
    if not cplx:
        L, R = L.real.copy(order="F"), R.real.copy(order="F")
    op = G.MatVecMPOS([W], L, R, dtype=dtype)
    try:
        Ht = op.get_matrix()  # MatVecMPO::get_matrix orientation (transpose of the applied operator)
        H = op.get_matrix(transpose=False)
        assert rel(Ht, H.T) == 0.0
        assert rel(H, O.dense_heff(W.astype(dtype), L.astype(dtype), R.astype(dtype))) <= 1e-12
        # dense branch: lowest eigenpair equals the iterative solver's
        w, V = G.dense_eigh(H)
        wref = np.linalg.eigvalsh(H)
        assert np.max(np.abs(w - wref)) <= 1e-12 * np.max(np.abs(wref))
        e_it, _, _ = op.eigs(nev=1, tol=1e-12)
        assert abs(e_it[0] - w[0]) <= 1e-10 * np.max(np.abs(wref))
        r = H @ V[:, 0] - w[0] * V[:, 0]
        assert np.linalg.norm(r) <= 1e-11 * np.max(np.abs(wref))
    finally:
        op.close()


@pytest.mark.parametrize("tag", ["1A", "1B", "2A", "2B"])
def test_dense_eigh_reference_fixtures(tag):
    print(f"[case] dense_eigh fixture {tag}", file=sys.stderr, flush=True)
    fx = np.load(GOLDEN / "lapacke_matrices.npz")
    A = fx[f"{tag}_matrix"]
    A = 0.5 * (A + A.T)  # fixtures are symmetric to print precision
    w, V = G.dense_eigh(A)
    ref = np.sort(fx[f"{tag}_eigvals"].ravel())
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(w - ref)) <= 1e-9 * scale
    assert np.max(np.abs(w - np.linalg.eigvalsh(A))) <= 1e-12 * scale
    assert np.linalg.norm(A @ V - V * w) <= 1e-11 * scale * np.sqrt(len(w))


def test_dense_eigh_values_only_and_bad_args():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((40, 40)) + 1j * rng.standard_normal((40, 40))
    A = X + X.conj().T
    w, V = G.dense_eigh(A, vectors=False)
    assert V is None
    assert np.max(np.abs(w - np.linalg.eigvalsh(A))) <= 1e-12 * np.max(np.abs(w))
    with pytest.raises(DmrgError):
        G.heff_dense(np.zeros((3, 3, 2, 2)), np.zeros((4, 4, 2)), np.zeros((4, 4, 3)))  # DL mismatch


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
@pytest.mark.parametrize("chi,D", [(1, 3), (7, 3), (5, 9), (33, 25)])
def test_env_set_edge(dtype, chi, D):
    rng = np.random.default_rng(chi * 100 + D)
    edge = rng.standard_normal(D).astype(dtype)
    if dtype == np.complex128:
        edge = edge + 1j * rng.standard_normal(D)
    blk = G.env_set_edge(chi, edge)
    ref = np.zeros((chi, chi, D), dtype=dtype)
    for i in range(chi):
        ref[i, i, :] = edge
    assert np.array_equal(blk, ref)


def test_env_edges_give_full_chain_energy():
    """With the ising MPO's left edge (last row) and right edge (first column), the env-contracted energy of a
    full-chain state equals <psi|H|psi> of the dense Hamiltonian (EnvBase.cpp:254-285, MpoSite.cpp:450-547)."""
    L, rng = 4, np.random.default_rng(2)
    h = rng.uniform(-1, 1, L)
    W = O.ising_tf_rf_mpo_bulk(1.0, 1.0 + h[1])  # a bulk site, D = 3
    eL = np.zeros(3, dtype=np.complex128)
    eL[2] = 1.0
    eR = np.zeros(3, dtype=np.complex128)
    eR[0] = 1.0
    Lb, Rb = G.env_set_edge(1, eL), G.env_set_edge(1, eR)
    th = (rng.standard_normal((2, 1, 1)) + 1j * rng.standard_normal((2, 1, 1))).astype(np.complex128)
    th /= np.linalg.norm(th)
    e = G.expectation_value(np.asfortranarray(th), np.asfortranarray(W), Lb, Rb)
    Hsite = -(1.0 + h[1]) * O.SX  # single site: only the field term survives the edges
    v = th.ravel()
    assert abs(e - np.vdot(v, Hsite.T @ v)) <= 1e-14


def check_svd(A, res, lim, rank_max):
    S_ref = np.linalg.svd(A, compute_uv=False)
    S = res["S"]
    assert np.max(np.abs(S - S_ref)) <= 1e-12 * S_ref[0]
    r = res["rank"]
    r_ref, terr_ref = OS._rank(S_ref, lim, rank_max)
    assert r == r_ref
    assert abs(res["truncation_error"] - np.linalg.norm(S[r:])) <= 1e-14 * S_ref[0] + 1e-15
    assert abs(res["discarded_weight"] - np.sum(S[r:] ** 2)) <= 1e-13 * S_ref[0] ** 2
    U, Vt = res["U"], res["Vt"]
    assert U.shape == (A.shape[0], r) and Vt.shape == (r, A.shape[1])
    assert np.linalg.norm(U.conj().T @ U - np.eye(r)) <= 1e-11 * np.sqrt(r)
    assert np.linalg.norm(Vt @ Vt.conj().T - np.eye(r)) <= 1e-11 * np.sqrt(r)
    # truncated reconstruction error equals the discarded norm (Eckart-Young)
    err = np.linalg.norm(A - (U * S[:r]) @ Vt)
    assert abs(err - np.linalg.norm(S_ref[r:])) <= 1e-11 * S_ref[0]


@pytest.mark.parametrize("method", [0, 3], ids=["gesvd", "jacobi"])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("m,n", [(1, 1), (1, 7), (9, 1), (2, 2), (30, 30), (31, 31), (50, 32), (33, 70), (64, 40), (37, 90), (300, 200),
                                 (600, 520), (513, 700), (640, 448), (600, 30), (20, 1200)])
def test_svd_truncate_random(m, n, cplx, method):
    print(f"[case] svd m={m} n={n} cplx={cplx} method={method}", file=sys.stderr, flush=True)
    rng = np.random.default_rng(m * 1000 + n)
    A = rng.standard_normal((m, n)) + (1j * rng.standard_normal((m, n)) if cplx else 0)
    # graded spectrum so that truncation actually drops something
    u, s, vt = np.linalg.svd(A, full_matrices=False)
    s = np.exp(-0.5 * np.arange(len(s)))
    A = np.asfortranarray((u * s) @ vt)
    for lim, rank_max in ((1e-14, 0), (1e-3, 0), (1e-14, 5)):
        check_svd(A, G.svd_truncate(A, lim, rank_max, method), lim, rank_max)


@pytest.mark.parametrize("key", ["m6", "m96", "m0", "m181", "m174", "m5", "m101", "m7"])
def test_svd_reference_matrices(key):
    """The reference's SVD test inputs (tests/svd/testmatrices/svdmatrix_*.h5, converted by
    tests/golden/make_svd_fixtures.py)."""
    print(f"[case] svd fixture {key}", file=sys.stderr, flush=True)
    A = np.asfortranarray(np.load(GOLDEN / "svd_matrices.npz")[key])
    for method in (0, 3):
        check_svd(A, G.svd_truncate(A, 1e-10, 0, method), 1e-10, 0)


def test_svd_rank_deficient_and_rejected_routines():
    """Exact zeros and ties at the rounding level come back in descending order (QR and Jacobi); rocSOLVER's
    gesdd / gesvdj paths (method 1, 2) are refused."""
    rng = np.random.default_rng(17)
    X = rng.standard_normal((80, 6)) + 1j * rng.standard_normal((80, 6))
    A = X @ X.conj().T  # 80 x 80, rank 6
    A = np.asfortranarray(np.hstack([A, A[:, :3]]))  # 80 x 83, rank 6
    for method in (0, 3):
        res = G.svd_truncate(A, 1e-8, 0, method)
        assert np.all(np.diff(res["S"]) <= 0)
        assert res["rank"] == 6
        check_svd(A, res, 1e-8, 0)
    for method in (1, 2):
        with pytest.raises(DmrgError) as e:
            G.svd_truncate(A, 1e-14, 0, method)
        assert e.value.status == 7


@pytest.mark.parametrize("cplx", [False, True])
def test_svd_block_jacobi_rank_deficient(cplx):
    """Block pairs with exactly dependent and zero columns (500 x 420 of rank 30, 20 zero columns): the block
    one-sided Jacobi path returns the same singular values and a valid factorisation as gesvd."""
    print(f"[case] block jacobi rank deficient cplx={cplx}", file=sys.stderr, flush=True)
    rng = np.random.default_rng(23)
    X = rng.standard_normal((500, 30)) + (1j * rng.standard_normal((500, 30)) if cplx else 0)
    Y = rng.standard_normal((30, 400)) + (1j * rng.standard_normal((30, 400)) if cplx else 0)
    A = np.asfortranarray(np.hstack([X @ Y, np.zeros((500, 20))]))
    for method in (0, 3):
        check_svd(A, G.svd_truncate(A, 1e-10, 0, method), 1e-10, 0)
