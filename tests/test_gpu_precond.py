"""GPU tests of the preconditioners of the H_eff operator (MatVecMPOS::MultPc, matvec_mpos.cpp:1333-1416).

Jacobi (jcb = 1): y = x / (diag(H) - shift). Block-Jacobi (jcb > 1, LU): y_k = (H_kk - shift I)^-1 x_k on the
diagonal blocks at offsets k*jcb. Both compared with numpy on the oracle's explicit H_eff (<= 1e-10 relative).
The eigensolver with either preconditioner must reach the same ground energy as without (<= 1e-10 |H|).
"""
import sys

import numpy as np
import pytest

import dmrg_amd as G
from dmrg_amd.workloads import random_heff_problem
from oracle import heff as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
@pytest.mark.parametrize("jcb", [1, 16, 64, 240, 1000])
def test_multpc_matches_dense(dtype, jcb):
    print(f"[case] MultPc dtype={np.dtype(dtype).name} jcb={jcb}", file=sys.stderr, flush=True)
    cplx = dtype == np.complex128
    th, W, L, R = random_heff_problem(2, 12, 10, 3, 3, seed=21, complex_=cplx)
    H = O.dense_heff(W.astype(dtype), L.astype(dtype), R.astype(dtype))
    n = H.shape[0]
    op = G.MatVecMPOS([W], L, R, dtype=dtype)
    try:
        op.set_jcb_block_size(jcb)
        rng = np.random.default_rng(4)
        X = rng.standard_normal((n, 2)) + (1j * rng.standard_normal((n, 2)) if cplx else 0)
        shifts = [-7.5, 3.25]
        Y = op.MultPc(X, shifts)
        for c in range(2):
            # the block preconditioner is built once, with the first shift of the first call (readyCalcPc)
            sh = shifts[c] if jcb == 1 else shifts[0]
            ref = np.empty(n, dtype=dtype)
            b = min(jcb, n)
            for off in range(0, n, b):
                e = min(off + b, n)
                blk = H[off:e, off:e] - sh * np.eye(e - off)
                ref[off:e] = np.linalg.solve(blk, X[off:e, c])
            assert rel(Y[:, c], ref) <= 1e-10
    finally:
        op.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_eigs_jacobi_ground_state(mixed):
    print(f"[case] eigs Jacobi mixed={mixed}", file=sys.stderr, flush=True)
    th, W, L, R = random_heff_problem(2, 40, 40, 5, 5, seed=8)
    op = G.MatVecMPOS([W], L, R, dtype=np.complex128)
    try:
        w0, _, i0 = op.eigs(nev=1, tol=1e-12)
        w1, v1, i1 = op.eigs(nev=1, tol=1e-12, precond=1, mixed=mixed)
        assert abs(w1[0] - w0[0]) <= 1e-10 * i0["anorm"]
        r = op.MultAx(v1[:, 0]) - w1[0] * v1[:, 0]
        assert np.linalg.norm(r) <= 1e-10 * i0["anorm"]
    finally:
        op.close()


@pytest.mark.parametrize("jcb", [32, 128])
@pytest.mark.parametrize("mixed", [False, True])
def test_eigs_block_jacobi_interior(jcb, mixed):
    """Block-Jacobi built at the target shift (the setting of the reference's xDMRG solves): the eigenvalue
    closest to an interior shift, checked against the dense spectrum."""
    print(f"[case] eigs block-Jacobi jcb={jcb} mixed={mixed}", file=sys.stderr, flush=True)
    th, W, L, R = random_heff_problem(2, 20, 20, 5, 5, seed=8)
    op = G.MatVecMPOS([W], L, R, dtype=np.complex128)
    try:
        H = op.get_matrix(transpose=False)
        w = np.linalg.eigvalsh(H)
        sigma = w[6] + 0.1 * (w[5] - w[6])  # closest to w[6]
        w1, v1, i1 = op.eigs(nev=1, target="CLOSEST", shift=sigma, tol=1e-12, precond=2, jcb=jcb, mixed=mixed)
        assert abs(w1[0] - w[6]) <= 1e-10 * np.max(np.abs(w))
        r = H @ v1[:, 0] - w1[0] * v1[:, 0]
        assert np.linalg.norm(r) <= 1e-10 * np.max(np.abs(w))
        print(f"[info] matvecs {i1['matvecs_fp64']}+{i1['matvecs_fp32']}", file=sys.stderr, flush=True)
    finally:
        op.close()


def llt_rule_reference(H, jcb, shift, X):
    """The block preconditioner with Factorization::LLT (matvec_mpos.cpp:1143-1190): a block B with Re B(0,0) <= 0
    is Cholesky-factorised negated, and when that succeeds the solve applies -B^-1; every other block (and one
    whose Cholesky factorisation fails, through LDLT / LU) applies B^-1."""
    n = H.shape[0]
    Y = np.empty_like(X)
    for off in range(0, n, jcb):
        e = min(off + jcb, n)
        B = H[off:e, off:e] - shift * np.eye(e - off)
        sign = 1.0
        if B[0, 0].real <= 0 and np.linalg.eigvalsh(-0.5 * (B + B.conj().T)).min() > 0:
            sign = -1.0
        Y[off:e] = sign * np.linalg.solve(B, X[off:e])
    return Y


@pytest.mark.parametrize("kind", ["indefinite", "negative", "half-negative"])
@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_block_jacobi_llt_sign_rule(kind, dtype):
    """Block-Jacobi with the LLT factorization against the reference's rule, on an operator whose blocks are
    indefinite (no block changes sign), negative definite (every block applies -B^-1) and, with R = diag(r) of
    mixed signs, blocks of either kind (jcb = d chiL, one block per column b). LU, LDLT and QR give B^-1."""
    print(f"[case] block-Jacobi LLT {kind} {np.dtype(dtype).name}", file=sys.stderr, flush=True)
    cplx = dtype == np.complex128
    rng = np.random.default_rng(31)
    d, cL, cR = 2, 6, 8
    if kind == "indefinite":
        th, W, L, R = random_heff_problem(d, cL, cR, 3, 3, seed=21, complex_=cplx)
        jcb, shift = 24, 0.5
    else:
        f = (lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)) if cplx else (lambda *s: rng.standard_normal(s))
        A = f(cL, cL)
        L = (A @ A.conj().T + np.eye(cL)).reshape(cL, cL, 1)
        W = np.eye(d).reshape(1, 1, d, d)
        if kind == "negative":
            Bm = f(cR, cR)
            R = -(Bm @ Bm.conj().T + np.eye(cR)).reshape(cR, cR, 1)
            jcb, shift = 24, 0.5
        else:
            R = np.diag(np.where(np.arange(cR) < cR // 2, 1.0, -1.0) * (1.0 + rng.random(cR))).reshape(cR, cR, 1)
            jcb, shift = d * cL, 0.0
        L, W, R = (np.asfortranarray(x.astype(dtype)) for x in (L, W, R))
    H = O.dense_heff(W.astype(dtype), L.astype(dtype), R.astype(dtype))
    n = H.shape[0]
    X = rng.standard_normal((n, 1)) + (1j * rng.standard_normal((n, 1)) if cplx else 0)
    op = G.MatVecMPOS([W], L, R, dtype=dtype)
    try:
        op.set_jcb_block_size(jcb)
        op.set_factorization("LLT")
        Y = op.MultPc(X, [shift])
        ref = llt_rule_reference(H, jcb, shift, X)
        assert rel(Y[:, 0], ref[:, 0]) <= 1e-10
        flips = sum(1 for off in range(0, n, jcb) if not np.allclose(ref[off:off + jcb], np.linalg.solve(
            H[off:off + jcb, off:off + jcb] - shift * np.eye(min(jcb, n - off)), X[off:off + jcb])))
        assert flips == {"indefinite": 0, "negative": n // jcb, "half-negative": cR // 2}[kind]
        for fact in ("LU", "LDLT", "QR", "NONE", "ILUT"):
            op.set_factorization(fact)
            Y2 = op.MultPc(X, [shift])
            ref2 = np.concatenate([np.linalg.solve(H[o:o + jcb, o:o + jcb] - shift * np.eye(min(jcb, n - o)), X[o:o + jcb])
                                   for o in range(0, n, jcb)])
            assert rel(Y2[:, 0], ref2[:, 0]) <= 1e-10
        with pytest.raises(G.DmrgError) as e:
            G._lib.call("dmrg_matvec_set_factorization", op.h, 7)  # not an eig::Factorization value
        assert e.value.status == 7
    finally:
        op.close()
