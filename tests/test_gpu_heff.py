"""GPU parity of the H_eff hot path against the oracle, through the C ABI.

Tolerances: FP64 (real and complex) relative 2-norm error <= 1e-12 against the oracle;
the FP32 matrix-core path (FP32 inputs, FP32 accumulation) <= 5e-5 relative.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import dmrg_amd as G
from dmrg_amd import DmrgError, _lib
from dmrg_amd.workloads import random_heff_problem
from oracle import heff as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 5e-5


@pytest.fixture(scope="module", autouse=True)
def device():
    n = C.c_int()
    _lib.call("dmrg_device_count", C.byref(n))
    assert n.value > 0, "no GPU visible"
    info = G.device_info()
    print(f"[gpu] {info}", file=sys.stderr, flush=True)
    return info


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def general(rng, d, cL, cR, DL, DR, complex_):
    f = (lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)) if complex_ else (lambda *s: rng.standard_normal(s))
    return [np.asfortranarray(x) for x in (f(d, cL, cR), f(DL, DR, d, d), f(cL, cL, DL), f(cR, cR, DR))]


SHAPES = [
    (1, 1, 1, 1, 1), (2, 1, 8, 1, 3), (2, 8, 1, 3, 1), (2, 7, 5, 3, 3), (2, 16, 16, 5, 5), (2, 33, 17, 5, 5),
    (2, 17, 33, 5, 5), (1, 9, 13, 2, 4), (4, 20, 31, 8, 8), (3, 50, 70, 4, 6), (2, 64, 64, 3, 3),
    (2, 100, 93, 9, 9), (2, 128, 128, 5, 5), (4, 64, 48, 8, 8), (2, 40, 40, 25, 25),
]


@pytest.mark.parametrize("complex_", [False, True], ids=["fp64", "cx64"])
@pytest.mark.parametrize("shape", SHAPES, ids=[str(s) for s in SHAPES])
def test_matvec_parity(shape, complex_):
    rng = np.random.default_rng(hash(shape) % 2**32)
    th, W, L, R = general(rng, *shape, complex_)
    ref = O.matrix_vector_product(th, W, L, R)
    out = G.matrix_vector_product(th, W, L, R)
    assert out.dtype == th.dtype
    assert rel(out, ref) <= TOL64, rel(out, ref)


@pytest.mark.parametrize("chi", [128, 256, 512])
def test_matvec_parity_isolated_workload(chi):
    """The isolated single-site workload: theta (2, chi, chi) cx64, Hermitian env blocks (chi, chi, 5), seed 1."""
    th, W, L, R = random_heff_problem(2, chi, chi, 5, 5, seed=1)
    ref = O.matrix_vector_product(th, W, L, R)
    out = G.matrix_vector_product(th, W, L, R)
    assert rel(out, ref) <= TOL64


@pytest.mark.parametrize("complex_", [False, True], ids=["fp32", "cx32"])
@pytest.mark.parametrize("shape", [(2, 7, 5, 3, 3), (2, 64, 64, 5, 5), (2, 130, 120, 5, 5), (2, 48, 48, 25, 25)])
def test_matvec_fp32_parity(shape, complex_):
    rng = np.random.default_rng(7)
    th, W, L, R = general(rng, *shape, complex_)
    ref = O.matrix_vector_product(th, W, L, R)
    lo = np.complex64 if complex_ else np.float32
    out = G.matrix_vector_product(*(x.astype(lo) for x in (th, W, L, R)))
    assert out.dtype == np.dtype(lo)
    assert rel(out, ref) <= TOL32


def test_matvec_dims_mismatch_raises():
    rng = np.random.default_rng(1)
    th, W, L, R = general(rng, 2, 4, 5, 3, 3, True)
    with pytest.raises(DmrgError) as e:
        G.matrix_vector_product(th, W, L[:3, :3], R)
    assert e.value.status == 1


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("complex_", [False, True], ids=["fp64", "cx64"])
@pytest.mark.parametrize("shape", [(2, 1, 1, 1, 3), (2, 5, 7, 3, 3), (2, 16, 16, 5, 5), (2, 37, 29, 3, 3),
                                   (4, 24, 40, 8, 8), (2, 64, 64, 9, 9), (2, 128, 96, 5, 5)])
def test_env_parity(shape, side, complex_):
    d, cL, cR, DL, DR = shape
    rng = np.random.default_rng(17)
    th, W, L, R = general(rng, d, cL, cR, DL, DR, complex_)
    other = L if side == "L" else R
    ref = O.env_build_block(side, other, th, W)
    out = G.env_build_block(side, other, th, W)
    assert out.shape == ref.shape
    assert rel(out, ref) <= TOL64


def test_env_dims_mismatch_raises():
    rng = np.random.default_rng(1)
    th, W, L, R = general(rng, 2, 4, 5, 3, 3, True)
    with pytest.raises(DmrgError) as e:
        G.env_build_block("L", R, th, W)
    assert e.value.status == 1


@pytest.mark.parametrize("complex_", [False, True])
def test_expectation_value_and_dots(complex_):
    rng = np.random.default_rng(5)
    th, W, L, R = general(rng, 2, 30, 20, 5, 5, complex_)
    ev = G.expectation_value(th, W, L, R)
    ref = O.expectation_value(th, W, L, R)
    assert abs(ev - ref) <= TOL64 * abs(ref) * 10
    y = rng.standard_normal(th.shape) + (1j * rng.standard_normal(th.shape) if complex_ else 0)
    y = np.asfortranarray(y.astype(th.dtype))
    assert abs(G.dot(th, y) - np.vdot(th, y)) <= 1e-13 * np.linalg.norm(th) * np.linalg.norm(y)
    assert abs(G.nrm2(th) - np.linalg.norm(th)) <= 1e-13 * np.linalg.norm(th)


def test_dot_large_and_empty():
    rng = np.random.default_rng(8)
    x = rng.standard_normal(3_000_001) + 1j * rng.standard_normal(3_000_001)
    y = rng.standard_normal(3_000_001) + 1j * rng.standard_normal(3_000_001)
    assert abs(G.dot(x, y) - np.vdot(x, y)) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(y)
    # deterministic: same result twice
    assert G.dot(x, y) == G.dot(x, y)


# ------------------------------------------------------------------------------ MatVecMPOS

@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_matvec_op_two_site_merge(dtype):
    rng = np.random.default_rng(31)
    cplx = dtype == np.complex128
    W1 = rng.standard_normal((3, 4, 2, 2)) + (1j * rng.standard_normal((3, 4, 2, 2)) if cplx else 0)
    W2 = rng.standard_normal((4, 5, 2, 2)) + (1j * rng.standard_normal((4, 5, 2, 2)) if cplx else 0)
    L = rng.standard_normal((12, 12, 3)) + (1j * rng.standard_normal((12, 12, 3)) if cplx else 0)
    R = rng.standard_normal((9, 9, 5)) + (1j * rng.standard_normal((9, 9, 5)) if cplx else 0)
    op = G.MatVecMPOS([W1, W2], L, R, dtype=dtype)
    assert op.shape_mps == (4, 12, 9) and op.rows() == 4 * 12 * 9
    Wm = O.merge_mpos([W1.astype(complex), W2.astype(complex)])
    x = rng.standard_normal((op.rows(), 3)) + (1j * rng.standard_normal((op.rows(), 3)) if cplx else 0)
    y = op.MultAx(x.astype(dtype))
    for j in range(3):
        th = x[:, j].reshape(op.shape_mps, order="F").astype(dtype)
        ref = O.matrix_vector_product(th, Wm.astype(dtype), L.astype(dtype), R.astype(dtype))
        assert rel(y[:, j], ref.reshape(-1, order="F")) <= TOL64
    assert op.num_mv == 3
    op.close()


def test_matvec_op_real_rejects_complex_mpo():
    W = np.zeros((1, 1, 2, 2), dtype=complex)
    W[0, 0] = O.SY
    with pytest.raises(DmrgError):
        G.MatVecMPOS([W], np.ones((3, 3, 1)), np.ones((3, 3, 1)), dtype=np.float64)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_matvec_op_set_shift_and_diagonal(dtype):
    th, W, L, R = random_heff_problem(2, 20, 24, 4, 3, seed=3, complex_=(dtype == np.complex128))
    op = G.MatVecMPOS([W], L, R, dtype=dtype)
    diag = op.get_diagonal()
    H = O.dense_heff(W, L, R)
    np.testing.assert_allclose(diag.reshape(-1, order="F"), np.diag(H), rtol=1e-12, atol=1e-12 * np.abs(H).max())
    sigma = 0.37
    op.set_shift(sigma)
    W2 = W.copy()
    for s in range(W.shape[2]):
        W2[-1, 0, s, s] -= sigma  # (sigma_old - sigma_new)/1 on the bottom-left corner
    ref = O.matrix_vector_product(th, W2, L, R)
    out = op.MultAx(th.reshape(-1, order="F")).reshape(th.shape, order="F")
    assert rel(out, ref) <= TOL64
    op.set_shift(1.5)  # readyShift: a second shift is ignored (matvec_mpos.cpp:1437)
    out2 = op.MultAx(th.reshape(-1, order="F")).reshape(th.shape, order="F")
    assert rel(out2, ref) <= TOL64
    op.close()


def test_matvec_op_device_block_and_fp32():
    th, W, L, R = random_heff_problem(2, 96, 80, 5, 5, seed=4)
    op = G.MatVecMPOS([W], L, R, dtype=np.complex128)
    n = op.rows()
    rng = np.random.default_rng(4)
    X = np.asfortranarray(rng.standard_normal((n, 4)) + 1j * rng.standard_normal((n, 4)))
    dx = G.DeviceArray.from_numpy(X)
    dy = G.DeviceArray((n, 4), np.complex128)
    op.MultAx_device(dx, dy, nvec=4)
    Y = dy.to_numpy()
    for j in range(4):
        ref = O.matrix_vector_product(X[:, j].reshape(th.shape, order="F"), W, L, R)
        assert rel(Y[:, j], ref.reshape(-1, order="F")) <= TOL64
    dx32 = G.DeviceArray.from_numpy(X.astype(np.complex64))
    dy32 = G.DeviceArray((n, 4), np.complex64)
    op.MultAx_device(dx32, dy32, nvec=4, fp32=True)
    Y32 = dy32.to_numpy()
    for j in range(4):
        ref = O.matrix_vector_product(X[:, j].reshape(th.shape, order="F"), W, L, R)
        assert rel(Y32[:, j], ref.reshape(-1, order="F")) <= TOL32
    op.close()


# ------------------------------------------------------------------------------ every kernel variant

@pytest.fixture
def plan_override():
    yield lambda cfg, ks: _lib.call("dmrg_gemm_plan_override", cfg, ks)
    _lib.call("dmrg_gemm_plan_override", -1, 0)


@pytest.mark.parametrize("cfg", range(8), ids=["64x64", "32x64", "64x32", "16x128", "128x64", "32x32", "smallk", "32x32k"])
@pytest.mark.parametrize("ks", [1, 3, 8])
def test_every_tile_config_matvec_and_env(cfg, ks, plan_override):
    """Each tile configuration, with and without K slicing, on aligned (chi=64) and ragged shapes."""
    plan_override(cfg, ks)
    rng = np.random.default_rng(cfg * 10 + ks)
    for shape in [(2, 64, 64, 5, 5), (2, 37, 29, 3, 4), (4, 48, 80, 8, 8)]:
        for cplx in (False, True):
            th, W, L, R = general(rng, *shape, cplx)
            assert rel(G.matrix_vector_product(th, W, L, R), O.matrix_vector_product(th, W, L, R)) <= TOL64
            assert rel(G.env_build_block("L", L, th, W), O.env_build_block("L", L, th, W)) <= TOL64
            assert rel(G.env_build_block("R", R, th, W), O.env_build_block("R", R, th, W)) <= TOL64
        th, W, L, R = general(rng, *shape, True)
        out = G.matrix_vector_product(*(x.astype(np.complex64) for x in (th, W, L, R)))
        assert rel(out, O.matrix_vector_product(th, W, L, R)) <= TOL32


@pytest.mark.parametrize("cfg", [0, 1, 2], ids=["64x64", "32x64", "64x32"])
@pytest.mark.parametrize("nsplit,ks2", [(64, 2), (128, 4), (192, 8), (96, 3)])
def test_hybrid_plans(cfg, nsplit, ks2):
    """Hybrid plans: columns [0, nsplit) in one launch, the tail sliced over K, on aligned and ragged shapes."""
    _lib.call("dmrg_gemm_hybrid_override", cfg, nsplit, ks2)
    try:
        rng = np.random.default_rng(cfg * 100 + nsplit + ks2)
        for shape in [(2, 64, 64, 5, 5), (2, 96, 40, 3, 4), (4, 48, 80, 8, 8), (2, 130, 120, 5, 5)]:
            for cplx in (False, True):
                th, W, L, R = general(rng, *shape, cplx)
                assert rel(G.matrix_vector_product(th, W, L, R), O.matrix_vector_product(th, W, L, R)) <= TOL64
                assert rel(G.env_build_block("R", R, th, W), O.env_build_block("R", R, th, W)) <= TOL64
            th, W, L, R = general(rng, *shape, True)
            out = G.matrix_vector_product(*(x.astype(np.complex64) for x in (th, W, L, R)))
            assert rel(out, O.matrix_vector_product(th, W, L, R)) <= TOL32
    finally:
        _lib.call("dmrg_gemm_plan_override", -1, 0)


def test_autotuner_caches_plans():
    _lib.call("dmrg_autotune_reset")
    _lib.call("dmrg_autotune_min_macs", 0)  # tune these small legs too
    try:
        th, W, L, R = random_heff_problem(2, 96, 96, 5, 5, seed=9)
        a = G.matrix_vector_product(th, W, L, R)
        n = _lib.load().dmrg_autotune_entries()
        assert n >= 3
        b = G.matrix_vector_product(th, W, L, R)
        assert _lib.load().dmrg_autotune_entries() == n
        np.testing.assert_array_equal(a, b)  # same plan -> bitwise identical
        assert rel(a, O.matrix_vector_product(th, W, L, R)) <= TOL64
    finally:
        _lib.call("dmrg_autotune_min_macs", 1 << 29)


def test_small_gemms_are_not_tuned():
    """Below the threshold (default 2^29 multiply-adds) the size heuristic plans the legs and nothing is tuned."""
    _lib.call("dmrg_autotune_reset")
    th, W, L, R = random_heff_problem(2, 96, 96, 5, 5, seed=9)
    out = G.matrix_vector_product(th, W, L, R)
    assert _lib.load().dmrg_autotune_entries() == 0
    assert rel(out, O.matrix_vector_product(th, W, L, R)) <= TOL64


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_matvec_op_multbx(dtype):
    """MultBx (matvec_mpos.cpp:986-1019): without a B operator the output is left as it is; with B set on other env
    blocks (the variance blocks of MatVecMPOS(mpos, enva, envb)) it is the oracle matvec on those blocks, counted
    in num_mv; a B operator of another theta shape is DMRG_ERR_DIMS."""
    print(f"[case] multbx {np.dtype(dtype).name}", file=sys.stderr, flush=True)
    th, W, L, R = random_heff_problem(2, 12, 10, 4, 4, seed=31, complex_=True)
    _, W2, L2, R2 = random_heff_problem(2, 12, 10, 3, 5, seed=32, complex_=True)
    if dtype == np.float64:
        th, W, L, R, W2, L2, R2 = (np.ascontiguousarray(a.real) for a in (th, W, L, R, W2, L2, R2))
    op = G.MatVecMPOS([W], L, R, dtype=dtype)
    try:
        x = th.reshape(-1, order="F").astype(dtype)
        y0 = np.full_like(x, 3.0)
        assert np.array_equal(op.MultBx(x, out=y0.copy()), y0)
        op.set_b([W2], L2, R2)
        nmv = op.num_mv
        yb = op.MultBx(x)
        ref = O.matrix_vector_product(th.astype(dtype), W2.astype(dtype), L2.astype(dtype), R2.astype(dtype)).reshape(-1, order="F")
        assert np.linalg.norm(yb - ref) <= 1e-12 * np.linalg.norm(ref)
        assert op.num_mv == nmv + 1
        ya = op.MultAx(x)
        refa = O.matrix_vector_product(th.astype(dtype), W.astype(dtype), L.astype(dtype), R.astype(dtype)).reshape(-1, order="F")
        assert np.linalg.norm(ya - refa) <= 1e-12 * np.linalg.norm(refa)
        _, W3, L3, R3 = random_heff_problem(2, 12, 11, 3, 3, seed=33, complex_=True)
        with pytest.raises(DmrgError):
            op.set_b([W3], L3, R3)
    finally:
        op.close()


# ------------------------------------------------------------------------------ MPO chains (> 5 MPOs)

CHAINS = [([2] * 6, [3, 4, 2, 5, 3, 4, 2], 9, 13), ([2, 3, 1, 2, 2, 1, 2], [1, 3, 2, 2, 4, 1, 3, 2], 17, 11),
          ([2] * 8, [3] * 9, 24, 24)]


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
@pytest.mark.parametrize("chain", CHAINS, ids=["6mpo", "7mpo-ragged", "8mpo"])
def test_chain_matvec_and_per_mpo_shift(chain, dtype):
    """Lists of more than 5 MPOs are applied one MPO at a time (contraction_tblis.cpp:307-420), against the
    oracle's sequential restatement; set_shift adds (sigma_old - sigma_new)/N on each MPO's bottom-left block
    (matvec_mpos.cpp:1439-1462) — the MPOs here are general (not lower-triangular), where that differs from a
    shift of the merged MPO. FP32 matrix cores, a 3-vector block, the diagonal and the lowest eigenvalue of the
    shifted operator ride along."""
    ds, Ds, cL, cR = chain
    cplx = dtype == np.complex128
    print(f"[case] chain {len(ds)} MPOs cplx={cplx}", file=sys.stderr, flush=True)
    rng = np.random.default_rng(len(ds) * 10 + cplx)
    f = (lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)) if cplx else (lambda *s: rng.standard_normal(s))
    mpos = [np.asfortranarray(f(Ds[i], Ds[i + 1], ds[i], ds[i])) for i in range(len(ds))]
    d = int(np.prod(ds))
    th = np.asfortranarray(f(d, cL, cR))
    L, R = np.asfortranarray(f(cL, cL, Ds[0])), np.asfortranarray(f(cR, cR, Ds[-1]))
    op = G.MatVecMPOS(mpos, L, R, dtype=dtype)
    try:
        ref = O.matrix_vector_product_chain(th, mpos, L, R)
        out = op.MultAx(th.reshape(-1, order="F")).reshape(th.shape, order="F")
        assert rel(out, ref) <= TOL64, rel(out, ref)
        sigma = 0.37
        op.set_shift(sigma)
        shifted = O.shift_mpos(mpos, 0.0, sigma)
        ref_s = O.matrix_vector_product_chain(th, shifted, L, R)
        X = np.asfortranarray(np.stack([th.reshape(-1, order="F")] + [f(d * cL * cR) for _ in range(2)], axis=1))
        dx, dy = G.DeviceArray.from_numpy(X.astype(dtype)), G.DeviceArray(X.shape, dtype)
        op.MultAx_device(dx, dy, nvec=3)
        Y = dy.to_numpy()
        assert rel(Y[:, 0].reshape(th.shape, order="F"), ref_s) <= TOL64
        for j in (1, 2):
            refj = O.matrix_vector_product_chain(X[:, j].reshape(th.shape, order="F"), shifted, L, R)
            assert rel(Y[:, j].reshape(th.shape, order="F"), refj) <= TOL64
        lo = np.complex64 if cplx else np.float32
        dx32, dy32 = G.DeviceArray.from_numpy(X.astype(lo)), G.DeviceArray(X.shape, lo)
        op.MultAx_device(dx32, dy32, nvec=3, fp32=True)
        assert rel(dy32.to_numpy()[:, 0].reshape(th.shape, order="F"), ref_s) <= TOL32
        for a in (dx, dy, dx32, dy32):
            a.close()
        # diag H[(s,a,b),(s,a,b)] = sum_{w,w'} L[a,a,w] W[w,w',s,s] R[b,b,w'] of the merged shifted MPO
        Wm = O.merge_mpos(shifted)
        ref_d = np.einsum("aw,wvs,bv->sab", np.einsum("aaw->aw", L), np.einsum("wvss->wvs", Wm),
                          np.einsum("bbv->bv", R)).reshape(-1, order="F")
        diag = op.get_diagonal().reshape(-1, order="F")
        assert np.abs(diag - ref_d).max() <= 1e-12 * np.abs(ref_d).max() * 10
    finally:
        op.close()


def test_chain_eigs_hermitian():
    """Ground state of a Hermitian 6-MPO chain operator (each MPO block Hermitian with real coefficients, Hermitian
    env blocks) through the Davidson solver on the chain path, against the dense lowest eigenvalue."""
    rng = np.random.default_rng(21)
    ds, D, chi = [2] * 6, 3, 6
    mpos = []
    for _ in ds:
        W = np.zeros((D, D, 2, 2), dtype=np.complex128, order="F")
        for a in range(D):
            for b in range(D):
                A = rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2))
                W[a, b] = (A + A.conj().T) / 2
        mpos.append(W)
    from dmrg_amd.workloads import _herm_blocks
    L, R = _herm_blocks(rng, chi, D, True), _herm_blocks(rng, chi, D, True)
    op = G.MatVecMPOS(mpos, L, R, dtype=np.complex128)
    try:
        H = O.dense_heff(O.merge_mpos(mpos), L, R)
        # The operator applied is H^T (out(s') = sum_s W[.., s, s'] theta(s)); same spectrum.
        e_ref = np.linalg.eigvalsh(0.5 * (H + H.conj().T))[0]
        w, _, info = op.eigs(nev=1, tol=1e-12, seed=3)
        assert info["nconv"] == 1
        assert abs(w[0] - e_ref) <= 1e-10 * np.abs(H).max() * H.shape[0] ** 0.5
    finally:
        op.close()
