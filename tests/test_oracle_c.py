"""CPU tests that pin the oracle's C restatement of the matvec and of the environment update
(oracle/heff_c.c) against brute force and against the numpy restatement (oracle/heff.py), on both contraction
branches and both sides.

The inputs are general (non-Hermitian, complex, ragged sizes) so that a transposed operand or a swapped
index in one of the three contractions fails the brute-force check.
"""
import numpy as np
import pytest

from oracle import heff as O
from oracle import heff_c as OC

pytestmark = pytest.mark.skipif(not OC.host_supported(), reason="the C oracle needs AVX2 and FMA")


@pytest.fixture(scope="module", autouse=True)
def built_library():
    if not OC.LIB_PATH.exists():
        from oracle.build import build

        build()


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def general_problem(rng, d, cL, cR, DL, DR, complex_=True):
    f = (lambda *s: rand_c(rng, *s)) if complex_ else (lambda *s: rng.standard_normal(s).astype(complex))
    return f(d, cL, cR), f(DL, DR, d, d), f(cL, cL, DL), f(cR, cR, DR)


@pytest.mark.parametrize("complex_", [False, True])
@pytest.mark.parametrize("shape", [(1, 1, 1, 1, 1), (2, 3, 2, 2, 3), (2, 2, 4, 3, 1), (4, 3, 3, 2, 2),
                                   (3, 2, 5, 1, 2), (2, 5, 7, 3, 4)])
def test_c_matvec_brute_force(shape, complex_):
    d, cL, cR, DL, DR = shape
    rng = np.random.default_rng(21)
    th, W, L, R = general_problem(rng, d, cL, cR, DL, DR, complex_)
    H = O.dense_heff(W, L, R)
    ref = (H @ th.reshape(-1, order="F")).reshape(th.shape, order="F")
    for order in ("L", "R", None):
        out = OC.matrix_vector_product(th, W, L, R, order=order)
        np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_c_matvec_literal_sum():
    """One output element as the explicit 6-index sum."""
    rng = np.random.default_rng(4)
    d, cL, cR, DL, DR = 2, 3, 4, 2, 3
    th, W, L, R = general_problem(rng, d, cL, cR, DL, DR)
    sp, ap, bp = 1, 2, 3
    acc = sum(L[a, ap, w] * W[w, wp, s, sp] * R[b, bp, wp] * th[s, a, b]
              for s in range(d) for a in range(cL) for b in range(cR) for w in range(DL) for wp in range(DR))
    for order in ("L", "R"):
        out = OC.matrix_vector_product(th, W, L, R, order=order)
        assert abs(out[sp, ap, bp] - acc) < 1e-12 * max(1, abs(acc))


@pytest.mark.parametrize("shape", [(2, 130, 129, 5, 5), (2, 67, 200, 3, 4), (4, 96, 80, 8, 8), (2, 256, 256, 5, 5)])
def test_c_matvec_matches_numpy_oracle(shape):
    """Sizes past every block edge of the C GEMM (4x4 tiles, 128 x 64 x 128 blocks), both branches."""
    d, cL, cR, DL, DR = shape
    rng = np.random.default_rng(cL + cR)
    th, W, L, R = general_problem(rng, d, cL, cR, DL, DR)
    ref = O.matrix_vector_product(th, W, L, R)
    for order in ("L", "R", None):
        out = OC.matrix_vector_product(th, W, L, R, order=order, threads=4)
        assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)


def test_c_matvec_repeat_and_shrink():
    """The kept workspace: a large call, then a smaller one, then the large one again give fresh results."""
    rng = np.random.default_rng(8)
    big = general_problem(rng, 2, 90, 70, 5, 5)
    small = general_problem(rng, 3, 11, 13, 2, 3)
    ref_big, ref_small = O.matrix_vector_product(*big), O.matrix_vector_product(*small)
    for args, ref in ((big, ref_big), (small, ref_small), (big, ref_big)):
        out = OC.matrix_vector_product(*args)
        assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)


def test_c_matvec_rejects_mismatch():
    rng = np.random.default_rng(1)
    th, W, L, R = general_problem(rng, 2, 4, 5, 2, 3)
    with pytest.raises(ValueError):
        OC.matrix_vector_product(th, W, R, R)


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("shape", [(2, 3, 4, 2, 3), (4, 5, 2, 3, 1), (1, 1, 1, 1, 1), (3, 6, 7, 2, 4)])
def test_c_env_literal_sum(side, shape):
    """Every element of the new block as the explicit sum over (s, s', a, a', w) of EnvBase.cpp:109-112 (L) or
    over (s, s', b, b', w') of :146-149 (R)."""
    d, cL, cR, DL, DR = shape
    rng = np.random.default_rng(sum(shape))
    th, W, L, R = general_problem(rng, d, cL, cR, DL, DR)
    if side == "L":
        # new[b, b2, w2] = sum L[a, a2, w] conj(th[s2, a2, b2]) W[w, w2, s, s2] th[s, a, b]
        ref = np.einsum("xyw,tyc,wvst,sxb->bcv", L, th.conj(), W, th)
        out = OC.env_build_block("L", L, th, W)
    else:
        # new[a, a2, w] = sum R[b, b2, w2] conj(th[s2, a2, b2]) W[w, w2, s, s2] th[s, a, b]
        ref = np.einsum("byv,tcy,wvst,sab->acw", R, th.conj(), W, th)
        out = OC.env_build_block("R", R, th, W)
    assert out.shape == ref.shape
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("shape", [(2, 130, 129, 5, 5), (2, 67, 200, 3, 4), (4, 96, 80, 8, 8)])
def test_c_env_matches_numpy_oracle(side, shape):
    d, cL, cR, DL, DR = shape
    rng = np.random.default_rng(cL * 3 + cR)
    th, W, L, R = general_problem(rng, d, cL, cR, DL, DR)
    other = L if side == "L" else R
    ref = O.env_build_block(side, other, th, W)
    out = OC.env_build_block(side, other, th, W, threads=4)
    assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)
