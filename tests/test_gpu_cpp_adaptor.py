"""GPU test of the C++ drop-in (include/dmrg_amd/matvec_mpos.hpp): the compiled driver calls MatVecMPOS<T>
through a PRIMME-style matrixMatvec function pointer (strided block of vectors), MultAx(T*, T*), and the free
matrix_vector_product / expectation_value; results are compared with the oracle (1e-12 relative)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from dmrg_amd.workloads import random_heff_problem
from oracle import heff as O

pytestmark = pytest.mark.gpu
BIN = Path(__file__).resolve().parent / "cpp" / "bin" / "heff_cpp_driver"


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", [(2, 16, 12, 3, 3), (2, 64, 64, 5, 5), (4, 20, 33, 8, 8)])
def test_cpp_adaptor(tmp_path, cplx, shape):
    assert BIN.exists(), "build first (__graft_entry__.build())"
    d, cL, cR, DL, DR = shape
    th, W, L, R = random_heff_problem(d, cL, cR, DL, DR, seed=5, complex_=cplx)
    n, nvec, ld = d * cL * cR, 3, d * cL * cR + 5
    dt = np.complex128 if cplx else np.float64
    rng = np.random.default_rng(1)
    X = np.zeros((ld, nvec), dtype=dt, order="F")
    X[:n] = rng.standard_normal((n, nvec)) + (1j * rng.standard_normal((n, nvec)) if cplx else 0)
    X[:n, 0] = th.reshape(-1, order="F")
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([int(cplx), d, cL, cR, DL, DR, nvec, ld], dtype=np.int64).tofile(f)
        for a in (W, L, R):
            np.asfortranarray(a.astype(np.complex128)).ravel(order="F").tofile(f)
        X.ravel(order="F").tofile(f)
    r = subprocess.run([str(BIN), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, dtype=np.uint8)
    isz = np.dtype(dt).itemsize
    off = 0
    Y = np.frombuffer(raw[off:off + ld * nvec * isz].tobytes(), dtype=dt).reshape(ld, nvec, order="F")
    off += ld * nvec * isz
    y1 = np.frombuffer(raw[off:off + n * isz].tobytes(), dtype=dt)
    off += n * isz
    y2 = np.frombuffer(raw[off:off + n * isz].tobytes(), dtype=dt)
    off += n * isz
    ev = np.frombuffer(raw[off:off + 16].tobytes(), dtype=np.float64)
    off += 16
    nmv = np.frombuffer(raw[off:off + 8].tobytes(), dtype=np.int64)[0]
    for j in range(nvec):
        ref = O.matrix_vector_product(X[:n, j].reshape(d, cL, cR, order="F"), W, L, R).reshape(-1, order="F")
        assert np.linalg.norm(Y[:n, j] - ref) <= 1e-12 * np.linalg.norm(ref)
        assert np.all(Y[n:, j] == 0)  # rows beyond n of each column untouched
    ref0 = O.matrix_vector_product(th, W, L, R).reshape(-1, order="F")
    assert np.linalg.norm(y1 - ref0) <= 1e-12 * np.linalg.norm(ref0)
    assert np.linalg.norm(y2 - ref0) <= 1e-12 * np.linalg.norm(ref0)
    evr = O.expectation_value(th, W, L, R)
    assert abs(complex(ev[0], ev[1]) - evr) <= 1e-12 * abs(evr) * 10
    assert nmv == nvec + 1
    # Jacobi preconditioner callback: y_i = x_i / (diag(H_eff) - shift_i)
    Ypc = np.frombuffer(raw[off + 8:off + 8 + ld * nvec * isz].tobytes(), dtype=dt).reshape(ld, nvec, order="F")
    off += 8 + ld * nvec * isz
    ysh = np.frombuffer(raw[off:off + n * isz].tobytes(), dtype=dt)
    off += n * isz
    npc = np.frombuffer(raw[off:off + 8].tobytes(), dtype=np.int64)[0]
    diag = heff_diagonal(W, L, R)
    for j, sh in enumerate([0.3, -0.7, 1.1][:nvec]):
        # backward error: (diag - shift) y = x, robust to small denominators
        resid = (diag - sh) * Ypc[:n, j] - X[:n, j]
        assert np.linalg.norm(resid) <= 1e-12 * np.linalg.norm(X[:n, j]) * np.max(np.abs(diag - sh))
    assert npc == nvec
    # set_shift(0.25) adds (0 - 0.25) I to the MPO's bottom-left block W[DL-1, 0] (matvec_mpos.cpp:1433-1472)
    Ws = np.array(W, dtype=np.complex128, order="F")
    Ws[DL - 1, 0] += -0.25 * np.eye(d)
    if not cplx:
        Ws = Ws.real.copy(order="F")
    ref_sh = O.matrix_vector_product(th, Ws, L, R).reshape(-1, order="F")
    assert np.linalg.norm(ysh - ref_sh) <= 1e-12 * np.linalg.norm(ref_sh)


def heff_diagonal(W, L, R):
    """diag H_eff[(s,a,b),(s,a,b)] = sum_{w,w'} L[a,a,w] W[w,w',s,s] R[b,b,w'] (from the brute-force definition)."""
    Ld = np.einsum("aaw->aw", L)
    Rd = np.einsum("bbv->bv", R)
    Wd = np.einsum("wvss->wvs", W)
    return np.einsum("aw,wvs,bv->sab", Ld, Wd, Rd).reshape(-1, order="F")


def test_diagonal_helper_matches_brute_force():
    th, W, L, R = random_heff_problem(2, 3, 4, 3, 2, seed=9, complex_=True)
    assert np.allclose(heff_diagonal(W, L, R), np.diag(O.dense_heff(W, L, R)), rtol=0, atol=1e-13)
