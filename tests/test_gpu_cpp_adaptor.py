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


def jcb_rule(size, n):
    """MatVecMPOS::set_jcbMaxBlockSize (matvec_mpos.cpp:1483-1505), restated in Python."""
    jcb = min(max(size, 1), n)
    newsize = jcb
    best = (n % newsize, newsize)
    while newsize >= jcb // 2:
        rem = n % newsize
        if rem > best[0] or rem == 0:
            best = (rem, newsize)
        if rem == 0 or rem > 4 * newsize // 5:
            break
        newsize -= 2
    if best[1] != jcb:
        jcb = min(max(best[1], jcb // 2), n)
    return jcb


def test_jcb_rule_cases():
    assert jcb_rule(64, 240) == 62 and jcb_rule(1, 240) == 1 and jcb_rule(60, 240) == 60 and jcb_rule(1000, 240) == 240


SOLVER_BIN = Path(__file__).resolve().parent / "cpp" / "bin" / "solver_templates"


@pytest.mark.parametrize("cplx", [False, True])
def test_cpp_adaptor_in_solver_templates(tmp_path, cplx):
    """The adaptor compiled inside stand-ins of eig::solver::eigs_primme and eig::solver_arpack (their member
    expressions on the operator: set_jcbMaxBlockSize(std::optional<long>), preconditioner, num_op, t_multOPv;
    set_mode, set_side, get_form, &MatrixType::MultAx, reverse communication, the num_mv field). Each stand-in's
    Lanczos reaches the lowest eigenvalue of the dense H_eff (1e-9 |H|); num_mv counts the vectors applied; the
    block size follows the reference's rule (64 -> 62 for n = 240); the PRIMME preconditioner callback with
    factorization = LLT applies the block rule of the reference."""
    assert SOLVER_BIN.exists(), "build first (__graft_entry__.build())"
    d, cL, cR, DL, DR = 2, 12, 10, 3, 3
    th, W, L, R = random_heff_problem(d, cL, cR, DL, DR, seed=8, complex_=cplx)
    # Identity blocks at the MPO's edge channels (as on a chain), so that set_shift(s) shifts H_eff by -s I.
    L, R = L.copy(order="F"), R.copy(order="F")
    L[:, :, -1], R[:, :, 0] = np.eye(cL), np.eye(cR)
    n = d * cL * cR
    dt = np.complex128 if cplx else np.float64
    x = np.random.default_rng(3).standard_normal(n) + (1j * np.random.default_rng(4).standard_normal(n) if cplx else 0)
    H = O.dense_heff(W.astype(dt), L.astype(dt), R.astype(dt))
    w = np.linalg.eigvalsh(0.5 * (H + H.conj().T))
    lev = n // 2
    sig_in = w[lev] + 0.3 * (w[lev + 1] - w[lev])  # inside the spectrum, nearer level lev
    sig_lo = w[0] - 1.0                      # below it: H - sig_lo is positive definite
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([int(cplx), d, cL, cR, DL, DR, 64], dtype=np.int64).tofile(f)
        for a in (W, L, R):
            np.asfortranarray(a.astype(np.complex128)).ravel(order="F").tofile(f)
        x.astype(dt).tofile(f)
        np.array([sig_in, sig_lo], dtype=np.float64).tofile(f)
    r = subprocess.run([str(SOLVER_BIN), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    raw = np.fromfile(out, dtype=np.uint8)
    E = np.frombuffer(raw[:24].tobytes(), dtype=np.float64)
    k = np.frombuffer(raw[24:72].tobytes(), dtype=np.int64)
    off = 72 + n * np.dtype(dt).itemsize
    ypc = np.frombuffer(raw[72:off].tobytes(), dtype=dt)
    esi = np.frombuffer(raw[off:off + 24].tobytes(), dtype=np.float64)
    ksi = np.frombuffer(raw[off + 24:off + 40].tobytes(), dtype=np.int64)
    scale = np.abs(w).max()
    # Shift-invert (FactorOP / MultOPv): the lowest eigenvalue of (H - s)^-1 is 1 / (w[lev] - s) for the level just
    # below an interior shift, and 1 / (w[-1] - s) below the spectrum; MultAx of the shifted operator gives w[0] - s.
    assert abs(1.0 / esi[0] + sig_in - w[lev]) <= 1e-9 * scale, (esi[0], w[lev - 1:lev + 2], sig_in)
    assert abs(esi[1] - (w[0] - sig_lo)) <= 1e-9 * scale, (esi[1], w[0] - sig_lo)
    assert abs(1.0 / esi[2] + sig_lo - w[-1]) <= 1e-9 * scale, (esi[2], w[-1])
    assert ksi[0] > 0 and ksi[1] > 0, ksi
    for e in E:
        assert abs(e - w[0]) <= 1e-9 * scale, (E, w[0])
    mv_stub, mv_field, mv_arpack, jcb, steps_primme, steps_arpack = (int(v) for v in k)
    assert mv_stub == mv_field == steps_primme and mv_arpack == steps_arpack
    assert jcb == jcb_rule(64, n) == 62
    from test_gpu_precond import llt_rule_reference
    ref = llt_rule_reference(H, jcb, 0.5, x.astype(dt).reshape(n, 1))[:, 0]
    assert np.linalg.norm(ypc - ref) <= 1e-10 * np.linalg.norm(ref)
