"""GPU parity at the sizes of the benchmark workloads (through the C ABI, against the oracle or numpy).

- H^2 matvec (squared MPO D^2 = 25 and squared env blocks, chi in {256, 512}, seed 2): FP64 relative 2-norm error
  <= 1e-12 against the oracle, FP32 matrix cores <= 5e-5; and H2 theta = H (H theta) through the oracle.
- Two-site matvec (theta (4, 1024, 1024), l-bit-style MPO of bond dim 8 with 1- to 3-body couplings, seed 3):
  FP64 <= 1e-12, FP32 <= 5e-5 against the oracle.
- Truncated SVD of the 2048 x 2048 two-site matrix to rank 1024 against numpy: singular values within
  1e-12 S[0], the truncation error |S[1024:]|_2 and the discarded weight sum S[1024:]^2 within 1e-10 relative.
- Serving requests drawn from the 512-request stream (L in {16, 20, 24, 32}, chi_max in {64, 128, 256},
  w in {0.5, 2, 6}, mixed precision for problems of size >= 131072): each result equals the same request run
  alone (|dE| <= 1e-10 |E|, mid-chain entropy within 1e-8) and the free-fermion ground energy (1e-9 relative).
"""
import sys

import numpy as np
import pytest

import dmrg_amd as G
from dmrg_amd.workloads import (ising_fields, random_heff_problem, squared_problem, theta_as_matrix,
                                two_site_problem)
from oracle import heff as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 5e-5


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def gpu_matvec(op, th, fp32=False):
    n = op.rows()
    dt = (np.complex64 if fp32 else np.complex128)
    dx = G.DeviceArray.from_numpy(th.reshape(n, order="F").astype(dt))
    dy = G.DeviceArray((n,), dt)
    try:
        op.MultAx_device(dx, dy, fp32=fp32)
        return dy.to_numpy().reshape(th.shape, order="F")
    finally:
        dx.close()
        dy.close()


@pytest.mark.parametrize("chi", [256, 512])
def test_h2_matvec_at_workload_size(chi):
    print(f"[case] H^2 matvec chi={chi}", file=sys.stderr, flush=True)
    th, W, L, R = random_heff_problem(2, chi, chi, 5, 5, seed=2)
    W2, L2, R2 = squared_problem(th, W, L, R)
    ref = O.matrix_vector_product(th, W2, L2, R2)
    # the squared operator is H applied twice (pins the construction through the oracle's own matvec)
    hh = O.matrix_vector_product(O.matrix_vector_product(th, W, L, R), W, L, R)
    assert rel(ref, hh) <= 1e-12
    op = G.MatVecMPOS([W2], L2, R2, dtype=np.complex128)
    try:
        assert rel(gpu_matvec(op, th), ref) <= TOL64
        assert rel(gpu_matvec(op, th, fp32=True), ref) <= TOL32
    finally:
        op.close()


def test_two_site_matvec_chi1024():
    print("[case] two-site matvec chi=1024 D=8", file=sys.stderr, flush=True)
    th, mpos, L, R = two_site_problem(1024, 8, seed=3)
    op = G.MatVecMPOS(mpos, L, R, dtype=np.complex128)
    try:
        ref = O.matrix_vector_product(th, O.merge_mpos(mpos), L, R)
        assert rel(gpu_matvec(op, th), ref) <= TOL64
        assert rel(gpu_matvec(op, th, fp32=True), ref) <= TOL32
    finally:
        op.close()


def test_truncated_svd_2048():
    """The 2048 x 2048 matrix of a two-site theta (4, 1024, 1024) with a graded column scaling (so that the kept
    and the discarded parts differ by orders of magnitude), truncated to rank 1024 by the one-sided Jacobi SVD
    and by gesvd."""
    print("[case] truncated SVD 2048 x 2048", file=sys.stderr, flush=True)
    th, _, _, _ = two_site_problem(1024, 8, seed=3)
    M = theta_as_matrix(th)
    M = np.asfortranarray(M * np.exp(-np.arange(M.shape[1]) / 400.0))
    S_ref = np.linalg.svd(M, compute_uv=False)
    tail = S_ref[1024:]
    for method in (3, 0):
        res = G.svd_truncate(M, 1e-14, 1024, method)
        S = res["S"]
        assert res["rank"] == 1024
        assert np.abs(S - S_ref).max() <= 1e-12 * S_ref[0], (method, np.abs(S - S_ref).max() / S_ref[0])
        assert abs(res["truncation_error"] - np.linalg.norm(tail)) <= 1e-10 * np.linalg.norm(tail)
        assert abs(res["discarded_weight"] - np.sum(tail ** 2)) <= 1e-10 * np.sum(tail ** 2)
        U, Vt = res["U"], res["Vt"]
        assert np.linalg.norm(U.conj().T @ U - np.eye(1024)) <= 1e-10
        assert np.linalg.norm(Vt @ Vt.conj().T - np.eye(1024)) <= 1e-10


def config4_requests():
    """Six requests drawn from the serving stream's grid, covering every L, chi_max and disorder width."""
    grid = [(16, 64, 0.5), (20, 128, 2.0), (24, 256, 6.0), (32, 64, 2.0), (32, 256, 0.5), (20, 256, 6.0)]
    return [{"L": L, "chi_max": chi, "h_rand": ising_fields(L, w, seed=100 + i), "arrival_s": 0.05 * i, "seed": 100 + i}
            for i, (L, chi, w) in enumerate(grid)]


def test_serve_config4_requests_alone_vs_mix():
    print("[case] serve config-4 requests", file=sys.stderr, flush=True)
    reqs = config4_requests()
    res, st = G.serve(reqs, workers=4, mixed=True)
    assert st["completed"] == len(reqs) and st["failed"] == 0, (st, res)
    assert st["peak_concurrency"] >= 2
    for r, q in zip(res, reqs):
        e0 = O.ising_free_fermion_ground_energy(1.0, 1.0, q["h_rand"])
        assert abs(r["energy"] - e0) <= 1e-9 * abs(e0), (q["L"], q["chi_max"], r["energy"], e0)
        alone, _ = G.serve([dict(q, arrival_s=0.0)], workers=1, mixed=True)
        assert abs(alone[0]["energy"] - r["energy"]) <= 1e-10 * abs(e0)
        assert abs(alone[0]["entropy_mid"] - r["entropy_mid"]) <= 1e-8
