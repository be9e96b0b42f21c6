"""Finite DMRG sweeps on the GPU against the free-fermion closed form of ising_tf_rf.

Bars: |E - E_exact| <= 1e-10 |E| when the bond dimension holds the ground state exactly (chi >= 2^(L/2)),
and 1e-9 |E| otherwise (truncation at chi_max); entropies within 1e-8 (exact) / 1e-6 (truncated);
energy variance below the requested 1e-12.
"""
import numpy as np
import pytest

import dmrg_amd as G
from dmrg_amd.workloads import ising_fields
from oracle import heff as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("expand", [0, None])
@pytest.mark.parametrize("L,chi,cplx", [(6, 8, False), (8, 16, False), (8, 16, True)])
def test_sweep_exact_bond_dimension(L, chi, cplx, expand):
    """Two-site growth (expand_mode 0) and the default subspace expansion (43) reach the exact ground state."""
    h = ising_fields(L, 1.0, seed=L)
    r = G.fdmrg_ising_tf_rf(L, h, chi_max=chi, complex_=cplx, expand_mode=expand)
    e0 = O.ising_free_fermion_ground_energy(1.0, 1.0, h)
    assert abs(r["energy"] - e0) <= 1e-10 * abs(e0), (r["energy"], e0, r)
    assert r["converged"] == 1 and r["variance"] < 1e-12
    np.testing.assert_allclose(r["entropies"], O.ising_free_fermion_entropies(1.0, 1.0, h), atol=1e-8)


def test_sweep_l16_chi64_seed0():
    """L=16, J=1, g=1, h_i ~ U[-1,1] (seed 0), chi_max=64, single-site sweeps to variance < 1e-12, FP64 real, with
    the default growth (the reference's subspace expansion)."""
    L = 16
    h = ising_fields(L, 1.0, seed=0)
    r = G.fdmrg_ising_tf_rf(L, h, chi_max=64)
    e0 = O.ising_free_fermion_ground_energy(1.0, 1.0, h)
    assert abs(r["energy"] - e0) <= 1e-9 * abs(e0), (r["energy"], e0)
    assert r["variance"] < 1e-12
    np.testing.assert_allclose(r["entropies"], O.ising_free_fermion_entropies(1.0, 1.0, h), atol=1e-6)
    print(f"L=16 chi=64: E={r['energy']:.15f} exact={e0:.15f} var={r['variance']:.2e} sweeps={r['sweeps']} "
          f"time={r['seconds']:.2f}s eigs={r['t_eigs']:.2f}s env={r['t_env']:.2f}s svd={r['t_svd']:.2f}s")


def test_sweep_mixed_matches_fp64():
    L = 12
    h = ising_fields(L, 1.0, seed=3)
    r64 = G.fdmrg_ising_tf_rf(L, h, chi_max=32)
    rmx = G.fdmrg_ising_tf_rf(L, h, chi_max=32, mixed=True, mixed_min_n=1)
    assert abs(rmx["energy"] - r64["energy"]) <= 1e-10 * abs(r64["energy"])
    assert rmx["variance"] < 1e-12 and rmx["matvecs_fp32"] > 0
    np.testing.assert_allclose(rmx["entropies"], r64["entropies"], atol=1e-8)


@pytest.mark.parametrize("cplx", [False, True])
def test_sweep_dense_branch_matches_iterative(cplx):
    """eig_max_size = 4096 (the reference's setting: dense solves for problems up to 4096, here every solve of an
    L=10, chi=32 chain) gives the energy and entropies of the iterative sweep and the free-fermion closed form."""
    L = 10
    h = ising_fields(L, 1.0, seed=5)
    rit = G.fdmrg_ising_tf_rf(L, h, chi_max=32, complex_=cplx)
    rde = G.fdmrg_ising_tf_rf(L, h, chi_max=32, complex_=cplx, eig_max_size=4096)
    e0 = O.ising_free_fermion_ground_energy(1.0, 1.0, h)
    assert abs(rde["energy"] - e0) <= 1e-10 * abs(e0), (rde["energy"], e0)
    assert abs(rde["energy"] - rit["energy"]) <= 1e-10 * abs(e0)
    assert rde["variance"] < 1e-12 and rde["matvecs"] < rit["matvecs"]
    np.testing.assert_allclose(rde["entropies"], rit["entropies"], atol=1e-8)
    np.testing.assert_allclose(rde["entropies"], O.ising_free_fermion_entropies(1.0, 1.0, h), atol=1e-8)


@pytest.mark.parametrize("L,chi,cplx,mode", [(8, 16, False, 43), (8, 16, True, 43), (16, 64, False, 43), (12, 32, False, 41),
                                             (12, 32, False, 42)])
def test_sweep_subspace_expansion(L, chi, cplx, mode):
    """Single-site sweeps whose bonds grow only by the reference's default subspace expansion (expand_mode H1 | H2 |
    NSITE | FORWARD = 43; also H1 or H2 alone), from the product state: the free-fermion energy and entropies (exact
    when chi >= 2^(L/2), truncated otherwise) and variance below 1e-12. Unsupported modes are refused."""
    import sys

    print(f"[case] sweep expansion L={L} chi={chi} cplx={cplx} mode={mode}", file=sys.stderr, flush=True)
    h = ising_fields(L, 1.0, seed=L + 1)
    r = G.fdmrg_ising_tf_rf(L, h, chi_max=chi, complex_=cplx, expand_mode=mode, max_sweeps=40)
    e0 = O.ising_free_fermion_ground_energy(1.0, 1.0, h)
    exact = chi >= 2 ** (L // 2)
    assert r["expansions"] > 0 and r["converged"] == 1, r
    assert abs(r["energy"] - e0) <= (1e-10 if exact else 1e-9) * abs(e0), (r["energy"], e0, r["sweeps"])
    assert r["variance"] < 1e-12
    np.testing.assert_allclose(r["entropies"], O.ising_free_fermion_entropies(1.0, 1.0, h), atol=1e-8 if exact else 1e-6)
    print(f"[info] sweeps {r['sweeps']} expansions {r['expansions']} t_expand {r['t_expand']:.3f}s total {r['seconds']:.2f}s")
    with pytest.raises(G.DmrgError):
        G.fdmrg_ising_tf_rf(L, h, chi_max=chi, expand_mode=1 | 2 | 4)  # SSITE
