"""CPU: the oracle's finite-DMRG sweep against the free-fermion closed form."""
import numpy as np
import pytest

from dmrg_amd.workloads import ising_fields
from oracle import heff as O
from oracle import sweep as S


@pytest.mark.parametrize("L,chi", [(6, 8), (10, 32)])
def test_oracle_sweep_exact(L, chi):
    h = ising_fields(L, 1.0, seed=L)
    r = S.fdmrg_ising(L, h, chi_max=chi)
    e0 = O.ising_free_fermion_ground_energy(1.0, 1.0, h)
    assert abs(r["energy"] - e0) <= 1e-10 * abs(e0)
    assert r["variance"] < 1e-12
    np.testing.assert_allclose(r["entropies"], O.ising_free_fermion_entropies(1.0, 1.0, h), atol=1e-8)


def test_rank_rule():
    s = np.array([0.8, 0.5, 0.3, 1e-8, 1e-16])
    r, te = S._rank(s, 1e-14, 0)
    assert r == 4 and abs(te - 1e-16) < 1e-30      # |S[4:]| = 1e-16 < lim is discarded
    r, te = S._rank(s, 1e-14, 2)
    assert r == 2 and abs(te - np.linalg.norm(s[2:])) < 1e-15
