"""GPU parity of the excited-state path: the folded-spectrum executor (optimize_folded_spectrum.cpp:235-342)
and xDMRG sweeps on the folded spectrum (xdmrg.cpp:297-341, 487-501), against exact diagonalisation of the
ising_tf_rf chain (oracle.heff.ising_dense).

Tolerances: the energy equals the exact eigenvalue closest to the target within 1e-9 * |H|, the variance
<(H-E_t)^2> - (E-E_t)^2 is below 1e-8 (an eigenstate, within what the folded problem resolves).
"""
import sys

import numpy as np
import pytest

import dmrg_amd as G
from oracle import heff as O

pytestmark = pytest.mark.gpu


def chain(L, seed):
    h = np.random.default_rng(seed).uniform(-1, 1, L)
    return h, O.ising_dense(L, 1.0, 1.0, h)


@pytest.mark.parametrize("frac", [0.3, 0.5, 0.7])
def test_folded_spectrum_executor_full_chain(frac):
    """H and (H - E_s)^2 of a whole 6-site chain as one merged MPO each (chi = 1 envs): the executor returns the
    eigenvector whose energy is closest to E_s. (Larger chains put the two closest folded levels within ~1e-4 of
    the folded norm, where any iterative solver may stop at the second one; the sweep sends such sizes to the
    dense solver, as the reference does up to eig_max_size.)"""
    print(f"[case] folded executor frac={frac}", file=sys.stderr, flush=True)
    L = 6
    h, Hd = chain(L, 7)
    w = np.linalg.eigvalsh(Hd)
    j = int(frac * (len(w) - 1))
    Es = w[j] + 0.1 * (w[j + 1] - w[j])  # w[j] is the closest level, the next folded level 81x further
    one = np.ones((1, 1, 1))
    W1 = O.merge_mpos(O.ising_tf_rf_mpos(L, 1.0, 1.0, h))
    W2 = O.merge_mpos([O.mpo_squared(m) for m in O.ising_tf_rf_mpos(L, 1.0, 1.0, h, e_shift=Es / L)])
    hop = G.MatVecMPOS([W1], one, one, dtype=np.float64)
    h2op = G.MatVecMPOS([W2], one, one, dtype=np.float64)
    try:
        x0 = np.random.default_rng(1).standard_normal(hop.rows())
        E, var, vec, info = h2op.eigs_folded_spectrum(hop, x0, Es, target="CLOSEST", shift=0.0, tol=1e-13)
        k = int(np.argmin(np.abs(w - Es)))
        assert abs(E - w[k]) <= 1e-9 * np.abs(w).max(), (E, w[k], info)
        assert var <= 1e-8, var
        r = Hd.T @ vec - E * vec
        assert np.linalg.norm(r) <= 1e-4 * np.linalg.norm(vec)
    finally:
        hop.close()
        h2op.close()


@pytest.mark.parametrize("frac,cplx", [(0.25, False), (0.5, False), (0.5, True)])
def test_xdmrg_finds_eigenstate_near_target(frac, cplx):
    """xDMRG on L=8 with chi_max=16 (exact bond dimensions): the sweeps end in the eigenstate closest to the
    target energy."""
    print(f"[case] xdmrg frac={frac} cplx={cplx}", file=sys.stderr, flush=True)
    L = 8
    h, Hd = chain(L, 11)
    w = np.linalg.eigvalsh(Hd)
    j = int(frac * (len(w) - 1))
    Et = w[j] + 0.1 * (w[j + 1] - w[j])
    r = G.fdmrg_ising_tf_rf(L, h, chi_max=16, xdmrg=True, energy_target=Et, var_tol=1e-10, eigs_tol=1e-13,
                            max_sweeps=30, two_site_sweeps=20, seed=3, complex_=cplx)
    k = int(np.argmin(np.abs(w - Et)))
    assert abs(r["variance"]) <= 1e-8, r
    assert abs(r["energy"] - w[k]) <= 1e-9 * np.abs(w).max(), (r["energy"], w[k], w[max(0, k - 2):k + 3])
