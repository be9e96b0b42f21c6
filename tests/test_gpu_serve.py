"""Serving: independent ising_tf_rf solves admitted by HBM footprint on one GPU.

Each result must match the same request run alone (|dE| <= 1e-10 |E|, mid-chain entropy within 1e-8) and the
free-fermion ground energy (1e-9 relative: chi_max may truncate)."""
import numpy as np
import pytest

import dmrg_amd as G
from dmrg_amd.workloads import ising_fields
from oracle import heff as O

pytestmark = pytest.mark.gpu


def make_requests():
    reqs = []
    Ls, chis, ws = [8, 10, 12], [16, 32], [0.5, 2.0, 6.0]
    for i in range(8):
        L, chi, w = Ls[i % 3], chis[i % 2], ws[i % 3]
        reqs.append({"L": L, "chi_max": chi, "h_rand": ising_fields(L, w, seed=100 + i), "arrival_s": 0.02 * i,
                     "seed": 100 + i, "w": w})
    return reqs


def test_serve_mix_matches_alone_and_exact():
    reqs = make_requests()
    res, st = G.serve(reqs, workers=4)
    assert st["completed"] == len(reqs) and st["failed"] == 0, (st, res)
    assert st["peak_concurrency"] >= 2
    assert st["p99_latency_s"] >= st["p50_latency_s"] > 0
    for r, q in zip(res, reqs):
        e0 = O.ising_free_fermion_ground_energy(1.0, 1.0, q["h_rand"])
        assert abs(r["energy"] - e0) <= 1e-9 * abs(e0), (r, e0)
        alone, _ = G.serve([dict(q, arrival_s=0.0)], workers=1)
        assert abs(alone[0]["energy"] - r["energy"]) <= 1e-10 * abs(e0)
        assert abs(alone[0]["entropy_mid"] - r["entropy_mid"]) <= 1e-8
        assert r["latency_s"] >= r["finish_s"] - r["start_s"] - 1e-9


def test_serve_budget_serialises():
    reqs = make_requests()[:4]
    fp = max(G._lib.load().dmrg_request_footprint(q["L"], q["chi_max"], 0) for q in reqs)
    res, st = G.serve(reqs, workers=4, hbm_budget_bytes=fp)
    assert st["completed"] == 4 and st["peak_concurrency"] == 1 and st["peak_admitted_bytes"] <= fp
    # no two requests overlap in time
    iv = sorted((r["start_s"], r["finish_s"]) for r in res)
    assert all(iv[k][1] <= iv[k + 1][0] + 1e-6 for k in range(len(iv) - 1))


def test_serve_rejects_oversized_request():
    q = make_requests()[0]
    with pytest.raises(G.DmrgError):
        G.serve([q], workers=1, hbm_budget_bytes=1 << 20)


@pytest.mark.parametrize("complex_", [False, True])
def test_serve_batched_eigensolves_match_alone_and_exact(complex_):
    """Requests whose energy solves run together in the shared batched eigensolver (grouped matvec launches,
    problems joining and leaving between Davidson steps) give the energies and entropies of the same requests
    run alone without batching (|dE| <= 1e-10 |E|, mid-chain entropy within 1e-8) and the free-fermion energy."""
    import sys
    print(f"[case] serve batched eigs complex={complex_}", file=sys.stderr, flush=True)
    reqs = make_requests()
    res, st = G.serve(reqs, workers=4, batch_eigs=True, complex_=complex_)
    assert st["completed"] == len(reqs) and st["failed"] == 0, (st, res)
    assert st["batch_steps"] > 0 and st["peak_batch"] >= 2, st
    for r, q in zip(res, reqs):
        e0 = O.ising_free_fermion_ground_energy(1.0, 1.0, q["h_rand"])
        assert abs(r["energy"] - e0) <= 1e-9 * abs(e0), (r, e0)
        alone, _ = G.serve([dict(q, arrival_s=0.0)], workers=1, complex_=complex_)
        assert abs(alone[0]["energy"] - r["energy"]) <= 1e-10 * abs(e0)
        assert abs(alone[0]["entropy_mid"] - r["entropy_mid"]) <= 1e-8
