"""Grouped matvecs (dmrg_matvec_group_*): the matvecs of several operators of uneven shapes as one launch per
contraction leg, direct and replayed from a HIP graph, against each operator's own matvec and the oracle.

Tolerances: FP64/CX64 relative 2-norm error <= 1e-12 per operator (oracle and single matvec); FP32/CX32 <= 1e-5.
Graph replays are checked one by one, with eager work and synchronisations between them.
"""
import sys

import numpy as np
import pytest

import dmrg_amd as G
from dmrg_amd.workloads import random_heff_problem
from oracle import heff as O

pytestmark = pytest.mark.gpu

SHAPES = [(2, 24, 24, 5, 5), (2, 40, 17, 3, 4), (2, 9, 33, 5, 2), (4, 16, 20, 8, 8), (2, 1, 1, 3, 3), (2, 64, 64, 5, 5),
          (2, 128, 96, 5, 5)]


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def problems(cplx, shapes=SHAPES, seed=3):
    out = []
    for i, (d, cL, cR, DL, DR) in enumerate(shapes):
        th, W, L, R = random_heff_problem(d, cL, cR, DL, DR, seed=seed + i, complex_=cplx)
        if not cplx:
            L, R = np.ascontiguousarray(L.real), np.ascontiguousarray(R.real)
        out.append((th, W, L, R))
    return out


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("cplx", [False, True])
def test_group_matches_single_and_oracle(cplx, graph):
    print(f"[case] group cplx={cplx} graph={graph}", file=sys.stderr, flush=True)
    dt = np.complex128 if cplx else np.float64
    probs = problems(cplx)
    ops = [G.MatVecMPOS([W], L, R, dtype=dt) for th, W, L, R in probs]
    xs = [G.DeviceArray.from_numpy(th.astype(dt).reshape(-1, order="F")) for th, *_ in probs]
    ys = [G.DeviceArray((x.shape[0],), dt) for x in xs]
    grp = G.MatvecGroup(ops, xs, ys, graph=graph)
    try:
        for rep in range(3):
            for y in ys:
                y.zero()
            grp.run()
            G._lib.call("dmrg_synchronize")
            for (th, W, L, R), op, x, y in zip(probs, ops, xs, ys):
                got = y.to_numpy()
                ref = O.matrix_vector_product(th.astype(dt), W.astype(dt), L.astype(dt), R.astype(dt)).reshape(-1, order="F")
                assert rel(got, ref) <= 1e-12, (th.shape, rep)
                single = G.DeviceArray((x.shape[0],), dt)
                op.MultAx_device(x, single)  # eager work between the replays
                assert rel(got, single.to_numpy()) <= 1e-12
                single.close()
        # new input in the same buffers: the next run (replay) uses it
        th0 = probs[0][0] * 0 + 1.0 / np.sqrt(probs[0][0].size)
        G._lib.call("dmrg_memcpy_htod", xs[0].ptr, np.ascontiguousarray(th0.astype(dt).reshape(-1, order="F")).ctypes.data_as(
            __import__("ctypes").c_void_p), xs[0].nbytes)
        grp.run()
        G._lib.call("dmrg_synchronize")
        th, W, L, R = probs[0]
        ref = O.matrix_vector_product(th0.astype(dt), W.astype(dt), L.astype(dt), R.astype(dt)).reshape(-1, order="F")
        assert rel(ys[0].to_numpy(), ref) <= 1e-12
        assert grp.flop > 0
    finally:
        grp.close()
        for a in xs + ys:
            a.close()
        for op in ops:
            op.close()


@pytest.mark.parametrize("cplx", [False, True])
def test_group_fp32_matches_single_fp32(cplx):
    dt, dt32 = (np.complex128, np.complex64) if cplx else (np.float64, np.float32)
    probs = problems(cplx, SHAPES[:4] + SHAPES[5:], seed=9)
    ops = [G.MatVecMPOS([W], L, R, dtype=dt) for th, W, L, R in probs]
    xs = [G.DeviceArray.from_numpy(th.astype(dt32).reshape(-1, order="F")) for th, *_ in probs]
    ys = [G.DeviceArray((x.shape[0],), dt32) for x in xs]
    grp = G.MatvecGroup(ops, xs, ys, fp32=True, graph=True)
    try:
        grp.run()
        G._lib.call("dmrg_synchronize")
        for (th, W, L, R), op, x, y in zip(probs, ops, xs, ys):
            single = G.DeviceArray((x.shape[0],), dt32)
            op.MultAx_device(x, single, fp32=True)
            ref = O.matrix_vector_product(th.astype(dt), W.astype(dt), L.astype(dt), R.astype(dt)).reshape(-1, order="F")
            assert rel(y.to_numpy(), single.to_numpy()) <= 1e-5
            assert rel(y.to_numpy(), ref) <= 1e-5
            single.close()
    finally:
        grp.close()
        for a in xs + ys:
            a.close()
        for op in ops:
            op.close()


def test_group_rejects_mixed_dtypes():
    th, W, L, R = random_heff_problem(2, 8, 8, 3, 3, seed=1)
    a = G.MatVecMPOS([W], L, R, dtype=np.complex128)
    b = G.MatVecMPOS([W], L.real.copy(), R.real.copy(), dtype=np.float64)
    x1, y1 = G.DeviceArray((128,), np.complex128), G.DeviceArray((128,), np.complex128)
    x2, y2 = G.DeviceArray((128,), np.float64), G.DeviceArray((128,), np.float64)
    try:
        with pytest.raises(G.DmrgError):
            G.MatvecGroup([a, b], [x1, x2], [y1, y2])
    finally:
        for o in (a, b, x1, y1, x2, y2):
            o.close()
