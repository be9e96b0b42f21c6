"""The allocation-free matvec entry point (dmrg_matrix_vector_product_noalloc), through the C ABI, against the
allocating one (dmrg_matrix_vector_product) and the oracle: FP64/CX64 relative 2-norm error <= 1e-12, CX32
<= 5e-5, on calls in a row that share the fixed device scratch, each later call smaller than the first."""
import ctypes as C
import sys

import numpy as np
import pytest

import dmrg_amd as G
from dmrg_amd import _lib
from dmrg_amd.workloads import random_heff_problem
from oracle import heff as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def noalloc_matvec(th, W, L, R):
    lib = _lib.load()
    fn = lib.dmrg_matrix_vector_product_noalloc
    vp, i64p = C.c_void_p, C.POINTER(C.c_int64)
    fn.restype, fn.argtypes = C.c_int, [C.c_int, vp, vp, i64p, vp, i64p, vp, i64p, vp, i64p, vp]
    m, w, l, r = (G.DeviceArray.from_numpy(np.asfortranarray(x)) for x in (th, W, L, R))
    out = G.DeviceArray(m.shape, m.dtype)
    dims = lambda a: (C.c_int64 * len(a.shape))(*a.shape)
    try:
        _lib.check(fn(m.code, out.ptr, m.ptr, dims(m), w.ptr, dims(w), l.ptr, dims(l), r.ptr, dims(r), None),
                   "dmrg_matrix_vector_product_noalloc")
        return out.to_numpy()
    finally:
        for a in (m, w, l, r, out):
            a.close()


CASES = [  # (d, chiL, chiR, DL, DR, dtype): the first fills most of the scratch, the later ones are smaller
    (2, 512, 512, 5, 5, np.complex128),
    (2, 37, 53, 3, 4, np.float64),
    (4, 64, 48, 8, 8, np.complex64),
    (2, 100, 93, 9, 9, np.complex128),
]


def test_noalloc_matches_allocating_entry_point():
    max_bytes = _lib.load().dmrg_matrix_vector_product_noalloc_max_bytes
    max_bytes.restype = C.c_int64
    assert max_bytes() == 128 << 20
    for i, (d, cL, cR, DL, DR, dt) in enumerate(CASES):
        print(f"[case] noalloc call {i}: {(d, cL, cR, DL, DR)} {np.dtype(dt).name}", file=sys.stderr, flush=True)
        th, W, L, R = random_heff_problem(d, cL, cR, DL, DR, seed=30 + i, complex_=np.iscomplexobj(dt(0)))
        th, W, L, R = (x.astype(dt) for x in (th, W, L, R))
        out = noalloc_matvec(th, W, L, R)
        ref_gpu = G.matrix_vector_product(th, W, L, R)
        ref = O.matrix_vector_product(*(x.astype(np.complex128) for x in (th, W, L, R)))
        tol = 5e-5 if dt == np.complex64 else 1e-12
        assert out.dtype == np.dtype(dt)
        assert rel(out, ref_gpu) <= tol and rel(out, ref) <= tol, (rel(out, ref_gpu), rel(out, ref))


def test_noalloc_rejects_problems_beyond_its_scratch():
    """chi = 512 with D = 25 needs 420 MB of intermediates: refused with DMRG_ERR_ARG and a message that names the
    sizes, before anything is launched."""
    th, W, L, R = random_heff_problem(2, 512, 512, 25, 25, seed=1)
    with pytest.raises(G.DmrgError) as e:
        noalloc_matvec(th, W, L, R)
    assert e.value.status == 2 and "bytes" in str(e.value)
