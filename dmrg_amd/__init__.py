"""dmrg_amd: MI355X-native H_eff hot path of DMRG++ (matvec, environment updates, eigensolver).

Python mirror of the C ABI (include/dmrg_amd.h), used by the tests and the benchmark.
Names follow the reference's C++ API (tools::common::contraction::matrix_vector_product,
EnvBase::build_block, MatVecMPOS<T>). Tensors are numpy arrays with the reference's
index order and column-major (Fortran) memory layout:
    mps (d, chiL, chiR), env (chi, chi, D), mpo (DL, DR, d, d).
All compute runs in libdmrg_amd.so on the GPU; nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import DMRG_CX32, DMRG_CX64, DMRG_FP32, DMRG_FP64, DmrgError, call, load

__all__ = [
    "DeviceArray", "MatVecMPOS", "matrix_vector_product", "env_build_block", "expectation_value",
    "dot", "nrm2", "device_info", "DmrgError", "dtype_code", "heff_dense", "dense_eigh", "env_set_edge",
    "svd_truncate",
]

_CODE = {np.dtype(np.float64): DMRG_FP64, np.dtype(np.complex128): DMRG_CX64,
         np.dtype(np.float32): DMRG_FP32, np.dtype(np.complex64): DMRG_CX32}
_NP = {v: k for k, v in _CODE.items()}


def dtype_code(dt) -> int:
    return _CODE[np.dtype(dt)]


def _dims(shape) -> C.Array:
    return (C.c_int64 * len(shape))(*[int(s) for s in shape])


class DeviceArray:
    """A column-major tensor in device memory (owned; freed on close/GC)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        call("dmrg_malloc", C.byref(p), max(self.nbytes, 0))
        self.ptr = p

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.asfortranarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            call("dmrg_memcpy_htod", d.ptr, a.ctypes.data_as(C.c_void_p), d.nbytes)
        return d

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order="F")
        if self.nbytes:
            call("dmrg_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes)
        return out

    def zero(self) -> None:
        if self.nbytes:
            call("dmrg_memset", self.ptr, 0, self.nbytes)

    @property
    def code(self) -> int:
        return _CODE[self.dtype]

    def close(self) -> None:
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            try:
                call("dmrg_free", self.ptr)
            finally:
                self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _dev(a) -> DeviceArray:
    return a if isinstance(a, DeviceArray) else DeviceArray.from_numpy(np.asarray(a))


def device_info() -> dict:
    name = C.create_string_buffer(256)
    cus = C.c_int()
    mem = C.c_size_t()
    call("dmrg_device_info", name, 256, C.byref(cus), C.byref(mem))
    return {"name": name.value.decode(), "cu_count": cus.value, "hbm_bytes": mem.value}


def matrix_vector_product(mps, mpo, envL, envR, stream=None):
    """out = H_eff mps (tools::common::contraction::matrix_vector_product, contraction_tblis.cpp:213-292).

    numpy inputs -> numpy output; DeviceArray inputs -> DeviceArray output."""
    on_dev = isinstance(mps, DeviceArray)
    m, w, l, r = _dev(mps), _dev(mpo), _dev(envL), _dev(envR)
    out = DeviceArray(m.shape, m.dtype)
    call("dmrg_matrix_vector_product", m.code, out.ptr, m.ptr, _dims(m.shape), w.ptr, _dims(w.shape), l.ptr,
         _dims(l.shape), r.ptr, _dims(r.shape), stream)
    if on_dev:
        return out
    res = out.to_numpy()
    out.close()
    return res


def env_build_block(side: str, otherblock, mps, mpo, stream=None):
    """EnvBase::build_block (EnvBase.cpp:67-151): the new L (side 'L') or R (side 'R') block."""
    on_dev = isinstance(mps, DeviceArray)
    o, m, w = _dev(otherblock), _dev(mps), _dev(mpo)
    if side == "L":
        shape = (m.shape[2], m.shape[2], w.shape[1])
    elif side == "R":
        shape = (m.shape[1], m.shape[1], w.shape[0])
    else:
        raise ValueError("side must be 'L' or 'R'")
    out = DeviceArray(shape, m.dtype)
    call("dmrg_env_build_block", m.code, C.c_char(side.encode()), out.ptr, o.ptr, _dims(o.shape), m.ptr,
         _dims(m.shape), w.ptr, _dims(w.shape), stream)
    if on_dev:
        return out
    res = out.to_numpy()
    out.close()
    return res


def env_expansion_term(side: str, block, mps, mpo, alpha: float = 1.0, stream=None):
    """Subspace-expansion term alpha * (block . mps . mpo) (EnvBase::get_expansion_term, EnvBase.cpp:303-361).
    side 'L': shape (d, chiL, chiR*DR); side 'R': shape (d, chiL*DL, chiR) in the reference's storage order."""
    on_dev = isinstance(mps, DeviceArray)
    m = _dev(mps)
    b, w = _dev(block), _dev(mpo)
    d, cL, cR = m.shape
    if side == "L":
        shape = (d, cL, cR * w.shape[1])
    elif side == "R":
        shape = (d, cL * w.shape[0], cR)
    else:
        raise ValueError("side must be 'L' or 'R'")
    out = DeviceArray(shape, m.dtype)
    call("dmrg_env_expansion_term", m.code, C.c_char(side.encode()), out.ptr, b.ptr, _dims(b.shape), m.ptr,
         _dims(m.shape), w.ptr, _dims(w.shape), float(alpha), stream)
    if on_dev:
        return out
    res = out.to_numpy()
    out.close()
    return res


def expectation_value(mps, mpo, envL, envR, stream=None) -> complex:
    """<mps|H_eff|mps> (contraction_tblis.cpp:30-97)."""
    m, w, l, r = _dev(mps), _dev(mpo), _dev(envL), _dev(envR)
    res = (C.c_double * 2)()
    call("dmrg_expectation_value", m.code, m.ptr, _dims(m.shape), w.ptr, _dims(w.shape), l.ptr, _dims(l.shape),
         r.ptr, _dims(r.shape), res, stream)
    return complex(res[0], res[1]) if np.iscomplexobj(np.empty(0, m.dtype)) else res[0]


def dot(x, y, stream=None):
    """sum conj(x) y (contract_mps_mps_overlap, contraction_tblis.cpp:518-533)."""
    a, b = _dev(x), _dev(y)
    res = (C.c_double * 2)()
    n = int(np.prod(a.shape))
    call("dmrg_dot", a.code, n, a.ptr, b.ptr, res, stream)
    return complex(res[0], res[1]) if a.dtype.kind == "c" else res[0]


def nrm2(x, stream=None) -> float:
    a = _dev(x)
    res = C.c_double()
    call("dmrg_nrm2", a.code, int(np.prod(a.shape)), a.ptr, C.byref(res), stream)
    return res.value


def heff_dense(mpo, envL, envR, transpose: bool = False, stream=None) -> np.ndarray:
    """Explicit H_eff (n x n): contract_mpo_env (TensorsFinite.cpp:119-142) so that H @ theta.ravel('F') is the
    matvec; transpose=True gives MatVecMPO::get_matrix's orientation (matvec_mpo.cpp:313-332)."""
    w, l, r = _dev(mpo), _dev(envL), _dev(envR)
    n = w.shape[2] * l.shape[0] * r.shape[0]
    out = DeviceArray((n, n), w.dtype)
    try:
        call("dmrg_heff_dense", w.code, out.ptr, n, w.ptr, _dims(w.shape), l.ptr, _dims(l.shape), r.ptr,
             _dims(r.shape), int(transpose), stream)
        return out.to_numpy()
    finally:
        out.close()


def dense_eigh(A, vectors: bool = True, stream=None):
    """All eigenpairs of a Hermitian matrix on the GPU (rocSOLVER syevd/heevd; the reference's dense
    eig::solver::eig branch, solver.cpp:205-245). Returns (w ascending, V or None)."""
    a = DeviceArray.from_numpy(np.asarray(A))
    n = a.shape[0]
    w = np.empty(n, dtype=np.float64)
    try:
        call("dmrg_dense_eigh", a.code, n, a.ptr, n, w.ctypes.data_as(C.POINTER(C.c_double)), int(vectors), stream)
        return w, (a.to_numpy() if vectors else None)
    finally:
        a.close()


def env_set_edge(chi: int, edge, stream=None) -> np.ndarray:
    """EnvBase::set_edge_dims (EnvBase.cpp:254-285): block (chi, chi, D) with block[i, i, :] = edge."""
    e = np.ascontiguousarray(edge)
    out = DeviceArray((chi, chi, e.shape[0]), e.dtype)
    try:
        call("dmrg_env_set_edge", out.code, out.ptr, chi, e.shape[0], e.ctypes.data_as(C.c_void_p), stream)
        return out.to_numpy()
    finally:
        out.close()


def svd_truncate(A, trunc_lim: float = 1e-14, rank_max: int = 0, method: int = 3, stream=None):
    """Truncated SVD on the GPU (svd::solver, svd_lapacke.cpp:82-330; rank rule svd.cpp:257-280); method 3 is the
    one-sided Jacobi SVD (gesvj), 0 rocSOLVER's gesvd.

    Returns dict(U (m x r), S (all singular values), Vt (r x n), rank, truncation_error = |S[r:]|_2,
    discarded_weight = sum S[r:]^2). A may be a numpy matrix or a DeviceArray."""
    a = _dev(A)
    m, n = a.shape
    k = min(m, n)
    U, Vt = DeviceArray((m, k), a.dtype), DeviceArray((k, n), a.dtype)
    S = np.empty(k, dtype=np.float64)
    rank, terr, disc = C.c_int64(), C.c_double(), C.c_double()
    try:
        call("dmrg_svd_truncate", a.code, m, n, a.ptr, trunc_lim, rank_max, method, U.ptr,
             S.ctypes.data_as(C.POINTER(C.c_double)), Vt.ptr, C.byref(rank), C.byref(terr), C.byref(disc), stream)
        r = rank.value
        u = U.to_numpy()[:, :r] if r else U.to_numpy()[:, :0]
        vt_flat = Vt.to_numpy().ravel(order="F")[: r * n]
        vt = vt_flat.reshape((r, n), order="F")
        return {"U": u, "S": S, "Vt": vt, "rank": r, "truncation_error": terr.value, "discarded_weight": disc.value}
    finally:
        U.close()
        Vt.close()
        if a is not A:
            a.close()


def fdmrg_ising_tf_rf(L: int, h_rand, J: float = 1.0, h_tran: float = 1.0, chi_max: int = 64, max_sweeps: int = 20,
                      two_site_sweeps: int = 10, var_tol: float = 1e-12, trunc_lim: float = 1e-14, eigs_tol: float = 1e-12,
                      mixed: bool = False, complex_: bool = False, seed: int = 0, mixed_min_n: int = 0,
                      xdmrg: bool = False, energy_target: float = 0.0, eig_max_size: int = 0, expand_mode=None,
                      stream=None) -> dict:
    """Finite DMRG ground state of the random transverse-field Ising chain on the GPU
    (fdmrg::run_algorithm with ising_tf_rf, fdmrg.cpp:170-376), or with xdmrg the eigenstate found by minimising
    <(H - energy_target)^2> (xdmrg.cpp:297-341, folded spectrum). expand_mode: EnvExpandMode bits; None keeps the
    library default 43 (H1 | H2 | NSITE | FORWARD, the reference's default: the bonds grow by subspace expansion in
    single-site sweeps); 0 grows them with two-site sweeps first.
    Returns energies, variance, entropies, timings."""
    h = np.ascontiguousarray(np.asarray(h_rand, dtype=np.float64))
    assert h.shape == (L,)
    cfg = _lib.FdmrgConfig()
    load().dmrg_fdmrg_default_config(C.byref(cfg))
    cfg.L, cfg.J, cfg.h_tran = L, J, h_tran
    cfg.h_rand = h.ctypes.data_as(C.POINTER(C.c_double))
    cfg.chi_max, cfg.max_sweeps, cfg.two_site_sweeps = chi_max, max_sweeps, two_site_sweeps
    cfg.var_tol, cfg.trunc_lim, cfg.eigs_tol = var_tol, trunc_lim, eigs_tol
    cfg.mixed, cfg.complex_, cfg.seed, cfg.mixed_min_n = int(mixed), int(complex_), seed, mixed_min_n
    cfg.xdmrg, cfg.energy_target = int(xdmrg), float(energy_target)
    cfg.eig_max_size = int(eig_max_size)
    if expand_mode is not None:
        cfg.expand_mode = int(expand_mode)
    res = _lib.FdmrgResult()
    call("dmrg_fdmrg_ising_tf_rf", C.byref(cfg), C.byref(res), stream)
    out = {k: getattr(res, k) for k, _ in res._fields_ if k not in ("entropies", "chi")}
    out["entropies"] = np.array(res.entropies[: L + 1])
    out["chi"] = np.array(res.chi[: L + 1])
    return out


def serve(requests, workers: int = 4, hbm_budget_bytes: int = 64 << 30, mixed: bool = False, complex_: bool = False,
          var_tol: float = 1e-12, eigs_tol: float = 1e-12, max_sweeps: int = 20, two_site_sweeps: int = 10,
          J: float = 1.0, h_tran: float = 1.0, batch_eigs: bool = False, expand_mode: int = 0):
    """Run independent ising_tf_rf ground-state requests as they arrive. requests: list of dicts with keys
    L, chi_max, h_rand, arrival_s, seed. With batch_eigs the requests' FP64 energy solves run together in one
    batched eigensolver (grouped matvec launches). expand_mode: the sweeps' EnvExpandMode (0: two-site growth;
    43: the reference's subspace expansion). Returns (per-request results, stats)."""
    n = len(requests)
    fields = [np.ascontiguousarray(np.asarray(r["h_rand"], dtype=np.float64)) for r in requests]
    arr = (_lib.Request * max(n, 1))()
    for i, r in enumerate(requests):
        arr[i].L, arr[i].chi_max = int(r["L"]), int(r["chi_max"])
        arr[i].h_rand = fields[i].ctypes.data_as(C.POINTER(C.c_double))
        arr[i].arrival_s, arr[i].seed = float(r.get("arrival_s", 0.0)), int(r.get("seed", 0))
    cfg = _lib.ServeConfig(workers, hbm_budget_bytes, int(mixed), int(complex_), var_tol, eigs_tol, max_sweeps,
                           two_site_sweeps, J, h_tran, int(batch_eigs), int(expand_mode))
    out = (_lib.RequestResult * max(n, 1))()
    st = _lib.ServeStats()
    call("dmrg_serve", C.byref(cfg), arr, n, out, C.byref(st))
    res = [{k: getattr(out[i], k) for k, _ in out[i]._fields_} for i in range(n)]
    stats = {k: getattr(st, k) for k, _ in st._fields_}
    return res, stats


def eigs_batched(ops, tol: float = 1e-12, target: str = "SA", x0s=None, ncv: int = 0, max_matvecs: int = 100000,
                 seed: int = 12345, mixed: bool = False, stream=None):
    """Lowest (SA) or highest (LA) eigenpair of each operator, solved together in lockstep (dmrg_eigs_batched);
    mixed: an FP32 phase on the FP32 matrix cores, then FP64 to tol. Returns a list of (eigenvalue, eigenvector
    numpy, result dict)."""
    k = len(ops)
    cfg = _lib.EigsConfig()
    load().dmrg_eigs_default_config(C.byref(cfg))
    cfg.nev, cfg.ncv, cfg.target = 1, ncv, {"SA": 0, "LA": 1}[target]
    cfg.tol, cfg.max_matvecs, cfg.seed, cfg.mixed = tol, max_matvecs, seed, int(mixed)
    outs = [DeviceArray((o.rows(),), o.dtype) for o in ops]
    ins = [None] * k
    if x0s is not None:
        ins = [DeviceArray.from_numpy(np.asarray(x, dtype=o.dtype).reshape(-1, order="F")) if x is not None else None
               for x, o in zip(x0s, ops)]
    arr_ops = (C.c_void_p * k)(*[o.h.value for o in ops])
    arr_x = (C.c_void_p * k)(*[(x.ptr.value if x is not None else None) for x in ins])
    arr_y = (C.c_void_p * k)(*[y.ptr.value for y in outs])
    res = (_lib.EigsResult * k)()
    try:
        call("dmrg_eigs_batched", arr_ops, k, C.byref(cfg), arr_x, arr_y, res, stream)
        out = []
        for i in range(k):
            info = {f: getattr(res[i], f) for f, _ in res[i]._fields_ if f not in ("eigval", "rnorm")}
            info["rnorm"] = res[i].rnorm[0]
            out.append((res[i].eigval[0], outs[i].to_numpy(), info))
        return out
    finally:
        for a in outs + [x for x in ins if x is not None]:
            a.close()


def serve_eigs(ops, arrivals=None, hbm_budget_bytes: int = 64 << 30, max_batch: int = 0, tol: float = 1e-12,
               target: str = "SA", x0s=None, ncv: int = 0, max_matvecs: int = 100000, seed: int = 12345,
               mixed: bool = False, stream=None):
    """Eigensolve requests on resident operators with continuous batching (dmrg_serve_eigs): request i arrives
    arrivals[i] seconds after the call (default: all at once), joins the running batch when the HBM budget and
    max_batch allow, and leaves when converged. Returns (list of (eigenvalue, eigenvector numpy, info dict), stats)."""
    k = len(ops)
    arrivals = [0.0] * k if arrivals is None else [float(a) for a in arrivals]
    if len(arrivals) != k:
        raise ValueError("one arrival time per operator")
    cfg = _lib.EigsConfig()
    load().dmrg_eigs_default_config(C.byref(cfg))
    cfg.nev, cfg.ncv, cfg.target = 1, ncv, {"SA": 0, "LA": 1}[target]
    cfg.tol, cfg.max_matvecs, cfg.seed, cfg.mixed = tol, max_matvecs, seed, int(mixed)
    outs = [DeviceArray((o.rows(),), o.dtype) for o in ops]
    ins = [None] * k
    if x0s is not None:
        ins = [DeviceArray.from_numpy(np.asarray(x, dtype=o.dtype).reshape(-1, order="F")) if x is not None else None
               for x, o in zip(x0s, ops)]
    reqs = (_lib.EigRequest * k)()
    for i in range(k):
        reqs[i].op, reqs[i].evec, reqs[i].arrival_s = ops[i].h.value, outs[i].ptr.value, arrivals[i]
        reqs[i].x0 = ins[i].ptr.value if ins[i] is not None else None
    res = (_lib.EigRequestResult * k)()
    st = _lib.ServeStats()
    try:
        call("dmrg_serve_eigs", C.byref(cfg), reqs, k, int(hbm_budget_bytes), int(max_batch), res, C.byref(st), stream)
        out = []
        for i in range(k):
            r = res[i].res
            info = {f: getattr(r, f) for f, _ in r._fields_ if f not in ("eigval", "rnorm")}
            info["rnorm"] = r.rnorm[0]
            info.update({f: getattr(res[i], f) for f, _ in res[i]._fields_ if f != "res"})
            out.append((r.eigval[0], outs[i].to_numpy(), info))
        return out, {f: getattr(st, f) for f, _ in st._fields_}
    finally:
        for a in outs + [x for x in ins if x is not None]:
            a.close()


class MatvecGroup:
    """Matvecs of several operators of uneven shapes as one launch per contraction leg (dmrg_matvec_group_*),
    optionally replayed from a HIP graph: y_i = H_i x_i for DeviceArrays xs[i], ys[i]. Keep the operators and
    the vectors alive and unchanged while the group is used."""

    def __init__(self, ops, xs, ys, fp32: bool = False, graph: bool = True):
        k = len(ops)
        assert k == len(xs) == len(ys) and k >= 1
        self._keep = (list(ops), list(xs), list(ys))
        arr_ops = (C.c_void_p * k)(*[o.h.value for o in ops])
        arr_x = (C.c_void_p * k)(*[x.ptr.value for x in xs])
        arr_y = (C.c_void_p * k)(*[y.ptr.value for y in ys])
        self.h = C.c_void_p()
        call("dmrg_matvec_group_create", arr_ops, k, arr_x, arr_y, int(fp32), int(graph), C.byref(self.h))
        self.flop = load().dmrg_matvec_group_flop(self.h)

    def run(self, stream=None) -> None:
        call("dmrg_matvec_group_run", self.h, stream)

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            load().dmrg_matvec_group_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def optimally_mixed_block(h1: "MatVecMPOS", h2: "MatVecMPOS", theta, mode: int = 3, stream=None):
    """get_optimally_mixed_block_1site (env.cpp:96-239) with h1 the energy operator and h2 the operator on the
    variance envs. mode bit 1 = H1, bit 2 = H2. Returns (mixed theta, (alpha_mps, alpha_h1v, alpha_h2v),
    opt_val); numpy theta gives a numpy result of theta's shape, a DeviceArray a DeviceArray."""
    on_dev = isinstance(theta, DeviceArray)
    t = theta if on_dev else DeviceArray.from_numpy(np.asarray(theta, dtype=h1.dtype))
    out = DeviceArray(t.shape, t.dtype)
    alphas = (C.c_double * 3)()
    val = C.c_double()
    try:
        call("dmrg_optimally_mixed_block", h1.h, h2.h, t.ptr, int(mode), out.ptr, alphas, C.byref(val), stream)
    finally:
        if not on_dev:
            t.close()
    if on_dev:
        return out, tuple(alphas), val.value
    res = out.to_numpy()
    out.close()
    return res, tuple(alphas), val.value


class MatVecMPOS:
    """Device-resident effective Hamiltonian (MatVecMPOS<T>, matvec_mpos.h:23-180).

    mpos: list of (DL, DR, d, d) cx64/fp64 arrays (merged when there are several);
    envL (chiL, chiL, DL), envR (chiR, chiR, DR); dtype np.float64 or np.complex128."""

    def __init__(self, mpos, envL, envR, dtype=np.complex128):
        self.dtype = np.dtype(dtype)
        mpos = [np.asfortranarray(np.asarray(m, dtype=np.complex128)) for m in mpos]
        envL = np.asfortranarray(np.asarray(envL, dtype=np.complex128))
        envR = np.asfortranarray(np.asarray(envR, dtype=np.complex128))
        self._keep = (mpos, envL, envR)
        ptrs = (C.c_void_p * len(mpos))(*[m.ctypes.data for m in mpos])
        dims = (C.c_int64 * (4 * len(mpos)))(*[int(x) for m in mpos for x in m.shape])
        h = C.c_void_p()
        call("dmrg_matvec_create", C.byref(h), _CODE[self.dtype], len(mpos), ptrs, dims, envL.ctypes.data_as(C.c_void_p),
             _dims(envL.shape), envR.ctypes.data_as(C.c_void_p), _dims(envR.shape))
        self.h = h
        shape = (C.c_int64 * 3)()
        call("dmrg_matvec_shape_mps", self.h, shape)
        self.shape_mps = tuple(shape)
        self._keep = None

    def rows(self) -> int:
        return int(load().dmrg_matvec_rows(self.h))

    @property
    def num_mv(self) -> int:
        return int(load().dmrg_matvec_num_mv(self.h))

    def MultAx(self, x: np.ndarray) -> np.ndarray:
        """Host path (the PRIMME/ARPACK callback): x is (n,) or (n, nvec) column-major."""
        x = np.asfortranarray(np.asarray(x, dtype=self.dtype))
        x2 = x.reshape(self.rows(), -1, order="F")
        y = np.empty_like(x2, order="F")
        call("dmrg_matvec_multax", self.h, x2.ctypes.data_as(C.c_void_p), x2.shape[0], y.ctypes.data_as(C.c_void_p),
             y.shape[0], x2.shape[1])
        return y.reshape(x.shape, order="F")

    def set_b(self, mpos, envL, envR) -> None:
        """The B operator of MatVecMPOS(mpos, enva, envb) (matvec_mpos.h:116-120), same layout as the constructor."""
        mpos = [np.asfortranarray(np.asarray(m, dtype=np.complex128)) for m in mpos]
        envL = np.asfortranarray(np.asarray(envL, dtype=np.complex128))
        envR = np.asfortranarray(np.asarray(envR, dtype=np.complex128))
        ptrs = (C.c_void_p * len(mpos))(*[m.ctypes.data for m in mpos])
        dims = (C.c_int64 * (4 * len(mpos)))(*[int(x) for m in mpos for x in m.shape])
        call("dmrg_matvec_set_b", self.h, len(mpos), ptrs, dims, envL.ctypes.data_as(C.c_void_p), _dims(envL.shape),
             envR.ctypes.data_as(C.c_void_p), _dims(envR.shape))

    def MultBx(self, x: np.ndarray, out: np.ndarray | None = None) -> np.ndarray:
        """MultBx (matvec_mpos.cpp:986-1019) on host arrays; with no B operator set, out (zeros when not given) is
        returned unwritten, as the reference leaves its output untouched."""
        x = np.asfortranarray(np.asarray(x, dtype=self.dtype))
        x2 = x.reshape(self.rows(), -1, order="F")
        y = np.zeros_like(x2, order="F") if out is None else np.asfortranarray(out).reshape(x2.shape, order="F")
        call("dmrg_matvec_multbx", self.h, x2.ctypes.data_as(C.c_void_p), x2.shape[0], y.ctypes.data_as(C.c_void_p),
             y.shape[0], x2.shape[1])
        return y.reshape(x.shape, order="F")

    def MultBx_device(self, dx: DeviceArray, dy: DeviceArray, nvec: int = 1, ld: int | None = None, stream=None) -> None:
        ld = ld or self.rows()
        call("dmrg_matvec_multbx_device", self.h, dx.ptr, ld, dy.ptr, ld, nvec, stream)

    def MultAx_device(self, dx: DeviceArray, dy: DeviceArray, nvec: int = 1, ld: int | None = None, stream=None,
                      fp32: bool = False) -> None:
        ld = ld or self.rows()
        fn = "dmrg_matvec_multax_device_fp32" if fp32 else "dmrg_matvec_multax_device"
        call(fn, self.h, dx.ptr, ld, dy.ptr, ld, nvec, stream)

    def get_matrix(self, transpose: bool = True, stream=None) -> np.ndarray:
        """Dense operator matrix (MatVecMPO::get_matrix orientation by default, matvec_mpo.cpp:313-332)."""
        n = self.rows()
        out = DeviceArray((n, n), self.dtype)
        try:
            call("dmrg_matvec_get_matrix", self.h, out.ptr, n, int(transpose), stream)
            return out.to_numpy()
        finally:
            out.close()

    def set_jcb_block_size(self, size: int) -> None:
        """set_jcbMaxBlockSize: 1 = Jacobi, > 1 = block-Jacobi preconditioner in MultPc."""
        call("dmrg_matvec_set_jcb_block_size", self.h, size)

    def set_factorization(self, factorization: str) -> None:
        """MatVecMPOS::factorization of the block-Jacobi blocks (an eig::Factorization name): 'LLT' keeps the
        reference's sign rule, every other name applies the exact inverse of each block ('LU' is the default)."""
        call("dmrg_matvec_set_factorization", self.h, {"NONE": 0, "LLT": 1, "LDLT": 2, "LU": 3, "QR": 4, "ILUT": 5, "ILDLT": 6}[factorization])

    def MultPc(self, x: np.ndarray, shifts) -> np.ndarray:
        """Preconditioner callback on host vectors (MultPc, matvec_mpos.cpp:1333-1361): column i is
        preconditioned with shifts[i]."""
        x = np.asfortranarray(np.asarray(x, dtype=self.dtype))
        x2 = x.reshape(self.rows(), -1, order="F")
        y = np.empty_like(x2, order="F")
        sh = np.ascontiguousarray(np.broadcast_to(np.asarray(shifts, dtype=np.float64), (x2.shape[1],)))
        call("dmrg_matvec_precond", self.h, x2.ctypes.data_as(C.c_void_p), x2.shape[0], y.ctypes.data_as(C.c_void_p),
             y.shape[0], x2.shape[1], sh.ctypes.data_as(C.POINTER(C.c_double)))
        return y.reshape(x.shape, order="F")

    def set_shift(self, shift: complex) -> None:
        call("dmrg_matvec_set_shift", self.h, float(np.real(shift)), float(np.imag(shift)))

    def FactorOP(self, factorization: str = "LU") -> None:
        """MatVecMPOS::FactorOP (matvec_mpos.cpp:887-913): factor the dense H_eff (with its shift) on the device for
        shift-invert; an eig::Factorization name: 'LLT' (Cholesky), 'LU' or 'LDLT' (pivoted LU), 'NONE'."""
        call("dmrg_matvec_factor_op", self.h, {"NONE": 0, "LLT": 1, "LDLT": 2, "LU": 3, "QR": 4}[factorization], None)

    def isReadyFactorOp(self) -> bool:
        return bool(load().dmrg_matvec_factor_ready(self.h))

    def MultOPv(self, x: np.ndarray) -> np.ndarray:
        """y = (H_eff)^-1 x for host vectors (columns), with the factor of FactorOP (matvec_mpos.cpp:936-975)."""
        x = np.asfortranarray(np.asarray(x, dtype=self.dtype))
        x2 = x.reshape(self.rows(), -1, order="F")
        y = np.empty_like(x2, order="F")
        call("dmrg_matvec_mult_opv", self.h, x2.ctypes.data_as(C.c_void_p), x2.shape[0], y.ctypes.data_as(C.c_void_p),
             y.shape[0], x2.shape[1])
        return y.reshape(x.shape, order="F")

    def get_diagonal(self) -> np.ndarray:
        out = DeviceArray(self.shape_mps, self.dtype)
        call("dmrg_matvec_get_diagonal", self.h, out.ptr, None)
        res = out.to_numpy()
        out.close()
        return res

    def eigs(self, nev: int = 1, target: str = "SA", tol: float = 1e-12, x0=None, mixed: bool = False,
             precond: int = 0, ncv: int = 0, min_restart: int = 0, shift: float = 0.0,
             max_matvecs: int = 100000, seed: int = 12345, jcb: int = 128, inner_max=None, inner_tol=None,
             projection: str = "RR", conv_test: str = "tol", stream=None):
        """Eigenpairs on the GPU (eig::solver::eigs with PRIMME, solver.cpp:491-509). Returns
        (eigenvalues, eigenvectors (n, nev) numpy, result dict). mixed: FP32 phase, then FP64 refinement whose
        expansions come from at most inner_max FP32 correction-equation CG steps (None: library default).
        projection 'refined' (target CLOSEST): refined Ritz vectors; conv_test 'saturation': the convergence rule
        of the reference's folded-spectrum convTestFun."""
        cfg = _lib.EigsConfig()
        load().dmrg_eigs_default_config(C.byref(cfg))
        cfg.projection = {"RR": 0, "refined": 1}[projection]
        cfg.conv_test = {"tol": 0, "saturation": 1}[conv_test]
        cfg.nev, cfg.ncv, cfg.min_restart = nev, ncv, min_restart
        cfg.target = {"SA": 0, "LA": 1, "CLOSEST": 2}[target]
        cfg.shift, cfg.tol, cfg.max_matvecs = shift, tol, max_matvecs
        cfg.precond, cfg.mixed, cfg.seed, cfg.jcb_block = int(precond), int(mixed), seed, jcb
        if inner_max is not None:
            cfg.inner_max = int(inner_max)
        if inner_tol is not None:
            cfg.inner_tol = float(inner_tol)
        n = self.rows()
        dx0 = None
        if x0 is not None:
            dx0 = DeviceArray.from_numpy(np.asarray(x0, dtype=self.dtype).reshape(n, order="F"))
        dev = DeviceArray((n, nev), self.dtype)
        res = _lib.EigsResult()
        call("dmrg_eigs", self.h, C.byref(cfg), dx0.ptr if dx0 is not None else None, dev.ptr, C.byref(res), stream)
        vecs = dev.to_numpy()
        dev.close()
        info = {k: getattr(res, k) for k, _ in res._fields_ if k not in ("eigval", "rnorm")}
        info["rnorm"] = list(res.rnorm[:nev])
        return np.array(res.eigval[:nev]), vecs, info

    def eigs_folded_spectrum(self, h: "MatVecMPOS", x0, energy_shift: float, target: str = "CLOSEST",
                             shift: float = 0.0, tol: float = 1e-12, precond: int = 0, ncv: int = 0,
                             max_matvecs: int = 100000, projection: str = "refined", jcb: int = 128,
                             mixed: bool = False, stream=None):
        """Folded-spectrum executor on this operator (the MPO^2 of H - energy_shift on the variance envs), h = H
        (optimize_folded_spectrum.cpp:235-342); by default with the refined projection of the reference's SM case
        (:295-299). Returns (energy, variance, eigenvector, result dict)."""
        cfg = _lib.EigsConfig()
        load().dmrg_eigs_default_config(C.byref(cfg))
        cfg.projection = {"RR": 0, "refined": 1}[projection]
        cfg.jcb_block, cfg.mixed = jcb, int(mixed)
        cfg.nev, cfg.ncv, cfg.target = 1, ncv, {"SA": 0, "LA": 1, "CLOSEST": 2}[target]
        cfg.shift, cfg.tol, cfg.max_matvecs, cfg.precond = shift, tol, max_matvecs, int(precond)
        n = self.rows()
        dx0 = DeviceArray.from_numpy(np.asarray(x0, dtype=self.dtype).reshape(n, order="F"))
        dev = DeviceArray((n, 1), self.dtype)
        res = _lib.EigsResult()
        e, v = C.c_double(), C.c_double()
        try:
            call("dmrg_eigs_folded_spectrum", h.h, self.h, C.byref(cfg), dx0.ptr, float(energy_shift), dev.ptr,
                 C.byref(e), C.byref(v), C.byref(res), stream)
            vec = dev.to_numpy()[:, 0]
        finally:
            dx0.close()
            dev.close()
        info = {k: getattr(res, k) for k, _ in res._fields_ if k not in ("eigval", "rnorm")}
        info["eigval"], info["rnorm"] = res.eigval[0], res.rnorm[0]
        return e.value, v.value, vec, info

    def time_legs(self, dx: DeviceArray, dy: DeviceArray, nrep: int, fp32: bool = False, stream=None):
        """Average durations (ms) over nrep device matvecs: [leg 1, leg 2, leg 3 main kernel, leg 3 split-K reduce]."""
        ms = (C.c_float * 4)()
        call("dmrg_matvec_time_legs", self.h, dx.ptr, dy.ptr, nrep, int(fp32), ms, stream)
        return [float(v) for v in ms]

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            call("dmrg_matvec_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
