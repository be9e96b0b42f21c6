"""ctypes binding of libdmrg_amd.so (the C ABI declared in include/dmrg_amd.h).

The shared library is built in-tree by dmrg_amd.build. Loading fails loudly when it is
missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libdmrg_amd.so"

DMRG_FP64, DMRG_CX64, DMRG_FP32, DMRG_CX32 = 0, 1, 2, 3

STATUS = {
    0: "DMRG_OK",
    1: "DMRG_ERR_DIMS",
    2: "DMRG_ERR_ARG",
    3: "DMRG_ERR_OOM",
    4: "DMRG_ERR_HIP",
    5: "DMRG_ERR_NAN",
    6: "DMRG_ERR_NOT_CONVERGED",
    7: "DMRG_ERR_UNSUPPORTED",
}

vp = C.c_void_p
i64 = C.c_int64
i64p = C.POINTER(C.c_int64)
dp = C.POINTER(C.c_double)
st = C.c_int

# name -> (restype, argtypes)
SIGNATURES: dict[str, tuple] = {
    "dmrg_last_error": (C.c_char_p, []),
    "dmrg_version": (C.c_int, []),
    "dmrg_device_count": (st, [C.POINTER(C.c_int)]),
    "dmrg_set_device": (st, [C.c_int]),
    "dmrg_device_info": (st, [C.c_char_p, C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_size_t)]),
    "dmrg_malloc": (st, [C.POINTER(vp), C.c_size_t]),
    "dmrg_free": (st, [vp]),
    "dmrg_memcpy_htod": (st, [vp, vp, C.c_size_t]),
    "dmrg_memcpy_dtoh": (st, [vp, vp, C.c_size_t]),
    "dmrg_memcpy_dtod": (st, [vp, vp, C.c_size_t]),
    "dmrg_memset": (st, [vp, C.c_int, C.c_size_t]),
    "dmrg_synchronize": (st, []),
    "dmrg_stream_create": (st, [C.POINTER(vp)]),
    "dmrg_stream_destroy": (st, [vp]),
    "dmrg_stream_synchronize": (st, [vp]),
    "dmrg_event_create": (st, [C.POINTER(vp)]),
    "dmrg_event_destroy": (st, [vp]),
    "dmrg_event_record": (st, [vp, vp]),
    "dmrg_event_elapsed_ms": (st, [C.POINTER(C.c_float), vp, vp]),
    "dmrg_matrix_vector_product": (st, [C.c_int, vp, vp, i64p, vp, i64p, vp, i64p, vp, i64p, vp]),
    "dmrg_matrix_vector_product_noalloc": (st, [C.c_int, vp, vp, i64p, vp, i64p, vp, i64p, vp, i64p, vp]),
    "dmrg_matrix_vector_product_noalloc_max_bytes": (i64, []),
    "dmrg_env_build_block": (st, [C.c_int, C.c_char, vp, vp, i64p, vp, i64p, vp, i64p, vp]),
    "dmrg_env_expansion_term": (st, [C.c_int, C.c_char, vp, vp, i64p, vp, i64p, vp, i64p, C.c_double, vp]),
    "dmrg_optimally_mixed_block": (st, [vp, vp, vp, C.c_int, vp, C.POINTER(C.c_double), C.POINTER(C.c_double), vp]),
    "dmrg_expectation_value": (st, [C.c_int, vp, i64p, vp, i64p, vp, i64p, vp, i64p, dp, vp]),
    "dmrg_dot": (st, [C.c_int, i64, vp, vp, dp, vp]),
    "dmrg_nrm2": (st, [C.c_int, i64, vp, dp, vp]),
    "dmrg_matvec_create": (st, [C.POINTER(vp), C.c_int, C.c_int, C.POINTER(vp), i64p, vp, i64p, vp, i64p]),
    "dmrg_matvec_destroy": (st, [vp]),
    "dmrg_matvec_rows": (i64, [vp]),
    "dmrg_matvec_shape_mps": (st, [vp, i64p]),
    "dmrg_matvec_num_mv": (i64, [vp]),
    "dmrg_matvec_multax": (st, [vp, vp, i64, vp, i64, C.c_int]),
    "dmrg_matvec_set_b": (st, [vp, C.c_int, C.POINTER(vp), i64p, vp, i64p, vp, i64p]),
    "dmrg_matvec_multbx": (st, [vp, vp, i64, vp, i64, C.c_int]),
    "dmrg_matvec_multbx_device": (st, [vp, vp, i64, vp, i64, C.c_int, vp]),
    "dmrg_matvec_multax_device": (st, [vp, vp, i64, vp, i64, C.c_int, vp]),
    "dmrg_matvec_multax_device_fp32": (st, [vp, vp, i64, vp, i64, C.c_int, vp]),
    "dmrg_matvec_set_shift": (st, [vp, C.c_double, C.c_double]),
    "dmrg_matvec_get_diagonal": (st, [vp, vp, vp]),
    "dmrg_matvec_precond": (st, [vp, vp, i64, vp, i64, C.c_int, dp]),
    "dmrg_matvec_set_jcb_block_size": (st, [vp, i64]),
    "dmrg_matvec_factor_op": (st, [vp, C.c_int, vp]),
    "dmrg_matvec_factor_ready": (C.c_int, [vp]),
    "dmrg_matvec_mult_opv": (st, [vp, vp, i64, vp, i64, C.c_int]),
    "dmrg_matvec_num_op": (i64, [vp]),
    "dmrg_matvec_set_factorization": (st, [vp, C.c_int]),
    "dmrg_host_heev": (st, [C.c_int, dp, dp, dp]),
    "dmrg_autotune_enable": (st, [C.c_int]),
    "dmrg_autotune_reset": (st, []),
    "dmrg_autotune_min_macs": (st, [i64]),
    "dmrg_autotune_entries": (C.c_int, []),
    "dmrg_gemm_plan_override": (st, [C.c_int, C.c_int]),
    "dmrg_gemm_hybrid_override": (st, [C.c_int, C.c_int, C.c_int]),
    "dmrg_matvec_time_legs": (st, [vp, vp, vp, C.c_int, C.c_int, C.POINTER(C.c_float), vp]),
    "dmrg_block_vector_time": (st, [C.c_int, i64, C.c_int, C.c_int, C.POINTER(C.c_float), dp]),
    "dmrg_heff_dense": (st, [C.c_int, vp, i64, vp, i64p, vp, i64p, vp, i64p, C.c_int, vp]),
    "dmrg_matvec_get_matrix": (st, [vp, vp, i64, C.c_int, vp]),
    "dmrg_dense_eigh": (st, [C.c_int, i64, vp, i64, dp, C.c_int, vp]),
    "dmrg_env_set_edge": (st, [C.c_int, vp, i64, i64, vp, vp]),
    "dmrg_svd_truncate": (st, [C.c_int, i64, i64, vp, C.c_double, i64, C.c_int, vp, dp, vp, C.POINTER(C.c_int64), dp, dp,
                               vp]),
}

class EigsConfig(C.Structure):
    _fields_ = [("nev", C.c_int), ("ncv", C.c_int), ("min_restart", C.c_int), ("target", C.c_int),
                ("shift", C.c_double), ("tol", C.c_double), ("max_matvecs", C.c_int64), ("precond", C.c_int),
                ("mixed", C.c_int), ("max_time_s", C.c_double), ("seed", C.c_uint64), ("jcb_block", C.c_int64),
                ("inner_max", C.c_int), ("inner_tol", C.c_double), ("projection", C.c_int), ("conv_test", C.c_int)]


class EigsResult(C.Structure):
    _fields_ = [("eigval", C.c_double * 16), ("rnorm", C.c_double * 16), ("nconv", C.c_int), ("iters", C.c_int),
                ("restarts", C.c_int), ("stagnated", C.c_int), ("matvecs_fp64", C.c_int64),
                ("matvecs_fp32", C.c_int64), ("seconds", C.c_double), ("anorm", C.c_double)]


MAX_L = 256


class FdmrgConfig(C.Structure):
    _fields_ = [("L", C.c_int), ("J", C.c_double), ("h_tran", C.c_double), ("h_rand", C.POINTER(C.c_double)),
                ("chi_max", C.c_int), ("max_sweeps", C.c_int), ("two_site_sweeps", C.c_int), ("var_tol", C.c_double),
                ("trunc_lim", C.c_double), ("eigs_tol", C.c_double), ("eigs_ncv", C.c_int), ("mixed", C.c_int),
                ("complex_", C.c_int), ("seed", C.c_uint64), ("mixed_min_n", C.c_int64), ("xdmrg", C.c_int),
                ("energy_target", C.c_double), ("eig_max_size", C.c_int64), ("expand_mode", C.c_int)]


class FdmrgResult(C.Structure):
    _fields_ = [("energy", C.c_double), ("variance", C.c_double), ("entropies", C.c_double * (MAX_L + 1)),
                ("chi", C.c_int64 * (MAX_L + 1)), ("sweeps", C.c_int), ("eigensolves", C.c_int), ("converged", C.c_int),
                ("matvecs", C.c_int64), ("matvecs_fp32", C.c_int64), ("seconds", C.c_double), ("t_eigs", C.c_double),
                ("t_env", C.c_double), ("t_svd", C.c_double), ("trunc_error_max", C.c_double), ("gflop", C.c_double),
                ("expansions", C.c_int), ("t_expand", C.c_double)]


class Request(C.Structure):
    _fields_ = [("L", C.c_int), ("chi_max", C.c_int), ("h_rand", C.POINTER(C.c_double)), ("arrival_s", C.c_double),
                ("seed", C.c_uint64)]


class RequestResult(C.Structure):
    _fields_ = [("energy", C.c_double), ("variance", C.c_double), ("entropy_mid", C.c_double), ("arrival_s", C.c_double),
                ("start_s", C.c_double), ("finish_s", C.c_double), ("latency_s", C.c_double), ("worker", C.c_int),
                ("status", C.c_int), ("sweeps", C.c_int), ("matvecs", C.c_int64), ("footprint_bytes", C.c_int64),
                ("gflop", C.c_double)]


class ServeConfig(C.Structure):
    _fields_ = [("workers", C.c_int), ("hbm_budget_bytes", C.c_int64), ("mixed", C.c_int), ("complex_", C.c_int),
                ("var_tol", C.c_double), ("eigs_tol", C.c_double), ("max_sweeps", C.c_int), ("two_site_sweeps", C.c_int),
                ("J", C.c_double), ("h_tran", C.c_double), ("batch_eigs", C.c_int), ("expand_mode", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("makespan_s", C.c_double), ("solves_per_s", C.c_double), ("p50_latency_s", C.c_double),
                ("p99_latency_s", C.c_double), ("mean_latency_s", C.c_double), ("completed", C.c_int), ("failed", C.c_int),
                ("peak_admitted_bytes", C.c_int64), ("peak_concurrency", C.c_int), ("matvecs", C.c_int64),
                ("batch_steps", C.c_int64), ("peak_batch", C.c_int), ("batch_step_s", C.c_double),
                ("batch_rebuild_s", C.c_double), ("batch_rr_s", C.c_double)]


class EigRequest(C.Structure):
    _fields_ = [("op", vp), ("x0", vp), ("evec", vp), ("arrival_s", C.c_double)]


class EigRequestResult(C.Structure):
    _fields_ = [("res", EigsResult), ("arrival_s", C.c_double), ("start_s", C.c_double), ("finish_s", C.c_double),
                ("latency_s", C.c_double), ("footprint_bytes", C.c_int64)]


SIGNATURES["dmrg_serve"] = (st, [C.POINTER(ServeConfig), C.POINTER(Request), C.c_int, C.POINTER(RequestResult),
                                 C.POINTER(ServeStats)])
SIGNATURES["dmrg_request_footprint"] = (i64, [C.c_int, C.c_int, C.c_int])
SIGNATURES["dmrg_eigs_folded_spectrum"] = (st, [vp, vp, C.POINTER(EigsConfig), vp, C.c_double, vp, C.POINTER(C.c_double),
                                                C.POINTER(C.c_double), C.POINTER(EigsResult), vp])
SIGNATURES["dmrg_eigs_batched"] = (st, [C.POINTER(vp), C.c_int, C.POINTER(EigsConfig), C.POINTER(vp), C.POINTER(vp),
                                         C.POINTER(EigsResult), vp])
SIGNATURES["dmrg_serve_eigs"] = (st, [C.POINTER(EigsConfig), C.POINTER(EigRequest), C.c_int, C.c_int64, C.c_int,
                                      C.POINTER(EigRequestResult), C.POINTER(ServeStats), vp])
SIGNATURES["dmrg_matvec_group_create"] = (st, [C.POINTER(vp), C.c_int, C.POINTER(vp), C.POINTER(vp), C.c_int, C.c_int,
                                               C.POINTER(vp)])
SIGNATURES["dmrg_matvec_group_run"] = (st, [vp, vp])
SIGNATURES["dmrg_matvec_group_flop"] = (C.c_double, [vp])
SIGNATURES["dmrg_matvec_group_destroy"] = (None, [vp])
SIGNATURES["dmrg_fdmrg_default_config"] = (None, [C.POINTER(FdmrgConfig)])
SIGNATURES["dmrg_fdmrg_ising_tf_rf"] = (st, [C.POINTER(FdmrgConfig), C.POINTER(FdmrgResult), vp])
SIGNATURES["dmrg_eigs_default_config"] = (None, [C.POINTER(EigsConfig)])
SIGNATURES["dmrg_eigs"] = (st, [vp, C.POINTER(EigsConfig), vp, vp, C.POINTER(EigsResult), vp])

_lib = None


class DmrgError(RuntimeError):
    def __init__(self, status: int, fn: str, msg: str):
        super().__init__(f"{fn} -> {STATUS.get(status, status)}: {msg}")
        self.status = status


def load() -> C.CDLL:
    """Load libdmrg_amd.so (built in-tree); raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or python -m dmrg_amd.build) first")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_LOCAL | getattr(os, "RTLD_NOW", 2))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, fn: str = "") -> None:
    if status != 0:
        msg = load().dmrg_last_error().decode(errors="replace")
        raise DmrgError(status, fn, msg)


def call(name: str, *args):
    """Call a status-returning entry point and raise DmrgError on failure."""
    fn = getattr(load(), name)
    r = fn(*args)
    check(r, name)
    return r
