"""Oracle (test infrastructure only): ctypes binding of the C restatement of the H_eff matvec
(heff_c.c, following contraction_tblis.cpp:213-292) and of the environment update (EnvBase.cpp:67-151).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle_heff.so"
_lib = None


def host_supported() -> bool:
    """The library is built for AVX2 + FMA; report whether this CPU has both."""
    try:
        flags = next(line for line in Path("/proc/cpuinfo").read_text().splitlines() if line.startswith("flags"))
    except (OSError, StopIteration):
        return False
    f = set(flags.split(":", 1)[1].split())
    return "avx2" in f and "fma" in f


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built (python oracle/build.py)")
        if not host_supported():
            raise RuntimeError("the oracle's C library needs a CPU with AVX2 and FMA")
        lib = ctypes.CDLL(str(LIB_PATH))
        p = ctypes.c_void_p
        lib.oracle_heff_matvec.argtypes = [p, p, p, p, p] + [ctypes.c_long] * 5 + [ctypes.c_int, ctypes.c_int]
        lib.oracle_heff_matvec.restype = ctypes.c_int
        lib.oracle_env_build_block.argtypes = [ctypes.c_int] + [p] * 4 + [ctypes.c_long] * 5 + [ctypes.c_int]
        lib.oracle_env_build_block.restype = ctypes.c_int
        _lib = lib
    return _lib


def _fortran_c128(x):
    return np.asfortranarray(x, dtype=np.complex128)


def matrix_vector_product(mps, mpo, envL, envR, order: str | None = None, threads: int = 0):
    """H_eff . mps in C (same contraction order as the reference; 'L' = theta.envL first, taken when
    chiL >= chiR). Complex128 in and out, Eigen column-major."""
    lib = load()
    if mps.shape[1] != envL.shape[0] or mps.shape[2] != envR.shape[0] or mps.shape[0] != mpo.shape[2] \
            or envL.shape[2] != mpo.shape[0] or envR.shape[2] != mpo.shape[1] \
            or envL.shape[0] != envL.shape[1] or envR.shape[0] != envR.shape[1] or mpo.shape[2] != mpo.shape[3]:
        raise ValueError("Dimension mismatch")
    th, W, L, R = (_fortran_c128(a) for a in (mps, mpo, envL, envR))
    d, chiL, chiR = th.shape
    DL, DR = W.shape[0], W.shape[1]
    out = np.zeros((d, chiL, chiR), dtype=np.complex128, order="F")
    o = -1 if order is None else (1 if order == "L" else 0)
    rc = lib.oracle_heff_matvec(th.ctypes.data, W.ctypes.data, L.ctypes.data, R.ctypes.data, out.ctypes.data,
                                d, chiL, chiR, DL, DR, o, int(threads))
    if rc != 0:
        raise RuntimeError(f"oracle_heff_matvec failed with code {rc}")
    return out


def env_build_block(side: str, other, mps, mpo, threads: int = 0):
    """EnvBase::build_block in C: the new left (side 'L', from the block L on the left of the site) or right
    ('R') environment block, complex128, Eigen column-major."""
    lib = load()
    th, W, B = (_fortran_c128(a) for a in (mps, mpo, other))
    d, chiL, chiR = th.shape
    DL, DR = W.shape[0], W.shape[1]
    if side == "L":
        if d != W.shape[2] or chiL != B.shape[0] or DL != B.shape[2] or B.shape[0] != B.shape[1]:
            raise ValueError("dimension mismatch")
        out = np.zeros((chiR, chiR, DR), dtype=np.complex128, order="F")
    elif side == "R":
        if d != W.shape[2] or chiR != B.shape[0] or DR != B.shape[2] or B.shape[0] != B.shape[1]:
            raise ValueError("dimension mismatch")
        out = np.zeros((chiL, chiL, DL), dtype=np.complex128, order="F")
    else:
        raise ValueError("side must be L or R")
    rc = lib.oracle_env_build_block(0 if side == "L" else 1, B.ctypes.data, th.ctypes.data, W.ctypes.data,
                                    out.ctypes.data, d, chiL, chiR, DL, DR, int(threads))
    if rc != 0:
        raise RuntimeError(f"oracle_env_build_block failed with code {rc}")
    return out
