"""CPU tests of the C-ABI boundary: the library loads and exports every declared symbol.

No compute call is made here (no GPU on the build host)."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "dmrg_amd.h"
LIB = ROOT / "dmrg_amd" / "libdmrg_amd.so"


def declared_symbols():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dmrg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for name in ("dmrg_matrix_vector_product", "dmrg_env_build_block", "dmrg_matvec_multax", "dmrg_matvec_create"):
        assert name in syms


def test_library_exports_all_declared_symbols():
    assert LIB.exists(), "build the library first (__graft_entry__.build())"
    lib = ctypes.CDLL(str(LIB))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"missing exports: {missing}"


def test_python_binding_covers_header():
    from dmrg_amd import _lib
    missing = [s for s in declared_symbols() if s not in _lib.SIGNATURES]
    assert not missing, f"ctypes signatures missing for: {missing}"


def test_no_device_needed_for_error_path():
    """Argument validation runs on the host: a null-argument call fails with DMRG_ERR_ARG, no GPU touched."""
    from dmrg_amd import _lib
    lib = _lib.load()
    assert lib.dmrg_matvec_rows(None) == -1
    r = lib.dmrg_matvec_set_shift(None, 0.0, 0.0)
    assert r == 2
    assert b"null" in lib.dmrg_last_error()


def test_exports_are_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (dmrg_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the C structs must have the C sizes (a mismatch corrupts memory at the boundary)."""
    import ctypes as C

    from dmrg_amd import _lib
    src = tmp_path / "sizes.c"
    names = ["dmrg_eigs_config", "dmrg_eigs_result", "dmrg_fdmrg_config", "dmrg_fdmrg_result", "dmrg_request",
             "dmrg_request_result", "dmrg_serve_config", "dmrg_serve_stats", "dmrg_eig_request", "dmrg_eig_request_result"]
    src.write_text('#include <stdio.h>\n#include "dmrg_amd.h"\nint main(void){\n' +
                   "".join(f'printf("%zu\\n", sizeof({n}));\n' for n in names) + "return 0;}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    mirrors = [_lib.EigsConfig, _lib.EigsResult, _lib.FdmrgConfig, _lib.FdmrgResult, _lib.Request, _lib.RequestResult,
               _lib.ServeConfig, _lib.ServeStats, _lib.EigRequest, _lib.EigRequestResult]
    assert sizes == [C.sizeof(m) for m in mirrors]
