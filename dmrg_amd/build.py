"""Build libdmrg_amd.so (all HIP sources, gfx950) in-tree with hipcc.

Object files go to dmrg_amd/_build/ and the shared library to dmrg_amd/libdmrg_amd.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libdmrg_amd.so"
INCLUDE = PKG.parent / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 VALU ops are disabled for all device code; the host pass warns that it
# does not know the feature, which is harmless.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                # Keep MFMA sources and accumulators in architectural VGPRs: avoids per-iteration
                # copies of the accumulators between VGPRs and AGPRs in the GEMM main loop.
                "-mllvm", "-amdgpu-mfma-vgpr-form"]
CXXFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", f"-I{INCLUDE}", "-Wall",
            "-Wno-unused-function", "-Wno-unused-variable"] + DEVICE_FLAGS
LDFLAGS = ["-shared", f"--offload-arch={ARCH}", "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocsolver",
           "-lrocblas", "-lpthread"]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *CXXFLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


CPP_DIR = PKG.parent / "tests" / "cpp"
# Host-only C++ programs exercising include/dmrg_amd/matvec_mpos.hpp against the built library.
CPP_DRIVERS = ["heff_cpp_driver", "solver_templates"]
CPP_DRIVER_SRC = CPP_DIR / "heff_cpp_driver.cpp"
CPP_DRIVER_BIN = CPP_DIR / "bin" / "heff_cpp_driver"


def build_cpp_driver() -> Path:
    """Builds every C++ driver under tests/cpp (g++, linked against libdmrg_amd.so)."""
    (CPP_DIR / "bin").mkdir(parents=True, exist_ok=True)
    for name in CPP_DRIVERS:
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", f"-I{INCLUDE}", str(CPP_DIR / f"{name}.cpp"), f"-L{PKG}",
               "-ldmrg_amd", "-Wl,-rpath,$ORIGIN/../../../dmrg_amd", "-o", str(CPP_DIR / "bin" / name)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"C++ driver build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return CPP_DRIVER_BIN


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    """Compile every HIP source for gfx950 and link libdmrg_amd.so. Always rebuilds."""
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, max(1, len(srcs)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *LDFLAGS, *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if CPP_DRIVER_SRC.exists():
        build_cpp_driver()
    if verbose:
        print(f"[dmrg_amd] built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
