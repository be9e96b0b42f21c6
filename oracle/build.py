"""Build the oracle's C restatement (test infrastructure only) into oracle/liboracle_heff.so with gcc.

The library is the checker and the host-core baseline; the GPU library never links or loads it.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "heff_c.c"
LIB = HERE / "liboracle_heff.so"
# AVX2 + FMA (x86-64-v3): every host this targets has them; heff_c.py checks the CPU flags before loading.
CFLAGS = ["-O3", "-mavx2", "-mfma", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-Wall", "-Wextra"]


def build() -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [os.environ.get("CC", "gcc"), *CFLAGS, str(SRC), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle C build failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
