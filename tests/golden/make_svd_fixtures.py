"""Extract a subset of the reference's SVD test matrices (tests/svd/testmatrices/svdmatrix_*.h5, dataset
`svdmatrix`, compound {real, imag} doubles) into tests/golden/svd_matrices.npz.

The reference keeps these as inputs only (its SVD test checks no outputs); the GPU test checks the
singular values against numpy's and the factorisation invariants. The HDF5 files are read as data with
the image's h5dump tool at full precision; nothing from the reference is executed.
Run from the repository root: python tests/golden/make_svd_fixtures.py
"""
import re
import subprocess
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/tests/svd/testmatrices")
OUT = Path(__file__).resolve().parent / "svd_matrices.npz"
H5DUMP = "/opt/conda/bin/h5dump"
# Ragged shapes: tiny, square, wide, tall, up to 70 x 172.
PICK = (6, 96, 0, 181, 174, 5, 101, 7)


def read_complex(path: Path, name: str) -> np.ndarray:
    txt = subprocess.run([H5DUMP, "-m", "%.17g", "-d", name, str(path)], capture_output=True, text=True, check=True).stdout
    dims = tuple(int(x) for x in re.search(r"SIMPLE \{ \( ([\d, ]+) \)", txt).group(1).split(","))
    body = txt[txt.index("DATA {") + 6:]
    body = re.sub(r"\(\d+(?:,\d+)*\):", " ", body)  # drop the "(i,j):" position markers
    vals = [float(v) for v in re.findall(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?", body)]
    n = int(np.prod(dims))
    z = np.array(vals[:2 * n], dtype=np.float64).reshape(n, 2)
    # h5dump prints row-major (C order) over the dataset dims
    return (z[:, 0] + 1j * z[:, 1]).reshape(dims)


def main():
    data = {f"m{k}": read_complex(SRC / f"svdmatrix_{k}.h5", "svdmatrix") for k in PICK}
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} with {len(data)} matrices: " + ", ".join(f"{k} {v.shape}" for k, v in data.items()))


if __name__ == "__main__":
    main()
