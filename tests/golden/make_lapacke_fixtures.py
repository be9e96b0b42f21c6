"""Extract the reference's dense eigen fixtures (tests/testmatrices/lapacke_matrix-{1A,1B,2A,2B}.h5:
datasets `matrix`, `eigvals`, `eigvecs`) into tests/golden/lapacke_matrices.npz.

The HDF5 files are read as data with the image's h5dump tool (full precision); nothing from the reference
is executed. Run from the repository root: python tests/golden/make_lapacke_fixtures.py
"""
import re
import subprocess
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/tests/testmatrices")
OUT = Path(__file__).resolve().parent / "lapacke_matrices.npz"
H5DUMP = "/opt/conda/bin/h5dump"


def read(path: Path, name: str) -> np.ndarray:
    txt = subprocess.run([H5DUMP, "-m", "%.17g", "-d", name, str(path)], capture_output=True, text=True, check=True).stdout
    dims = tuple(int(x) for x in re.search(r"SIMPLE \{ \( ([\d, ]+) \)", txt).group(1).split(","))
    body = txt[txt.index("DATA {") + 6:].split("}")[0]
    body = re.sub(r"\(\d+(?:,\d+)*\):", " ", body)  # drop the "(i,j):" position markers
    vals = [float(v) for v in body.replace(",", " ").split()]
    # h5dump prints row-major (C order) over the dataset dims
    return np.array(vals, dtype=np.float64).reshape(dims)


def main():
    data = {}
    for tag in ("1A", "1B", "2A", "2B"):
        f = SRC / f"lapacke_matrix-{tag}.h5"
        for ds in ("matrix", "eigvals", "eigvecs"):
            data[f"{tag}_{ds}"] = read(f, ds)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} with {len(data)} arrays")


if __name__ == "__main__":
    main()
