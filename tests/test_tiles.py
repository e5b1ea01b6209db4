"""Host checks of the width-64 row tiles (csrc/tiles.h): the checker in
tile_check.cpp is compiled with g++ and run on the refinery matrix and its
transpose (both iteration kernels' row sets) and on random matrices,
including ones with rows too long for the LDS budget."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import pdhg_amd as P

HERE = Path(__file__).resolve().parent
CSRC = HERE.parent / "optimized-gpu-accelerator_amd" / "csrc"


def write_csr(path, M):
    M = M.tocsr()
    M.sort_indices()
    with open(path, "wb") as f:
        np.array([M.shape[0], M.shape[1], M.nnz], np.int32).tofile(f)
        M.indptr.astype(np.int32).tofile(f)
        M.indices.astype(np.int32).tofile(f)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_tile_invariants(tmp_path):
    exe = tmp_path / "tile_check"
    subprocess.run(["g++", "-std=c++20", "-O2", f"-I{CSRC}", str(HERE / "tile_check.cpp"), "-o", str(exe)],
                   check=True, timeout=300)
    lp = P.gen_refinery(3)
    A = lp.scipy_csc()
    write_csr(tmp_path / "A.bin", A)
    write_csr(tmp_path / "AT.bin", A.T)
    out = subprocess.run([str(exe), str(tmp_path / "A.bin"), str(tmp_path / "AT.bin")], capture_output=True,
                         text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    assert out.stdout.strip().endswith("2 matrices from files, 300 random")
