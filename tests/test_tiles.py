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


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_xcd_placement_model(tmp_path):
    """L2 fills of the gathered vector per width-64 FP64 launch on the refinery
    matrix (xcd_placement.cpp): with the XCD remap, each XCD gathers a
    contiguous share of the rows, so the distinct (XCD, column) pairs stay
    near the distinct columns; without it they are several times more, as the
    FETCH_SIZE counters of the fixed-mapping runs show."""
    exe = tmp_path / "xcd_placement"
    subprocess.run(["g++", "-std=c++20", "-O2", f"-I{CSRC}", str(HERE / "xcd_placement.cpp"), "-o", str(exe)],
                   check=True, timeout=300)
    lp = P.gen_refinery(3)
    A = lp.scipy_csc()
    write_csr(tmp_path / "A.bin", A)
    write_csr(tmp_path / "AT.bin", A.T)
    out = subprocess.run([str(exe), str(tmp_path / "A.bin"), str(tmp_path / "AT.bin")], capture_output=True,
                         text=True, timeout=300)
    print(out.stdout)
    assert out.returncode == 0, out.stdout
    got = {}
    for line in out.stdout.strip().splitlines():
        f = line.split()
        got[Path(f[0]).stem] = {f[i]: int(f[i + 1]) for i in range(1, len(f), 2)}
    mb = lambda k: k * 512 / 1e6  # one column of 64 FP64 lanes
    for name, n in (("A", lp.n), ("AT", lp.m)):
        g = got[name]
        assert g["distinct"] == n  # every column of the refinery matrix has entries
        for kind in ("tiled", "onerow"):
            assert g["distinct"] <= g[f"{kind}_on"] < g[f"{kind}_off"] <= 8 * g["distinct"]
            assert g[f"{kind}_on"] <= 1.4 * g["distinct"]
        print(name, {k: round(mb(v), 1) for k, v in g.items()}, "MB")
    # The figures quoted with the FETCH_SIZE comparison (MB per launch).
    assert round(mb(got["A"]["tiled_off"] - got["A"]["tiled_on"]), 1) == 20.7
    assert round(mb(got["A"]["onerow_off"] - got["A"]["onerow_on"]), 1) == 60.7
