"""The reference solver's own pipeline running on this engine (GPU): its
solve() (presolve -> PDHG on the HIP engine -> crossover -> postsolve ->
certificate, src/api/solver.cpp:158-189,553-566), its family driver
compareFamilyStrategies (batched PDHG through solvePdhgBatch,
src/parametric/parametric.cpp:483-530) via its C API, and its `calibrate`
CLI verb (calibrateBandwidth, parametric.cpp:537-590), all built from the
reference's sources with the adapter in
optimized-gpu-accelerator_amd/integration/pdhg_hip_adapter.cpp."""
import gzip
import json
import os

import pytest

import integration_lib as I

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not I.available(), reason="integration build absent")]

MODELS = ["afiro", "sc50a", "sc105", "kb2", "stocfor1", "scagr7", "adlittle", "blend", "share2b", "recipe",
          "lotfi", "share1b", "boeing2", "israel"]


@pytest.mark.parametrize("name", MODELS)
def test_reference_solve_pipeline_on_hip_engine(gpu, name):
    """engine pdhg-gpu: the HIP engine's answer is polished by the reference's
    crossover into a certified vertex at the published Netlib optimum."""
    opt = I.netlib_optima()[name]
    st, obj, cert, j = I.solve(I.netlib_path(name), algorithm="pdhg-gpu", time_limit=120)
    tel = j["telemetry"]
    print(f"{name}: {j['engine']} {j['status']} obj {obj:.10g} (published {opt:.10g}), pdhg "
          f"{tel['pdhg']['iterations']} iterations on {tel['pdhg']['device']}")
    assert tel["pdhg"]["backend"] == "gpu" and tel["gpu"]["arch"].startswith("gfx950")
    assert st == 1 and cert == 1, (j["status"], j.get("certificate"))
    assert j["engine"] == "pdhg-gpu+crossover"
    assert abs(obj - opt) <= 1e-8 * (1 + abs(opt))


@pytest.mark.parametrize("name", ["afiro", "sc105", "adlittle", "share2b"])
def test_reference_solve_raw_pdhg_on_hip_engine(gpu, name):
    """pdhg-gpu without crossover: the engine's own answer at rel-KKT 1e-6."""
    opt = I.netlib_optima()[name]
    st, obj, cert, j = I.solve(I.netlib_path(name), algorithm="pdhg-gpu", pdhg_crossover=False, time_limit=120)
    assert j["engine"] == "pdhg-gpu" and j["telemetry"]["pdhg"]["relative_kkt"] <= 1e-6
    assert abs(obj - opt) <= 1e-4 * (1 + abs(opt)), (obj, opt)


def first_cost_column(path):
    """Name and cost of the first column with an objective entry (free MPS)."""
    rows_n = None
    sec = None
    with gzip.open(path, "rt") as f:
        for line in f:
            if not line.strip():
                continue
            if not line[0].isspace():
                sec = line.split()[0]
                continue
            t = line.split()
            if sec == "ROWS" and t[0] == "N" and rows_n is None:
                rows_n = t[1]
            elif sec == "COLUMNS" and len(t) >= 3 and t[1] == rows_n:
                return t[0], float(t[2])
    raise AssertionError("no cost column")


def test_reference_family_driver_on_hip_engine(gpu):
    """pramana_family (src/api/c_api.cpp:267-278): 16 cost cases of the
    planning model plan_6x4, batched PDHG on the HIP engine (one
    solvePdhgBatch) beside the reference's warm-started simplex; every PDHG
    case solved and accepted by the reference's safe-bound check."""
    path = os.path.join(I.GOLDEN, "plan_6x4.mps.gz")
    col, cost = first_cost_column(path)
    res = I.family(path, {"col": col, "kind": "cost", "from": 1.3 * cost, "to": 0.7 * cost, "cases": 16,
                          "allow_gpu": True})
    by = {s["strategy"]: s for s in res["strategies"]}
    print(json.dumps({k: {f: v.get(f) for f in ("seconds", "solved", "certified", "max_relative_error", "note")}
                      for k, v in by.items()}, indent=1))
    gpu_s = by["batched PDHG (GPU) 1e-6 + safe bounds"]
    assert gpu_s["solved"] == 16 and gpu_s["certified"] == 16
    assert gpu_s["max_relative_error"] <= 1e-4
    assert res["device"]  # the engine's device name, reported by the batch's stats


def test_reference_calibrate_verb_on_hip_engine(gpu, tmp_path):
    """`pramana calibrate` (calibrateBandwidth): the reference's own six sweep
    LPs, drawn from one Rng(7) stream, timed for 512 iterations with its CPU
    PDHG and with this engine."""
    out = tmp_path / "calib.json"
    rc, stdout, stderr = I.cli("calibrate", "--out", str(out), timeout=900)
    assert rc == 0, stderr[-2000:]
    j = json.loads(out.read_text())
    rows = j["pdhg_iteration_timing"]
    assert [r["rows"] for r in rows] == [1000, 4000, 16000, 64000, 256000, 1000000]
    assert j["gpu"]["available"] and j["gpu"]["arch"].startswith("gfx950")
    for r in rows:
        assert r["gpu"]["seconds_per_iteration"] > 0 and r["cpu"]["seconds_per_iteration"] > 0
        print(f"nnz {r['nnz']}: cpu {r['cpu']['seconds_per_iteration'] * 1e6:.1f} us/it, gpu "
              f"{r['gpu']['seconds_per_iteration'] * 1e6:.1f} us/it ({r['gpu']['effective_GBs']:.0f} GB/s)")
    assert rows[-1]["gpu"]["seconds_per_iteration"] < rows[-1]["cpu"]["seconds_per_iteration"]
    dest = os.environ.get("PDHG_CALIBRATION_OUT")
    if dest:
        with open(dest, "w") as f:
            json.dump(j, f, indent=1)
