"""Router re-fit for this machine (bench/fit_router.py's procedure, restated
in optimized-gpu-accelerator_amd/router_fit.py) and the reference's own
Router::decide (src/router/router.cpp:157-203) driven by the re-fitted model.

test_router_sweep is a measurement run, skipped unless PDHG_ROUTER_SWEEP_OUT
names the JSON file to write: every instance is solved by each of the
reference's four LP engines (dual simplex, IPM and CPU PDHG on the host,
pdhg-gpu on this engine, the PDHG engines followed by the reference's
crossover) through the reference CLI built over this engine, with the same
time limit for all; a run without a proof counts 4x the limit.
Its output is committed as profiles/router_model_mi355x.json.

test_committed_router_model_drives_the_reference_router loads that file into
the reference's router through --router-model and checks its choices."""
import json
import os
import subprocess

import numpy as np
import pytest

import integration_lib as I
import pdhg_amd as P
import router_fit as R
from mps_write import write_mps

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not I.available(), reason="integration build absent")]

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODEL = os.path.join(ROOT, "profiles", "router_model_mi355x.json")
NETLIB = ["afiro", "sc50a", "sc50b", "sc105", "sc205", "kb2", "stocfor1", "scagr7", "adlittle", "blend", "share2b",
          "recipe", "lotfi", "vtpbase", "share1b", "boeing2", "israel"]
LIMIT = 20.0
PROVEN = {"OPTIMAL", "INFEASIBLE", "UNBOUNDED"}


def generated(tmp):
    """The 'gen' family: refinery-style planning LPs and random LPs of several sizes."""
    out = {}
    for periods, streams in ((2, 12), (4, 20), (12, 40)):
        lp = P.gen_refinery(3, periods=periods, streams=streams)
        out[f"refinery_{periods}x{streams}"] = write_mps(lp, tmp / f"ref_{periods}.mps.gz", "REFINERY")
    for s in (1000, 4000):
        lp = P.gen_random_lp(1, s, 2 * s, 4.0, 0.0)
        out[f"random_{s}"] = write_mps(lp, tmp / f"rand_{s}.mps.gz", "RANDLP")
    out["plan_6x4"] = os.path.join(I.GOLDEN, "plan_6x4.mps.gz")
    return out


def run(path, engine, tmp):
    """One solve by `engine` under the time limit, with the reference's
    defaults: a first-order answer counts once the crossover has turned it
    into a certified vertex (without it solve() reports no proof).  A run
    that outlives the limit by far is stopped and counts as unproven."""
    js = tmp / "r.json"
    if js.exists():
        js.unlink()
    print(f"  {os.path.basename(str(path))} {engine}", flush=True)
    try:
        rc, out, err = I.cli(str(path), "--algo", engine, "--time", str(LIMIT), "--json", str(js), "--log", "0",
                             timeout=LIMIT * 3 + 30)
    except subprocess.TimeoutExpired:
        return None
    return json.loads(js.read_text()) if js.exists() else None


def test_router_sweep(gpu, tmp_path):
    dest = os.environ.get("PDHG_ROUTER_SWEEP_OUT")
    if not dest:
        pytest.skip("measurement run: set PDHG_ROUTER_SWEEP_OUT")
    inst = {n: I.netlib_path(n) for n in NETLIB}
    fam = {n: "netlib" for n in NETLIB}
    for n, p in generated(tmp_path).items():
        inst[n] = p
        fam[n] = "gen"
    records = []
    for n, path in inst.items():
        rec = {"instance": n, "family": fam[n], "times": {}, "status": {}}
        for e in R.ENGINES:
            j = run(path, e, tmp_path)
            if j is None:
                rec["status"][e] = "STOPPED"
                rec["times"][e] = R.censored_time(None, False, LIMIT)
                continue
            rec["features"] = j["telemetry"]["router"]["features"]
            rec["status"][e] = j["status"]
            rec["times"][e] = R.censored_time(j.get("seconds"), j["status"] in PROVEN, LIMIT)
            rec.setdefault("engine_label", {})[e] = j.get("engine")
        if "features" not in rec:
            continue
        print(n, json.dumps(rec["times"]), flush=True)
        records.append(rec)
    fit = R.fit_router(records)
    fit["records"] = records
    fit["time_limit_seconds"] = LIMIT
    fit["gpu"] = I.gpu_info()
    fit["host_nproc"] = os.cpu_count()
    with open(dest, "w") as f:
        json.dump(fit, f, indent=1)
    print(json.dumps(fit["validation"], indent=1))


def test_committed_router_model_drives_the_reference_router(gpu, tmp_path):
    if not os.path.exists(MODEL):
        pytest.skip("no committed router model")
    model = json.load(open(MODEL))
    lp = P.gen_refinery(3)
    big = write_mps(lp, tmp_path / "refinery.mps.gz", "REFINERY")
    for path in (I.netlib_path("afiro"), I.netlib_path("share1b"), big):
        js = tmp_path / "d.json"
        rc, out, err = I.cli(str(path), "--router-model", MODEL, "--json", str(js), "--log", "0", "--no-crossover",
                             "--time", "60", timeout=400)
        tel = json.loads(js.read_text())["telemetry"]["router"]
        assert tel["calibrated_model"] is True
        # In auto mode solve() offers the GPU engine only to LPs of 5e4
        # nonzeros or more (src/api/solver.cpp:285-291).
        probe = tel["features"]["nnz"] >= 50000
        for e in R.ENGINES:
            if e == "pdhg-gpu" and not probe:
                assert tel["predicted_seconds"][e] >= 1e300
                continue
            want = R.predicted_seconds(model["weights"][e], tel["features"])
            assert tel["predicted_seconds"][e] == pytest.approx(want, rel=1e-12)
        assert tel["engine"] == R.decide(model["weights"], tel["features"], gpu=probe)
        print(os.path.basename(str(path)), tel["engine"], tel["reason"])
