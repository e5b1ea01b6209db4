"""Writes tests/golden/c3_reference_cases.json: LPs of configs C2 and C3
solved by the reference's own CPU PDHG (solvePdhg, src/pdhg/pdhg.cpp:712-728,
compiled from the reference's sources into oracle/_ref by oracle/Makefile),
FP64, the reference's 1M-iteration cap.

* "cases": cases of the 64-case refinery price family (refinery seed 3,
  family seed 4, 64 cases, theta range 0.3) at rel-KKT 1e-6.  They are the
  six that the cold start solves in the fewest iterations (152k to 332k), so
  that the reference finishes them in about half an hour of CPU.
* "base": the refinery LP itself (seed 3, config C2) at rel-KKT 1e-4 and 1e-6.

For each: status, iterations, primal and dual objective, relative KKT and the
reference's family safe bound of its duals (implied bounds, then the
Neumaier-Shcherbina bound, src/parametric/parametric.cpp:505-525).  Entries
already in the file are kept, so an interrupted run resumes.

    python tests/golden/make_c3_reference.py [threads]
"""
import json
import os
import sys
import time
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "optimized-gpu-accelerator_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle_lib as O  # noqa: E402
import pdhg_amd as P  # noqa: E402

OUT = os.path.join(HERE, "c3_reference_cases.json")
CASES = (47, 16, 38, 9, 39, 45)
BASE_TOLS = (1e-4, 1e-6)


def case_lp(lp, cost):
    """The family's LP with case cost `cost` (arrays shared with lp)."""
    return types.SimpleNamespace(m=lp.m, n=lp.n, start=lp.start, index=lp.index, value=lp.value,
                                 c=np.ascontiguousarray(cost, dtype=np.float64), l=lp.l, u=lp.u, L=lp.L, U=lp.U,
                                 offset=lp.offset)


def record(lq, tol, threads):
    t = time.time()
    r, _, y = O.ref_solve(lq, tol=tol, max_iter=1_000_000)
    bound = O.ref_family_bound(lq, y)
    rec = {"status": r["status"], "iterations": r["iterations"], "objective": r["objective"],
           "dual_objective": r["dual_objective"], "relative_kkt": r["relative_kkt"], "safe_bound": bound,
           "threads": threads}
    return rec, time.time() - t


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else 6
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libpramana_ref.so missing: run `make -C oracle ref`")
    O.ref_set_threads(threads)
    lp = P.gen_refinery(3)
    _, costs = P.gen_family(lp, 4, 64, 0.3)
    out = json.load(open(OUT)) if os.path.exists(OUT) else {}
    out["workload"] = ("refinery seed 3 (m=13980 n=18480 nnz=555852); price family seed 4, 64 cases, range 0.3; "
                       "cap 1000000")
    out["solver"] = "reference solvePdhg (its own sources), FP64, host threads as recorded per entry"
    out.setdefault("cases", {})
    out.setdefault("base", {})

    def save():
        with open(OUT, "w") as f:
            json.dump(out, f, indent=1)

    for q in CASES:
        if str(q) in out["cases"]:
            continue
        rec, sec = record(case_lp(lp, costs[q]), 1e-6, threads)
        out["cases"][str(q)] = rec
        print(f"case {q}: {rec['iterations']} iterations, objective {rec['objective']:.12g}, "
              f"bound {rec['safe_bound']:.12g}, {sec:.0f} s", flush=True)
        save()
    for tol in BASE_TOLS:
        key = "%g" % tol
        if key in out["base"]:
            continue
        rec, sec = record(lp, tol, threads)
        out["base"][key] = rec
        print(f"base LP tol {key}: {rec['iterations']} iterations, objective {rec['objective']:.12g}, "
              f"{sec:.0f} s", flush=True)
        save()


if __name__ == "__main__":
    main()
