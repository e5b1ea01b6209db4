"""Writes tests/golden/plan_*.npz from the reference's refinery planning models.

Reads, as data, data/gen/plan_<T>x<P>.mps of the reference (maximisation
models, stored here in minimisation form) and the optimal objective the
reference published for each (results/gpu/planning_engines.csv, the row of
its dual simplex with status OPTIMAL; the LP optimum does not depend on the
presolve that run used).  The fixture's `optimum` is that objective in
minimisation form (negated).

    python tests/golden/make_plan_fixtures.py [/root/reference]
"""
import csv
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from make_netlib_fixtures import read_mps  # noqa: E402

MODELS = ("plan_6x4", "plan_10x12", "plan_20x12")


def published_optima(ref):
    out = {}
    with open(os.path.join(ref, "results", "gpu", "planning_engines.csv")) as f:
        for r in csv.DictReader(f):
            if r["config"] == "dual" and r["status"] == "OPTIMAL":
                out[r["model"]] = (float(r["objective"]), int(r["rows"]), int(r["cols"]), int(r["nnz"]))
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = os.path.dirname(os.path.abspath(__file__))
    opt = published_optima(ref)
    for name in MODELS:
        d = read_mps(os.path.join(ref, "data", "gen", name + ".mps"))
        sense = d.pop("sense")
        obj = opt[name][0]
        np.savez_compressed(os.path.join(out, f"{name}.npz"), optimum=sense * obj, published=obj, **d)
        print(name, d["m"], d["n"], int(d["start"][-1]), "published", opt[name], "sense", sense)


if __name__ == "__main__":
    main()
