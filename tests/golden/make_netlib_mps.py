"""Copies small Netlib models the reference vendors (data/netlib/*.mps.gz of
the reference, data files) into tests/golden/ and records two optimal
objectives per model in tests/golden/netlib_optima.json: the certified
dual-simplex objective of the reference's own benchmark record
(results/netlib/runs.csv, status OPTIMAL, certificate ACCEPTED) as "optimum",
and the value published in data/netlib/README.netlib (11 significant digits;
for scagr7 it differs from the certified optimum in the 7th digit) as
"published".  The integration tests run the reference's own solve()
pipeline on the models over this engine.

    python tests/golden/make_netlib_mps.py [/root/reference]
"""
import json
import os
import re
import shutil
import sys

MODELS = ["afiro", "sc50a", "sc50b", "sc105", "sc205", "kb2", "stocfor1", "scagr7", "adlittle", "blend",
          "share2b", "recipe", "lotfi", "vtpbase", "share1b", "boeing2", "israel"]


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    readme = open(os.path.join(ref, "data", "netlib", "README.netlib"), errors="replace").read()
    import csv

    certified = {}
    with open(os.path.join(ref, "results", "netlib", "runs.csv")) as f:
        for row in csv.reader(f):
            if len(row) > 10 and row[1] == "dual" and row[2] == "OPTIMAL" and row[10] == "ACCEPTED":
                certified[row[0]] = float(row[5])
    optima = {}
    for name in MODELS:
        src = os.path.join(ref, "data", "netlib", f"{name}.mps.gz")
        dst = os.path.join(here, f"netlib_{name}.mps.gz")
        if not os.path.exists(dst):
            shutil.copyfile(src, dst)
        key = "VTP.BASE" if name == "vtpbase" else name.upper()
        m = re.search(r"^\s*" + re.escape(key) + r"\s+\d+\s+\d+\s+\d+\s+\d+\s+(?:[A-Z]+\s+)?(-?[0-9.]+E[+-]\d+)",
                      readme, re.M)
        assert m, name
        optima[name] = {"optimum": certified[name], "published": float(m.group(1))}
    with open(os.path.join(here, "netlib_optima.json"), "w") as f:
        json.dump(optima, f, indent=1, sort_keys=True)
    print(json.dumps(optima, indent=1))


if __name__ == "__main__":
    main()
