"""Writes tests/golden/netlib_*.npz from the reference's vendored Netlib models.

Run in a container where the reference tree is mounted (it reads
data/netlib/*.mps.gz of the reference as data, with this file's own MPS
reader).  The fixtures are pure data: the CSC matrix, bounds, costs and the
known optimum the reference's own tests assert
(tests/cpp/test_solvers.cpp:27-28 of the reference).

    python tests/golden/make_netlib_fixtures.py [/root/reference]
"""
import gzip
import os
import sys

import numpy as np

KNOWN = {  # tests/cpp/test_solvers.cpp:27-28
    "afiro": -464.753142857,
    "sc50a": -64.5750770585,
    "adlittle": 225494.963162,
    "blend": -30.8121498458,
    "share2b": -415.732240741,
}
INF = float("inf")


def read_mps(path):
    """Minimal fixed/free MPS reader: OBJSENSE, ROWS, COLUMNS (MARKER lines
    skipped), RHS (objective RHS = -offset), RANGES, BOUNDS (UP LO FX FR MI PL
    BV).  A maximisation model is returned in minimisation form (cost and
    offset negated) with sense = -1."""
    opener = gzip.open if path.endswith(".gz") else open
    rows, rtype, obj = [], {}, None
    cols, colidx = [], {}
    entries = {}
    cost = {}
    rhs, ranges = {}, {}
    lo, up = {}, {}
    offset = 0.0
    sense = 1
    sec = None
    with opener(path, "rt") as f:
        for raw in f:
            line = raw.rstrip("\n")
            if not line.strip() or line.startswith("*"):
                continue
            if not line[0].isspace():
                sec = line.split()[0]
                continue
            t = line.split()
            if sec == "OBJSENSE":
                sense = -1 if t[0].upper().startswith("MAX") else 1
            elif sec == "ROWS":
                typ, name = t[0], t[1]
                if typ == "N":
                    if obj is None:
                        obj = name
                    continue
                rtype[name] = typ
                rows.append(name)
            elif sec == "COLUMNS":
                if "'MARKER'" in t:
                    continue
                c = t[0]
                if c not in colidx:
                    colidx[c] = len(cols)
                    cols.append(c)
                j = colidx[c]
                for k in range(1, len(t) - 1, 2):
                    r, v = t[k], float(t[k + 1])
                    if r == obj:
                        cost[j] = cost.get(j, 0.0) + v
                    elif r in rtype:
                        entries[(r, j)] = entries.get((r, j), 0.0) + v
            elif sec == "RHS":
                pairs = t[1:] if len(t) % 2 == 1 else t
                for k in range(0, len(pairs) - 1, 2):
                    r, v = pairs[k], float(pairs[k + 1])
                    if r == obj:
                        offset = -v
                    else:
                        rhs[r] = v
            elif sec == "RANGES":
                pairs = t[1:] if len(t) % 2 == 1 else t
                for k in range(0, len(pairs) - 1, 2):
                    ranges[pairs[k]] = float(pairs[k + 1])
            elif sec == "BOUNDS":
                typ = t[0]
                c = t[2] if len(t) >= 3 else t[1]
                v = float(t[3]) if len(t) >= 4 else 0.0
                j = colidx[c]
                if typ == "UP":
                    up[j] = v
                    if v < 0 and j not in lo:
                        lo[j] = -INF
                elif typ == "LO":
                    lo[j] = v
                elif typ == "FX":
                    lo[j] = up[j] = v
                elif typ == "FR":
                    lo[j], up[j] = -INF, INF
                elif typ == "MI":
                    lo[j] = -INF
                elif typ == "PL":
                    up[j] = INF
                elif typ == "BV":
                    lo[j], up[j] = 0.0, 1.0
                else:
                    raise ValueError(f"bound type {typ}")
    m, n = len(rows), len(cols)
    ridx = {r: i for i, r in enumerate(rows)}
    L, U = np.zeros(m), np.zeros(m)
    for i, r in enumerate(rows):
        b = rhs.get(r, 0.0)
        typ = rtype[r]
        if typ == "E":
            L[i] = U[i] = b
        elif typ == "L":
            L[i], U[i] = -INF, b
        else:
            L[i], U[i] = b, INF
        if r in ranges:
            R = ranges[r]
            if typ == "E":
                if R >= 0:
                    U[i] = b + abs(R)
                else:
                    L[i] = b - abs(R)
            elif typ == "L":
                L[i] = b - abs(R)
            else:
                U[i] = b + abs(R)
    c = sense * np.array([cost.get(j, 0.0) for j in range(n)])
    offset = sense * offset
    l = np.array([lo.get(j, 0.0) for j in range(n)])
    u = np.array([up.get(j, INF) for j in range(n)])
    trip = sorted(((j, ridx[r], v) for (r, j), v in entries.items() if v != 0.0))
    start = np.zeros(n + 1, dtype=np.int32)
    index = np.array([t[1] for t in trip], dtype=np.int32)
    value = np.array([t[2] for t in trip], dtype=np.float64)
    for j, _, _ in trip:
        start[j + 1] += 1
    start = np.cumsum(start).astype(np.int32)
    return dict(m=m, n=n, start=start, index=index, value=value, c=c, l=l, u=u, L=L, U=U, offset=offset,
                sense=sense)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    out = os.path.dirname(os.path.abspath(__file__))
    for name, opt in KNOWN.items():
        d = read_mps(os.path.join(ref, "data", "netlib", name + ".mps.gz"))
        assert d.pop("sense") == 1
        np.savez(os.path.join(out, f"netlib_{name}.npz"), optimum=opt, **d)
        print(name, d["m"], d["n"], int(d["start"][-1]))


if __name__ == "__main__":
    main()
