"""Free-format MPS writer for the tests' generated LPs (pdhg_amd.LP, minimisation
form), so that the reference solver built over this engine can read them.
Rows R<i>, columns C<j>; a row with both bounds finite and different becomes
an L row with a RANGES entry; |v| >= 1e20 is infinite."""
import gzip

import numpy as np

INF = 1e20


def write_mps(lp, path, name="GEN"):
    fin = lambda v: abs(v) < INF  # noqa: E731
    out = [f"NAME {name}", "ROWS", " N  OBJ"]
    kinds = []
    for i in range(lp.m):
        L, U = lp.L[i], lp.U[i]
        if fin(L) and fin(U):
            k = "E" if L == U else "L"
        elif fin(U):
            k = "L"
        elif fin(L):
            k = "G"
        else:
            k = "N"
            raise ValueError("free rows are not written")
        kinds.append(k)
        out.append(f" {k}  R{i}")
    out.append("COLUMNS")
    for j in range(lp.n):
        if lp.c[j] != 0:
            out.append(f"    C{j} OBJ {lp.c[j]:.17g}")
        for e in range(lp.start[j], lp.start[j + 1]):
            out.append(f"    C{j} R{lp.index[e]} {lp.value[e]:.17g}")
    out.append("RHS")
    if lp.offset != 0:
        out.append(f"    RHS OBJ {-lp.offset:.17g}")
    for i, k in enumerate(kinds):
        v = lp.L[i] if k == "G" else lp.U[i]
        if v != 0:
            out.append(f"    RHS R{i} {v:.17g}")
    ranged = [i for i, k in enumerate(kinds) if k == "L" and fin(lp.L[i])]
    if ranged:
        out.append("RANGES")
        for i in ranged:
            out.append(f"    RNG R{i} {lp.U[i] - lp.L[i]:.17g}")
    out.append("BOUNDS")
    for j in range(lp.n):
        lo, up = lp.l[j], lp.u[j]
        if not fin(lo) and not fin(up):
            out.append(f" FR BND C{j}")
            continue
        if fin(lo) and fin(up) and lo == up:
            out.append(f" FX BND C{j} {lo:.17g}")
            continue
        if not fin(lo):
            out.append(f" MI BND C{j}")
        elif lo != 0:
            out.append(f" LO BND C{j} {lo:.17g}")
        if fin(up):
            out.append(f" UP BND C{j} {up:.17g}")
    out.append("ENDATA")
    data = ("\n".join(out) + "\n").encode()
    with (gzip.open(path, "wb") if str(path).endswith(".gz") else open(path, "wb")) as f:
        f.write(data)
    return path
