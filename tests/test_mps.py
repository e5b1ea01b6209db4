"""The MPS reader (pdhg_read_mps): the LP part of the reference's readMps
(src/io/mps.cpp) followed by toMinimization (src/core/model.cpp:92-101).

* The five vendored Netlib models (tests/golden/netlib_*.mps.gz, data files of
  the reference) and the planning model plan_6x4 (a maximisation model) read
  into exactly the arrays of the committed fixtures, which the scripts in
  tests/golden wrote with their own (Python) reader.
* A hand-written file exercises every rule, against arrays worked out by hand:
  OBJSENSE MAX, extra N rows dropped, duplicate entries summed, explicit zeros
  dropped, integer markers skipped, RHS on the objective (offset), optional
  set names, sign-dependent RANGES on E rows, and the bound types, including
  a negative UP with no lower bound; the same model in fixed columns with
  blanks in its names reads into the same arrays.
* Unsupported input fails loudly, and afiro read from its file solves to the
  optimum the reference's tests assert (tests/cpp/test_solvers.cpp:27).
"""
import os

import numpy as np
import pytest

import oracle_lib as O
import pdhg_amd as P
from conftest import load_netlib, load_plan

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
INF = np.inf

TINY = """NAME          TINY
* comment line
OBJSENSE
    MAX
ROWS
 N  COST
 N  EXTRA
 E  R1
 E  R2
 L  R3
 G  R4
COLUMNS
    X1        COST      1.0          R1        1.0
    X1        R3        2.0          EXTRA     5.0
    X2        COST      -2.0         R1        1.0
    X2        R2        3.0          R2        1.0
    MARKER    'MARKER'  'INTORG'
    X3        R4        1.0          COST      0.5
    MARKER    'MARKER'  'INTEND'
    X4        R2        0.0
RHS
    RHS       COST      -7.0         R1        4.0
    RHS       R2        1.0          R3        6.0
    R4        2.0
RANGES
    RNG       R1        2.5          R2        -1.5
    RNG       R3        4.0          R4        3.0
BOUNDS
 UP BND       X1        -1.0
 LO BND       X2        -5.0
 UP BND       X2        -2.0
 FR BND       X3
 BV BND       X4
ENDATA
"""


@pytest.mark.parametrize("name", ["afiro", "sc50a", "adlittle", "blend", "share2b"])
def test_netlib_files_read_into_the_fixtures(name):
    lp, sense = P.read_mps(os.path.join(GOLDEN, f"netlib_{name}.mps.gz"))
    ref, _ = load_netlib(name)
    assert sense == 1 and (lp.m, lp.n) == (ref.m, ref.n)
    for f in ("start", "index", "value", "c", "l", "u", "L", "U"):
        assert np.array_equal(getattr(lp, f), getattr(ref, f)), f
    assert lp.offset == ref.offset


def test_planning_model_reads_into_its_fixture():
    """plan_6x4 (the reference's data/gen/plan_6x4.mps, gzipped here) is a
    maximisation model: costs and offset come back negated, as in the fixture
    tests/golden/make_plan_fixtures.py wrote, with sense -1."""
    lp, sense = P.read_mps(os.path.join(GOLDEN, "plan_6x4.mps.gz"))
    ref, _ = load_plan("plan_6x4")
    assert sense == -1 and (lp.m, lp.n) == (ref.m, ref.n)
    for f in ("start", "index", "value", "c", "l", "u", "L", "U"):
        assert np.array_equal(getattr(lp, f), getattr(ref, f)), f
    assert lp.offset == ref.offset


def test_every_rule_on_a_hand_written_file(tmp_path):
    path = tmp_path / "tiny.mps"
    path.write_text(TINY)
    lp, sense = P.read_mps(str(path))
    assert sense == -1 and (lp.m, lp.n) == (4, 4)
    assert lp.start.tolist() == [0, 2, 4, 5, 5]
    assert lp.index.tolist() == [0, 2, 0, 1, 3]
    assert lp.value.tolist() == [1.0, 2.0, 1.0, 4.0, 1.0]
    assert lp.c.tolist() == [-1.0, 2.0, -0.5, 0.0]  # maximisation: negated
    assert lp.offset == -7.0  # RHS -7 on the objective is +7, negated
    assert lp.L.tolist() == [4.0, -0.5, 2.0, 2.0]
    assert lp.U.tolist() == [6.5, 1.0, 6.0, 5.0]
    assert lp.l.tolist() == [-INF, -5.0, -INF, 0.0]
    assert lp.u.tolist() == [-1.0, -2.0, INF, 1.0]
    gz = tmp_path / "tiny.mps.gz"
    import gzip
    with gzip.open(gz, "wt") as f:
        f.write(TINY)
    lz, _ = P.read_mps(str(gz))
    for f in ("start", "index", "value", "c", "l", "u", "L", "U"):
        assert np.array_equal(getattr(lz, f), getattr(lp, f)), f


@pytest.mark.parametrize("edit,msg", [
    (("    X4        R2        0.0\n", "    X4        R9        1.0\n"), "unknown row R9"),
    (("ENDATA\n", "QUADOBJ\n    X1 X1 1.0\nENDATA\n"), "QUADOBJ is not supported"),
    ((" E  R1\n", " X  R1\n"), "row type X"),
    ((" BV BND       X4\n", " SC BND       X4        3.0\n"), "bound type SC is not supported"),
    ((" FR BND       X3\n", " FR BND       X9\n"), "unknown column X9"),
    (("    RHS       R2        1.0", "    RHS       R2        one"), "bad number 'one'"),
])
def test_unsupported_input_fails_loudly(tmp_path, edit, msg):
    path = tmp_path / "bad.mps"
    assert edit[0] in TINY
    path.write_text(TINY.replace(edit[0], edit[1]))
    with pytest.raises(RuntimeError, match=msg):
        P.read_mps(str(path))
    with pytest.raises(RuntimeError, match="cannot open"):
        P.read_mps(str(tmp_path / "missing.mps"))


def fixed_line(f1="", f2="", f3="", f4="", f5="", f6=""):
    """A fixed-format data line: fields in columns 2-3, 5-12, 15-22, 25-36, 40-47, 50-61."""
    return (" " + f1.ljust(2) + " " + f2.ljust(8) + "  " + f3.ljust(8) + "  " + f4.ljust(12) + "   " + f5.ljust(8)
            + "  " + f6.ljust(12)).rstrip() + "\n"


def test_fixed_format_with_blanks_in_names(tmp_path):
    """The hand-written model again, in fixed columns with blanks inside its
    row, column and set names (a ROWS line of more than two fields switches
    the file to fixed fields, as the reference does), and a value written
    with a Fortran D exponent: the same arrays as the free-format file."""
    R = {"R1": "ROW 1", "R2": "ROW 2", "R3": "ROW 3", "R4": "ROW 4", "COST": "COST", "EXTRA": "EXTRA"}
    X = {f"X{k}": f"COL {k}" for k in range(1, 5)}
    L = fixed_line
    text = "NAME          TINY FIXED\nOBJSENSE\n    MAX\nROWS\n"
    text += L("N", "COST") + L("N", "EXTRA")
    text += "".join(L(t, R[r]) for t, r in (("E", "R1"), ("E", "R2"), ("L", "R3"), ("G", "R4")))
    text += "COLUMNS\n"
    text += L("", X["X1"], "COST", "1.0", R["R1"], "1.0") + L("", X["X1"], R["R3"], "2.0", "EXTRA", "5.0")
    text += L("", X["X2"], "COST", "-2.0", R["R1"], "1.0") + L("", X["X2"], R["R2"], "3.0", R["R2"], "1.0")
    text += L("", "MARKER", "'MARKER'", "'INTORG'")
    text += L("", X["X3"], R["R4"], "1.0", "COST", "0.5") + L("", "MARKER", "'MARKER'", "'INTEND'")
    text += L("", X["X4"], R["R2"], "0.0")
    text += "RHS\n" + L("", "RHS SET", "COST", "-7.0", R["R1"], "0.4D+01") + L("", "RHS SET", R["R2"], "1.0", R["R3"], "6.0")
    text += L("", "RHS SET", R["R4"], "2.0")
    text += "RANGES\n" + L("", "RNG", R["R1"], "2.5", R["R2"], "-1.5") + L("", "RNG", R["R3"], "4.0", R["R4"], "3.0")
    text += "BOUNDS\n" + L("UP", "BND", X["X1"], "-1.0") + L("LO", "BND", X["X2"], "-5.0") + L("UP", "BND", X["X2"], "-2.0")
    text += L("FR", "BND", X["X3"]) + L("BV", "BND", X["X4"]) + "ENDATA\n"
    fixed, free = tmp_path / "fixed.mps", tmp_path / "free.mps"
    fixed.write_text(text)
    free.write_text(TINY)
    a, sa = P.read_mps(str(fixed))
    b, sb = P.read_mps(str(free))
    assert sa == sb == -1 and (a.m, a.n) == (b.m, b.n)
    for f in ("start", "index", "value", "c", "l", "u", "L", "U"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    assert a.offset == b.offset


def test_afiro_from_its_file_reaches_the_known_optimum():
    lp, _ = P.read_mps(os.path.join(GOLDEN, "netlib_afiro.mps.gz"))
    o, _, _, _ = O.oracle_solve(lp, tol=1e-7, max_iter=500_000)
    assert o[0]["status"] == 1
    assert abs(o[0]["objective"] - (-464.753142857)) <= 1e-5 * 465
