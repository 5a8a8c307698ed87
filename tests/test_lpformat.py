"""LP-format reader and writer (core/model/lp_parser.cpp) against the
reference's own LPParser and write_lp, compiled from its sources into the
test-side reference library: the same problem (matrix, senses, right-hand
sides, bounds, objective, names, column types) and the same errors on each
text, and the same written file byte for byte.  The reference's own
known-answer tests (tests/unit/test_model.cpp:267-309, the LP round trip of
tests/unit/test_roundtrip.cpp:44-60) are restated first; they run without
the reference build."""
import math
import os

import numpy as np
import pytest

import lpformat
import lpgen
import oracle

HAVE_REF = oracle.REF is not None and hasattr(oracle.REF, "ref_lp_parse")
need_ref = pytest.mark.skipif(not HAVE_REF, reason="reference build not present")

SIMPLE = """
Maximize
 OBJ: 1 x1 + 2 x2
Subject To
 C1: 1 x1 + 1 x2 <= 10
 C2: 2 x1 + 0 x2 >= 5
Bounds
 0 <= x1 <= 10
 0 <= x2
End
"""

MILP = """
Minimize
 OBJ: 1 x1 + 2 x2
Subject To
 C1: 1 x1 + 1 x2 <= 10
Bounds
 0 <= x1 <= 10
 0 <= x2
General
 x1
End
"""

# texts that reach the parser's corners: multi-line constraints and a
# right-hand side on the next line, unnamed rows, comments, duplicate terms,
# zero sums, constants, '2 * x', exponents, every bound form (and the ones it
# ignores), binaries, sections after END, and the error paths
CASES = {
    "simple": SIMPLE,
    "milp": MILP,
    "multiline": """\\ a comment
min
 obj: 3 x + 2 y - z
 + 4 w
st
 c1: x + y
   + z >= 2
 x - y <= 
 4
 c3: 2 x + 3 x - 5 x + w = 1
 -x - y >= -10
bounds
 x <= 4
 -2 <= z <= 3.5
 y >= 1
 w = 2
 z free
end
""",
    "quirks": """** star comment
MAXIMIZE
 profit: 2 * a + 3b + 1e2 c - 0.5 d + 7 + 2 a
SUBJECT TO
 r1: a + b + c + d <= 1e3
 r2: 2.5e-1 a - .5 b >= -3 x
 r3: a + a + a = 3
BOUNDS
 a <= 1e+06
 -inf <= b <= +inf
 .5 <= c
 d >= -inf
 0 <= nosuch <= 1
BINARIES
 c nosuch
GENERALS
 d
END
BOUNDS
 a <= 9
""",
    "after_end": """min
 obj: x
st
 c: x >= 1
end
this line stops the parse
bounds
 x <= 0
""",
    "unterminated": """min
 obj: x + y
st
 c1: x + y
 c2: x - y <= 1
""",
    "bad_rhs": """min
 obj: x
st
 c1: x <= abc
""",
    "bad_format": """min
 obj: x
st
 c1: x + y
""",
    "empty_rows": """min
 obj: 0 x
st
 c1: 0 x + 0 y <= 5
 c2: x - x >= 0
""",
}


def _same_lp(a, b):
    assert a.m == b.m and a.n == b.n and bool(a.maximize) == bool(b.maximize)
    np.testing.assert_array_equal(a.row_ptr, b.row_ptr)
    np.testing.assert_array_equal(a.col, b.col)
    np.testing.assert_array_equal(a.val, b.val)
    np.testing.assert_array_equal(a.sense, b.sense)
    np.testing.assert_array_equal(a.rhs, b.rhs)
    np.testing.assert_array_equal(a.lb, b.lb)
    np.testing.assert_array_equal(a.ub, b.ub)
    np.testing.assert_array_equal(a.obj, b.obj)


def test_reference_known_answers():
    r = lpformat.LPParser().parse_string(SIMPLE)
    assert r.success and r.problem.n == 2 and r.problem.m == 2 and r.problem.maximize
    r = lpformat.LPParser().parse_string(MILP)
    assert r.success and list(r.var_type) == [lpformat.INTEGER, lpformat.CONTINUOUS]
    assert r.problem.ub[0] == 10.0 and r.problem.lb[1] == 0.0 and math.isinf(r.problem.ub[1])


def test_corner_cases_by_hand():
    r = lpformat.LPParser().parse_string(CASES["multiline"])
    p = r.problem
    assert r.success and r.var_names == ["x", "y", "z", "w"] and r.con_names == ["c1", "c1", "c3", "c3"]
    assert list(p.obj) == [3.0, 2.0, -1.0, 4.0]
    assert list(p.sense) == [lpformat.GE, lpformat.LE, lpformat.EQ, lpformat.GE]
    assert list(p.rhs) == [2.0, 4.0, 1.0, -10.0]
    # c3: 2x + 3x - 5x sums to zero and is dropped; w stays
    assert list(p.col[p.row_ptr[2]:p.row_ptr[3]]) == [3]
    # 'z free' names no variable: z keeps [-2, 3.5]
    assert (p.lb[2], p.ub[2]) == (-2.0, 3.5) and (p.lb[3], p.ub[3]) == (2.0, 2.0)
    q = lpformat.LPParser().parse_string(CASES["quirks"])
    assert q.success
    a, b, c, d = (q.var_names.index(v) for v in "abcd")
    # '2 * a' reads as a with coefficient 1, then '+ 2 a' replaces it; 7 is dropped
    assert q.problem.obj[a] == 2.0 and q.problem.obj[b] == 3.0 and q.problem.obj[c] == 100.0
    # 'a <= 1e+06' is not a bound form; the BOUNDS header after END reopens
    # the section, so 'a <= 9' applies
    assert q.problem.ub[a] == 9.0
    assert q.problem.lb[b] == 0.0 and q.problem.ub[b] == math.inf  # '+inf' is not either
    assert q.problem.lb[c] == 0.0 and q.problem.ub[c] == 1.0 and q.var_type[c] == lpformat.BINARY
    assert q.problem.lb[d] == -math.inf and q.var_type[d] == lpformat.INTEGER
    assert lpformat.LPParser().parse_string(CASES["after_end"]).problem.ub[0] == math.inf
    for bad in ("unterminated", "bad_rhs", "bad_format"):
        assert not lpformat.LPParser().parse_string(CASES[bad]).success


def test_quadratic_and_sos_are_refused():
    for t in ("min\n obj: x + [ x ^ 2 ] / 2\nst\n c: x >= 1\n",
              "min\n obj: x\nst\n c: x >= 1\nquadratic\n x * x 2\n",
              "min\n obj: x + y\nst\n c: x + y >= 1\nsos\n S1 s1 x 1 y 2\n"):
        r = lpformat.LPParser().parse_string(t)
        assert not r.success and "not supported" in r.errors[-1]


@need_ref
@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("keep", [True, False])
def test_parse_matches_reference(name, keep):
    text = CASES[name]
    ref = oracle.ref_lp_parse(text, keep)
    p = lpformat.LPParser()
    p.set_keep_names(keep)
    r = p.parse_string(text)
    assert r.success == ref["success"]
    assert r.errors == ref["errors"]
    if ref["success"]:
        _same_lp(r.problem, ref["lp"])
        assert r.var_names == ref["var_names"] and r.con_names == ref["con_names"]
        np.testing.assert_array_equal(r.var_type, ref["var_type"])


def _random_text(seed):
    """A random LP text in the forms the reference reads: names, signs,
    coefficients with and without exponents, duplicate terms, rows split over
    lines, bounds of every form."""
    g = np.random.default_rng(seed)
    n, m = int(g.integers(1, 12)), int(g.integers(0, 9))
    names = ["v%d_%s" % (j, "ab"[j % 2]) for j in range(n)]

    def num():
        k = int(g.integers(0, 4))
        v = float(g.normal() * 10.0 ** int(g.integers(-3, 4)))
        return ("%.6g" % abs(v)) if k else ("%.3e" % abs(v))

    def expr():
        terms = []
        for _ in range(int(g.integers(1, 6))):
            sgn = " - " if g.random() < 0.4 else " + "
            co = "" if g.random() < 0.3 else num() + ("" if g.random() < 0.5 else " ")
            terms.append(sgn + co + names[int(g.integers(0, n))])
        return "".join(terms)

    lines = ["Maximize" if g.random() < 0.5 else "Minimize", " obj:" + expr(), "Subject To"]
    for i in range(m):
        e = expr()
        op = ["<=", ">=", "="][int(g.integers(0, 3))]
        rhs = ("-" if g.random() < 0.3 else "") + num()
        if g.random() < 0.3:
            cut = len(e) // 2
            while cut < len(e) and e[cut] != " ":
                cut += 1
            lines += [" r%d: %s" % (i, e[:cut]), "   %s %s %s" % (e[cut:], op, rhs)]
        else:
            lines.append(" r%d: %s %s %s" % (i, e, op, rhs) if g.random() < 0.8 else " %s %s %s" % (e, op, rhs))
    lines.append("Bounds")
    for j in range(n):
        k = int(g.integers(0, 6))
        a, b = sorted([float(g.integers(-5, 5)), float(g.integers(-5, 10))])
        lines.append([" %g <= %s <= %g" % (a, names[j], b), " %g <= %s" % (a, names[j]), " %s <= %g" % (names[j], b),
                      " %s >= %g" % (names[j], a), " %s = %g" % (names[j], a), " %s free" % names[j]][k])
    lines.append("End")
    return "\n".join(lines) + "\n"


@need_ref
@pytest.mark.parametrize("seed", range(40))
def test_random_texts_match_reference(seed):
    text = _random_text(seed)
    ref = oracle.ref_lp_parse(text)
    r = lpformat.LPParser().parse_string(text)
    assert r.success == ref["success"] and r.errors == ref["errors"], text
    if ref["success"]:
        _same_lp(r.problem, ref["lp"])
        assert r.var_names == ref["var_names"] and r.con_names == ref["con_names"]


@need_ref
@pytest.mark.parametrize("source", ["simple", "milp", "multiline", "quirks", "empty_rows", "random"])
def test_write_matches_reference(source, tmp_path):
    text = _random_text(7) if source == "random" else CASES[source]
    path = str(tmp_path / "ref.lp")
    oracle.ref_write_lp(text, path)
    r = lpformat.LPParser().parse_string(text)
    mine = lpformat.write_lp_string(r.problem, r.var_names, r.con_names, r.var_type)
    with open(path) as fh:
        assert mine == fh.read()


def test_round_trip_of_a_generated_lp(tmp_path):
    """write_lp then parse (tests/unit/test_roundtrip.cpp:44-60): the same
    sizes and bounds; coefficients within the %g format's 6 digits."""
    lp = lpgen.random_lp(5, 30, 60)
    lp.ub = np.where(np.arange(lp.n) % 3 == 0, 4.0, np.inf)
    path = str(tmp_path / "rt.lp")
    lpformat.write_lp(lp, path)
    back, res = lpformat.read_lp(path)
    assert back.m == lp.m and back.n == lp.n
    np.testing.assert_array_equal(back.lb, lp.lb)
    np.testing.assert_array_equal(back.ub, lp.ub)
    np.testing.assert_allclose(back.val, lp.val, rtol=5e-6)
    np.testing.assert_allclose(back.rhs, lp.rhs, rtol=5e-6)
    np.testing.assert_allclose(back.obj, lp.obj, rtol=5e-6)


@need_ref
@pytest.mark.parametrize("data", [
    b"Maximize\r obj: x + y\rSubject To\r c1: x + y <= 4\rEnd\r",            # bare CR separators
    b"\\ caf\xe9 au lait\nMaximize\n obj: x\nSubject To\n c1: x <= 1\nEnd\n",  # a Latin-1 byte
    b"Minimize\r\n obj: 2 x\r\nSubject To\r\n c1: x >= 1\r\nEnd\r\n",          # CR LF
], ids=["cr_only", "latin1_comment", "crlf"])
def test_parse_file_matches_reference_bytes(tmp_path, data):
    # parse_file reads the file's bytes as the reference does (lp_parser.cpp:12-23)
    path = tmp_path / "m.lp"
    path.write_bytes(data)
    ref = oracle.ref_lp_parse(data)
    r = lpformat.LPParser().parse_file(str(path))
    assert r.success == ref["success"]
    assert r.errors == ref["errors"]
    if ref["success"]:
        _same_lp(r.problem, ref["lp"])
        assert r.var_names == ref["var_names"] and r.con_names == ref["con_names"]
