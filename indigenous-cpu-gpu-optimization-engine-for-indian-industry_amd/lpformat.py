"""LP-format reader and writer for LP instances.

Reads the LP file format the way the reference's ``LPParser`` does
(core/model/lp_parser.cpp) and returns an :class:`lpgen.LP` in the
reference's ``Problem`` shape, ready for ``hypernova_hip.ipm_solve``:

* Lines are trimmed of `` \\t\\r\\n``; empty lines, lines starting with ``\\``
  and lines starting with ``**`` are skipped (lp_parser.cpp:42-44).
* A section header is a whole line, compared upper-cased: MINIMIZE / MAXIMIZE
  / MIN / MAX, SUBJECT TO / ST / S.T. / CONSTRAINTS, BOUNDS, GENERAL(S),
  BINARY / BINARIES, SEMI-CONTINUOUS / SEMI-CONT, SEMI-INTEGER / SEMI-INT,
  SOS, QUADRATIC / Q, END (:114-131).  A line after END that is not a header
  ends the parse (:53).
* Objective: each line, after its first ``:``, is appended; the linear part
  is parsed when the section ends, and a term's coefficient replaces, not
  adds to, an earlier one of the same variable (:138-186).
* Constraints accumulate over lines until they hold ``<=``, ``>=`` or ``=``;
  a line with ``:`` starts a named one (an unfinished one before it is an
  error).  ``<=`` / ``>=`` are searched first, then ``=``; an empty
  right-hand side keeps the constraint open; the right-hand side is read as
  ``std::stod`` reads it (its longest numeric prefix) (:63-81, :199-240).
* Terms: ``[sign] [coefficient] [name]`` scanned left to right; a
  coefficient without a name is dropped (constants), a character that starts
  neither is skipped, so ``2 * x`` reads as ``x`` with coefficient 1
  (:424-475).  Variables are numbered in order of first appearance, with
  bounds [0, +inf).
* Bounds: the reference's five regular expressions (two-sided, ``l <= x``,
  ``x <= u``, ``x >= l``, ``x = v``) with numbers ``-?(digits[.digits] | .digits
  | inf)``; any other line names no variable and is ignored (so ``x free``
  and ``x <= 1e+06`` change nothing) (:242-314).
* GENERAL / BINARY: whitespace-separated names; BINARY also sets [0, 1]
  (:316-344).
* Rows are then sorted by column with duplicates summed and exact zeros
  dropped (ProblemBuilder::build, core/model/problem.hpp:267-281).

Integer and binary columns are relaxed (their bounds kept) and reported in
``LPParseResult.var_type``.  Quadratic objective terms and SOS sets have no
place on the LP path: a file with them parses with an error naming them,
where the reference would build a QP or a MIP.

:func:`write_lp` writes what the reference's ``write_lp`` writes (:603-714):
numbers in C++'s default stream format (``%g``, 6 significant digits), every
column in the Bounds section, free columns as ``-inf <= x <= +inf``.  The
reader above does not take that line (``+inf`` is not in its number
pattern), nor ``%g`` exponents in bounds, so a round trip through the
reference's own pair loses those bounds; both sides here keep that
behaviour.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

import lpgen

LE, GE, EQ = 0, 1, 2  # HN_LE / HN_GE / HN_EQ (ConstraintSense order)
CONTINUOUS, INTEGER, BINARY = 0, 1, 2  # HN_VAR_* (VarType order)

_WS = " \t\n\v\f\r"  # C isspace
_TRIM = " \t\r\n"

_NUM = r"(-?(?:\d+\.?\d*|\.\d+|inf))"
_NAME = r"([a-zA-Z_][a-zA-Z0-9_]*)"
_B_TWO = re.compile(r"\s*" + _NUM + r"\s*<=\s*" + _NAME + r"\s*<=\s*" + _NUM + r"\s*", re.ASCII)
_B_LB = re.compile(r"\s*" + _NUM + r"\s*<=\s*" + _NAME + r"\s*", re.ASCII)
_B_UB = re.compile(r"\s*" + _NAME + r"\s*<=\s*" + _NUM + r"\s*", re.ASCII)
_B_GE = re.compile(r"\s*" + _NAME + r"\s*>=\s*" + _NUM + r"\s*", re.ASCII)
_B_EQ = re.compile(r"\s*" + _NAME + r"\s*=\s*" + _NUM + r"\s*", re.ASCII)

# the longest prefix strtod accepts (decimal, inf/infinity, nan), after
# leading white space
_STOD = re.compile(r"[ \t\n\v\f\r]*([+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan))",
                   re.ASCII | re.IGNORECASE)


def _stod(s: str) -> float:
    """std::stod: the longest numeric prefix; ValueError if there is none or
    it is out of double range."""
    mt = _STOD.match(s)
    if not mt:
        raise ValueError("stod")
    t = mt.group(1)
    v = float(t)
    if "inf" in t.lower() or "nan" in t.lower():
        return v
    mant = re.split(r"[eE]", t)[0]
    if math.isinf(v) or (abs(v) < 2.2250738585072014e-308 and re.search(r"[1-9]", mant)):
        raise ValueError("stod")  # overflow or underflow: std::out_of_range
    return v


def _trim(s: str) -> str:
    return s.strip(_TRIM)


_SECTIONS = {
    "MINIMIZE": "OBJ", "MAXIMIZE": "OBJ", "MIN": "OBJ", "MAX": "OBJ",
    "SUBJECT TO": "CON", "ST": "CON", "S.T.": "CON", "CONSTRAINTS": "CON",
    "BOUNDS": "BOUNDS", "GENERAL": "GENERAL", "GENERALS": "GENERAL",
    "BINARY": "BINARY", "BINARIES": "BINARY",
    "SEMI-CONTINUOUS": "SEMI", "SEMI-CONT": "SEMI", "SEMI-INTEGER": "SEMI", "SEMI-INT": "SEMI",
    "SOS": "SOS", "QUADRATIC": "QUAD", "Q": "QUAD", "END": "END",
}


@dataclass
class LPParseResult:
    """LPParseResult (lp_parser.hpp:10-15) with the problem as an lpgen.LP
    (None unless success) plus names and column types."""
    problem: Optional[lpgen.LP] = None
    var_names: List[str] = field(default_factory=list)
    con_names: List[str] = field(default_factory=list)
    var_type: Optional[np.ndarray] = None
    warnings: List[str] = field(default_factory=list)
    errors: List[str] = field(default_factory=list)
    success: bool = False


class LPParser:
    """LPParser (lp_parser.hpp:17-67): parse_file / parse_string /
    set_keep_names."""

    def __init__(self) -> None:
        self.keep_names = True

    def set_keep_names(self, keep: bool) -> None:
        self.keep_names = bool(keep)

    def parse_file(self, path: str) -> LPParseResult:
        # The reference reads the file's bytes unchanged into a string and
        # parses that (lp_parser.cpp:13-23): no newline translation, no
        # decoding.  Latin-1 maps every byte to one character, so a bare CR
        # stays inside its line and any byte value reads without error.
        try:
            with open(path, "rb") as fh:
                text = fh.read().decode("latin-1")
        except OSError:
            return LPParseResult(errors=["Cannot open file: " + path])
        return self.parse_string(text)

    # ---- state of one parse
    def _reset(self) -> None:
        self._names: List[str] = []
        self._lb: List[float] = []
        self._ub: List[float] = []
        self._obj: List[float] = []
        self._vtype: List[int] = []
        self._idx = {}
        self._rows: List[List[Tuple[int, float]]] = []
        self._sense: List[int] = []
        self._rhs: List[float] = []
        self._cnames: List[str] = []
        self._maximize = False
        self._obj_buf = ""
        self._con_acc = ""
        self._con_name = ""
        self._errors: List[str] = []

    def _add_variable(self, name: str) -> int:
        j = len(self._names)
        self._names.append(name if name else "x%d" % j)
        self._lb.append(0.0)
        self._ub.append(math.inf)
        self._obj.append(0.0)
        self._vtype.append(CONTINUOUS)
        return j

    def _linear(self, s: str) -> List[Tuple[int, float]]:
        """parse_linear_expr (lp_parser.cpp:424-475)."""
        terms = []
        pos, n = 0, len(s)
        while pos < n:
            while pos < n and s[pos] in _WS:
                pos += 1
            if pos >= n:
                break
            sign = 1.0
            if s[pos] == "+":
                pos += 1
            elif s[pos] == "-":
                sign = -1.0
                pos += 1
            while pos < n and s[pos] in _WS:
                pos += 1
            old = pos
            coeff = 1.0
            end = pos
            while end < n and (s[end].isdigit() and s[end].isascii() or s[end] == "."):
                end += 1
            if end < n and s[end] in "eE":
                e = end + 1
                if e < n and s[e] in "+-":
                    e += 1
                d = e
                while d < n and s[d].isascii() and s[d].isdigit():
                    d += 1
                if d > e:
                    end = d
            if end > pos:
                coeff = _stod(s[pos:end])
                pos = end
            while pos < n and s[pos] in _WS:
                pos += 1
            start = pos
            while pos < n and s[pos].isascii() and (s[pos].isalnum() or s[pos] == "_"):
                pos += 1
            name = s[start:pos]
            if name:
                j = self._idx.get(name)
                if j is None:
                    j = self._add_variable(name if self.keep_names else "")
                    self._idx[name] = j
                terms.append((j, sign * coeff))
            elif pos == old:
                pos += 1
        return terms

    def _finalize_obj(self) -> None:
        """finalize_obj (:138-186): the linear part outside [ ] groups."""
        if not self._obj_buf:
            return
        linear, cur, depth, quad = [], [], 0, False
        for ch in self._obj_buf:
            if ch == "[":
                depth += 1
                continue
            if ch == "]":
                if depth > 0:
                    depth -= 1
                    if depth == 0:
                        quad |= bool(_trim("".join(cur)))
                        cur = []
                continue
            (cur if depth > 0 else linear).append(ch)
        if depth > 0 and _trim("".join(cur)):
            quad = True
        for j, c in self._linear("".join(linear)):
            self._obj[j] = c
        if quad:
            self._errors.append("Quadratic objective terms are not supported on the LP path")
        self._obj_buf = ""

    def _finalize_constraint(self) -> None:
        """finalize_constraint (:199-240)."""
        if not self._con_acc and not self._con_name:
            return
        expr = self._con_acc
        sense, at = LE, -1
        for i in range(len(expr) - 1):
            if expr[i] == "<" and expr[i + 1] == "=":
                sense, at = LE, i
                break
            if expr[i] == ">" and expr[i + 1] == "=":
                sense, at = GE, i
                break
        if at < 0:
            at = expr.find("=")
            if at >= 0:
                sense = EQ
        if at < 0:
            self._errors.append("Invalid constraint format: " + expr)
            self._con_acc = self._con_name = ""
            return
        lhs = _trim(expr[:at])
        rhs_str = _trim(expr[at + (1 if sense == EQ else 2):])
        if not rhs_str:
            return  # stays open: the right-hand side may follow on the next line
        try:
            rhs = _stod(rhs_str)
        except ValueError:
            self._errors.append("Invalid constraint RHS: " + rhs_str)
            self._con_acc = self._con_name = ""
            return
        terms = self._linear(lhs)
        i = len(self._rows)
        self._rows.append(terms)
        self._sense.append(sense)
        self._rhs.append(rhs)
        self._cnames.append(self._con_name if (self.keep_names and self._con_name) else "c%d" % i)
        self._con_acc = self._con_name = ""

    def _find(self, name: str) -> int:
        for j, nm in enumerate(self._names):
            if nm == name:
                return j
        return -1

    def _bounds(self, line: str) -> None:
        """parse_bounds (:242-314)."""
        def num(t: str) -> float:
            t = _trim(t)
            lo = t.lower()
            if lo in ("inf", "+inf", "infinity"):
                return math.inf
            if lo in ("-inf", "-infinity"):
                return -math.inf
            return _stod(t)

        lb = ub = None
        fixed = False
        m = _B_TWO.fullmatch(line)
        if m:
            lb, name, ub = num(m.group(1)), m.group(2), num(m.group(3))
        elif _B_LB.fullmatch(line):
            m = _B_LB.fullmatch(line)
            lb, name = num(m.group(1)), m.group(2)
        elif _B_UB.fullmatch(line):
            m = _B_UB.fullmatch(line)
            name, ub = m.group(1), num(m.group(2))
        elif _B_GE.fullmatch(line):
            m = _B_GE.fullmatch(line)
            name, lb = m.group(1), num(m.group(2))
        elif _B_EQ.fullmatch(line):
            m = _B_EQ.fullmatch(line)
            name = m.group(1)
            lb = ub = num(m.group(2))
            fixed = True
        else:
            name = _trim(line)
        j = self._find(name)
        if j < 0:
            return
        if fixed:
            self._lb[j] = lb
            self._ub[j] = ub
        else:
            if lb is not None:
                self._lb[j] = lb
            if ub is not None:
                self._ub[j] = ub

    def _types(self, line: str, binary: bool) -> None:
        """parse_general / parse_binary (:316-344)."""
        for name in line.split():
            j = self._find(name)
            if j < 0:
                continue
            if binary:
                self._vtype[j] = BINARY
                self._lb[j], self._ub[j] = 0.0, 1.0
            else:
                self._vtype[j] = INTEGER

    def parse_string(self, text: str) -> LPParseResult:
        """parse_string (lp_parser.cpp:25-111)."""
        self._reset()
        section = None
        for num, raw in enumerate(text.split("\n"), start=1):
            t = _trim(raw)
            if not t or t[0] == "\\" or t.startswith("**"):
                continue
            up = t.upper()
            new = _SECTIONS.get(up)
            if new is not None:
                if section == "OBJ":
                    self._finalize_obj()
                if section == "CON":
                    self._finalize_constraint()
                if new == "OBJ":
                    self._maximize = up in ("MAXIMIZE", "MAX")
                section = new
                continue
            if section == "END":
                break
            try:
                if section == "OBJ":
                    c = t.find(":")
                    self._obj_buf += " " + (t[c + 1:] if c >= 0 else t)
                elif section == "CON":
                    if ":" in t:
                        if self._con_acc:
                            self._errors.append("Unterminated constraint before: " + t)
                            self._con_acc = self._con_name = ""
                        c = t.find(":")
                        self._con_name = _trim(t[:c])
                        self._con_acc = _trim(t[c + 1:])
                    else:
                        self._con_acc += " " + t
                    if "<=" in self._con_acc or ">=" in self._con_acc or "=" in self._con_acc:
                        self._finalize_constraint()
                elif section == "BOUNDS":
                    self._bounds(t)
                elif section == "GENERAL":
                    self._types(t, False)
                elif section == "BINARY":
                    self._types(t, True)
                elif section in ("SOS", "QUAD"):
                    self._errors.append("Line %d: %s sections are not supported on the LP path"
                                        % (num, "SOS" if section == "SOS" else "Quadratic"))
            except ValueError as e:
                self._errors.append("Line %d: %s" % (num, e))
        if section == "OBJ":
            self._finalize_obj()
        if section == "CON":
            self._finalize_constraint()
        res = LPParseResult(errors=list(self._errors))
        if res.errors:
            return res
        res.problem = self._build()
        res.var_names = list(self._names)
        res.con_names = list(self._cnames)
        res.var_type = np.asarray(self._vtype, dtype=np.int8)
        res.success = True
        return res

    def _build(self) -> lpgen.LP:
        """ProblemBuilder::build: rows sorted by column, duplicates summed in
        their file order, exact zeros dropped."""
        n, m = len(self._names), len(self._rows)
        rp, ci, va = [0], [], []
        for terms in self._rows:
            order = sorted(range(len(terms)), key=lambda k: terms[k][0])
            k = 0
            while k < len(order):
                j, s = terms[order[k]]
                k += 1
                while k < len(order) and terms[order[k]][0] == j:
                    s += terms[order[k]][1]
                    k += 1
                if s != 0.0:
                    ci.append(j)
                    va.append(s)
            rp.append(len(ci))
        return lpgen.LP(m=m, n=n, row_ptr=np.asarray(rp, dtype=np.int64), col=np.asarray(ci, dtype=np.int64),
                        val=np.asarray(va, dtype=np.float64), rhs=np.asarray(self._rhs, dtype=np.float64),
                        sense=np.asarray(self._sense, dtype=np.int8), lb=np.asarray(self._lb, dtype=np.float64),
                        ub=np.asarray(self._ub, dtype=np.float64), obj=np.asarray(self._obj, dtype=np.float64),
                        maximize=self._maximize)


def read_lp(path: str) -> Tuple[lpgen.LP, LPParseResult]:
    """Parses an LP file; raises ValueError with the parser's errors."""
    r = LPParser().parse_file(path)
    if not r.success:
        raise ValueError("; ".join(r.errors))
    return r.problem, r


def _g(v: float) -> str:
    """C++ ostream's default format for a double (printf %g)."""
    return "%g" % v


def write_lp_string(lp: lpgen.LP, var_names: Optional[List[str]] = None, con_names: Optional[List[str]] = None,
                    var_type: Optional[np.ndarray] = None) -> str:
    """write_lp (lp_parser.cpp:603-714) into a string.  Default names are the
    builder's x<j> / c<i>."""
    n, m = int(lp.n), int(lp.m)
    vn = list(var_names) if var_names is not None else ["x%d" % j for j in range(n)]
    cn = list(con_names) if con_names is not None else ["c%d" % i for i in range(m)]

    def terms(pairs) -> str:
        out, first = [], True
        for j, c in pairs:
            if c == 0.0:
                continue
            if not first:
                out.append(" + " if c > 0 else " - ")
            elif c < 0:
                out.append("-")
            a = abs(c)
            if a != 1.0:
                out.append(_g(a))
            out.append(" " + vn[j])
            first = False
        if first:
            out.append(" 0")
        return "".join(out)

    lines = ["\\* LP file written by HyperNova *\\", "", "Maximize" if lp.maximize else "Minimize"]
    lines.append(" OBJ:" + terms((j, float(lp.obj[j])) for j in range(n)))
    lines += ["", "Subject To"]
    rp, ci, va = lp.row_ptr, lp.col, lp.val
    for i in range(m):
        # the reference walks columns in order through get(i, j): the row's
        # entries sorted by column, duplicates summed
        row = {}
        for k in range(int(rp[i]), int(rp[i + 1])):
            row[int(ci[k])] = row.get(int(ci[k]), 0.0) + float(va[k])
        s = int(lp.sense[i])
        op = " <= " if s == LE else (" >= " if s == GE else " = ")
        lines.append(" " + cn[i] + ":" + terms(sorted(row.items())) + op + _g(float(lp.rhs[i])))
    lines += ["", "Bounds"]
    for j in range(n):
        lo, hi = float(lp.lb[j]), float(lp.ub[j])
        if lo > -math.inf and hi < math.inf:
            lines.append(" %s <= %s <= %s" % (_g(lo), vn[j], _g(hi)))
        elif lo > -math.inf:
            lines.append(" %s <= %s" % (_g(lo), vn[j]))
        elif hi < math.inf:
            lines.append(" %s <= %s" % (vn[j], _g(hi)))
        else:
            lines.append(" -inf <= %s <= +inf" % vn[j])
    if var_type is not None:
        gen = [vn[j] for j in range(n) if int(var_type[j]) == INTEGER]
        binr = [vn[j] for j in range(n) if int(var_type[j]) == BINARY]
        if gen:
            lines += ["", "General"] + [" " + x for x in gen]
        if binr:
            lines += ["", "Binary"] + [" " + x for x in binr]
    lines += ["", "End", ""]
    return "\n".join(lines)


def write_lp(lp: lpgen.LP, path: str, var_names: Optional[List[str]] = None,
             con_names: Optional[List[str]] = None, var_type: Optional[np.ndarray] = None) -> None:
    with open(path, "w") as fh:
        fh.write(write_lp_string(lp, var_names, con_names, var_type))
