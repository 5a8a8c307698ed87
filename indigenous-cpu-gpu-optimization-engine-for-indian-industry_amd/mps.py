"""MPS reader and writer for LP instances.

Reads the MPS format the way the reference's ``MPSParser`` interprets it
(core/model/mps_parser.cpp) and returns an :class:`lpgen.LP` in the
reference's ``Problem`` shape, ready for ``hypernova_hip.ipm_solve`` /
``ipm_solve_batch``:

* ROWS: the first ``N`` row is the objective; further ``N`` rows are free rows
  and are dropped.  ``E`` / ``L`` / ``G`` rows become constraints in file order.
* COLUMNS: entries by column; ``'MARKER'`` records (INTORG / INTEND) only
  delimit integer blocks, which this LP path relaxes (their bounds are kept).
  Rows of the constraint matrix are summed over duplicate entries, sorted by
  column, exact zeros dropped (ProblemBuilder::build, core/model/problem.hpp:
  265-281).
* RHS: sets a constraint's right-hand side; an entry naming the objective row
  is ignored, as the reference's RHS handler matches constraint names only
  (mps_parser.cpp:315-323).  Missing entries are 0.
* RANGES: the reference stores row ranges on its constraints
  (mps_parser.cpp:356-397) but its IPM builds the standard form from sense and
  rhs alone (interior_point.cpp:202-373 reads only variable bounds), so by
  default ranges are read and ignored, giving the problem the reference's IPM
  solves.  ``split_ranges=True`` applies the standard MPS meaning instead
  (L: [rhs - |R|, rhs], G: [rhs, rhs + |R|], E: [rhs, rhs + R] for R >= 0,
  [rhs + R, rhs] for R < 0) by splitting the row into a ``>=`` and a ``<=``
  row, since hn_lp has one sense per row.
* BOUNDS: LO, UP, FX, FR, MI, PL, BV, LI, UI as in mps_parser.cpp:404-441
  (UP sets only the upper bound, MI only the lower); default bounds [0, +inf).
  A bound line may leave the set name blank (dfl001 does); the column is then
  the second field.  The reference reads the fields by position
  (mps_parser.cpp:400-405), so on such lines it takes the column for the set
  name and drops the bound; dfl001's published optimum needs those bounds.
* OBJSENSE MAX / MAXIMIZE selects maximisation.

Fields are whitespace-separated (names without blanks); files may be gzipped.
"""
from __future__ import annotations

import gzip
import math
from typing import Dict, List, Optional, Tuple

import numpy as np

import lpgen

LE, GE, EQ = 0, 1, 2  # HN_LE / HN_GE / HN_EQ (ConstraintSense order)


class MpsError(ValueError):
    pass


def _open(path: str):
    return gzip.open(path, "rt") if path.endswith(".gz") else open(path, "r")


def read_mps(path: str, split_ranges: bool = False) -> Tuple[lpgen.LP, dict]:
    """Parses an MPS file.  Returns (lp, info) where info holds the instance
    name, row and column names, the integer-marked columns and the row
    ranges read."""
    name = ""
    section = None
    obj_row: Optional[str] = None
    row_sense: Dict[str, int] = {}
    row_order: List[str] = []
    free_rows = set()
    cols: Dict[str, int] = {}
    col_names: List[str] = []
    entries: List[Tuple[int, int, float]] = []  # (row index, column, value)
    obj: Dict[int, float] = {}
    rhs: Dict[str, float] = {}
    ranges: Dict[str, float] = {}
    lb: Dict[int, float] = {}
    ub: Dict[int, float] = {}
    integer = set()
    in_int = False
    maximize = False
    row_index: Dict[str, int] = {}

    def col_id(c: str) -> int:
        j = cols.get(c)
        if j is None:
            j = cols[c] = len(col_names)
            col_names.append(c)
            if in_int:
                integer.add(j)
        return j

    with _open(path) as f:
        for raw in f:
            line = raw.rstrip("\n")
            if not line.strip() or line.startswith("*"):
                continue
            tok = line.split()
            if not line[0].isspace():  # a section header
                head = tok[0].upper()
                if head == "NAME":
                    name = tok[1] if len(tok) > 1 else ""
                    section = None
                elif head in ("ROWS", "COLUMNS", "RHS", "RANGES", "BOUNDS", "OBJSENSE"):
                    section = head
                    if head == "OBJSENSE" and len(tok) > 1:
                        maximize = tok[1].upper().startswith("MAX")
                    if head == "COLUMNS":
                        row_index = {r: i for i, r in enumerate(row_order)}
                elif head == "ENDATA":
                    break
                else:
                    raise MpsError(f"{path}: unknown section {tok[0]}")
                continue
            if section == "OBJSENSE":
                maximize = tok[0].upper().startswith("MAX")
            elif section == "ROWS":
                kind, r = tok[0].upper(), tok[1]
                if kind == "N":
                    if obj_row is None:
                        obj_row = r
                    else:
                        free_rows.add(r)
                elif kind in ("E", "L", "G"):
                    row_sense[r] = {"E": EQ, "L": LE, "G": GE}[kind]
                    row_order.append(r)
                else:
                    raise MpsError(f"{path}: bad row type {tok[0]}")
            elif section == "COLUMNS":
                if len(tok) >= 3 and tok[1].strip("'").upper() == "MARKER":
                    kind = tok[2].strip("'").upper()
                    if kind == "INTORG":
                        in_int = True
                    elif kind == "INTEND":
                        in_int = False
                    continue
                j = col_id(tok[0])
                for r, v in zip(tok[1::2], tok[2::2]):
                    val = float(v)
                    if r == obj_row:
                        obj[j] = obj.get(j, 0.0) + val
                    elif r in row_index:
                        entries.append((row_index[r], j, val))
                    elif r in free_rows:
                        continue
                    else:
                        raise MpsError(f"{path}: column {tok[0]} names unknown row {r}")
            elif section in ("RHS", "RANGES"):
                # [set name] row value [row value]; the set name may be blank
                pairs = tok[1:] if len(tok) % 2 == 1 else tok
                target = rhs if section == "RHS" else ranges
                for r, v in zip(pairs[0::2], pairs[1::2]):
                    if r in row_sense:
                        target[r] = float(v)
            elif section == "BOUNDS":
                # type [set name] column [value]; the set name may be blank
                kind = tok[0].upper()
                if len(tok) >= 3 and tok[2] in cols:
                    c, rest = tok[2], tok[3:]
                elif len(tok) >= 2 and tok[1] in cols:
                    c, rest = tok[1], tok[2:]
                else:
                    raise MpsError(f"{path}: bound names no known column: {line.strip()}")
                j = cols[c]
                val = float(rest[0]) if rest else 0.0
                if kind == "LO":
                    lb[j] = val
                elif kind == "UP":
                    ub[j] = val
                elif kind == "FX":
                    lb[j] = ub[j] = val
                elif kind == "FR":
                    lb[j], ub[j] = -math.inf, math.inf
                elif kind == "MI":
                    lb[j] = -math.inf
                elif kind == "PL":
                    ub[j] = math.inf
                elif kind == "BV":
                    lb[j], ub[j] = 0.0, 1.0
                    integer.add(j)
                elif kind == "LI":
                    lb[j] = val
                    integer.add(j)
                elif kind == "UI":
                    ub[j] = val
                    integer.add(j)
                else:
                    raise MpsError(f"{path}: unsupported bound type {tok[0]}")
    if obj_row is None:
        raise MpsError(f"{path}: no objective row")

    n = len(col_names)
    # rows, with ranged rows split into a >= and a <= row
    out_rows: List[Tuple[int, int, float]] = []  # (source row, sense, rhs)
    for i, r in enumerate(row_order):
        s, b = row_sense[r], rhs.get(r, 0.0)
        if split_ranges and r in ranges:
            q = ranges[r]
            if s == LE:
                lo, hi = b - abs(q), b
            elif s == GE:
                lo, hi = b, b + abs(q)
            else:
                lo, hi = (b, b + q) if q >= 0 else (b + q, b)
            out_rows.append((i, GE, lo))
            out_rows.append((i, LE, hi))
        else:
            out_rows.append((i, s, b))
    per_row: Dict[int, Dict[int, float]] = {}
    for i, j, v in entries:
        d = per_row.setdefault(i, {})
        d[j] = d.get(j, 0.0) + v
    rp, ci, vv = [0], [], []
    for src, _, _ in out_rows:
        d = per_row.get(src, {})
        for j in sorted(d):
            if d[j] != 0.0:
                ci.append(j)
                vv.append(d[j])
        rp.append(len(ci))
    lp = lpgen.LP(m=len(out_rows), n=n, row_ptr=np.array(rp, np.int64), col=np.array(ci, np.int64),
                  val=np.array(vv, np.float64), rhs=np.array([b for _, _, b in out_rows], np.float64),
                  sense=np.array([s for _, s, _ in out_rows], np.int8),
                  lb=np.array([lb.get(j, 0.0) for j in range(n)], np.float64),
                  ub=np.array([ub.get(j, math.inf) for j in range(n)], np.float64),
                  obj=np.array([obj.get(j, 0.0) for j in range(n)], np.float64), maximize=maximize)
    info = {"name": name, "rows": row_order, "columns": col_names, "integer": sorted(integer),
            "constraint_nnz": len(entries), "ranges": dict(ranges)}
    return lp, info


def _num(v: float) -> str:
    return repr(float(v))


def write_mps(lp: lpgen.LP, path: str, name: str = "LP") -> None:
    """Writes `lp` as MPS (free-format fields, names R<i> / C<j>, numbers
    exact to the last bit); read_mps returns the same LP."""
    m, n = lp.m, lp.n
    sense = {LE: "L", GE: "G", EQ: "E"}
    # columns of the CSR rows
    by_col: List[List[Tuple[int, float]]] = [[] for _ in range(n)]
    for i in range(m):
        for k in range(int(lp.row_ptr[i]), int(lp.row_ptr[i + 1])):
            by_col[int(lp.col[k])].append((i, float(lp.val[k])))
    op = gzip.open if path.endswith(".gz") else open
    with op(path, "wt") as f:
        f.write(f"NAME          {name}\n")
        if lp.maximize:
            f.write("OBJSENSE\n    MAX\n")
        f.write("ROWS\n N  COST\n")
        for i in range(m):
            f.write(f" {sense[int(lp.sense[i])]}  R{i}\n")
        f.write("COLUMNS\n")
        for j in range(n):
            if lp.obj[j] != 0.0:
                f.write(f"    C{j}  COST  {_num(lp.obj[j])}\n")
            for i, v in by_col[j]:
                f.write(f"    C{j}  R{i}  {_num(v)}\n")
            if lp.obj[j] == 0.0 and not by_col[j]:
                f.write(f"    C{j}  COST  0.0\n")  # keeps an empty column declared
        f.write("RHS\n")
        for i in range(m):
            if lp.rhs[i] != 0.0:
                f.write(f"    RHS  R{i}  {_num(lp.rhs[i])}\n")
        f.write("BOUNDS\n")
        for j in range(n):
            lo, hi = float(lp.lb[j]), float(lp.ub[j])
            if lo == hi:
                f.write(f" FX BND  C{j}  {_num(lo)}\n")
                continue
            if lo == -math.inf and hi == math.inf:
                f.write(f" FR BND  C{j}\n")
                continue
            if lo == -math.inf:
                f.write(f" MI BND  C{j}\n")
            elif lo != 0.0:
                f.write(f" LO BND  C{j}  {_num(lo)}\n")
            if hi != math.inf:
                f.write(f" UP BND  C{j}  {_num(hi)}\n")
        f.write("ENDATA\n")
