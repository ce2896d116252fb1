#!/usr/bin/env python3
"""Regenerate ``quad_rules.inc``: the tetrahedron and triangle cubature rules
selected by the reference's ``AutoQuadSelector`` for orders 0..5.

The rules are published cubature data (Jaskowiec & Sukumar, IJNME 2020,
doi:10.1002/nme.6313 for the tetrahedra; rotationally / fully symmetric
triangle rules) that the reference transcribes as C++ constants in
``include/GaussQuadrature/{QuadTetLowerOrder.h,QuadTetHighOrder.h,
QuadTriLowerOrder.h,QuadTriHighOrder.h}``.  Parity of the DG residual needs
the *same doubles* the reference compiles, so this script reads those header
texts, evaluates every constant expression with C double semantics (a decimal
literal is rounded once to the nearest double, ``a/6.0`` and ``0.5*a`` are one
IEEE operation each) and writes them as C99 hex-float literals.

Selector mapping (``include/GaussQuadrature/GaussLegendre.h:84-111``,
``AutoQuadSelectorTri.h:1-22``):

  order  tet rule          tri rule
  0..2   Degree5Points15   Degree4Points7 (0,1) / Degree10Points24 (2)
  3      Degree7Points31   Degree10Points24
  4      Degree9Points57   Degree10Points24
  5      Degree11Points94  Degree13Points36

Run from the repository root in a container that holds the reference tree; it writes
``include/GaussQuadrature/quad_rules.inc``:
    python3 tests/golden/gen_quad_rules.py /root/reference
"""
import re
import sys
from pathlib import Path

TET_RULES = ["Degree5Points15", "Degree7Points31", "Degree9Points57", "Degree11Points94"]
TRI_RULES = ["Degree4Points7", "Degree10Points24", "Degree13Points36"]
TET_FILES = ["QuadTetLowerOrder.h", "QuadTetHighOrder.h"]
TRI_FILES = ["QuadTriLowerOrder.h", "QuadTriHighOrder.h"]


class CExpr:
    """Recursive-descent evaluator for + - * / ( ) over C numeric literals and
    previously defined names, with C semantics (int/int truncates)."""

    TOK = re.compile(r"\s*(?:(\d+\.\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?|\d+[eE][+-]?\d+)|(\d+)|([A-Za-z_]\w*)|(.))")

    def __init__(self, text, names):
        self.toks = []
        for m in self.TOK.finditer(text):
            if m.group(1):
                self.toks.append(("num", float(m.group(1))))
            elif m.group(2):
                self.toks.append(("num", int(m.group(2))))
            elif m.group(3):
                if m.group(3) not in names:
                    raise ValueError(f"unknown name {m.group(3)!r} in {text!r}")
                self.toks.append(("num", names[m.group(3)]))
            elif m.group(4).strip():
                self.toks.append(("op", m.group(4)))
        self.i = 0

    def peek(self):
        return self.toks[self.i] if self.i < len(self.toks) else (None, None)

    def take(self):
        t = self.peek()
        self.i += 1
        return t

    def parse(self):
        v = self.expr()
        if self.i != len(self.toks):
            raise ValueError("trailing tokens")
        return v

    def expr(self):
        v = self.term()
        while self.peek() in (("op", "+"), ("op", "-")):
            op = self.take()[1]
            r = self.term()
            v = v + r if op == "+" else v - r
        return v

    def term(self):
        v = self.unary()
        while self.peek() in (("op", "*"), ("op", "/")):
            op = self.take()[1]
            r = self.unary()
            if op == "*":
                v = v * r
            elif isinstance(v, int) and isinstance(r, int):
                v = int(v / r)
            else:
                v = float(v) / float(r)
        return v

    def unary(self):
        if self.peek() == ("op", "-"):
            self.take()
            return -self.unary()
        if self.peek() == ("op", "+"):
            self.take()
            return self.unary()
        kind, val = self.take()
        if kind == "num":
            return val
        if (kind, val) == ("op", "("):
            v = self.expr()
            if self.take() != ("op", ")"):
                raise ValueError("missing )")
            return v
        raise ValueError(f"unexpected token {val!r}")


def struct_body(text, name):
    m = re.search(r"struct\s+" + name + r"\s*\{", text)
    if not m:
        return None
    depth, i = 1, m.end()
    while depth:
        c = text[i]
        depth += c == "{"
        depth -= c == "}"
        i += 1
    return text[m.end():i - 1]


def split_top(s):
    """Split a brace-free comma list."""
    return [p.strip() for p in s.split(",") if p.strip()]


def parse_rule(body, dim):
    body = re.sub(r"//[^\n]*", "", body)
    names = {}
    for m in re.finditer(r"static\s+constexpr\s+Scalar\s+(\w+)\s*=\s*([^;]+);", body):
        names[m.group(1)] = float(CExpr(m.group(2), names).parse())
    npts = int(re.search(r"num_points\s*=\s*(\d+)", body).group(1))
    pm = re.search(r"points\s*=\s*\{\{(.*?)\}\};", body, re.S)
    groups = re.findall(r"\{([^{}]*)\}", pm.group(1))
    pts = [[float(CExpr(e, names).parse()) for e in split_top(g)] for g in groups]
    wm = re.search(r"weights\s*=\s*\{(.*?)\};", body, re.S)
    wts = [float(CExpr(e, names).parse()) for e in split_top(wm.group(1))]
    assert len(pts) == npts == len(wts), (len(pts), npts, len(wts))
    assert all(len(p) == dim for p in pts)
    return pts, wts


def find_rule(root, files, name, dim):
    for f in files:
        text = (root / "include" / "GaussQuadrature" / f).read_text()
        body = struct_body(text, name)
        if body is not None:
            return parse_rule(body, dim), f
    raise KeyError(name)


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    out = Path(__file__).resolve().parents[2] / "include" / "GaussQuadrature" / "quad_rules.inc"
    lines = [
        "/* Generated by gen_quad_rules.py -- do not edit.  Usable from C and C++:",
        " * each rule is <TAG>_N, <TAG>_PTS (brace list of points), <TAG>_W (weights).",
        " * Cubature rules selected by the reference's AutoQuadSelector",
        " * (include/GaussQuadrature/GaussLegendre.h:84-111, AutoQuadSelectorTri.h:1-22),",
        " * as the exact doubles the reference compiles, written as hex floats.",
        " * Tet weights sum to 1/6, triangle weights to 1/2. */",
        "",
    ]
    for kind, rules, files, dim in (("TET", TET_RULES, TET_FILES, 3), ("TRI", TRI_RULES, TRI_FILES, 2)):
        for name in rules:
            (pts, wts), src = find_rule(ref, files, name, dim)
            tag = f"FVDG_{kind}_{name.upper()}"
            lines.append(f"/* {name}: {len(pts)} points, from {src} */")
            lines.append(f"#define {tag}_N {len(pts)}")
            lines.append(f"#define {tag}_PTS {{ \\")
            for p in pts:
                lines.append("  {" + ", ".join(x.hex() for x in p) + "}, \\")
            lines.append("}")
            lines.append(f"#define {tag}_W {{ \\")
            for w in wts:
                lines.append("  " + w.hex() + ", \\")
            lines.append("}")
            lines.append("")
    out.write_text("\n".join(lines))
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
