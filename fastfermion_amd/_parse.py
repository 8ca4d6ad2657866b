"""poly(): parse a Pauli or Majorana expression into a polynomial.

Accepts the expression language of fastfermion.poly (fastfermion/_parse.py:33-136):
products of factors X<i>, Y<i>, Z<i> (Pauli) or m<i> (Majorana), real and
imaginary number literals (`2.5`, `1j`, `3e-2j`), `+ - * /`, parentheses and
implied multiplication by juxtaposition ("3.1 (X1 + Y1) Z1").  Fermi factors
(f<i>, f<i>^) are outside this build and raise NotImplementedError.
This is a recursive-descent parser of our own, not a port of the reference's
regex/ast rewriting.
"""
from __future__ import annotations

import re

from . import _ffcore as ff

_TOKEN = re.compile(r"\s*(?:(?P<num>(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?j?|j)|(?P<pauli>[XYZ]\d+)"
                    r"|(?P<maj>m\d+)|(?P<fermi>f\d+\^?)|(?P<op>[-+*/()]))")


def _tokens(e: str):
    pos = 0
    out = []
    e = e.strip()
    while pos < len(e):
        mt = _TOKEN.match(e, pos)
        if not mt or mt.end() == pos:
            raise ValueError(f"Invalid expression near {e[pos:pos + 10]!r}")
        pos = mt.end()
        kind = mt.lastgroup
        val = mt.group(kind)
        out.append((kind, val))
    return out


class _Parser:
    def __init__(self, toks):
        self.t = toks
        self.i = 0

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def take(self):
        tok = self.peek()
        self.i += 1
        return tok

    def expr(self):
        v = self.term()
        while self.peek()[1] in ("+", "-") and self.peek()[0] == "op":
            op = self.take()[1]
            r = self.term()
            v = v + r if op == "+" else v - r
        return v

    def term(self):
        v = self.unary()
        while True:
            kind, val = self.peek()
            if kind == "op" and val in ("*", "/"):
                self.take()
                r = self.unary()
                v = v * r if val == "*" else v / r
            elif kind in ("num", "pauli", "maj", "fermi") or (kind == "op" and val == "("):
                v = v * self.unary()  # implied multiplication
            else:
                return v

    def unary(self):
        kind, val = self.peek()
        if kind == "op" and val in ("+", "-"):
            self.take()
            v = self.unary()
            return v if val == "+" else -v
        return self.factor()

    def factor(self):
        kind, val = self.take()
        if kind == "num":
            if val == "j":
                return 1j
            return complex(val) if val.endswith("j") else (float(val) if ("." in val or "e" in val.lower()) else int(val))
        if kind == "pauli":
            return ff.PauliPolynomial([(int(val[1:]), val[0])])
        if kind == "maj":
            return ff.MajoranaPolynomial([int(val[1:])])
        if kind == "fermi":
            raise NotImplementedError("Fermi operators are not part of fastfermion_amd (propagation path only)")
        if kind == "op" and val == "(":
            v = self.expr()
            if self.take() != ("op", ")"):
                raise ValueError("Invalid expression: missing ')'")
            return v
        raise ValueError("Invalid expression")


def poly(e: str):
    """Parse an expression into a PauliPolynomial or MajoranaPolynomial.

    >>> poly("X0 Z2")
    >>> poly("3.1 (X1 + Y1) Z1")
    >>> poly("m1 m3 m2")
    """
    p = _Parser(_tokens(e))
    v = p.expr()
    if p.i != len(p.t):
        raise ValueError("Invalid expression")
    return v


__all__ = ["poly"]
