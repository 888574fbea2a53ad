"""Replays the reference planner's kernel plans on the hip variant.

The reference prints one plan per statement with `--plan` (README.md:55-69), for example
``(dot_f64 x (rand_f64 1000000 2))``: an s-expression of kernel names over bound names and
literals. ``Plan.run`` evaluates such expressions through ``Dispatch.call`` the way
``fused::reference`` evaluates patterns (src/kernels/scalar/fused.rs:403-458): atoms are
bindings or literals (integers are i64 scalars, numbers with a point or exponent are f64
scalars, ``true``/``false`` are bools), lists apply a kernel by name to evaluated arguments.
Values stay on the device between statements.
"""
from __future__ import annotations

import re
from typing import Dict, List, Union

from .kernels import Array, Dispatch, Scalar

Sx = Union[str, List["Sx"]]


def parse(text: str) -> Sx:
    """One s-expression; atoms are runs of characters other than parentheses and whitespace."""
    toks = text.replace("(", " ( ").replace(")", " ) ").split()
    pos = 0

    def at() -> Sx:
        nonlocal pos
        if pos >= len(toks):
            raise ValueError("unexpected end of plan")
        t = toks[pos]
        pos += 1
        if t == ")":
            raise ValueError("unexpected `)` in plan")
        if t == "(":
            items = []
            while True:
                if pos >= len(toks):
                    raise ValueError("missing `)` in plan")
                if toks[pos] == ")":
                    pos += 1
                    return items
                items.append(at())
        return t

    sx = at()
    if pos != len(toks):
        raise ValueError(f"trailing token `{toks[pos]}` in plan `{text}`")
    return sx


_INT = re.compile(r"^-?\d+$")
_FLOAT = re.compile(r"^-?(\d+\.\d*|\d*\.\d+|\d+)([eE][-+]?\d+)?$|^-?inf$|^NaN$")


class Plan:
    """Evaluates plan lines against one dispatch, keeping named bindings on the device."""

    def __init__(self, dispatch: Dispatch):
        self.d = dispatch
        self.env: Dict[str, Union[Array, Scalar]] = {}

    def literal(self, atom: str):
        if atom in self.env:
            return self.env[atom]
        if atom in ("true", "false"):
            return Scalar("bool", atom == "true")
        if _INT.match(atom):
            return Scalar("i64", int(atom))
        if _FLOAT.match(atom):
            return Scalar("f64", float(atom))
        raise KeyError(f"unbound plan atom `{atom}`")

    def eval(self, sx: Sx):
        if isinstance(sx, str):
            return self.literal(sx)
        if not sx or not isinstance(sx[0], str):
            raise ValueError("plan application needs a kernel name")
        args = [self.eval(a) for a in sx[1:]]
        return self.d.call(sx[0], args)

    def run(self, text: str, bind: str | None = None):
        v = self.eval(parse(text))
        if bind:
            self.env[bind] = v
        return v
