"""Replays the reference planner's kernel plans on the hip variant.

The reference prints one plan per statement with `--plan` (README.md:55-69), for example
``(dot_f64 x (rand_f64 1000000 2))``: an s-expression of kernel names over bound names and
literals. ``Plan.run`` evaluates such expressions through ``Dispatch.call`` the way
``fused::reference`` evaluates patterns (src/kernels/scalar/fused.rs:403-458): atoms are
bindings or literals (integers are i64 scalars, numbers with a point or exponent are f64
scalars, ``true``/``false`` are bools), lists apply a kernel by name to evaluated arguments.
Values stay on the device between statements. `slice` is a zero-copy view, and
``format_value`` prints a value the way the reference's `print` does, so replayed programs can
be compared with the expected outputs of tests/programs/*.dd.
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
_TYPED = re.compile(r"^(-?\d+(?:\.\d*)?)(i64|u64|i32|u8|f64|f32)$")
_FLOAT = re.compile(r"^-?(\d+\.\d*|\d*\.\d+|\d+)([eE][-+]?\d+)?$|^-?inf$|^NaN$")


def format_float(v: float) -> str:
    """Rust's `{:?}` for a float, as the reference prints scalars and array rows
    (src/value.rs:170-183): always a decimal point or exponent, `NaN`, `inf`."""
    if v != v:
        return "NaN"
    if v in (float("inf"), float("-inf")):
        return "inf" if v > 0 else "-inf"
    r = repr(float(v))
    if "e" in r:
        m, e = r.split("e")
        return f"{m}e{int(e)}"
    return r


def format_value(v) -> str:
    """A value as the reference's `print` writes it (src/value.rs:170-183, 621-660): scalars
    plainly, floats with `{:?}`, bools as true/false, arrays as `[a, b, c]`."""
    if isinstance(v, Scalar):
        if v.type in ("f64", "f32"):
            return format_float(v.value)
        if v.type == "bool":
            return "true" if v.value else "false"
        return str(v.value)
    rows = v.to_numpy()
    if v.type in ("f64", "f32"):
        items = [format_float(float(x)) for x in rows]
    elif v.type == "bool":
        items = ["true" if x else "false" for x in rows]
    else:
        items = [str(int(x)) for x in rows]
    return "[" + ", ".join(items) + "]"


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
        m = _TYPED.match(atom)
        if m:  # a literal with its type suffix, as the reference's source writes them (255u8)
            t = m.group(2)
            return Scalar(t, float(m.group(1)) if t in ("f64", "f32") else int(m.group(1)))
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
        if sx[0] == "slice":
            # a zero-copy view of rows start..end (Array::slice, src/value.rs:236-249)
            a, start, end = args
            return a.slice(start.value, end.value - start.value)
        return self.d.call(sx[0], args)

    def bind(self, name: str, rows, type: str | None = None):
        """Binds `name` to a device array of the given host rows (a program's array literal)."""
        import numpy as np
        self.env[name] = Array.from_numpy(self.d.ctx, np.asarray(rows), type)
        return self.env[name]

    def run(self, text: str, bind: str | None = None):
        v = self.eval(parse(text))
        if bind:
            self.env[bind] = v
        return v
