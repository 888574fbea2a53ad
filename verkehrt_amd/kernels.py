"""The host-side mirror of verkehrt's kernel-table API for the hip variant.

The reference exposes its kernels through ``Registry`` (src/kernels.rs:462-644) and
``Dispatch`` (src/kernels/dispatch.rs:25-265): kernels by name, ``Dispatch::call`` over whole
values, and tile-level reducers, scanners, selections and elementwise entries. This module
offers the same surface over the C ABI of ``libvk_hip.so``; values are ``Array`` (a
device-resident view in HBM) and ``Scalar``. Errors raise ``KernelError`` whose message is
the reference entry's exact text.
"""
from __future__ import annotations

import ctypes as C
import struct
from typing import Sequence, Union

import numpy as np

from . import native as N

import atexit

_ALIVE = [True]


@atexit.register
def _shutting_down():
    # Handles are not released during interpreter teardown: the order of finalizers is
    # arbitrary then, and the process exit returns every device allocation anyway.
    _ALIVE[0] = False


TYPES = ["f64", "f32", "i64", "u64", "i32", "u8", "bool"]
TYPE_CODE = {t: i for i, t in enumerate(TYPES)}
NP_DTYPE = {"f64": np.float64, "f32": np.float32, "i64": np.int64, "u64": np.uint64,
            "i32": np.int32, "u8": np.uint8, "bool": np.uint8}
CLASSES = ["elementwise", "reduction", "selection", "scan", "whole"]


class KernelError(Exception):
    """An entry's error. ``kind`` is "runtime" (a runtime error of the kernel, the
    reference's ``Err(String)``), "usage" (arguments the kernel does not accept) or
    "device" (a HIP failure)."""

    def __init__(self, message: str, kind: str):
        super().__init__(message)
        self.message = message
        self.kind = kind


def _check(rc: int, st: N.Status):
    if rc != 0:
        kind = {N.ERR_RUNTIME: "runtime", N.ERR_USAGE: "usage"}.get(st.code, "device")
        raise KernelError(st.msg.decode(), kind)


def type_of_dtype(dt) -> str:
    dt = np.dtype(dt)
    for t, d in NP_DTYPE.items():
        if t != "bool" and np.dtype(d) == dt:
            return t
    if dt == np.bool_:
        return "bool"
    raise TypeError(f"no verkehrt element type for {dt}")


class Scalar:
    """A scalar of an element type (Scalar, src/value.rs:116-130)."""

    __slots__ = ("type", "value")

    def __init__(self, type: str, value):
        self.type = type
        if type in ("f64", "f32"):
            self.value = float(value)
        elif type == "bool":
            self.value = bool(value)
        else:
            self.value = int(value)

    def bits(self) -> int:
        t, v = self.type, self.value
        if t == "f64":
            return struct.unpack("<Q", struct.pack("<d", v))[0]
        if t == "f32":
            return struct.unpack("<I", struct.pack("<f", v))[0]
        if t == "bool":
            return int(v)
        width = {"i64": 64, "u64": 64, "i32": 32, "u8": 8}[t]
        return v & ((1 << width) - 1)

    @staticmethod
    def from_c(cs: N.CScalar) -> "Scalar":
        return Scalar.from_bits(TYPES[cs.type], cs.bits)

    @staticmethod
    def from_bits(t: str, b: int) -> "Scalar":
        """The scalar of type t whose 64-bit word is b (the layout of vk_scalar.bits)."""
        if t == "f64":
            v = struct.unpack("<d", struct.pack("<Q", b))[0]
        elif t == "f32":
            v = struct.unpack("<f", struct.pack("<I", b & 0xFFFFFFFF))[0]
        elif t == "i64":
            v = b - (1 << 64) if b >> 63 else b
        elif t == "i32":
            b &= 0xFFFFFFFF
            v = b - (1 << 32) if b >> 31 else b
        elif t == "bool":
            v = b != 0
        else:
            v = b
        return Scalar(t, v)

    def __eq__(self, o):
        return isinstance(o, Scalar) and self.type == o.type and self.bits() == o.bits()

    def __repr__(self):
        return f"Scalar({self.type}, {self.value!r})"


class Context:
    """A device, its HIP stream and its HBM pool (vk_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = N.lib()
        h = C.c_void_p()
        st = N.Status()
        _check(self._lib.vk_ctx_create(device, C.byref(h), C.byref(st)), st)
        self.handle = h
        self.device = device

    def close(self):
        if self.handle and _ALIVE[0]:
            self._lib.vk_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        st = N.Status()
        _check(self._lib.vk_ctx_sync(self.handle, C.byref(st)), st)

    def stream(self) -> int:
        return self._lib.vk_ctx_stream(self.handle) or 0

    def trim(self):
        self._lib.vk_ctx_trim(self.handle)

    def pool_bytes(self):
        used = C.c_uint64()
        held = self._lib.vk_ctx_pool_bytes(self.handle, C.byref(used))
        return held, used.value

    def set_trace(self, on: bool):
        self._lib.vk_ctx_set_trace(self.handle, int(on))

    def take_trace(self) -> str:
        buf = C.create_string_buffer(1 << 20)
        self._lib.vk_ctx_trace_take(self.handle, buf, len(buf))
        return buf.value.decode()

    def last_us(self) -> float:
        return self._lib.vk_ctx_last_us(self.handle)


class Array:
    """An immutable device-resident array view (Array, src/value.rs:215-409)."""

    def __init__(self, ctx: Context, handle: int):
        self.ctx = ctx
        self.handle = C.c_void_p(handle) if isinstance(handle, int) else handle
        self._lib = N.lib()

    @staticmethod
    def from_numpy(ctx: Context, a, type: str | None = None) -> "Array":
        a = np.ascontiguousarray(a)
        t = type or type_of_dtype(a.dtype)
        a = np.ascontiguousarray(a, dtype=NP_DTYPE[t])
        h = C.c_void_p()
        st = N.Status()
        _check(N.lib().vk_array_from_host(ctx.handle, TYPE_CODE[t], a.ctypes.data_as(C.c_void_p),
                                          a.size, C.byref(h), C.byref(st)), st)
        return Array(ctx, h)

    @staticmethod
    def empty(ctx: Context, type: str, n: int) -> "Array":
        h = C.c_void_p()
        st = N.Status()
        _check(N.lib().vk_array_new(ctx.handle, TYPE_CODE[type], n, C.byref(h), C.byref(st)), st)
        return Array(ctx, h)

    @staticmethod
    def wrap(ctx: Context, type: str, device_ptr: int, n: int) -> "Array":
        h = C.c_void_p()
        st = N.Status()
        _check(N.lib().vk_array_wrap(ctx.handle, TYPE_CODE[type], C.c_void_p(device_ptr), n,
                                     C.byref(h), C.byref(st)), st)
        return Array(ctx, h)

    @property
    def type(self) -> str:
        return TYPES[self._lib.vk_array_type(self.handle)]

    def __len__(self) -> int:
        return self._lib.vk_array_len(self.handle)

    @property
    def nbytes(self) -> int:
        return self._lib.vk_array_bytes(self.handle)

    def device_ptr(self) -> int:
        return self._lib.vk_array_device_ptr(self.handle) or 0

    def slice(self, start: int, n: int) -> "Array":
        h = C.c_void_p()
        st = N.Status()
        _check(self._lib.vk_array_slice(self.handle, start, n, C.byref(h), C.byref(st)), st)
        return Array(self.ctx, h)

    def to_numpy(self) -> np.ndarray:
        out = np.empty(len(self), NP_DTYPE[self.type])
        st = N.Status()
        _check(self._lib.vk_array_to_host(self.ctx.handle, self.handle, out.ctypes.data_as(C.c_void_p),
                                          C.byref(st)), st)
        return out

    def release(self):
        if self.handle and _ALIVE[0]:
            self._lib.vk_array_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def __repr__(self):
        return f"Array<{self.type}>[{len(self)}]"


Value = Union[Array, Scalar]


# vk_arg and vk_value as numpy records (the layouts of N.CArg and N.CValue: 32 bytes each), so
# long argument lists of batched calls are filled column by column instead of field by field.
_ARG_DT = np.dtype({"names": ["kind", "rows", "stype", "bits"], "formats": ["<i4", "<u8", "<i4", "<u8"],
                    "offsets": [0, 8, 16, 24], "itemsize": 32})
_VALUE_DT = np.dtype({"names": ["kind", "array", "stype", "bits"], "formats": ["<i4", "<u8", "<i4", "<u8"],
                      "offsets": [0, 8, 16, 24], "itemsize": 32})


def _args(values: Sequence[Value]):
    n = len(values)
    if n < 64:  # short lists: field by field
        arr = (N.CArg * max(1, n))()
        for i, v in enumerate(values):
            if isinstance(v, Array):
                arr[i].kind = N.ROWS
                arr[i].rows = v.handle
            elif isinstance(v, Scalar):
                arr[i].kind = N.SPLAT
                arr[i].splat.type = TYPE_CODE[v.type]
                arr[i].splat.bits = v.bits()
            else:
                raise TypeError(f"argument {i} is neither Array nor Scalar: {v!r}")
        return arr
    kinds, rows, stypes, bits = [], [], [], []
    for i, v in enumerate(values):
        if isinstance(v, Array):
            kinds.append(N.ROWS)
            rows.append(v.handle.value or 0)
            stypes.append(0)
            bits.append(0)
        elif isinstance(v, Scalar):
            kinds.append(N.SPLAT)
            rows.append(0)
            stypes.append(TYPE_CODE[v.type])
            bits.append(v.bits())
        else:
            raise TypeError(f"argument {i} is neither Array nor Scalar: {v!r}")
    rec = np.zeros(max(1, n), _ARG_DT)
    if n:
        rec["kind"][:n] = kinds
        rec["rows"][:n] = rows
        rec["stype"][:n] = stypes
        rec["bits"][:n] = np.array(bits, dtype=np.uint64)
    return (N.CArg * max(1, n)).from_buffer(rec)  # keeps rec alive


class Registry:
    """The hip kernel table, by the reference's kernel names (Registry, src/kernels.rs:462)."""

    def __init__(self):
        self._lib = N.lib()

    def names(self):
        return [self._lib.vk_kernel_name(i).decode() for i in range(self._lib.vk_kernel_count())]

    def by_name(self, name: str):
        k = self._lib.vk_kernel_lookup(name.encode())
        return None if k < 0 else k

    def kernel_class(self, name: str) -> str:
        k = self.by_name(name)
        if k is None:
            raise KeyError(name)
        return CLASSES[self._lib.vk_kernel_class(k)]

    def variant(self) -> str:
        return self._lib.vk_variant_name().decode()

    def resolve(self, device: int = 0) -> "Dispatch":
        """A dispatch on one device (Registry::resolve, src/kernels.rs:628-644)."""
        return Dispatch(self, Context(device))


class Reducer:
    """A reduction state fed tile by tile (Reducer, src/kernels.rs:56-63)."""

    def __init__(self, ctx: Context, name: str):
        self.ctx = ctx
        self._lib = N.lib()
        h = C.c_void_p()
        st = N.Status()
        _check(self._lib.vk_reducer_new(ctx.handle, name.encode(), C.byref(h), C.byref(st)), st)
        self.handle = h

    def update(self, args: Sequence[Value]):
        a = _args(args)
        st = N.Status()
        _check(self._lib.vk_reducer_update(self.handle, a, len(args), C.byref(st)), st)

    def finish(self) -> Scalar:
        out = N.CScalar()
        st = N.Status()
        try:
            _check(self._lib.vk_reducer_finish(self.handle, C.byref(out), C.byref(st)), st)
        finally:
            self.free()
        return Scalar.from_c(out)

    def free(self):
        if self.handle and _ALIVE[0]:
            self._lib.vk_reducer_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Scanner:
    """A scan state fed tile by tile (Scanner, src/kernels.rs:78-85)."""

    def __init__(self, ctx: Context, name: str):
        self.ctx = ctx
        self._lib = N.lib()
        h = C.c_void_p()
        st = N.Status()
        _check(self._lib.vk_scanner_new(ctx.handle, name.encode(), C.byref(h), C.byref(st)), st)
        self.handle = h

    def update(self, args: Sequence[Value], out: Array):
        a = _args(args)
        st = N.Status()
        _check(self._lib.vk_scanner_update(self.handle, a, len(args), out.handle, C.byref(st)), st)

    def free(self):
        if self.handle and _ALIVE[0]:
            self._lib.vk_scanner_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Builder:
    """A device-side ArrayBuilder (src/value.rs:469-604) that selections append to."""

    def __init__(self, ctx: Context, type: str):
        self.ctx = ctx
        self._lib = N.lib()
        h = C.c_void_p()
        st = N.Status()
        _check(self._lib.vk_builder_new(ctx.handle, TYPE_CODE[type], C.byref(h), C.byref(st)), st)
        self.handle = h

    def __len__(self):
        return self._lib.vk_builder_len(self.handle)

    def finish(self) -> Array:
        h = C.c_void_p()
        st = N.Status()
        rc = self._lib.vk_builder_finish(self.handle, C.byref(h), C.byref(st))
        self.handle = None
        _check(rc, st)
        return Array(self.ctx, h)

    def __del__(self):
        try:
            if self.handle and _ALIVE[0]:
                self._lib.vk_builder_free(self.handle)
        except Exception:
            pass


class Dispatch:
    """Runs hip entries on one context (Dispatch, src/kernels/dispatch.rs:25-265)."""

    def __init__(self, registry: Registry, ctx: Context):
        self.registry = registry
        self.ctx = ctx
        self._lib = N.lib()

    def call(self, name: str, args: Sequence[Value]) -> Value:
        """Dispatch::call: one kernel over whole values."""
        a = _args(args)
        out = N.CValue()
        st = N.Status()
        _check(self._lib.vk_call(self.ctx.handle, name.encode(), a, len(args), C.byref(out),
                                 C.byref(st)), st)
        if out.kind == N.VALUE_SCALAR:
            return Scalar.from_c(out.scalar)
        return Array(self.ctx, out.array)

    def batch_call(self, name: str, items: Sequence[Sequence[Value]]) -> list:
        """Runs kernel `name` on every argument list of `items` with one launch per kernel
        phase (vk_batch_call); result i equals call(name, items[i])."""
        if not items:
            return []
        nargs = len(items[0])
        if any(len(it) != nargs for it in items):
            raise ValueError("every item needs the same number of arguments")
        flat = [v for it in items for v in it]
        a = _args(flat)
        rec = np.zeros(len(items), _VALUE_DT)
        out = (N.CValue * len(items)).from_buffer(rec)
        st = N.Status()
        _check(self._lib.vk_batch_call(self.ctx.handle, name.encode(), a, nargs, len(items), out,
                                       C.byref(st)), st)
        res = []
        for kind, arr, stype, b in zip(rec["kind"].tolist(), rec["array"].tolist(), rec["stype"].tolist(),
                                       rec["bits"].tolist()):
            if kind == N.VALUE_SCALAR:
                res.append(Scalar.from_bits(TYPES[stype], b))
            else:
                res.append(Array(self.ctx, arr))
        return res

    def batch_plan(self, inputs: Sequence[Sequence[Array]]) -> "BatchPlan":
        """A plan of batched statements over the items' input arrays (vk_bplan_new):
        inputs[i][k] is item i's input k."""
        return BatchPlan(self.ctx, inputs)

    def reducer(self, name: str) -> Reducer:
        return Reducer(self.ctx, name)

    def scanner(self, name: str) -> Scanner:
        return Scanner(self.ctx, name)

    def builder(self, type: str) -> Builder:
        return Builder(self.ctx, type)

    def select_tile(self, name: str, args: Sequence[Value], base: int, out: Builder):
        a = _args(args)
        st = N.Status()
        _check(self._lib.vk_select_tile(self.ctx.handle, name.encode(), a, len(args), base,
                                        out.handle, C.byref(st)), st)

    def region_reduce(self, kernels, root: str, root_args, inputs: Sequence[Value]):
        """A fused region (vk_region_reduce): `kernels` is a list of (name, sources) in
        post-order, `root` a reduction over `root_args`; a source is ("in", i) for inputs[i] or
        ("k", j) for kernel j's result. Returns (result, fused) where fused tells whether the
        region ran as one device pass."""
        ks = (N.CRKernel * max(1, len(kernels)))()
        names = []
        def src(s):
            kind, idx = s
            return N.CSrc(N.SRC_INPUT if kind == "in" else N.SRC_KERNEL, idx)
        for i, (name, srcs) in enumerate(kernels):
            names.append(name.encode())
            ks[i].name = names[-1]
            ks[i].nargs = len(srcs)
            for j, s_ in enumerate(srcs[:3]):
                ks[i].args[j] = src(s_)
        ra = (N.CSrc * max(1, len(root_args)))(*[src(s_) for s_ in root_args])
        a = _args(inputs)
        out = N.CScalar()
        fused = C.c_int32()
        st = N.Status()
        _check(self._lib.vk_region_reduce(self.ctx.handle, ks, len(kernels), root.encode(), ra, len(root_args), a,
                                          len(inputs), C.byref(out), C.byref(fused), C.byref(st)), st)
        return Scalar.from_c(out), bool(fused.value)

    def select_into(self, name: str, args: Sequence[Value], out: Array) -> int:
        """The selection written into the caller's array `out` with no device allocation
        (vk_select_into); returns the kept row count."""
        a = _args(args)
        kept = C.c_uint64()
        st = N.Status()
        _check(self._lib.vk_select_into(self.ctx.handle, name.encode(), a, len(args), out.handle,
                                        C.byref(kept), C.byref(st)), st)
        return kept.value

    def elementwise_tile(self, name: str, args: Sequence[Value], out: Array):
        a = _args(args)
        st = N.Status()
        _check(self._lib.vk_elementwise_tile(self.ctx.handle, name.encode(), a, len(args),
                                             out.handle, C.byref(st)), st)

    # conveniences
    def array(self, a, type: str | None = None) -> Array:
        return Array.from_numpy(self.ctx, a, type)

    def sync(self):
        self.ctx.sync()


class Input:
    """A plan argument: the item's plan input number `k`."""

    def __init__(self, k: int):
        self.k = k


class Stmt:
    """A plan argument: the item's array result of plan statement `s`."""

    def __init__(self, s: int):
        self.s = s


class BatchPlan:
    """Batched statements over fixed resident arrays (vk_bplan_*): add statements, compile
    (optionally autotuning tile shapes; the launch sequence, with tiles planned on the device,
    is captured into one hipGraph), then run (one graph replay, no host read-back) and read
    results; every result equals the single call on that item alone."""

    def __init__(self, ctx: Context, inputs: Sequence[Sequence[Array]]):
        self.ctx = ctx
        self._lib = N.lib()
        self.nitems = len(inputs)
        self.ninputs = len(inputs[0]) if inputs else 0
        if any(len(it) != self.ninputs for it in inputs):
            raise ValueError("every item needs the same number of inputs")
        flat = (C.c_void_p * max(1, self.nitems * self.ninputs))()
        for i, it in enumerate(inputs):
            for k, a in enumerate(it):
                flat[i * self.ninputs + k] = a.handle
        h = C.c_void_p()
        st = N.Status()
        _check(self._lib.vk_bplan_new(ctx.handle, flat, self.ninputs, self.nitems, C.byref(h), C.byref(st)), st)
        self.handle = h
        self.kinds = []

    def add(self, name: str, args: Sequence) -> int:
        refs = (N.CBref * max(1, len(args)))()
        for i, a in enumerate(args):
            if isinstance(a, Input):
                refs[i].kind, refs[i].index = N.BREF_INPUT, a.k
            elif isinstance(a, Stmt):
                refs[i].kind, refs[i].index = N.BREF_STMT, a.s
            elif isinstance(a, Scalar):
                refs[i].kind = N.BREF_SCALAR
                refs[i].scalar.type = TYPE_CODE[a.type]
                refs[i].scalar.bits = a.bits()
            else:
                raise TypeError(f"plan argument {a!r}")
        out = C.c_size_t()
        st = N.Status()
        _check(self._lib.vk_bplan_add(self.handle, name.encode(), refs, len(args), C.byref(out), C.byref(st)), st)
        return out.value

    def compile(self, tune: bool = False):
        st = N.Status()
        _check(self._lib.vk_bplan_compile(self.handle, 1 if tune else 0, C.byref(st)), st)

    def run(self):
        st = N.Status()
        _check(self._lib.vk_bplan_run(self.handle, C.byref(st)), st)

    def result(self, stmt: int, item: int):
        out = N.CValue()
        st = N.Status()
        _check(self._lib.vk_bplan_result(self.handle, stmt, item, C.byref(out), C.byref(st)), st)
        if out.kind == N.VALUE_SCALAR:
            return Scalar.from_c(out.scalar)
        return Array(self.ctx, out.array)

    def bits(self, stmt: int) -> np.ndarray:
        """Statement `stmt`'s result for every item as u64 words (vk_bplan_counts): a
        selection's kept row counts, or each item's scalar bits."""
        out = np.empty(self.nitems, np.uint64)
        st = N.Status()
        _check(self._lib.vk_bplan_counts(self.handle, stmt, out.ctypes.data_as(C.POINTER(C.c_uint64)),
                                         C.byref(st)), st)
        return out

    def tuning(self) -> str:
        n = self._lib.vk_bplan_tuning(self.handle, None, 0)
        buf = C.create_string_buffer(n + 1)
        self._lib.vk_bplan_tuning(self.handle, buf, n + 1)
        return buf.value.decode()

    def close(self):
        if self.handle and _ALIVE[0]:
            self._lib.vk_bplan_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
