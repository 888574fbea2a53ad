"""TEST INFRASTRUCTURE, NOT PRODUCT CODE.

ctypes wrapper over ``oracle/lib/libvk_oracle.so``, the plain-C restatement of verkehrt's
scalar kernel entries (``oracle/vk_oracle.c``). Only ``tests/``, ``__graft_entry__.smoke``
and the CPU-baseline leg of ``bench.py`` import this module, as the checker.

Each wrapper returns numpy arrays or Python scalars and mirrors the error behaviour of the
reference entry it restates (the message strings are produced here from the C return codes,
with the exact texts of the reference).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libvk_oracle.so")

CMPS = {"eq": 0, "ne": 1, "lt": 2, "le": 3, "gt": 4, "ge": 5}

_lib = None


class OracleError(Exception):
    """A runtime error of a reference entry, carrying its exact message."""


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"oracle library missing at {LIB_PATH}; run oracle/build.py (or __graft_entry__.build())"
            )
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    u64, i64, dbl, vp, i32 = C.c_uint64, C.c_int64, C.c_double, C.c_void_p, C.c_int
    sig = {
        "vko_splitmix64": (u64, [u64]),
        "vko_rand_f64": (None, [u64, i64, vp]),
        "vko_rand_i64": (None, [u64, i64, vp]),
        "vko_canonical_sum_f64": (dbl, [vp, u64]),
        "vko_canonical_product_f64": (dbl, [vp, u64]),
        "vko_canonical_sum_f32": (C.c_float, [vp, u64]),
        "vko_canonical_fold_f32": (C.c_uint32, [i32, vp, vp, u64, i32, C.c_float, C.c_float]),
        "vko_fused_sum_int": (u64, [i32, i32, vp, vp, u64, i32, u64, u64]),
        "vko_canonical_fold_f64": (i32, [i32, vp, vp, u64, i32, dbl, dbl, vp, i32, i32, vp]),
        "vko_dot_f64": (dbl, [vp, vp, u64]),
        "vko_sumsq_f64": (dbl, [vp, u64]),
        "vko_condsum_f64": (dbl, [vp, u64, i32, dbl, dbl]),
        "vko_sum_i64": (i64, [vp, u64]),
        "vko_sum_u64": (u64, [vp, u64]),
        "vko_sum_i32": (i64, [vp, u64]),
        "vko_sum_u8": (u64, [vp, u64]),
        "vko_product_int": (u64, [vp, C.c_uint32, u64]),
        "vko_count_cmp_f64": (i64, [vp, u64, i32, dbl]),
        "vko_count_cmp_i64": (i64, [vp, u64, i32, i64]),
        "vko_count_bool": (i64, [vp, u64]),
        "vko_count_cmp_u64": (i64, [vp, u64, i32, u64]),
        "vko_reduce_extreme_u64": (i32, [vp, u64, i32, vp]),
        "vko_arg_extreme_u64": (i32, [vp, u64, i32, vp]),
        "vko_min2_f64": (dbl, [dbl, dbl]),
        "vko_max2_f64": (dbl, [dbl, dbl]),
        "vko_reduce_extreme_f64": (i32, [vp, u64, i32, vp]),
        "vko_reduce_extreme_i64": (i32, [vp, u64, i32, vp]),
        "vko_arg_extreme_f64": (i32, [vp, u64, i32, vp]),
        "vko_arg_extreme_i64": (i32, [vp, u64, i32, vp]),
        "vko_filter_f64": (u64, [vp, u64, i32, dbl, vp]),
        "vko_filter_i64": (u64, [vp, u64, i32, i64, vp]),
        "vko_compress": (u64, [vp, C.c_uint32, vp, u64, vp]),
        "vko_where": (u64, [vp, u64, u64, vp]),
        "vko_scan_add_i64": (None, [vp, u64, vp]),
        "vko_scan_by_add_i64": (None, [vp, vp, u64, vp]),
        "vko_scan_extreme_f64": (None, [vp, vp, u64, i32, vp]),
        "vko_scan_extreme_i64": (None, [vp, vp, u64, i32, vp]),
        "vko_scan_extreme_u64": (None, [vp, vp, u64, i32, vp]),
        "vko_scan_add_f64": (None, [vp, u64, vp]),
        "vko_gather": (i32, [vp, C.c_uint32, u64, vp, u64, vp, vp]),
        "vko_histogram": (i32, [vp, u64, i64, vp, vp]),
        "vko_reduce_into_add_i64": (i32, [vp, vp, u64, i64, vp, vp]),
        "vko_reduce_into_extreme_f64": (i32, [vp, vp, u64, i64, i32, vp, vp]),
        "vko_reduce_into_extreme_i64": (i32, [vp, vp, u64, i64, i32, vp, vp]),
        "vko_reduce_into_add_f64": (i32, [vp, vp, u64, i64, vp, vp]),
        "vko_argsort_u64": (None, [vp, u64, vp]),
        "vko_argsort_i64": (None, [vp, u64, vp]),
        "vko_argsort_i32": (None, [vp, u64, vp]),
        "vko_argsort_u8": (None, [vp, u64, vp]),
        "vko_argsort_f64": (None, [vp, u64, vp]),
        "vko_group_id_u64": (None, [vp, u64, vp]),
        "vko_group_id_pair_u64": (None, [vp, vp, u64, vp]),
        "vko_ew_f64": (None, [i32, vp, vp, u64, vp]),
        "vko_ew_f32": (None, [i32, vp, vp, u64, vp]),
        "vko_ew_int": (None, [i32, i32, vp, vp, u64, vp]),
        "vko_cast": (None, [i32, i32, vp, u64, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ----------------------------------------------------------------------- RNG

def splitmix64(x: int) -> int:
    return lib().vko_splitmix64(x & (2**64 - 1))


def rand_f64(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib().vko_rand_f64(n, seed, _p(out))
    return out


def rand_i64(n: int, seed: int) -> np.ndarray:
    out = np.empty(n, np.int64)
    lib().vko_rand_i64(n, seed, _p(out))
    return out


# -------------------------------------------------------- canonical reductions

def canonical_sum_f64(x) -> float:
    x = _c(x, np.float64)
    return lib().vko_canonical_sum_f64(_p(x), x.size)


def canonical_product_f64(x) -> float:
    x = _c(x, np.float64)
    return lib().vko_canonical_product_f64(_p(x), x.size)


def canonical_sum_f32(x) -> np.float32:
    x = _c(x, np.float32)
    return np.float32(lib().vko_canonical_sum_f32(_p(x), x.size))


def canonical_fold_f32(kind: str, x, y=None, cmp="gt", c=0.0, e=0.0) -> np.float32:
    """The canonical f32 fold (accumulation in f32) of row values of ``kind``: sum, dot, sumsq,
    condsum or mul (the product)."""
    kinds = {"sum": 0, "dot": 1, "sumsq": 2, "condsum": 3, "mul": 4}
    x = _c(x, np.float32)
    y = _c(y if y is not None else x, np.float32)
    if x.size != y.size:
        raise OracleError(f"length mismatch: {x.size} vs {y.size}")
    bits = lib().vko_canonical_fold_f32(kinds[kind], _p(x), _p(y), x.size, CMPS[cmp], float(np.float32(c)),
                                        float(np.float32(e)))
    return np.array([bits], np.uint32).view(np.float32)[0]


def fused_sum_int(kind: str, x, y=None, cmp="gt", c=0, e=0) -> int:
    """dot, sumsq or condsum over an i64, u64 or i32 array: wrapping row values in the element
    type, widened, summed with wrapping; the result as a Python int of the sum type (i64 for
    i64 and i32, u64 for u64)."""
    kinds = {"dot": 1, "sumsq": 2, "condsum": 3}
    x = np.ascontiguousarray(x)
    t = {np.dtype(np.int64): 2, np.dtype(np.uint64): 3, np.dtype(np.int32): 4}[x.dtype]
    y = np.ascontiguousarray(y if y is not None else x)
    if y.dtype != x.dtype or y.size != x.size:
        raise OracleError(f"length mismatch: {x.size} vs {y.size}")
    mask = (1 << 64) - 1
    bits = lib().vko_fused_sum_int(kinds[kind], t, _p(x), _p(y), x.size, CMPS[cmp], int(c) & mask, int(e) & mask)
    if x.dtype == np.uint64:
        return bits
    return bits - (1 << 64) if bits >> 63 else bits


def canonical_fold_f64(kind: str, x, y=None, cmp="gt", c=0.0, e=0.0, cuts=(), mul=False) -> float:
    """The canonical fold of row values of ``kind`` (sum, dot, sumsq, condsum) fed in
    tiles ending at ``cuts``; raises OracleError for a partial block before an update."""
    kinds = {"sum": 0, "dot": 1, "sumsq": 2, "condsum": 3}
    x = _c(x, np.float64)
    y = _c(y if y is not None else x, np.float64)
    cuts_a = np.asarray(cuts, np.uint64)
    out = np.zeros(1, np.float64)
    r = lib().vko_canonical_fold_f64(kinds[kind], _p(x), _p(y), x.size, CMPS[cmp], c, e,
                                     _p(cuts_a) if len(cuts) else None, len(cuts), int(mul), _p(out))
    if r != 0:
        raise OracleError("canonical update after a partial block")
    return float(out[0])


def dot_f64(x, y) -> float:
    x, y = _c(x, np.float64), _c(y, np.float64)
    if x.size != y.size:
        raise OracleError(f"length mismatch: {x.size} vs {y.size}")
    return lib().vko_dot_f64(_p(x), _p(y), x.size)


def sumsq_f64(x) -> float:
    x = _c(x, np.float64)
    return lib().vko_sumsq_f64(_p(x), x.size)


def condsum_f64(x, cmp, c, e) -> float:
    x = _c(x, np.float64)
    return lib().vko_condsum_f64(_p(x), x.size, CMPS[cmp], c, e)


# --------------------------------------------------------- integer reductions

def sum_int(x) -> int:
    x = np.ascontiguousarray(x)
    f = {np.dtype(np.int64): lib().vko_sum_i64, np.dtype(np.uint64): lib().vko_sum_u64,
         np.dtype(np.int32): lib().vko_sum_i32, np.dtype(np.uint8): lib().vko_sum_u8}[x.dtype]
    return f(_p(x), x.size)


def product_int(x) -> int:
    """reduce_mul over an integer array: the wrapping product in the element type."""
    x = np.ascontiguousarray(x)
    bits = lib().vko_product_int(_p(x), x.dtype.itemsize, x.size)
    if x.dtype.kind == "i" and bits >> (8 * x.dtype.itemsize - 1):
        return bits - (1 << (8 * x.dtype.itemsize))
    return bits


def count_cmp(x, cmp: str, c) -> int:
    x = np.ascontiguousarray(x)
    if x.dtype == np.float64:
        return lib().vko_count_cmp_f64(_p(x), x.size, CMPS[cmp], float(c))
    if x.dtype == np.uint64:
        return lib().vko_count_cmp_u64(_p(x), x.size, CMPS[cmp], int(c))
    x = _c(x, np.int64)
    return lib().vko_count_cmp_i64(_p(x), x.size, CMPS[cmp], int(c))


def count_bool(m) -> int:
    m = _c(m, np.uint8)
    return lib().vko_count_bool(_p(m), m.size)


# ------------------------------------------------------------------- extremes

def min2(a: float, b: float) -> float:
    return lib().vko_min2_f64(a, b)


def max2(a: float, b: float) -> float:
    return lib().vko_max2_f64(a, b)


def reduce_extreme(x, want_max: bool):
    x = np.ascontiguousarray(x)
    what = "max" if want_max else "min"
    if x.dtype == np.float64:
        out = np.zeros(1, np.float64)
        r = lib().vko_reduce_extreme_f64(_p(x), x.size, int(want_max), _p(out))
    elif x.dtype == np.uint64:
        out = np.zeros(1, np.uint64)
        r = lib().vko_reduce_extreme_u64(_p(x), x.size, int(want_max), _p(out))
    else:
        x = _c(x, np.int64)
        out = np.zeros(1, np.int64)
        r = lib().vko_reduce_extreme_i64(_p(x), x.size, int(want_max), _p(out))
    if r != 0:
        raise OracleError(f"reduce_{what} of an empty array")
    return out[0]


def arg_extreme(x, want_max: bool) -> int:
    x = np.ascontiguousarray(x)
    out = np.zeros(1, np.int64)
    if x.dtype == np.float64:
        r = lib().vko_arg_extreme_f64(_p(x), x.size, int(want_max), _p(out))
    elif x.dtype == np.uint64:
        r = lib().vko_arg_extreme_u64(_p(x), x.size, int(want_max), _p(out))
    else:
        x = _c(x, np.int64)
        r = lib().vko_arg_extreme_i64(_p(x), x.size, int(want_max), _p(out))
    if r != 0:
        raise OracleError(f"{'argmax' if want_max else 'argmin'} of an empty array")
    return int(out[0])


# ----------------------------------------------------------------- selections

def filter_cmp(x, cmp: str, c) -> np.ndarray:
    x = np.ascontiguousarray(x)
    out = np.empty(max(x.size, 1), x.dtype)
    if x.dtype == np.float64:
        k = lib().vko_filter_f64(_p(x), x.size, CMPS[cmp], float(c), _p(out))
    else:
        x = _c(x, np.int64)
        k = lib().vko_filter_i64(_p(x), x.size, CMPS[cmp], int(c), _p(out))
    return out[:k].copy()


def compress(x, mask) -> np.ndarray:
    x = np.ascontiguousarray(x)
    mask = _c(mask, np.uint8)
    if x.size != mask.size:
        raise OracleError(f"length mismatch: {x.size} values vs {mask.size} mask rows")
    out = np.empty(max(x.size, 1), x.dtype)
    k = lib().vko_compress(_p(x), x.dtype.itemsize, _p(mask), x.size, _p(out))
    return out[:k].copy()


def where(mask, base: int = 0) -> np.ndarray:
    mask = _c(mask, np.uint8)
    out = np.empty(max(mask.size, 1), np.int64)
    k = lib().vko_where(_p(mask), mask.size, base, _p(out))
    return out[:k].copy()


# ---------------------------------------------------------------------- scans

def scan_add_i64(x) -> np.ndarray:
    x = _c(x, np.int64)
    out = np.empty(x.size, np.int64)
    lib().vko_scan_add_i64(_p(x), x.size, _p(out))
    return out


def scan_by_add_i64(x, starts) -> np.ndarray:
    x, s = _c(x, np.int64), _c(starts, np.uint8)
    out = np.empty(x.size, np.int64)
    lib().vko_scan_by_add_i64(_p(x), _p(s), x.size, _p(out))
    return out


def scan_extreme(x, want_max: bool, starts=None) -> np.ndarray:
    x = np.ascontiguousarray(x)
    s = _c(starts, np.uint8) if starts is not None else None
    out = np.empty(x.size, x.dtype)
    if x.dtype == np.float64:
        lib().vko_scan_extreme_f64(_p(x), _p(s), x.size, int(want_max), _p(out))
    elif x.dtype == np.uint64:
        lib().vko_scan_extreme_u64(_p(x), _p(s), x.size, int(want_max), _p(out))
    else:
        lib().vko_scan_extreme_i64(_p(_c(x, np.int64)), _p(s), x.size, int(want_max), _p(out))
    return out


def scan_add_f64(x) -> np.ndarray:
    x = _c(x, np.float64)
    out = np.empty(x.size, np.float64)
    lib().vko_scan_add_f64(_p(x), x.size, _p(out))
    return out


# --------------------------------------------------------------------- gather

def gather(xs, idx) -> np.ndarray:
    xs = np.ascontiguousarray(xs)
    idx = _c(idx, np.int64)
    out = np.empty(max(idx.size, 1), xs.dtype)
    bad = np.zeros(1, np.int64)
    r = lib().vko_gather(_p(xs), xs.dtype.itemsize, xs.size, _p(idx), idx.size, _p(out), _p(bad))
    if r != 0:
        raise OracleError(f"gather index {int(bad[0])} out of range for length {xs.size}")
    return out[: idx.size].copy()


# -------------------------------------------------------------------- buckets

def _bucket_error(r, bad, buckets):
    if r == -2:
        raise OracleError(f"negative bucket count {buckets}")
    raise OracleError(f"bucket key {int(bad[0])} out of range for {buckets} buckets")


def histogram(keys, buckets: int) -> np.ndarray:
    keys = _c(keys, np.int64)
    out = np.zeros(max(buckets, 1), np.int64)
    bad = np.zeros(1, np.int64)
    r = lib().vko_histogram(_p(keys), keys.size, buckets, _p(out), _p(bad))
    if r:
        _bucket_error(r, bad, buckets)
    return out[:buckets].copy()


def reduce_into(keys, vals, buckets: int, op: str = "add") -> np.ndarray:
    keys = _c(keys, np.int64)
    vals = np.ascontiguousarray(vals)
    bad = np.zeros(1, np.int64)
    out = np.zeros(max(buckets, 1), vals.dtype)
    if keys.size != vals.size:
        # checked_keys runs before the length check, index.rs:199-207
        histogram(keys, buckets)
        raise OracleError(f"length mismatch: {keys.size} keys for {vals.size} values")
    if vals.dtype == np.int64 and op == "add":
        r = lib().vko_reduce_into_add_i64(_p(keys), _p(vals), keys.size, buckets, _p(out), _p(bad))
    elif vals.dtype == np.float64 and op == "add":
        r = lib().vko_reduce_into_add_f64(_p(keys), _p(vals), keys.size, buckets, _p(out), _p(bad))
    elif vals.dtype == np.float64 and op in ("min", "max"):
        r = lib().vko_reduce_into_extreme_f64(_p(keys), _p(vals), keys.size, buckets,
                                              int(op == "max"), _p(out), _p(bad))
    elif vals.dtype == np.int64 and op in ("min", "max"):
        r = lib().vko_reduce_into_extreme_i64(_p(keys), _p(vals), keys.size, buckets,
                                              int(op == "max"), _p(out), _p(bad))
    else:
        raise NotImplementedError(f"reduce_into {op} on {vals.dtype}")
    if r:
        _bucket_error(r, bad, buckets)
    return out[:buckets].copy()


# -------------------------------------------------------------------- argsort

def argsort(x) -> np.ndarray:
    x = np.ascontiguousarray(x)
    out = np.empty(x.size, np.int64)
    f = {np.dtype(np.uint64): lib().vko_argsort_u64, np.dtype(np.int64): lib().vko_argsort_i64,
         np.dtype(np.int32): lib().vko_argsort_i32, np.dtype(np.uint8): lib().vko_argsort_u8,
         np.dtype(np.float64): lib().vko_argsort_f64}[x.dtype]
    f(_p(x), x.size, _p(out))
    return out


# ------------------------------------------------------------------- group_id

def _bits(x) -> np.ndarray:
    x = np.ascontiguousarray(x)
    if x.dtype in (np.int64, np.uint64):
        return x.view(np.uint64)
    if x.dtype == np.int32:
        return x.view(np.uint32).astype(np.uint64)
    return x.astype(np.uint64)


def group_id(x, y=None) -> np.ndarray:
    a = _bits(x)
    out = np.empty(a.size, np.int64)
    if y is None:
        lib().vko_group_id_u64(_p(a), a.size, _p(out))
    else:
        b = _bits(y)
        if a.size != b.size:
            raise OracleError(f"length mismatch: {a.size} vs {b.size}")
        lib().vko_group_id_pair_u64(_p(a), _p(b), a.size, _p(out))
    return out


# ---------------------------------------------------------------- elementwise

TYPE_CODE = {"f64": 0, "f32": 1, "i64": 2, "u64": 3, "i32": 4, "u8": 5, "bool": 6}
NP = {"f64": np.float64, "f32": np.float32, "i64": np.int64, "u64": np.uint64, "i32": np.int32,
      "u8": np.uint8, "bool": np.uint8}
FLOAT_OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "min": 4, "max": 5, "neg": 6, "abs": 7, "sqrt": 8}
INT_OPS = {"add": 0, "sub": 1, "mul": 2, "div_const": 3, "min": 4, "max": 5, "neg": 6, "abs": 7,
           "and": 8, "or": 9, "xor": 10, "not": 11}


def _rows(v, t, n):
    a = np.asarray(v, dtype=NP[t])
    return np.ascontiguousarray(np.broadcast_to(a, (n,)) if a.ndim == 0 else a)


def elementwise(op: str, t: str, a, b=None) -> np.ndarray:
    """The elementwise kernel `<op>_<t>` over rows (scalars broadcast to the rows' length);
    comparisons return 0/1 bytes. Raises OracleError for div_const by zero."""
    n = max(np.size(a), np.size(b) if b is not None else 0)
    if op in ("eq", "ne", "lt", "le", "gt", "ge"):
        x, y = _rows(a, t, n), _rows(b, t, n)
        if t in ("f64", "f32"):
            x = x.astype(np.float64)
            y = y.astype(np.float64)
        else:
            x = [int(v) for v in x]
            y = [int(v) for v in y]
        f = {"eq": lambda p, q: p == q, "ne": lambda p, q: p != q, "lt": lambda p, q: p < q,
             "le": lambda p, q: p <= q, "gt": lambda p, q: p > q, "ge": lambda p, q: p >= q}[op]
        return np.array([1 if f(float(p) if t in ("f64", "f32") else p, float(q) if t in ("f64", "f32") else q)
                         else 0 for p, q in zip(x, y)], np.uint8)
    x = _rows(a, t, n)
    y = _rows(b, t, n) if b is not None else None
    out = np.empty(n, NP[t])
    if t == "f64":
        lib().vko_ew_f64(FLOAT_OPS[op], _p(x), _p(y), n, _p(out))
    elif t == "f32":
        lib().vko_ew_f32(FLOAT_OPS[op], _p(x), _p(y), n, _p(out))
    else:
        if op == "div_const" and (y is None or not np.all(y != 0)):
            raise OracleError("div_const: division by zero")
        lib().vko_ew_int(INT_OPS[op], TYPE_CODE[t], _p(x), _p(y), n, _p(out))
    return out


def select(cond, t: str, a, b) -> np.ndarray:
    """select_<t>: rows of `a` where cond is nonzero, else rows of `b` (elementwise.rs:105-120)."""
    n = max(np.size(cond), np.size(a), np.size(b))
    c, x, y = _rows(cond, "u8", n), _rows(a, t, n), _rows(b, t, n)
    out = np.empty(n, NP[t])
    for i in range(n):
        out[i] = x[i] if c[i] else y[i]
    return out


def cast(x, src: str, dst: str) -> np.ndarray:
    """cast_<src>_<dst> with Rust `as` semantics."""
    x = np.ascontiguousarray(x, dtype=NP[src])
    out = np.empty(x.size, NP[dst])
    lib().vko_cast(TYPE_CODE[src], TYPE_CODE[dst], _p(x), x.size, _p(out))
    return out


# ----------------------------------------------- best x86 CPU path (baseline only)

BEST_PATH = os.path.join(HERE, "lib", "libvk_cpu_best.so")
_best = None


def best():
    """ctypes handle of oracle/cpu_best.c, the restated x86-64-v4 backend timed as the CPU
    baseline. Checked against the oracle functions above by tests/test_cpu_best.py."""
    global _best
    if _best is None:
        L = C.CDLL(BEST_PATH)
        u64, i64, dbl, vp = C.c_uint64, C.c_int64, C.c_double, C.c_void_p
        for name, res, args in [
            ("vkb_sum_f64", dbl, [vp, u64]), ("vkb_dot_f64", dbl, [vp, vp, u64]),
            ("vkb_count_gt_f64", i64, [vp, u64, dbl]), ("vkb_min_f64", dbl, [vp, u64]),
            ("vkb_argmax_f64", i64, [vp, u64]), ("vkb_scan_add_i64", None, [vp, u64, vp]),
            ("vkb_filter_gt_f64", u64, [vp, u64, dbl, vp]), ("vkb_compress_f64", u64, [vp, vp, u64, vp]),
            ("vkb_gather_f64", C.c_int, [vp, u64, vp, u64, vp, vp]),
            ("vkb_scan_by_add_i64", None, [vp, vp, u64, vp]),
            ("vkb_argsort_64", C.c_int, [C.c_int, vp, u64, vp]),
            ("vkb_group_id_64", C.c_int, [vp, u64, vp]),
            ("vkb_reduce_into_add_f64", C.c_int, [vp, vp, u64, u64, vp]),
        ]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _best = L
    return _best
