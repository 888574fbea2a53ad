"""CPU checks of the drop-in boundary: the C-ABI library loads without a GPU, exports every
function include/vk_hip.h declares, and its kernel table holds the reference's kernel names
of the hot path. No compute call is made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vk_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from verkehrt_amd import native
    lib = ctypes.CDLL(native.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 35
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares exactly the header's functions
    assert sorted(native.SIGNATURES) == names


HOT_PATH = [
    "dot_f64", "reduce_add_f64", "reduce_add_i64", "count_gt_f64", "filter_gt_f64",
    "compress_f64", "compress_i64", "compress_bool", "where", "gather_f64", "gather_i64",
    "scan_add_i64", "scan_by_add_i64", "reduce_min_f64", "reduce_max_f64", "argmin_f64",
    "argmax_f64", "argsort_u64", "argsort_i64", "argsort_i32", "argsort_u8", "group_id_i64",
    "group_id2_i64_i64", "reduce_into_add_i64", "histogram", "rand_f64", "rand_i64",
    "length_f64", "sumsq_f64", "condsum_gt_f64", "count",
]


def test_kernel_table_has_the_hot_path():
    from verkehrt_amd import Registry
    reg = Registry()
    names = set(reg.names())
    missing = [n for n in HOT_PATH if n not in names]
    assert not missing, missing
    assert reg.variant() == "isa=hip"
    assert reg.kernel_class("dot_f64") == "reduction"
    assert reg.kernel_class("filter_gt_f64") == "selection"
    assert reg.kernel_class("scan_by_add_i64") == "scan"
    assert reg.kernel_class("argsort_u64") == "whole"
    assert reg.kernel_class("gt_f64") == "elementwise"
    assert reg.by_name("no_such_kernel") is None


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from verkehrt_amd import native
    monkeypatch.setattr(native, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(native, "_lib", None)
    with pytest.raises(RuntimeError, match="hip backend not built"):
        native.lib()
