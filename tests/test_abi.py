"""CPU checks of the C ABI: the library loads and exports every entry point
declared in include/ttgpu.h, and the ctypes table covers the header."""
import ctypes
import os
import re

import pytest

from treetime_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    with open(os.path.join(ROOT, "include", "ttgpu.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tt_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = header_functions()
    assert "tt_marginal" in names and "tt_joint" in names and len(names) >= 20


def test_library_exports_every_header_symbol():
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in header_functions():
        assert hasattr(lib, name), name


def test_ctypes_table_matches_header():
    assert sorted(_native.SIGNATURES) == header_functions()


def test_version_and_error_without_gpu():
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("library not built")
    lib = _native.load()
    assert lib.tt_version() >> 16 == 1
    # invalid arguments are rejected before any device call
    assert lib.tt_tree_create(None, None, None) == _native.TT_EINVAL
    assert b"NULL" in lib.tt_last_error()
