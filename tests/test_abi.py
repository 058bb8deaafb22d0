"""CPU-side checks of the C ABI library: it loads without a GPU and exports
every entry point include/bilu.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

import bilu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "bilu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bilu_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(bilu.LIB_PATH):
        pytest.fail("libbilu.so is not built (run __graft_entry__.build())")
    return bilu.lib()


def test_header_lists_the_exports():
    assert header_symbols() == sorted(bilu.EXPORTS)


def test_library_exports_every_declared_symbol(lib):
    for name in header_symbols():
        assert hasattr(lib, name), name
        assert ctypes.cast(getattr(lib, name), ctypes.c_void_p).value


def test_host_only_entry_points(lib):
    o = bilu.Opts()
    lib.bilu_opts_default(ctypes.byref(o))
    assert o.precond == bilu.ILU0 and o.max_iter == 500 and o.tol == 1e-11
    assert o.tol_mode == bilu.TOL_ABSOLUTE and o.merge == bilu.MERGE_AUTO and o.parilu_sweeps == 5
    assert lib.bilu_version().decode().startswith("bilu")
    assert lib.bilu_status_string(3).decode() == "missing diagonal"
    # invalid arguments are rejected before any device work
    assert lib.bilu_pattern_create(0, 0, None, None, None) == 1
    assert lib.bilu_solve(None, 1, None, None, None, None, None, None, None, None, None) == 1
