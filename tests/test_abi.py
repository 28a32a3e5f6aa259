"""The C-ABI library loads and exports every entry point include/hornet_hip.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import subprocess

import pytest

from hornet_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exists():
    assert os.path.exists(_lib.LIB_PATH), "build with __graft_entry__.build()"


def test_exports_every_header_symbol():
    declared = _lib.header_symbols()
    assert len(declared) >= 20
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing


def test_python_signatures_cover_header():
    declared = set(_lib.header_symbols())
    assert declared <= set(_lib.SIGNATURES), declared - set(_lib.SIGNATURES)


def test_exports_are_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for s in _lib.header_symbols():
        assert s in names, s  # unmangled, i.e. extern "C"


def test_status_strings_without_device():
    lib = _lib.load()
    assert lib.hn_status_string(0) == b"ok"
    assert b"capacity" in lib.hn_status_string(5)
    assert lib.hn_version().startswith(b"hornet-hip")


def test_invalid_args_rejected_without_device():
    lib = _lib.load()
    # null out-pointer and negative sizes fail before touching the device
    assert lib.hn_graph_create_csr(-1, 0, None, None, 0, 0.0, None, None) == _lib.HN_E_INVALID_ARG
    assert lib.hn_graph_nv(None, None) == _lib.HN_E_INVALID_ARG
    assert lib.hn_bfs_topdown(None, 0, None, None, None) == _lib.HN_E_INVALID_ARG
    assert b"null" in lib.hn_last_error()
