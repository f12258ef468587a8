"""The C-ABI library loads and exports every entry point its header declares."""
import ctypes
import os
import re

from beta1_amd.capi import LIB_PATH, SIGNATURES, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "beta1_mhd.h")).read()
    return sorted(set(re.findall(r"\b(mhd_[a-z0-9_]+)\s*\(", text)))


def test_library_exists_and_loads():
    assert os.path.exists(LIB_PATH)
    assert lib() is not None


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(LIB_PATH)
    names = declared_symbols()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(declared_symbols()) == set(SIGNATURES)
