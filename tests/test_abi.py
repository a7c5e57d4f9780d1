"""The C ABI library: built, loadable, exporting every declared symbol; host-only calls."""

from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

from hafnium_amd import _lib
from oracle import layout_ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hfjk.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"HFJK_API\s+[\w\s\*]*?\b(hfjk_\w+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_symbols()
    for needed in ("hfjk_create", "hfjk_coulomb", "hfjk_exchange_occ", "hfjk_exchange_density", "hfjk_bind"):
        assert needed in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "build libhfjk.so first (python -m hafnium_amd._build)"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\sT\s(hfjk_\w+)", out.stdout))
    assert set(declared_symbols()) <= exported
    # Nothing but the C ABI leaks out of the library.
    assert exported == set(declared_symbols())


def test_binding_covers_the_header():
    assert {name for name, _, _ in _lib.SIGNATURES} == set(declared_symbols())


def test_host_only_calls():
    lib = _lib.load()
    assert lib.hfjk_version() >= 100
    for n in (1, 31, 32, 33, 100, 1500):
        assert lib.hfjk_packed_elems(n) == layout_ref.packed_elems(n)
    assert lib.hfjk_packed_elems(0) == 0
    # A null context is rejected with an input error, not a crash.
    status = lib.hfjk_bind(None, 3, 4, ctypes.c_void_p(1), None, None)
    assert status == _lib.ERR_INPUT
    assert b"null context" in lib.hfjk_last_error()
    with pytest.raises(Exception, match="null context"):
        _lib.check(status)


def test_precision_names():
    assert _lib.precision_code("fp64") == _lib.FP64
    assert _lib.precision_code("bf16x3") == _lib.BF16X3
    with pytest.raises(ValueError):
        _lib.precision_code("fp16")
