"""The C-ABI library loads and exports every entry point include/*.h declares.

CPU-only: nothing here calls into the GPU.
"""
import ctypes
import os
import re

from ising_engine_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            text = open(os.path.join(inc, f)).read()
            names |= set(re.findall(r"\b(ige_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    assert set(_lib.EXPORTED) == declared_symbols()


def test_binding_loads_without_gpu():
    lib = _lib.load()
    assert b"gfx950" in lib.ige_version()
