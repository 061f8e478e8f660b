"""The C ABI library loads and exports every entry point the header
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from akitsushima_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "akitsushima_amd.h")) as fh:
        src = fh.read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ak_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_abi_version_and_device_count_without_gpu():
    L = _lib.lib()
    assert L.ak_abi_version() == 1
    assert _lib.device_count() >= 0
