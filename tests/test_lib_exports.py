"""The C-ABI library loads without a GPU and exports every entry point declared in
include/dhip.h (no compute calls here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "dhip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dhip_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from drops_amd import dhip as D
    L = D.load()
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(D.EXPORTS) == syms


def test_library_is_gfx950_code_object():
    from drops_amd import dhip as D
    data = open(D.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_missing_library_fails_loudly(tmp_path):
    import importlib
    from drops_amd import dhip as D
    import pytest
    saved = D._lib
    try:
        D._lib = None
        with pytest.raises(RuntimeError):
            D.load(str(tmp_path / "libdhip.so"))
    finally:
        D._lib = saved
