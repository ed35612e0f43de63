"""The C-ABI library loads and exports every entry point include/srcwa.h declares (no GPU calls)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "srcwa.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(srcwa_\w+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from stablercwa_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_prototypes_cover_header():
    from stablercwa_amd import _lib
    assert set(declared_symbols()) <= set(_lib.PROTOTYPES), sorted(set(declared_symbols()) - set(_lib.PROTOTYPES))


def test_version_string():
    from stablercwa_amd import _lib
    assert b"gfx950" in _lib.load().srcwa_version()
