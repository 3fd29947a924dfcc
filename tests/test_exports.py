"""The C-ABI library loads and exports every function include/abea.h
declares (no GPU needed: nothing is called)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "abea.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(abea_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    from abea_amd import _lib
    names = declared_functions()
    assert len(names) >= 20
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    assert sorted(_lib.EXPORTS) == names


def test_binding_declares_every_export():
    """The ctypes binding gives every export an explicit signature."""
    from abea_amd import _lib
    L = _lib.lib()
    for n in _lib.EXPORTS:
        assert getattr(L, n).argtypes is not None, n
