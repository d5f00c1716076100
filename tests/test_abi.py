"""The C-ABI library loads and exports every symbol that include/*.h declares
(no GPU needed: nothing is called)."""
import ctypes
import os
import re

import pytest

from ratbite_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            src = open(os.path.join(ROOT, "include", f)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(rbh_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared()) <= set(L.declared_symbols())
    L.lib()
    assert not L.MISSING


def test_record_layouts():
    assert L.MATERIAL_DTYPE.itemsize == 152
    assert L.TEXTURE_DTYPE.itemsize == 48
    assert L.EMISSIVE_TRI_DTYPE.itemsize == 68
    assert L.PACKED_NODE_DTYPE.itemsize == 32


def test_host_only_entry_rejects_bad_arguments():
    assert L.lib().rbh_light_tree_build_host(None, None, 3, None, None, None) == L.RBH_ERR_INVALID
    assert L.lib().rbh_light_tree_build_host(None, None, 0, None, None, None) == 0
