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


ABI_C = os.path.join(ROOT, "tests", "abi_c")


def _abi_binary():
    import subprocess
    subprocess.run(["make", "-s", "-C", ABI_C], check=True)
    return os.path.join(ABI_C, "_build", "abi_check")


def test_c_caller_host_entry_points():
    """A C program compiled against include/ratbite_hip.h alone: record and
    guide sizes as documented, host-only light-tree builds write no more than
    the documented capacities."""
    import subprocess
    r = subprocess.run([_abi_binary(), "host"], capture_output=True, text=True)
    assert r.returncode == 0 and "abi_check: ok" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
def test_c_caller_every_entry_point():
    """The same C program calls every entry point of the header on the GPU
    with buffers of exactly the documented sizes plus a guard band, and checks
    the bands after every call that writes into caller memory."""
    import subprocess
    print("abi_check gpu", flush=True)
    r = subprocess.run([_abi_binary(), "gpu"], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0 and "abi_check: ok" in r.stdout, r.stdout + r.stderr
