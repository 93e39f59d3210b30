"""CPU checks of the C ABI: libmdx.so loads and exports every entry point
that include/*.h declares (no device calls are made here)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(mdx_\w+)\s*\(", text, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    from lammps_amd import mdx
    lib = ctypes.CDLL(mdx.LIB_PATH)
    names = declared_symbols()
    assert "mdx_compute" in names and "mdx_run" in names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(mdx.exported_symbols()) <= names


def test_version_and_error_string_without_gpu():
    from lammps_amd import mdx
    L = mdx.lib()
    assert L.mdx_version() == 1
    assert isinstance(L.mdx_last_error(), bytes)
