"""The C-ABI library loads and exports every entry point include/mhfem.h
declares (no compute: runs without a GPU)."""
import ctypes
import os
import re

from mhfem_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(mhfem_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    assert declared_functions() <= set(capi.SIGNATURES), declared_functions() - set(capi.SIGNATURES)
    capi.lib()  # every binding resolves


def test_create_rejects_bad_arguments():
    # argument checks run before any device call
    L = capi.lib()
    assert not L.mhfem_create(None, None)
    assert b"null" in L.mhfem_last_error(None)
