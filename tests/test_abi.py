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


STRUCTS = {"mhfem_mesh_desc": "MeshDesc", "mhfem_model_desc": "ModelDesc", "mhfem_solver_config": "SolverConfig",
           "mhfem_solve_stats": "SolveStats", "mhfem_timers": "Timers", "mhfem_guess_config": "GuessConfig",
           "mhfem_guess_stats": "GuessStats", "mhfem_time_function": "TimeFunction"}


def test_struct_layouts_match_the_python_mirror(tmp_path):
    """Size and every field offset of each struct in include/mhfem.h, as the C
    compiler lays them out, equal the ctypes mirror's."""
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "mhfem.h"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        st = getattr(capi, pyname)
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in st._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, pyname in STRUCTS.items():
        st = getattr(capi, pyname)
        assert got[(cname, "size")] == ctypes.sizeof(st), cname
        for field, _ in st._fields_:
            assert got[(cname, field)] == getattr(st, field).offset, (cname, field)
