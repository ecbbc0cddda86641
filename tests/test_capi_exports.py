"""The C-ABI library loads and exports every function include/*.h declares."""
import ctypes
import glob
import os
import re

from pigfem_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(pgh?_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_headers_declare_entry_points():
    names = declared_functions()
    assert {"pg_asm_create", "pg_asm_assemble", "pg_asm_assemble_device", "pg_asm_destroy",
            "pg_asm_last_error", "pgh_build_rve", "pgh_get_desc"} <= names


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(_ffi.LIB_PATH)
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_ffi_signatures_cover_header():
    assert declared_functions() <= set(_ffi.SIGNATURES), declared_functions() - set(_ffi.SIGNATURES)


def test_struct_layouts():
    assert ctypes.sizeof(_ffi.PgCutRec) == 136
    assert ctypes.sizeof(_ffi.PgMaterial) == 8 + 36 * 8 + 7 * 8 + 8
