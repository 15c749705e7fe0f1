"""The C-ABI library builds, loads without a GPU, and exports every entry
point that include/igr.h declares (no compute calls here)."""
import ctypes as C
import os
import re

from igr_sensitivities_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names += re.findall(r"^\s*(?:int|const char\*)\s+(igr_\w+)\s*\(", txt, flags=re.M)
    return names


def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    names = header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n


def test_python_prototypes_cover_header():
    assert set(header_functions()) == set(_lib.PROTOTYPES)


def test_loader_sets_prototypes():
    L = _lib.lib()
    assert L.igr_create.restype is C.c_int
