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


def test_wrappers_refuse_wrong_sizes():
    # the C side reads and writes exactly the counts it expects through raw
    # pointers; the Python mirror refuses anything else before the call
    import numpy as np
    import pytest
    from igr_sensitivities_amd import core
    assert core._sized(np.zeros((2, 3)), 6, "x").shape == (2, 3)
    with pytest.raises(ValueError):
        core._sized(np.zeros(5), 6, "x")
    np.testing.assert_array_equal(core._prm4((0.4, 0.6, 0.2)), [0.4, 0.6, 0.2, 0.0])
    assert core._prm4(None) is None
    with pytest.raises(ValueError):
        core._prm4(np.zeros(5))
