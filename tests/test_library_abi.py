"""The C-ABI library loads and exports every entry point include/godar_gpu.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from godar_amd import _lib
from godar_amd.params import GodarParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "godar_gpu.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(godar_\w+)\s*\(", txt, re.M)))


def test_header_declares_bound_functions():
    names = declared()
    assert set(names) == set(_lib.EXPORTS), set(names) ^ set(_lib.EXPORTS)


def test_library_exports_every_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libgodar_amd.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared():
        assert hasattr(L, name), name


def test_params_struct_matches_header():
    """ctypes GodarParams mirrors struct godar_params field by field."""
    txt = open(HEADER).read()
    body = txt[txt.index("typedef struct godar_params {"):txt.index("} godar_params;")]
    fields = []
    for decl in re.findall(r"(double|int32_t|uint64_t)\s+([^;]+);", body):
        fields += [(decl[0], nm.strip()) for nm in decl[1].split(",")]
    cmap = {"double": ctypes.c_double, "int32_t": ctypes.c_int32, "uint64_t": ctypes.c_uint64}
    assert [(n, cmap[t]) for t, n in fields] == list(GodarParams._fields_)


def test_params_default_matches_c():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libgodar_amd.so not built")
    L = _lib.load()
    p = GodarParams()
    L.godar_params_default(ctypes.byref(p))
    q = GodarParams()
    for name, _ in GodarParams._fields_:
        assert getattr(p, name) == getattr(q, name), name


def test_library_fails_loudly_when_missing(tmp_path):
    with pytest.raises(RuntimeError):
        _lib._lib = None
        try:
            _lib.load(str(tmp_path / "missing.so"))
        finally:
            _lib._lib = None
