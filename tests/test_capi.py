"""CPU-side checks of the C ABI boundary: the shared library loads and
exports every function include/lesgo_amd.h declares (no GPU calls)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "lesgo_amd.h")).read()
    return sorted(set(re.findall(r"\b(lg_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    import ctypes
    from lesgo_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    assert set(declared()) == set(_lib.SIGNATURES), "ctypes signatures out of sync with header"


def test_field_ids_match_header():
    from lesgo_amd import _lib
    src = open(os.path.join(ROOT, "include", "lesgo_amd.h")).read()
    body = src[src.index("LG_U = 0"):src.index("LG_NFIELDS")]
    names = re.findall(r"LG_([A-Z0-9_]+)", body)
    assert len(names) == len(_lib.FIELDS)
    for h, f in zip(names, _lib.FIELDS):
        assert h.replace("_", "") == f.upper().replace("_", ""), (h, f)
