"""The C-ABI library loads on a host without a GPU, exports every entry point
include/gample_amd.h declares, and rejects bad arguments before touching the
device."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "gample_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s+(gp_\w+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol(gp):
    L = gp.lib()
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) <= set(gp.EXPORTS) | set(names)


def test_argument_validation_without_device(gp):
    L = gp.lib()
    bad = gp.make_config(16, 16, 0.1, 0.1, order=2)
    h = C.c_void_p()
    assert L.gp_ctx_create(C.byref(bad), C.byref(h)) == -1
    assert b"WeightingOrder" in L.gp_last_error()
    bad = gp.make_config(16, 16, 0.1, 0.1, order=3, boris=3)
    assert L.gp_ctx_create(C.byref(bad), C.byref(h)) == -1
    bad = gp.make_config(1, 16, 0.1, 0.1, order=3)
    assert L.gp_ctx_create(C.byref(bad), C.byref(h)) == -1
    assert L.gp_push(None, 0, 1e-12, 0) == -1
    assert L.gp_sync(None) == -1


def test_missing_library_fails_loudly(monkeypatch):
    import importlib
    import gample_amd
    monkeypatch.setattr(gample_amd, "LIB_PATH", "/nonexistent/libgample_amd.so")
    monkeypatch.setattr(gample_amd, "_lib", None)
    with pytest.raises(ImportError):
        gample_amd.lib()
