"""CPU checks of the C ABI: libhbi.so loads, exports every function that
include/hbi.h declares, and rejects bad arguments without touching a GPU."""
import ctypes as C
import os
import re

import pytest

from hbi_amd import lib as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "hbi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hbi_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = H.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(H.SYMBOLS)


def test_version_and_argument_errors_without_gpu():
    L = H.lib()
    assert L.hbi_version() >= 1
    out = C.c_void_p()
    rc = L.hbi_hmat_build(None, None, C.byref(out))
    assert rc == -1 and b"NULL" in L.hbi_last_error()
    g = H.Geom()
    g.kind, g.n = 9, 10
    o = H.HmatOpts(leaf=15, eta=2.0, eps_aca=1e-4, kmax=1000)
    assert L.hbi_hmat_build(C.byref(g), C.byref(o), C.byref(out)) == -1
    assert b"kind" in L.hbi_last_error()
    assert L.hbi_rk_create(None, None, None, *([None] * 8), C.byref(out)) == -1
    assert L.hbi_hmat_destroy(None) == 0 and L.hbi_rk_destroy(None) == 0


def test_no_cpu_fallback_when_library_missing(monkeypatch, tmp_path):
    monkeypatch.setattr(H, "LIBPATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(H, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        H.lib()
