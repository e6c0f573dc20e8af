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
    g.kind, g.n = 14, 10
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


def test_ctypes_structs_match_the_c_layout(tmp_path):
    """Size and every field offset of the ctypes mirrors equal the C
    structs of include/hbi.h (a short mirror would let the library write past
    the Python-owned buffer)."""
    import subprocess
    structs = {"hbi_geom": H.Geom, "hbi_hmat_opts": H.HmatOpts, "hbi_hmat_stats": H.HmatStats,
               "hbi_rk_opts": H.RkOpts, "hbi_rk_info": H.RkInfo, "hbi_rk_out_opts": H.RkOutOpts,
               "hbi_rk_event": H.RkEvent}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "hbi.h"', 'int main(void) {']
    for cname, py in structs.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines += ['return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for ln in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        a, b, c = ln.split()
        got[(a, b)] = int(c)
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
