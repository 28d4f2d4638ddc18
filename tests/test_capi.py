"""The C-ABI library loads and exports every symbol include/lapyhip.h declares
(no compute calls: these run without a GPU)."""

import ctypes
import os
import re

import pytest

from lapy_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "lapyhip.h")) as f:
        txt = f.read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(lh_[a-z_0-9]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "lh_assemble" in syms and "lh_eigs" in syms and "lh_pcg" in syms
    assert len(syms) >= 20


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="liblapyhip.so not built")
def test_library_exports_every_declared_symbol():
    h = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(h, s)]
    assert not missing, missing


def test_binding_table_matches_header():
    assert sorted(_lib.SIGNATURES) == declared_symbols()


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="liblapyhip.so not built")
def test_ctypes_signatures_bind():
    L = _lib.lib()
    for name in _lib.SIGNATURES:
        assert getattr(L, name).argtypes is not None


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="liblapyhip.so not built")
def test_context_creation_without_gpu_fails_loudly():
    """Without a HIP device the default context raises (no CPU fallback) and
    does not dead-lock on the loader lock."""
    import threading

    L = _lib.lib()
    if L.lh_device_count() > 0:
        pytest.skip("a GPU is visible")
    box = {}

    def run():
        try:
            _lib.default_context()
        except RuntimeError as e:
            box["err"] = str(e)

    th = threading.Thread(target=run, daemon=True)
    th.start()
    th.join(20)
    assert not th.is_alive(), "default_context() hung"
    assert "no HIP device" in box.get("err", "")


def test_stats_struct_layout():
    # lh_stats: 4 int32, 5 double, 5 int32 (+ 4 bytes of padding to 8)
    assert ctypes.sizeof(_lib.Stats) == 4 * 4 + 5 * 8 + 5 * 4 + 4
    assert ctypes.sizeof(_lib.EigOpts) == 4 * 4 + 8 + 8 + 4 * 2 + 8 + 4 * 2
    assert ctypes.sizeof(_lib.PcgOpts) == 4 * 2 + 8 + 4 * 4


def test_struct_layout_matches_header(tmp_path):
    """Compile the header with the host C compiler and compare every struct's
    size and field offsets with the ctypes mirror."""
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    structs = {"lh_stats": _lib.Stats, "lh_eig_opts": _lib.EigOpts, "lh_pcg_opts": _lib.PcgOpts,
               "lh_tune_entry": _lib.TuneEntry}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "lapyhip.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {(a, b): int(c) for a, b, c in (ln.split() for ln in out if ln)}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
