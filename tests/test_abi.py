"""The C ABI library: it loads, exports every entry point include/euchip.h
declares, and the product path refuses to run without it.  CPU only: no call
here computes anything on a device."""
import ctypes as C
import os
import re

import pytest

from euc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "euchip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(euchip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ["euchip_ctx_create", "euchip_render", "euchip_render_async", "euchip_buffer_create",
                 "euchip_mesh_create", "euchip_buffer_clear", "euchip_synchronize"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_status_strings():
    L = _lib.lib()
    for code, name in _lib.STATUS.items():
        assert L.euchip_status_string(code).decode()
    assert L.euchip_status_string(1).decode() == "size mismatch"


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libeuchip.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(FileNotFoundError):
        _lib.lib()


def test_null_arguments_are_rejected():
    """Argument validation happens before any device work."""
    L = _lib.lib()
    assert L.euchip_ctx_create(0, None) == 3  # INVALID_ARG
    assert L.euchip_ctx_destroy(None) == 3
    assert L.euchip_render(None, None, None, None, None, None) == 3
    assert L.euchip_buffer_destroy(None) == 3
    assert L.euchip_mesh_destroy(None) == 3


def test_cpp_mirror_header_compiles(tmp_path):
    """include/euchip.hpp (the C++ mirror of euc's Pipeline surface) is self-contained C++17."""
    import subprocess
    src = tmp_path / "t.cpp"
    src.write_text('#include "euchip.hpp"\nint main() { return euc::DepthMode::less_write().uses_depth() ? 0 : 1; }\n')
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", inc, str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
