"""The C-ABI library (no GPU calls): it loads, exports every entry point include/m2c_hip.h
declares with C linkage, and the Python front end fails loudly when the library is missing."""
import ctypes
import os
import re

import pytest

import m2c_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "m2c_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(m2c_[a-z0-9_]+)\s*\(", txt)))


def library_path():
    if not os.path.exists(M.LIB_PATH):
        from m2c_amd import build
        build.build()
    return M.LIB_PATH


def test_header_declares_the_entry_points():
    names = declared_symbols()
    for n in ["m2c_ctx_create", "m2c_ctx_destroy", "m2c_upload_state", "m2c_download_state",
              "m2c_compute_residual", "m2c_time_step_size", "m2c_advance_dt", "m2c_run",
              "m2c_get_status", "m2c_flux_batch", "m2c_riemann_batch", "m2c_upload_levelset",
              "m2c_download_levelset", "m2c_varfcn_batch", "m2c_interface_flux_batch", "m2c_set_kchunk",
              "m2c_set_embedded_faces"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(library_path())
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    # C linkage: no mangled names needed
    lib.m2c_version.restype = ctypes.c_char_p
    assert lib.m2c_version().startswith(b"m2c_hip")


def test_front_end_fails_loudly_without_library(monkeypatch):
    monkeypatch.setattr(M, "_lib", None)
    monkeypatch.setattr(M, "LIB_PATH", os.path.join(ROOT, "m2c_amd", "no_such_library.so"))
    with pytest.raises(ImportError):
        M.lib()


def test_cpp_host_layer_compiles():
    import subprocess
    inc = os.path.join(ROOT, "m2c_amd", "host")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", inc, "-x", "c++", "-"],
                       input='#include "m2c_host.hpp"\nint main() { return 0; }\n', capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
