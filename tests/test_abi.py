"""CPU checks of the C-ABI library: it loads and exports every entry point that
include/femus_hip.h declares (no compute calls without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "femus_hip.h")
LIB = os.path.join(ROOT, "exfemus_amd", "libfemus_hip.so")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int64_t|int|void|char|double)\s*\**\s*(fm_[a-z0-9_]+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared()
    assert "fm_mg_solve" in names and "fm_heat_genmatrhs" in names and len(names) >= 30


def test_library_exports_all_declared_symbols():
    if not os.path.exists(LIB):
        pytest.skip("libfemus_hip.so not built")
    lib = C.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_mirror_lists_every_export():
    from exfemus_amd import femus as F
    assert sorted(F.EXPORTS) == declared()


def test_error_path_without_device():
    """fm_ctx_create reports a status code (no abort) when it cannot use a device index."""
    if not os.path.exists(LIB):
        pytest.skip("libfemus_hip.so not built")
    lib = C.CDLL(LIB)
    lib.fm_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
    h = C.c_void_p()
    rc = lib.fm_ctx_create(-1, C.byref(h))
    assert rc < 0


def test_cpp_host_layer_compiles():
    """include/femus_hip.hpp (MGSolDA / MGSolT / MGSolNS over the C ABI) compiles with the host compiler,
    together with the application-style driver that the GPU tests run."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([gxx, "-fsyntax-only", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(root, "include"),
                        os.path.join(root, "tests", "cpp", "mgsol_driver.cpp")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
