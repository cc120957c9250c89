"""The reference-side binding in the repository's integration guide (the MGSolT subclass a FEMuS maintainer would
add) compiles and links against include/femus_hip.h and the built library, with a stand-in MGSolverT.h declaring
the reference members it uses (tests/cpp/stub).  The code block is taken from the guide itself, so the document
cannot drift from what compiles."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "tests", "cpp", "_build")


def test_binding_compiles_and_links(tmp_path):
    lib = os.path.join(ROOT, "exfemus_amd", "libfemus_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libfemus_hip.so not built")
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = txt[txt.index("## Binding a maintainer would add inside FEMuS"):]
    code = re.search(r"```cpp\n(.*?)```", sec, flags=re.S).group(1)
    assert "class MGSolT_HIP" in code and "fm_mesh_create" in code
    src = tmp_path / "binding.cpp"
    src.write_text(code + "\nint main() {\n  MGMesh m;\n  MGSolT_HIP s(m, 3);\n  (void)s;\n  return 0;\n}\n")
    exe = tmp_path / "binding"
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "tests", "cpp", "stub"),
           "-I", os.path.join(ROOT, "include"), str(src), "-L", os.path.join(ROOT, "exfemus_amd"), "-lfemus_hip",
           "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
