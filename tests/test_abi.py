"""CPU checks of the C-ABI library: it loads and exports every entry point that
include/femus_hip.h declares (no compute calls without a GPU)."""
import ctypes as C
import os
import re

import numpy as np

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


def test_error_messages_name_the_failing_call():
    """Every failing call leaves its own message for fm_last_error, never an older one (host-only mesh entry
    points with bad arguments: no device needed)."""
    if not os.path.exists(LIB):
        pytest.skip("libfemus_hip.so not built")
    lib = C.CDLL(LIB)
    lib.fm_last_error.restype = C.c_char_p
    lib.fm_mesh_create_box.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                       C.c_int, C.POINTER(C.c_void_p)]
    lib.fm_mesh_info.argtypes = [C.c_void_p, C.c_void_p]
    lib.fm_mesh_level.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    nx0 = (C.c_int * 3)(0, 1, 1)
    lo = (C.c_double * 3)(0, 0, 0)
    hi = (C.c_double * 3)(1, 1, 1)
    h = C.c_void_p()
    assert lib.fm_mesh_create_box(2, nx0, lo, hi, 3, C.byref(h)) == -1
    assert b"fm_mesh_create_box" in lib.fm_last_error()
    assert lib.fm_mesh_info(None, None) == -1
    assert lib.fm_last_error() == b"fm_mesh_info: invalid argument"
    assert lib.fm_mesh_level(None, 0, None, None, None) == -1
    assert lib.fm_last_error() == b"fm_mesh_level: invalid argument"


def test_wrapper_checks_array_lengths():
    """The Python mirror hands the library only arrays of the length it reads (no device needed)."""
    import numpy as np
    from exfemus_amd.femus import _vec
    v = _vec(np.arange(6, dtype=np.int32).reshape(2, 3), 6, "x")
    assert v.dtype == np.float64 and v.flags["C_CONTIGUOUS"] and v.shape == (6,)
    v = _vec(np.arange(12.0).reshape(3, 4).T, 12, "x")     # non-contiguous input is copied
    assert v.flags["C_CONTIGUOUS"] and np.array_equal(v, np.arange(12.0).reshape(3, 4).T.reshape(-1))
    for bad in (np.zeros(5), np.zeros(7), []):
        with pytest.raises(ValueError):
            _vec(bad, 6, "x")


def test_cpp_host_layer_checks_vector_sizes(tmp_path):
    """femus_hip::sized (used by every setter of the C++ host layer before it passes a pointer) reports a
    vector of the wrong length through the error handler as FM_EINVAL and lets the caller skip the call."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    src = tmp_path / "sized.cpp"
    src.write_text(r'''
#include "femus_hip.hpp"
#include <cstdio>
int main() {
  int code = 0; std::string where, msg;
  femus_hip::set_error_handler([&](int c, const std::string &w, const std::string &m) { code = c; where = w; msg = m; });
  const bool ok = femus_hip::sized(std::vector<double>(4), 4, "a");
  const bool bad = femus_hip::sized(std::vector<double>(3), 4, "set_x");
  std::printf("%d %d %d %s|%s\n", ok, bad, code, where.c_str(), msg.c_str());
  return 0;
}
''')
    exe = tmp_path / "sized"
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=30).stdout.strip()
    assert out == "1 0 -1 set_x|expected 4 values, got 3", out


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


def test_caller_mesh_through_the_host_abi():
    """fm_mesh_create takes a caller's hierarchy in MGMesh's layout without a device: a box hierarchy read back
    through fm_mesh_conn / fm_mesh_coords, its elements renumbered, goes back in and comes out unchanged; an
    inconsistent hierarchy (a vertex outside the level's Q1 nodes, a missing child) is refused with a message."""
    if not os.path.exists(LIB):
        pytest.skip("libfemus_hip.so not built")
    from exfemus_amd import femus as F
    box = F.Mesh(3, [2, 1, 2], [0, 0, 0], [1, 1, 1], 2)
    rng = np.random.default_rng(0)
    counts = [box.level(L) for L in range(2)]
    off, emap, eneigh, par, chd, perms = [0], [], [], [], [], []
    for L in range(2):
        conn, neigh, parent = box.conn(L)
        ne = conn.shape[0]
        p = rng.permutation(ne); inv = np.argsort(p); perms.append(p)
        emap.append(conn[inv]); nb = neigh[inv]
        eneigh.append(np.where(nb >= 0, p[np.maximum(nb, 0)] + off[-1], -1))
        if L:
            par.append(perms[0][parent[inv]]); chd.append(inv - parent[inv] * 8)
        else:
            par.append(np.full(ne, -1)); chd.append(np.zeros(ne, int))
        off.append(off[-1] + ne)
    args = (3, box.coords(), off, np.concatenate(emap), np.concatenate(eneigh), [c[1] for c in counts],
            [c[2] for c in counts], np.concatenate(par), np.concatenate(chd))
    m = F.Mesh.from_arrays(*args)
    assert (m.dim, m.nlev, m.N) == (3, 2, box.N)
    np.testing.assert_array_equal(m.coords(), box.coords())
    for L in range(2):
        assert m.level(L) == box.level(L)
        c, nb, pa = m.conn(L)
        np.testing.assert_array_equal(c, emap[L])
        np.testing.assert_array_equal(nb, np.where(eneigh[L] >= 0, eneigh[L] - off[L], -1))
        if L:
            np.testing.assert_array_equal(pa, par[L])
    # without parent / child arrays (the reference's mesh file has none) they are inferred from the nested nodes
    m2 = F.Mesh.from_arrays(*args[:7])
    np.testing.assert_array_equal(m2.conn(1)[2], par[1])
    bad = list(args)
    bad_map = np.concatenate(emap).copy()
    bad_map[0] = counts[0][1] - 1          # level-0 vertex beyond the level's Q1 nodes
    bad[3] = bad_map
    with pytest.raises(F.FemusError) as e:
        F.Mesh.from_arrays(*bad)
    assert "vertex" in str(e.value) or "no element" in str(e.value)
    bad = list(args)
    ch = np.concatenate(chd).copy()
    sib = np.where(par[1] == par[1][0])[0]  # level-1 elements of one parent
    ch[off[1] + sib[1]] = ch[off[1] + sib[0]]   # two of them claim the same child slot
    bad[8] = ch
    with pytest.raises(F.FemusError) as e:
        F.Mesh.from_arrays(*bad)
    assert "same child" in str(e.value)
