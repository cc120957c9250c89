"""The library's host-side mesh unit against the oracle, bit for bit, without a device.

The mesh hierarchy goes through the library's own ABI: the fm_mesh_* entry points are host-only.  The DoF
maps, the sparsity pattern of A and both transfer operators of every level come from tests/cpp/maps_dump,
which links the library's exfemus_amd/csrc/gencase.cpp and calls it as fm_system_init does.  Compared with
oracle/ (restated from USER_APPL/gencase/src/MGGenCase.C:63-758, 1370-2566 and src/MGSolverDA.C:880-948,
1785-2444):
  - coordinates, connectivity, neighbours and parents;
  - node_dof and the DoF counts;
  - the row pointers and columns of A;
  - the rows, columns and values of R (with the top-level Dirichlet mask) and of P.
Heat (one Q2 variable) and Navier-Stokes (dim Q2 + one Q1) on the cases of the GPU parity tests.  The GPU
tests check the same maps after the upload, and the BC masks, through a device context."""
import functools
import os
import subprocess

import numpy as np
import pytest

from exfemus_amd import configs

HERE = os.path.dirname(os.path.abspath(__file__))
DUMP = os.path.join(HERE, "cpp", "_build", "maps_dump")

CASES = [
    # name, dim, nx0, nlev, hi, nq2, nq1, bc
    ("heat2d", 2, [2, 2], 3, [1, 1], 1, 0, configs.heat_dirichlet_all),
    ("heat3d", 3, [2, 2, 2], 3, [1, 1, 1], 1, 0, configs.heat_dirichlet_all),
    ("heat3d_ragged", 3, [1, 2, 3], 2, [1, 1, 1], 1, 0, configs.heat_dirichlet_all),
    ("cavity2d", 2, [2, 2], 3, [1, 1], 2, 1, functools.partial(configs.cavity_bc, dim=2)),
    ("cavity3d", 3, [1, 1, 1], 3, [1, 1, 1], 3, 1, functools.partial(configs.cavity_bc, dim=3)),
    ("channel3d", 3, [4, 2, 2], 2, [4, 1, 1], 3, 1, configs.channel_bc),
]


def _read(path):
    raw = open(path, "rb").read()
    pos = 0

    def take(dtype):
        nonlocal pos
        n = int(np.frombuffer(raw, np.int64, 1, pos)[0])
        pos += 8
        a = np.frombuffer(raw, dtype, n, pos)
        pos += n * np.dtype(dtype).itemsize
        return a
    return take, lambda: pos == len(raw)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_host_maps_bitexact(case, oracle, tmp_path):
    from exfemus_amd import femus as F
    name, dim, nx, nlev, hi, nq2, nq1, bc = case
    if not os.path.exists(DUMP):
        pytest.fail("maps_dump not built: run __graft_entry__.build()")
    lo = [0.0] * dim
    pm = F.Mesh(dim, nx, lo, hi, nlev)
    om = oracle.Mesh(dim, nx, lo, hi, nlev)
    # mesh through the library's ABI (host-only entry points)
    assert pm.N == om.N
    np.testing.assert_array_equal(pm.coords(), om.coords())
    for L in range(nlev):
        assert pm.level(L) == om.level(L)
        for a, b in zip(pm.conn(L), om.conn(L)):
            np.testing.assert_array_equal(a, b)
    # DoF maps, A's pattern, R and P from the library's mesh unit
    top = nlev - 1
    dofs = [oracle.init_dof(om, L, nq2, nq1) for L in range(nlev)]
    b0, _ = oracle.genbc(om, nq2, nq1, dofs[top][0], bc)
    bfile, ofile = tmp_path / "bc0.bin", tmp_path / "maps.bin"
    np.ascontiguousarray(b0, np.int32).tofile(bfile)
    h = (list(hi) + [1.0, 1.0])[:3]
    nz = nx[2] if dim == 3 else 1
    r = subprocess.run([DUMP, str(dim), str(nx[0]), str(nx[1]), str(nz), str(nlev), str(nq2), str(nq1),
                        *map(str, h), str(bfile), str(ofile)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    take, done = _read(ofile)
    for L in range(nlev):
        od, ndof = dofs[L]
        np.testing.assert_array_equal(take(np.int32), od)
        assert int(take(np.int64)[0]) == ndof
        rp, ci = oracle.system_pattern(om, L, nq2, nq1, od)
        np.testing.assert_array_equal(take(np.int64), rp)
        np.testing.assert_array_equal(take(np.int32), ci)
        if L < top:
            rrp, rci, rv = oracle.system_rest(om, L, nq2, nq1, od, dofs[L + 1][0], dofs[top][0], b0)
            np.testing.assert_array_equal(take(np.int64), rrp)
            np.testing.assert_array_equal(take(np.int32), rci)
            np.testing.assert_array_equal(take(np.float64), rv)
        if L > 0:
            prp, pci, pv = oracle.system_prol(om, L, nq2, nq1, od, dofs[L - 1][0])
            np.testing.assert_array_equal(take(np.int64), prp)
            np.testing.assert_array_equal(take(np.int32), pci)
            np.testing.assert_array_equal(take(np.float64), pv)
    assert done()
