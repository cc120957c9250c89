"""Solution output / restart input / XDMF files (exfemus_amd.solio), on the CPU with the oracle's DoF maps
(bit-identical to the library's).  Pinned by: the Q1 interpolation reproduces multilinear fields exactly at
every Q2 node; every linear cell has positive volume and the cells tile the box; print -> read round trip
restores the system vector bit for bit (Q2 everywhere, Q1 at vertices); the XDMF files parse and address the
right bytes."""
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from exfemus_amd import femus, solio


@pytest.mark.parametrize("dim,nx,nlev", [(2, [2, 3], 2), (3, [1, 2, 1], 2)])
def test_linear_cells_tile_the_box(dim, nx, nlev):
    m = femus.Mesh(dim, nx, [0] * dim, [1.0, 2.0, 0.5][:dim], nlev)
    L = nlev - 1
    cells = solio.linear_cells(m, L)
    nel = m.level(L)[0]
    assert cells.shape == (nel * 2 ** dim, 2 ** dim)
    X = m.coords()[:, :m.level(L)[1]]
    P = X[:, cells]                                          # (dim, ncell, nv)
    if dim == 2:
        # shoelace area of counter-clockwise quads: positive, summing to the box area
        x, y = P[0], P[1]
        area = 0.5 * np.sum(x * np.roll(y, -1, axis=1) - np.roll(x, -1, axis=1) * y, axis=1)
        assert np.all(area > 0)
        np.testing.assert_allclose(area.sum(), 2.0, rtol=1e-13)
    else:
        # axis-aligned hexes: corner 0 is the min corner, corner 6 the max corner, volumes sum to the box
        ext = P[:, :, 6] - P[:, :, 0]
        assert np.all(ext > 0)
        np.testing.assert_allclose(np.prod(ext, axis=0).sum(), 1.0, rtol=1e-13)
        # bottom face counter-clockwise, top face above it
        np.testing.assert_allclose(P[2, :, 4:], P[2, :, 6:7].repeat(4, axis=1))


@pytest.mark.parametrize("dim", [2, 3])
def test_q1_interpolation_exact_for_multilinear(dim):
    m = femus.Mesh(dim, [2] * dim, [0] * dim, [1] * dim, 2)
    L = 1
    nn = m.level(L)[1]
    X = m.coords()[:, :nn]
    f = 1.5 + 0.7 * X[0] - 0.3 * X[1] + 0.2 * X[0] * X[1]
    if dim == 3:
        f = f + 0.9 * X[2] - 0.4 * X[0] * X[1] * X[2]
    got = solio.q1_to_q2_nodes(m, L, f)                      # only the vertex values of f are read
    np.testing.assert_allclose(got, f, rtol=0, atol=1e-14)


def test_print_read_round_trip_and_xmf(oracle, tmp_path):
    dim, nx, nlev = 2, [2, 2], 2
    m = femus.Mesh(dim, nx, [0, 0], [1, 1], nlev)
    om = oracle.Mesh(dim, nx, [0, 0], [1, 1], nlev)
    L = nlev - 1
    nd, n = oracle.init_dof(om, L, dim, 1)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n)
    names = ["u", "v", "p"]
    ref = [2.0, 2.0, 0.5]
    base = str(tmp_path / "sol.0001")
    mbase = str(tmp_path / "mesh")
    solio.print_mesh(m, L, mbase)
    solio.print_u(m, L, base, [(names, nd, x, dim, 1, ref)])
    y = solio.read_u(m, L, base, names, nd, dim, 1, n, ref)
    np.testing.assert_array_equal(y, x)                      # Q2 at all nodes, Q1 at vertices, exact scaling
    # the Q1 dataset at the vertices is the pressure itself, elsewhere its interpolant
    data = solio._read(base)
    N = m.N
    pd = nd[2 * N:2 * N + m.level(L)[1]]
    np.testing.assert_allclose(data["p"][pd >= 0], x[pd[pd >= 0]] * 0.5, rtol=0, atol=0)
    xmf = str(tmp_path / "sol.0001.xmf")
    solio.print_xmf(xmf, 0.25, m, L, mbase, base, names)
    solio.print_time_xmf(str(tmp_path / "time.xmf"), [xmf])
    root = ET.parse(xmf).getroot()
    items = root.findall(".//DataItem")
    assert len(items) == 1 + dim + len(names)
    for it in root.findall(".//Attribute"):
        di = it.find("DataItem")
        raw = np.fromfile(os.path.join(tmp_path, di.text.strip()), dtype="<f8", offset=int(di.get("Seek")),
                          count=int(di.get("Dimensions")))
        np.testing.assert_array_equal(raw, data[it.get("Name")])
    ET.parse(str(tmp_path / "time.xmf"))
