"""Preprocessor output: the mesh file and the Matrix / Prol / Rest files that exfemus_amd.gencase_io writes
(gencase's print_mesh_h5 and print_op_h5 datasets, USER_APPL/gencase/src/MGGenCase.C:993-1364) against the
oracle's restatement, bit-exact, on the 2D 32x32 Quad9 three-level hierarchy of the cavity and on the 3D 64^3
Hex27 five-level hierarchy of the heat benchmark.  The restatement itself is pinned first by what the mesh
fixes: counts from the connectivity, neighbours by brute-force face matching, the boundary sides by geometry
(face node sets, outward vertex cycle, edge midpoints, face centre), the piecewise blocks by brute force.
Then the files are read back (MGMesh::read_c's view) into a hierarchy equal to the one written, also for a
renumbered caller hierarchy."""
import os

import numpy as np
import pytest

from exfemus_amd import gencase_io
from exfemus_amd.solio import _read
from oracle import gencase_files as G

C1 = (2, [8, 8], 3)        # 32x32 Quad9 on the top level
C2 = (3, [4, 4, 4], 5)     # 64^3 Hex27 on the top level

# libMesh side numbering: side -> (axis, direction) of the face in the element's reference frame
SIDE_DIR = {2: [(1, -1), (0, 1), (1, 1), (0, -1)], 3: [(2, -1), (1, -1), (0, 1), (1, 1), (0, -1), (2, 1)]}


@pytest.fixture(scope="module")
def c1(oracle, femus):
    dim, nx, nlev = C1
    return oracle.Mesh(dim, nx, [0.0] * dim, [1.0] * dim, nlev), femus.Mesh(dim, nx, [0.0] * dim, [1.0] * dim, nlev)


def _faces(xyz, conn, dim):
    """Per element and libMesh side, the local node slots on that face (box elements: extreme coordinate)."""
    X = xyz[:, conn]                                           # (dim, ne, n2)
    out = []
    for axis, sgn in SIDE_DIR[dim]:
        ext = X[axis].max(axis=1) if sgn > 0 else X[axis].min(axis=1)
        out.append(np.isclose(X[axis], ext[:, None]))          # (ne, n2)
    return out


# ---------------------------------------------------------------------------------------------------------------
# the restatement, pinned
# ---------------------------------------------------------------------------------------------------------------
def test_oracle_mesh_datasets_pinned(oracle, c1):
    om, _ = c1
    d = G.mesh_datasets(om)
    dim, nlev = om.dim, om.nlev
    xyz = om.coords()
    for l in range(nlev):
        conn, nb, par = om.conn(l)
        used = np.unique(conn)
        # the first levels' nodes are the first ids: NDxLEV / OFF_ND count them, MAP<l> numbers exactly them
        assert d["/NODES/MAP/NDxLEV"][l] == len(used) == used.max() + 1 == d["/NODES/MAP/OFF_ND"][l + 1]
        assert np.array_equal(np.flatnonzero(d[f"/NODES/MAP/MAP{l}"] >= 0), used)
        # OFF_ND1 counts the level's vertices
        assert d["/NODES/MAP/OFF_ND1"][l + 1] == len(np.unique(conn[:, :om.n1]))
        # neighbours: the element sharing the face's node set, as a global id; -1 on the domain boundary
        faces = _faces(xyz, conn, dim)
        key = {}
        for e in range(conn.shape[0]):
            for s in range(om.nsides):
                key.setdefault(frozenset(conn[e][faces[s][e]]), []).append(e)
        off = d["/ELEMS/FEM0/OFF_EL"][l]
        neig = d["/ELEMS/FEM0/EL_NEIG"][off * om.nsides:(off + conn.shape[0]) * om.nsides].reshape(-1, om.nsides)
        for e in range(conn.shape[0]):
            for s in range(om.nsides):
                other = [f for f in key[frozenset(conn[e][faces[s][e]])] if f != e]
                assert neig[e, s] == (other[0] + off if other else -1)
        assert np.array_equal(d[f"/ELEMS/FEM0/MSH{l}"], conn.reshape(-1))
    assert d["/NODES/MAP/NDxLEV"][2 * nlev] == d["/NODES/MAP/OFF_ND1"][1]
    assert np.array_equal(np.flatnonzero(d[f"/NODES/MAP/MAP{nlev}"] >= 0), np.unique(om.conn(0)[0][:, :om.n1]))


@pytest.mark.parametrize("case", [C1, (3, [2, 3, 2], 2)])
def test_oracle_boundary_sides_pinned(oracle, case):
    """Every boundary side: its nodes are the element's nodes on that face, the first 2^(dim-1) are the face's
    corners in a cycle whose normal points out of the element, the next ones are the midpoints of consecutive
    corners, a 3D side ends with the face centre; the count is the box's boundary face count."""
    dim, nx, nlev = case
    om = oracle.Mesh(dim, nx, [0.0] * dim, [1.0] * dim, nlev)
    xyz = om.coords()
    nb1 = 2 if dim == 2 else 4
    for l in range(nlev):
        conn, nb, _ = om.conn(l)
        bd = G.boundary(om, l)
        k = 2 ** l
        faces_expected = 2 * (nx[0] + nx[1]) * k if dim == 2 else \
            2 * (nx[0] * nx[1] + nx[1] * nx[2] + nx[0] * nx[2]) * k * k
        assert bd.shape[0] == faces_expected
        faces = _faces(xyz, conn, dim)
        q = 0
        for e in range(conn.shape[0]):
            centre = xyz[:, conn[e]].mean(axis=1)
            for s in range(om.nsides):
                if nb[e, s] != -1:
                    continue
                side = bd[q]; q += 1
                assert set(side) == set(conn[e][faces[s][e]])
                P = xyz[:, side]
                V = P[:, :nb1]
                for i in range(1 if dim == 2 else nb1):           # midpoints of consecutive corners
                    assert np.allclose(P[:, nb1 + i], 0.5 * (V[:, i] + V[:, (i + 1) % nb1]))
                fc = V.mean(axis=1)
                out = fc - centre                                  # outward direction
                if dim == 2:
                    t = V[:, 1] - V[:, 0]
                    assert np.dot([t[1], -t[0]], out) > 0          # the corner order keeps the element on the left
                else:
                    assert np.allclose(P[:, 8], fc)
                    n = np.cross(V[:, 1] - V[:, 0], V[:, 2] - V[:, 0])
                    assert np.dot(n, out) > 0
                    # a cycle, not a crossing: opposite corners are the face diagonals
                    assert np.allclose(0.5 * (V[:, 0] + V[:, 2]), fc) and np.allclose(0.5 * (V[:, 1] + V[:, 3]), fc)
        assert q == bd.shape[0]


def test_oracle_piecewise_blocks_pinned(oracle, c1):
    om, _ = c1
    for L in range(om.nlev):
        conn, _, _ = om.conn(L)
        ne, nq2, nq1 = om.level(L)
        for r, c in [(2, 0), (1, 0), (0, 2), (0, 1), (0, 0)]:
            ln, pos = G.piecewise_block(om, L, r, c)
            rows = [pos[ln[i]:ln[i + 1]] for i in range(len(ln) - 1)]
            if r == 0 and c == 0:
                assert all(list(rw) == [e] for e, rw in enumerate(rows))
            elif r == 0:
                ncol = conn.shape[1] if c == 2 else om.n1
                # an element row: its nodes (corners only for Q1) in local order
                assert all(np.array_equal(rw, conn[e, :ncol]) for e, rw in enumerate(rows))
            else:
                nr = om.n2 if r == 2 else om.n1
                for i, rw in enumerate(rows):
                    assert np.array_equal(rw, np.flatnonzero((conn[:, :nr] == i).any(axis=1)))
        if L > 0:
            ln, pos, val = G.piecewise_transfer(om, L, rest=False)
            assert np.array_equal(pos, om.conn(L)[2]) and np.all(val == 1.0)
        if L + 1 < om.nlev:
            # the children of a coarse element sit in its corners in child order: child i at corner bits of i
            ln, pos, val = G.piecewise_transfer(om, L, rest=True)
            xyz = om.coords()
            cf = om.conn(L + 1)[0]
            for e in range(om.level(L)[0]):
                X = xyz[:, conn[e]]
                mid = 0.5 * (X.min(axis=1) + X.max(axis=1))
                for i, f in enumerate(pos[ln[e]:ln[e + 1]]):
                    cc = xyz[:, cf[f]].mean(axis=1)
                    assert all((cc[d] > mid[d]) == bool((i >> d) & 1) for d in range(om.dim))


# ---------------------------------------------------------------------------------------------------------------
# the product's files against the restatement
# ---------------------------------------------------------------------------------------------------------------
def _expected_matrix(oracle, om, L, r, c):
    if r and c:
        return oracle.mat_block(om, L, r, c)
    return G.piecewise_block(om, L, r, c)


def _expected_transfer(oracle, om, L, k, rest):
    if k == 0:
        return G.piecewise_transfer(om, L, rest)
    return (oracle.rest_block if rest else oracle.prol_block)(om, L, k)


def _rows(om, L, k):
    ne, nq2, nq1 = om.level(L)
    return {2: nq2, 1: nq1, 0: ne}[k]


def _check_mesh_file(om, base):
    got = _read(base)
    exp = G.mesh_datasets(om)
    assert list(got) == list(exp.keys())
    for name, a in exp.items():
        b = got[name]
        assert b.dtype == (np.float64 if a.dtype.kind == "f" else np.int32), name
        assert np.array_equal(b, a), name


def _check_operator_files(oracle, om, outdir, levels):
    for L in levels:
        got = _read(os.path.join(outdir, f"Matrix{L}"))
        names = []
        for r, c in gencase_io.MATRIX_BLOCKS:
            m = f"{r}{c}"
            ln, pos = _expected_matrix(oracle, om, L, r, c)
            assert np.array_equal(got[f"DIM{m}"], [_rows(om, L, r), _rows(om, L, c)])
            assert np.array_equal(got[f"LEN{m}"], ln), (L, m)
            assert np.array_equal(got[f"POS{m}"], pos), (L, m)
            assert not got[f"OFFLEN{m}"].any() and len(got[f"OFFLEN{m}"]) == len(ln)
            names += [f"DIM{m}", f"POS{m}", f"LEN{m}", f"OFFLEN{m}"]
        assert list(got) == names
        for rest, other, fname in [(False, L - 1, f"Prol{L - 1}_{L}"), (True, L + 1, f"Rest{L + 1}_{L}")]:
            if not 0 <= other < om.nlev:
                continue
            got = _read(os.path.join(outdir, fname))
            for k in gencase_io.TRANSFER_MODES:
                m = f"0{k}"
                ln, pos, val = _expected_transfer(oracle, om, L, k, rest)
                assert np.array_equal(got[f"DIM{m}"], [_rows(om, L, k), _rows(om, other, k)])
                assert np.array_equal(got[f"LEN{m}"], ln), (fname, m)
                assert np.array_equal(got[f"POS{m}"], pos), (fname, m)
                assert np.array_equal(got[f"VAL{m}"], val), (fname, m)
                assert not got[f"OFFLEN{m}"].any()


def test_files_cavity_32x32(oracle, c1, tmp_path):
    om, m = c1
    bases = gencase_io.write_gencase(m, str(tmp_path))
    assert len(bases) == 1 + m.nlev + 2 * (m.nlev - 1)
    _check_mesh_file(om, bases[0])
    _check_operator_files(oracle, om, str(tmp_path), range(m.nlev))


def test_files_heat_64cubed(oracle, femus, tmp_path):
    dim, nx, nlev = C2
    om = oracle.Mesh(dim, nx, [0.0] * dim, [1.0] * dim, nlev)
    m = femus.Mesh(dim, nx, [0.0] * dim, [1.0] * dim, nlev)
    # the mesh file in full; the operator files of the four lower levels (the top level's 22 block alone
    # holds ~1.4e8 entries: it is compared through the transfers into it and the mesh it is built from)
    base = gencase_io.write_gencase(m, str(tmp_path), operators=False)[0]
    _check_mesh_file(om, base)
    for L in range(nlev - 1):
        gencase_io._write(os.path.join(str(tmp_path), f"Matrix{L}"), gencase_io.matrix_datasets(m, L))
        if L > 0:
            gencase_io._write(os.path.join(str(tmp_path), f"Prol{L - 1}_{L}"),
                              gencase_io.transfer_datasets(m, L, rest=False))
        gencase_io._write(os.path.join(str(tmp_path), f"Rest{L + 1}_{L}"), gencase_io.transfer_datasets(m, L, rest=True))
    top = nlev - 1
    gencase_io._write(os.path.join(str(tmp_path), f"Prol{top - 1}_{top}"), gencase_io.transfer_datasets(m, top, False))
    _check_operator_files(oracle, om, str(tmp_path), range(nlev - 1))
    got = _read(os.path.join(str(tmp_path), f"Prol{top - 1}_{top}"))
    for k in (2, 1):
        ln, pos, val = oracle.prol_block(om, top, k)
        assert np.array_equal(got[f"LEN0{k}"], ln) and np.array_equal(got[f"POS0{k}"], pos)
        assert np.array_equal(got[f"VAL0{k}"], val)


# ---------------------------------------------------------------------------------------------------------------
# reading the files back
# ---------------------------------------------------------------------------------------------------------------
def _same_mesh(a, b):
    assert (a.dim, a.nlev, a.N) == (b.dim, b.nlev, b.N)
    np.testing.assert_array_equal(a.coords(), b.coords())
    for L in range(a.nlev):
        assert a.level(L) == b.level(L)
        for x, y in zip(a.conn(L), b.conn(L)):
            np.testing.assert_array_equal(x, y)


def test_read_mesh_roundtrip(c1, tmp_path):
    _, m = c1
    base = gencase_io.write_gencase(m, str(tmp_path), operators=False)[0]
    _same_mesh(gencase_io.read_mesh(base), m)


def test_read_mesh_renumbered_caller_hierarchy(femus, tmp_path):
    """A hierarchy numbered by another generator (elements of every level and each level's node block
    permuted, curved) is written and read back unchanged, parents inferred again from the nested nodes."""
    dim, nx, nlev = 3, [2, 1, 2], 3
    box = femus.Mesh(dim, nx, [0.0] * dim, [1.0] * dim, nlev)
    rng = np.random.default_rng(3)
    counts = [box.level(L) for L in range(nlev)]
    cuts = [0, counts[0][2], counts[0][1]] + [counts[L][1] for L in range(1, nlev)]
    new_id = np.arange(box.N)
    for a, b in zip(cuts[:-1], cuts[1:]):
        new_id[a:b] = a + rng.permutation(b - a)
    xyz = np.zeros((dim, box.N))
    x = box.coords()
    xyz[:, new_id] = np.stack([x[0], x[1] + 0.2 * x[0] ** 2, x[2] + 0.3 * x[0] * x[1]])
    off, emap, eneigh = [0], [], []
    for L in range(nlev):
        conn, nb, _ = box.conn(L)
        p = rng.permutation(conn.shape[0]); inv = np.argsort(p)
        emap.append(new_id[conn[inv]])
        eneigh.append(np.where(nb[inv] >= 0, p[np.maximum(nb[inv], 0)] + off[-1], -1))
        off.append(off[-1] + conn.shape[0])
    m = femus.Mesh.from_arrays(dim, xyz, off, np.concatenate(emap), np.concatenate(eneigh),
                               [c[1] for c in counts], [c[2] for c in counts])
    base = gencase_io.write_gencase(m, str(tmp_path), operators=False)[0]
    d = _read(base)
    np.testing.assert_array_equal(d["/ELEMS/FEM0/EL_NEIG"], np.concatenate(eneigh).reshape(-1))
    _same_mesh(gencase_io.read_mesh(base), m)
