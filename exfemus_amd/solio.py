"""Solution output, restart input and XDMF time files.

The reference writes, per time step, the nodal solution of every system on one level into
sol.<step>.h5 (MGSolDA::print_u, src/MGSolverDA.C:1219-1330), reads it back for a restart
(MGSolDA::read_u, src/MGSolverDA.C:1527-1684), and describes mesh and solution for a viewer
in XDMF files (MGEquationsSystem::print_soln_xmf, src/MGEquationsSystem.C:657-722;
MGMesh::print_xmf_topology / print_xmf_geometry, src/MGMesh.C:438-482; the time file of
src/MGTimeLoop.C).  What is written follows the reference:
  - a Q2 variable: its value at every node of the level, times its reference value;
  - a Q1 variable: interpolated element by element from the element's vertices to all of its
    Q2 nodes (the vertex-to-node table MGGeomEl::Prol), times its reference value;
  - the mesh for viewing: every Q2 element cut into 2^dim linear cells (Quadrilateral /
    Hexahedron topology) over the Q2 nodes, coordinates as X1, X2 (, X3).
HDF5 is not available to this build, so the datasets of one file go, in order, into one raw
little-endian binary file (<base>.bin) with a JSON index of dataset names, offsets and sizes
(<base>.json) in place of the HDF5 dataset directory; the XDMF files address the datasets by
byte offset (Format="Binary", Seek=...), which XDMF readers support.
"""
import json
import os

import numpy as np


def _local_index(mesh, L):
    """Per element of level L, the position (0, 1, 2) of each Q2 node along each axis inside the element
    (box meshes: element edges are axis-aligned)."""
    conn = mesh.conn(L)[0]
    xyz = mesh.coords()
    X = xyz[:, conn]                                  # (dim, nel, n2)
    lo = X.min(axis=2, keepdims=True)
    hi = X.max(axis=2, keepdims=True)
    t = (X - lo) / (hi - lo)                          # 0, 0.5, 1
    idx = np.rint(2.0 * t).astype(np.int64)
    return conn, np.moveaxis(idx, 0, -1)              # (nel, n2), (nel, n2, dim)


def linear_cells(mesh, L):
    """Connectivity of the 2^dim linear cells of each Q2 element of level L (XDMF Quadrilateral / Hexahedron
    node order: counter-clockwise bottom face, then top face)."""
    dim = mesh.dim
    conn, idx = _local_index(mesh, L)
    nel, n2 = conn.shape
    key = (idx * np.array([1, 3, 9][:dim])).sum(axis=2)          # local lattice id 0 .. 3^dim - 1
    lat = np.empty((nel, 3 ** dim), np.int64)
    lat[np.arange(nel)[:, None], key] = conn
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)][:2 ** dim]
    cells = []
    for c in np.ndindex(*([2] * dim)):
        ids = []
        for corner in corners:
            p = [c[d] + corner[d] for d in range(dim)]
            ids.append(sum(p[d] * [1, 3, 9][d] for d in range(dim)))
        cells.append(lat[:, ids])
    return np.stack(cells, axis=1).reshape(nel * 2 ** dim, 2 ** dim).astype(np.int32)


def q1_to_q2_nodes(mesh, L, vertex_values):
    """Nodal field on the level's Q2 nodes from a field given at the Q1 (vertex) nodes: on every element the
    Q1 interpolant of its vertex values at its Q2 nodes (print_u's Prol table; a continuous Q1 field gives the
    same value from every element sharing a node)."""
    dim = mesh.dim
    conn, idx = _local_index(mesh, L)
    out = np.zeros(mesh.level(L)[1])
    # vertex slots of the element node list (the first 2^dim in libMesh order, the same in every element)
    vslots = np.flatnonzero(np.all((idx[0] == 0) | (idx[0] == 2), axis=1))
    s = idx / 2.0                                                    # local position in [0, 1]^dim
    vpos = s[:, vslots, :]                                           # (nel, 2^dim, dim)
    vv = vertex_values[conn[:, vslots]]                              # (nel, 2^dim)
    vals_at = np.zeros(conn.shape)
    for a in range(len(vslots)):
        # Q1 shape function of vertex a at every node: prod_d (s_d if the vertex is at 1 else 1 - s_d)
        w = np.ones(conn.shape)
        for d in range(dim):
            w *= np.where(vpos[:, a, d][:, None] > 0.5, s[:, :, d], 1.0 - s[:, :, d])
        vals_at += w * vv[:, a][:, None]
    out[conn] = vals_at
    return out


def nodal_fields(mesh, L, node_dof, x, nq2, nq1, refvalue=None):
    """The datasets of print_u for one system: a list of nodal arrays over the level's Q2 nodes, one per
    variable (Q2 variables first, then Q1), each scaled by its reference value."""
    N = mesh.N
    nn = mesh.level(L)[1]
    ref = np.ones(nq2 + nq1) if refvalue is None else np.asarray(refvalue, np.float64)
    x = np.asarray(x, np.float64)
    out = []
    for iv in range(nq2):
        d = node_dof[iv * N:iv * N + nn]
        if np.any(d < 0):
            raise ValueError(f"Q2 variable {iv} has no DoF at some level-{L} node")
        out.append(x[d] * ref[iv])
    for iv in range(nq2, nq2 + nq1):
        d = node_dof[iv * N:iv * N + nn]
        vv = np.zeros(nn)
        ok = d >= 0
        vv[ok] = x[d[ok]]
        out.append(q1_to_q2_nodes(mesh, L, vv) * ref[iv])
    return out


def _write(base, named):
    """Raw datasets + JSON index: {name: {"offset": bytes, "count": n, "type": "f8"|"i4", "shape": [...]}}."""
    index = {}
    off = 0
    with open(base + ".bin", "wb") as f:
        for name, arr in named:
            a = np.ascontiguousarray(arr)
            a = a.astype("<f8") if a.dtype.kind == "f" else a.astype("<i4")
            f.write(a.tobytes())
            index[name] = {"offset": off, "count": int(a.size), "type": "f8" if a.dtype.kind == "f" else "i4",
                           "shape": list(a.shape)}
            off += a.nbytes
    with open(base + ".json", "w") as f:
        json.dump(index, f, indent=1)
    return index


def _read(base):
    index = json.load(open(base + ".json"))
    raw = np.fromfile(base + ".bin", dtype=np.uint8)
    out = {}
    for name, e in index.items():
        dt = np.dtype("<f8" if e["type"] == "f8" else "<i4")
        out[name] = raw[e["offset"]:e["offset"] + e["count"] * dt.itemsize].view(dt).reshape(e["shape"])
    return out


def print_mesh(mesh, L, base):
    """Coordinates of the level's nodes (X1, X2, X3) and the linear-cell connectivity (MSH0CONN)."""
    nn = mesh.level(L)[1]
    xyz = mesh.coords()
    named = [(f"X{d + 1}", xyz[d, :nn]) for d in range(mesh.dim)]
    named.append(("MSH0CONN", linear_cells(mesh, L)))
    return _write(base, named)


def print_u(mesh, L, base, systems):
    """Solution file of one time step: systems = [(names, node_dof, x, nq2, nq1, refvalue or None), ...];
    writes every variable of every system (print_u per system, one file per step as in the reference)."""
    named = []
    for names, node_dof, x, nq2, nq1, ref in systems:
        fields = nodal_fields(mesh, L, node_dof, x, nq2, nq1, ref)
        if len(names) != len(fields):
            raise ValueError("one name per variable")
        named += list(zip(names, fields))
    return _write(base, named)


def read_u(mesh, L, base, names, node_dof, nq2, nq1, ndof, refvalue=None):
    """Restart: the system vector (reference DoF order) from a solution file written by print_u on the same
    level: Q2 variables at every node, Q1 variables at their vertex nodes, divided by the reference value."""
    data = _read(base)
    N = mesh.N
    nn = mesh.level(L)[1]
    ref = np.ones(nq2 + nq1) if refvalue is None else np.asarray(refvalue, np.float64)
    x = np.zeros(ndof)
    for iv, name in enumerate(names):
        vals = np.asarray(data[name], np.float64)
        if vals.shape != (nn,):
            raise ValueError(f"dataset {name}: {vals.shape} values, level {L} has {nn} nodes")
        d = node_dof[iv * N:iv * N + nn]
        ok = d >= 0
        x[d[ok]] = vals[ok] / ref[iv]
    return x


def print_xmf(path, time, mesh, L, mesh_base, sol_base, names):
    """XDMF file of one time step: linear cells over the Q2 nodes, nodal attributes (print_soln_xmf)."""
    dim = mesh.dim
    mi = json.load(open(mesh_base + ".json"))
    si = json.load(open(sol_base + ".json"))
    rel = lambda p: os.path.relpath(p, os.path.dirname(os.path.abspath(path)) or ".")
    nn = mesh.level(L)[1]
    ncell, nv = mi["MSH0CONN"]["shape"]
    topo = "Quadrilateral" if dim == 2 else "Hexahedron"

    def item(base, e, dims, ntype, prec):
        return (f'<DataItem Format="Binary" Endian="Little" NumberType="{ntype}" Precision="{prec}" '
                f'Dimensions="{dims}" Seek="{e["offset"]}">{rel(base + ".bin")}</DataItem>\n')
    out = ['<?xml version="1.0" ?>\n<Xdmf Version="2.0">\n<Domain>\n<Grid Name="Mesh" GridType="Uniform">\n',
           f'<Time Value="{time:.17g}" />\n',
           f'<Topology TopologyType="{topo}" NumberOfElements="{ncell}">\n',
           item(mesh_base, mi["MSH0CONN"], f"{ncell} {nv}", "Int", 4), '</Topology>\n',
           f'<Geometry GeometryType="{"X_Y" if dim == 2 else "X_Y_Z"}">\n']
    for d in range(dim):
        out.append(item(mesh_base, mi[f"X{d + 1}"], f"{nn}", "Float", 8))
    out.append('</Geometry>\n')
    for name in names:
        out.append(f'<Attribute Name="{name}" AttributeType="Scalar" Center="Node">\n')
        out.append(item(sol_base, si[name], f"{nn}", "Float", 8))
        out.append('</Attribute>\n')
    out.append('</Grid>\n</Domain>\n</Xdmf>\n')
    with open(path, "w") as f:
        f.write("".join(out))


def print_time_xmf(path, step_files):
    """Temporal collection over the per-step XDMF files (the reference's time file)."""
    out = ['<?xml version="1.0" ?>\n<Xdmf Version="2.0" xmlns:xi="http://www.w3.org/2001/XInclude">\n<Domain>\n',
           '<Grid Name="TimeSeries" GridType="Collection" CollectionType="Temporal">\n']
    base = os.path.dirname(os.path.abspath(path))
    for f in step_files:
        out.append(f'<xi:include href="{os.path.relpath(f, base)}" xpointer="xpointer(//Xdmf/Domain/Grid)" />\n')
    out.append('</Grid>\n</Domain>\n</Xdmf>\n')
    with open(path, "w") as f:
        f.write("".join(out))
