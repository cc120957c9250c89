"""Restatement of what gencase writes into its files, from the oracle's mesh hierarchy.

TEST INFRASTRUCTURE ONLY.  Imported by tests/ -- never by the product package exfemus_amd.

Restates USER_APPL/gencase/src/MGGenCase.C:
  - print_mesh_h5 (:993-1303) with the counts and maps of printMesh (:582-694): the node offsets OFF_ND /
    OFF_ND1 are cumulative counts of the nodes by first level (lev) and by first vertex level (levp)
    (:596-650), NDxLEV counts the nodes up to each level plus the coarse vertices (:604-615), MAP<l> numbers
    the nodes of the first l levels (:678-694); BC 0 and MAT 1 everywhere (:282-293, no MATBC_INTERFACE);
    EL_NEIG holds global element ids (v_inv_el, :1178-1194); the boundary mesh lists, over each level's
    elements in order, the sides without neighbour with libMesh's build_side node order (:358-380, :555-570);
  - the piecewise-constant pattern blocks compute_mat_qk / _kq / _lk / _kl / _kk (:1855-2307) and transfers
    compute_prol_kk (:2392-2448), compute_res_kk (:2568-2631), NDOF_K = 1.
The node / node blocks and the Q2 / Q1 transfers are the C oracle's om_mat_block / om_prol_block /
om_rest_block.

libMesh 0.9.4 is not vendored: its Quad9 / Hex27 side_nodes_map (the node order of build_side) is written here
from libMesh's published element definitions.  The tests pin it by what the geometry fixes -- each side's node
set is the face of MGGeomEl's _surf_top table, its vertices run round the face with the outward normal, the
edge midpoints follow the vertices edge by edge, the face centre comes last -- which leaves only the choice of
the first vertex of each side unpinned.
"""
import numpy as np

# libMesh 0.9.4 Quad9::side_nodes_map / Hex27::side_nodes_map (side s -> local nodes of build_side(s))
LIBMESH_SIDE_NODES = {
    2: [[0, 1, 4], [1, 2, 5], [2, 3, 6], [3, 0, 7]],
    3: [[0, 3, 2, 1, 11, 10, 9, 8, 20], [0, 1, 5, 4, 8, 13, 16, 12, 21], [1, 2, 6, 5, 9, 14, 17, 13, 22],
        [2, 3, 7, 6, 10, 15, 18, 14, 23], [3, 0, 4, 7, 11, 12, 19, 15, 24], [4, 5, 6, 7, 16, 17, 18, 19, 25]],
}


def boundary(om, L):
    """Boundary sides of level L, [sides][nb2] node ids (printMesh's side loop, MGGenCase.C:358-380, ordered by
    level, :555-570)."""
    conn, nb, _ = om.conn(L)
    table = np.array(LIBMESH_SIDE_NODES[om.dim])
    out = []
    for e in range(conn.shape[0]):
        for s in range(om.nsides):
            if nb[e, s] == -1:
                out.append(conn[e, table[s]])
    return np.array(out, np.int32).reshape(-1, table.shape[1])


def mesh_datasets(om):
    """print_mesh_h5's datasets (name -> array) for one subdomain."""
    dim, nlev, N = om.dim, om.nlev, om.N
    n2, n1 = om.n2, om.n1
    nb2, nb1 = (3, 2) if dim == 2 else (9, 4)
    _, lev, levp = om.node_keys()
    # node offsets by (subdomain, level) and by (subdomain, vertex level): cumulative counts (:596-650)
    off_nd = np.concatenate([[0], np.cumsum([np.sum(lev == l) for l in range(nlev)])]).astype(np.int32)
    off_nd1 = np.concatenate([[0], np.cumsum([np.sum(levp == l) for l in range(nlev)])]).astype(np.int32)
    ndxlev = np.zeros(2 * nlev + 1, np.int32)
    for l in range(nlev):
        ndxlev[l] = np.sum(lev <= l)               # nodes present on level l (:609)
    ndxlev[2 * nlev] = np.sum(levp == 0)           # coarse vertices (:615)
    d = {"/DFLS": np.array([dim, 2, nlev, 1], np.int32),
         "/FEM": np.array([n2] * nlev + [nb2] * nlev, np.int32)}
    xyz = om.coords()
    for k in range(3):
        d[f"/NODES/COORD/X{k + 1}"] = xyz[k] if k < dim else np.zeros(N)
    d["/NODES/COORD/BC"] = np.zeros(N, np.int32)
    d["/NODES/MAP/OFF_ND"] = off_nd
    d["/NODES/MAP/OFF_ND1"] = off_nd1
    d["/NODES/MAP/NDxLEV"] = ndxlev
    for l in range(nlev):                          # g_indexL[l]: the nodes below off_nd[l+1] numbered in order
        m = np.full(N, -1, np.int32)
        m[:off_nd[l + 1]] = np.arange(off_nd[l + 1])
        d[f"/NODES/MAP/MAP{l}"] = m
    m = np.full(N, -1, np.int32)                   # g_indexL[nlev]: the coarse vertices
    m[:off_nd1[1] - off_nd1[0]] = np.arange(off_nd1[1] - off_nd1[0])
    d[f"/NODES/MAP/MAP{nlev}"] = m
    levels = [om.level(l) for l in range(nlev)]
    nel = np.array([lv[0] for lv in levels], np.int32)
    off_el = np.concatenate([[0], np.cumsum(nel)]).astype(np.int32)
    conns, neighs = [], []
    for l in range(nlev):
        c, nb, _ = om.conn(l)
        conns.append(c.reshape(-1))
        neighs.append(np.where(nb >= 0, nb + off_el[l], -1).reshape(-1))
    d["/ELEMS/FEM0/NDOFEM"] = np.array([n2, 0, n1], np.int32)
    d["/ELEMS/FEM0/NExLEV"] = nel
    d["/ELEMS/FEM0/OFF_EL"] = off_el
    d["/ELEMS/FEM0/MSH"] = np.concatenate(conns).astype(np.int32)
    d["/ELEMS/SUB/MAT"] = np.ones(off_el[-1], np.int32)
    d["/ELEMS/FEM0/EL_NEIG"] = np.concatenate(neighs).astype(np.int32)
    for l in range(nlev):
        d[f"/ELEMS/FEM0/MSH{l}"] = conns[l].astype(np.int32)
        d[f"/ELEMS/SUB/MAT{l}"] = np.ones(nel[l], np.int32)
    bds = [boundary(om, l) for l in range(nlev)]
    nbd = np.array([b.shape[0] for b in bds], np.int32)
    d["/ELEMS/FEM1/NDOFEM"] = np.array([nb2, 0, nb1], np.int32)
    d["/ELEMS/FEM1/NExLEV"] = nbd
    d["/ELEMS/FEM1/MSH"] = np.concatenate([b.reshape(-1) for b in bds]).astype(np.int32)
    d["/ELEMS/FEM1/OFF_EL"] = np.concatenate([[0], np.cumsum(nbd)]).astype(np.int32)
    for l in range(nlev):
        d[f"/ELEMS/FEM1/MSH{l}"] = bds[l].reshape(-1).astype(np.int32)
    return d


def _csr(rows):
    ln = np.zeros(len(rows) + 1, np.int32)
    ln[1:] = np.cumsum([len(r) for r in rows])
    pos = np.array([c for r in rows for c in r], np.int32)
    return ln, pos


def piecewise_block(om, L, rkind, ckind):
    """compute_mat_qk (20), _lk (10), _kq (02), _kl (01), _kk (00) of level L: (LEN, POS)."""
    conn, _, _ = om.conn(L)
    ne, nq2, nq1 = om.level(L)
    if rkind == 0:
        if ckind == 0:
            return _csr([[e] for e in range(ne)])
        ncols = nq2 if ckind == 2 else nq1
        # the element's n2 nodes mapped through the column level's node map (-1 dropped), first occurrence
        return _csr([[int(c) for c in conn[e] if c < ncols] for e in range(ne)])
    nr = om.n2 if rkind == 2 else om.n1
    nrows = nq2 if rkind == 2 else nq1
    rows = [[] for _ in range(nrows)]
    for e in range(ne):                            # elements in order; a row node appends its element once
        for k in range(nr):
            r = conn[e, k]
            if e not in rows[r][-1:]:
                rows[r].append(e)
    return _csr(rows)


def piecewise_transfer(om, L, rest):
    """compute_prol_kk (rest False: fine element -> its parent) / compute_res_kk (rest True: coarse element ->
    its children in child order), values 1: (LEN, POS, VAL)."""
    if not rest:
        _, _, par = om.conn(L)
        ln, pos = _csr([[int(p)] for p in par])
    else:
        _, _, par = om.conn(L + 1)
        nch = 2 ** om.dim
        ne = om.level(L)[0]
        kids = [[None] * nch for _ in range(ne)]
        # the oracle appends the children of a parent in child order (libMesh child(i) = embedding i)
        for f, p in enumerate(par):
            kids[p][f - p * nch] = f
        ln, pos = _csr(kids)
    return ln, pos, np.ones(len(pos))
