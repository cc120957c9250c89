"""The preprocessor's output files: the mesh file and the Matrix / Prol / Rest operator files of gencase.

gencase (USER_APPL/gencase/src/MGGenCase.C) refines a coarse mesh, orders elements and nodes, and writes
  - the mesh file (print_mesh_h5, :993-1303), read back by MGMesh::read_c (src/MGMesh.C:493-899):
      /DFLS [dim, 2 families, levels, subdomains], /FEM [n2 x levels, nb2 x levels],
      /NODES/COORD/X1..X3, /NODES/COORD/BC, /NODES/MAP/OFF_ND, OFF_ND1, NDxLEV, MAP0 .. MAP<levels>,
      /ELEMS/FEM0/{NDOFEM, NExLEV, OFF_EL, MSH, EL_NEIG, MSH<l>}, /ELEMS/SUB/{MAT, MAT<l>},
      /ELEMS/FEM1/{NDOFEM, NExLEV, MSH, OFF_EL, MSH<l>} (the boundary mesh);
  - per level L the matrix pattern file Matrix<L> (compute_matrix, :1370-1551) with the blocks
    22, 21, 12, 11, 20, 02, 10, 01, 00 (row kind, column kind; 2 = Q2 nodes, 1 = Q1 nodes, 0 = elements),
    each as DIM<rc> [rows, columns], POS<rc>, LEN<rc>, OFFLEN<rc> (print_op_h5, :1307-1364);
  - per level pair the prolongation Prol<L-1>_<L> (compute_prol, :1556-1660) and the restriction
    Rest<L+1>_<L> (compute_rest, :1665-1776) with the modes 02, 01, 00, each as DIM, POS, VAL, LEN, OFFLEN.
One subdomain is written (subdomain offsets of length levels + 1, OFFLEN all zero: no entry lies outside the
subdomain), no boundary or material file is read (BC = 0 on every node, MAT = 1 on every element: gencase's
defaults without MATBC_INTERFACE, MGGenCase.C:282-293).  The index sets and values come from the native library
(fm_gencase_matrix_block, fm_gencase_transfer_block, fm_gencase_boundary); this module lays them out.
HDF5 is not available to this build, so each file is one raw little-endian binary file (<base>.bin) with a JSON
index of dataset names, offsets and sizes (<base>.json), the container of exfemus_amd.solio.
"""
import os

import numpy as np

from . import femus
from .solio import _read, _write

MATRIX_BLOCKS = [(2, 2), (2, 1), (1, 2), (1, 1), (2, 0), (0, 2), (1, 0), (0, 1), (0, 0)]
TRANSFER_MODES = [2, 1, 0]


def _mode(a, b=None):
    return f"0{a}" if b is None else f"{a}{b}"


def mesh_datasets(mesh):
    """The datasets of the mesh file in print_mesh_h5's order, as (name, array) pairs."""
    dim, nlev, N = mesh.dim, mesh.nlev, mesh.N
    n2, n1 = mesh.n2, mesh.n1
    nb2, nb1 = (3, 2) if dim == 2 else (9, 4)
    levels = [mesh.level(L) for L in range(nlev)]
    nel = np.array([lv[0] for lv in levels], np.int32)
    nq2 = np.array([lv[1] for lv in levels], np.int32)
    nq1 = np.array([lv[2] for lv in levels], np.int32)
    xyz = mesh.coords()
    conns, neighs = [], []
    off_el = np.concatenate([[0], np.cumsum(nel)]).astype(np.int32)
    for L in range(nlev):
        c, nb, _ = mesh.conn(L)
        conns.append(c.reshape(-1))
        neighs.append(np.where(nb >= 0, nb + off_el[L], -1).reshape(-1))
    bds = [mesh.boundary(L) for L in range(nlev)]
    nbd = np.array([b.shape[0] for b in bds], np.int32)
    out = [("/DFLS", np.array([dim, 2, nlev, 1], np.int32)),
           ("/FEM", np.array([n2] * nlev + [nb2] * nlev, np.int32))]
    for d in range(3):
        out.append((f"/NODES/COORD/X{d + 1}", xyz[d] if d < dim else np.zeros(N)))
    out.append(("/NODES/COORD/BC", np.zeros(N, np.int32)))
    out.append(("/NODES/MAP/OFF_ND", np.concatenate([[0], nq2]).astype(np.int32)))
    out.append(("/NODES/MAP/OFF_ND1", np.concatenate([[0], nq1]).astype(np.int32)))
    out.append(("/NODES/MAP/NDxLEV", np.concatenate([nq2, np.zeros(nlev, np.int32), nq1[:1]]).astype(np.int32)))
    ids = np.arange(N, dtype=np.int32)
    for L in range(nlev):
        out.append((f"/NODES/MAP/MAP{L}", np.where(ids < nq2[L], ids, -1).astype(np.int32)))
    out.append((f"/NODES/MAP/MAP{nlev}", np.where(ids < nq1[0], ids, -1).astype(np.int32)))
    out.append(("/ELEMS/FEM0/NDOFEM", np.array([n2, 0, n1], np.int32)))
    out.append(("/ELEMS/FEM0/NExLEV", nel))
    out.append(("/ELEMS/FEM0/OFF_EL", off_el))
    out.append(("/ELEMS/FEM0/MSH", np.concatenate(conns).astype(np.int32)))
    out.append(("/ELEMS/SUB/MAT", np.ones(off_el[-1], np.int32)))
    out.append(("/ELEMS/FEM0/EL_NEIG", np.concatenate(neighs).astype(np.int32)))
    for L in range(nlev):
        out.append((f"/ELEMS/FEM0/MSH{L}", conns[L].astype(np.int32)))
        out.append((f"/ELEMS/SUB/MAT{L}", np.ones(nel[L], np.int32)))
    out.append(("/ELEMS/FEM1/NDOFEM", np.array([nb2, 0, nb1], np.int32)))
    out.append(("/ELEMS/FEM1/NExLEV", nbd))
    out.append(("/ELEMS/FEM1/MSH", np.concatenate([b.reshape(-1) for b in bds]).astype(np.int32)))
    out.append(("/ELEMS/FEM1/OFF_EL", np.concatenate([[0], np.cumsum(nbd)]).astype(np.int32)))
    for L in range(nlev):
        out.append((f"/ELEMS/FEM1/MSH{L}", bds[L].reshape(-1).astype(np.int32)))
    return out


def matrix_datasets(mesh, L, blocks=MATRIX_BLOCKS):
    """The datasets of Matrix<L> (the node / element pattern blocks), in compute_matrix's order."""
    out = []
    for r, c in blocks:
        ln, pos = mesh.gencase_matrix_block(L, r, c)
        m = _mode(r, c)
        out += [(f"DIM{m}", np.array([mesh._rows(L, r), mesh._rows(L, c)], np.int32)), (f"POS{m}", pos),
                (f"LEN{m}", ln), (f"OFFLEN{m}", np.zeros_like(ln))]
    return out


def transfer_datasets(mesh, L, rest, modes=TRANSFER_MODES):
    """The datasets of Prol<L-1>_<L> (rest False) or Rest<L+1>_<L> (rest True)."""
    out = []
    other = L + 1 if rest else L - 1
    for k in modes:
        ln, pos, val = mesh.gencase_transfer_block(L, k, rest)
        m = _mode(k)
        out += [(f"DIM{m}", np.array([mesh._rows(L, k), mesh._rows(other, k)], np.int32)), (f"POS{m}", pos),
                (f"VAL{m}", val), (f"LEN{m}", ln), (f"OFFLEN{m}", np.zeros_like(ln))]
    return out


def write_gencase(mesh, outdir, mesh_name="mesh", matrix="Matrix", prol="Prol", rest="Rest", operators=True):
    """Write the mesh file and (operators=True) every Matrix / Prol / Rest file of the hierarchy into outdir
    (the file-name stems of the reference's BASEMESH, F_MATRIX, F_PROL and F_REST keys).  Returns the bases."""
    os.makedirs(outdir, exist_ok=True)
    bases = [os.path.join(outdir, mesh_name)]
    _write(bases[0], mesh_datasets(mesh))
    if not operators:
        return bases
    for L in range(mesh.nlev):
        b = os.path.join(outdir, f"{matrix}{L}")
        _write(b, matrix_datasets(mesh, L))
        bases.append(b)
    for L in range(1, mesh.nlev):
        b = os.path.join(outdir, f"{prol}{L - 1}_{L}")
        _write(b, transfer_datasets(mesh, L, rest=False))
        bases.append(b)
    for L in range(mesh.nlev - 1):
        b = os.path.join(outdir, f"{rest}{L + 1}_{L}")
        _write(b, transfer_datasets(mesh, L, rest=True))
        bases.append(b)
    return bases


def read_datasets(base):
    """All datasets of one written file, by name."""
    return _read(base)


def read_mesh(base):
    """MGMesh::read_c on a mesh file (src/MGMesh.C:509-670): the hierarchy as a Mesh (fm_mesh_create) built from
    DFLS, NDxLEV, X1..X<dim>, OFF_EL, MSH and EL_NEIG; the level's Q1 node counts are NDxLEV[2 levels] on the
    coarse level and the Q2 counts of the level below elsewhere (MGSolDA::init_dof's Q1 numbering)."""
    d = _read(base)
    dim, fam, nlev, nsub = (int(v) for v in d["/DFLS"])
    if fam != 2 or nsub != 1:
        raise ValueError(f"{base}: {fam} element families and {nsub} subdomains; one volume + one boundary family "
                         "on one subdomain expected")
    nd = d["/NODES/MAP/NDxLEV"]
    nq2 = [int(v) for v in nd[:nlev]]
    nq1 = [int(nd[2 * nlev])] + nq2[:-1]
    xyz = np.stack([d[f"/NODES/COORD/X{k + 1}"] for k in range(dim)])
    return femus.Mesh.from_arrays(dim, xyz, d["/ELEMS/FEM0/OFF_EL"], d["/ELEMS/FEM0/MSH"],
                                  d["/ELEMS/FEM0/EL_NEIG"], nq2, nq1)
