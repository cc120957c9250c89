"""Python mirror of the FEMuS operator interface over the HIP C ABI.

The classes follow the reference's MGSolBase / MGSolDA / MGSolT / MGSolNS
interface (include/MGSolverBase.h:24-389, include/MGSolverDA.h:55-376,
template_appl/nse/SRC/MGSolverT.h, MGSolverNS.h): same method names
(GenBc, GenIc, init, GenMatRhs, MGSolve, MGStep-driven cycles, MGTimeStep),
same argument meaning, and -- instead of abort() -- a FemusError carrying the
library's status code and message.  All compute goes through
exfemus_amd/libfemus_hip.so on the GPU; there is no CPU fallback: loading
fails loudly when the library is missing.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libfemus_hip.so")

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
BCFN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p)

FM_PREC_FP64, FM_PREC_FP32_IR, FM_PREC_FP32_FGMRES, FM_PREC_FP64_FGMRES = 0, 1, 2, 3
FM_SMOOTH_CHEBYSHEV, FM_SMOOTH_JACOBI, FM_SMOOTH_VANKA, FM_SMOOTH_CHEBYSHEV4 = 0, 1, 2, 3
FM_ENOCONV = -3

# storage format of the SELL SpMV
SPMV_FORMAT = ("SELL-64 sigma4096 slab-grouped node order; columns per slice as 2-bit lane-predicted residuals or "
               "8-, 16- or 32-bit offsets")

# Exported entry points (the C ABI of include/femus_hip.h)
EXPORTS = [
    "fm_ctx_create", "fm_ctx_destroy", "fm_last_error", "fm_mesh_create_box", "fm_mesh_destroy",
    "fm_mesh_info", "fm_mesh_level", "fm_mesh_coords", "fm_mesh_conn", "fm_system_create",
    "fm_system_destroy", "fm_system_ndof", "fm_system_node_dof", "fm_system_genbc", "fm_system_get_bc",
    "fm_system_init", "fm_system_csr_nnz", "fm_system_csr_get", "fm_system_set_xold", "fm_system_get_x",
    "fm_system_set_x", "fm_system_get_rhs", "fm_system_set_rhs", "fm_system_set_elem_coef",
    "fm_system_set_velocity", "fm_heat_genmatrhs", "fm_ns_genmatrhs", "fm_spmv", "fm_residual",
    "fm_restrict", "fm_prolong", "fm_get_res", "fm_mg_setup", "fm_mg_solve", "fm_vanka_patch", "fm_adjoint_solve",
    "fm_heat_grad_coef", "fm_bench_spmv", "fm_sync", "fm_heat_genmat_transpose", "fm_ns_genmat_transpose", "fm_ns_genjac",
    "fm_get_adjoint_csr", "fm_heat_dissipation", "fm_bench_stream", "fm_ns_dissipation", "fm_ns_grad_coef",
    "fm_system_set_bcdata", "fm_heat_bc_rhs", "fm_ns_bc_rhs", "fm_ns_pressure_drop",
    "fm_ns_set_buoyancy", "fm_heat_nusselt", "fm_heat_coupling", "fm_adjoint_spmv", "fm_set_sell_split", "fm_set_sell_offsets", "fm_system_sell_info",
    "fm_system_set_vanka_store", "fm_mg_release", "fm_system_couple", "fm_system_advance", "fm_heat_nusselt_noalloc",
    "fm_mesh_create", "fm_system_set_pattern", "fm_system_set_transfer", "fm_system_get_x_dev", "fm_system_set_x_dev",
    "fm_system_get_rhs_dev", "fm_system_set_rhs_dev", "fm_gencase_matrix_block", "fm_gencase_transfer_block",
    "fm_gencase_boundary", "fm_set_assembly_gather", "fm_system_set_matrix", "fm_system_set_matrix_dev",
]


class FemusError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"femus_hip error {code}: {msg}")
        self.code = code


class HeatParams(C.Structure):
    _fields_ = [("alpha0", C.c_double), ("rhocp", C.c_double), ("dt", C.c_double),
                ("time_coeff", C.c_double), ("source", C.c_double), ("upwind", C.c_double),
                ("vel_const", C.c_double * 3)]


class NSParams(C.Structure):
    _fields_ = [("IRe", C.c_double), ("dt", C.c_double), ("time_coeff", C.c_double),
                ("upwind", C.c_double), ("IFr", C.c_double), ("dirg", C.c_double * 3),
                ("src", C.c_double * 3)]


class MGOpts(C.Structure):
    _fields_ = [("precision", C.c_int), ("smoother", C.c_int), ("nu_pre", C.c_int), ("nu_post", C.c_int),
                ("correction", C.c_int), ("max_iter", C.c_int), ("fgmres_restart", C.c_int),
                ("tol", C.c_double), ("cheb_lo", C.c_double), ("cheb_hi", C.c_double),
                ("omega", C.c_double), ("vanka_colored", C.c_int), ("gamma", C.c_int), ("use_graph", C.c_int),
                ("coarse_mult", C.c_int)]


class MGReport(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("rel_resid", C.c_double),
                ("hist", C.c_double * 256), ("seconds", C.c_double)]


_LIB = None


def lib():
    """Load libfemus_hip.so (raises if it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise ImportError(f"{LIBPATH} missing: run __graft_entry__.build() (hipcc, gfx950)")
    L = C.CDLL(LIBPATH)
    vp = C.c_void_p
    pp = C.POINTER(C.c_void_p)
    L.fm_last_error.restype = C.c_char_p
    L.fm_set_sell_split.argtypes = [C.c_int]
    L.fm_set_sell_offsets.argtypes = [C.c_int]
    L.fm_set_assembly_gather.argtypes = [C.c_int]
    L.fm_system_sell_info.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    L.fm_system_set_vanka_store.argtypes = [C.c_void_p, C.c_int]
    L.fm_mg_release.argtypes = [C.c_void_p, C.c_int]
    L.fm_system_couple.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    L.fm_system_advance.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    L.fm_ctx_create.argtypes = [C.c_int, pp]
    L.fm_ctx_destroy.argtypes = [vp]
    L.fm_sync.argtypes = [vp]
    L.fm_mesh_create_box.argtypes = [C.c_int, i32p, f64p, f64p, C.c_int, pp]
    L.fm_mesh_destroy.argtypes = [vp]
    L.fm_mesh_create.argtypes = [C.c_int, C.c_int, C.c_int, f64p, i32p, i32p, i32p, i32p, i32p, vp, vp, pp]
    L.fm_system_set_pattern.argtypes = [vp, C.c_int, vp, vp]
    L.fm_system_set_transfer.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
    for fn in ("fm_system_get_x_dev", "fm_system_set_x_dev", "fm_system_get_rhs_dev", "fm_system_set_rhs_dev",
               "fm_system_set_matrix", "fm_system_set_matrix_dev"):
        getattr(L, fn).argtypes = [vp, C.c_int, C.c_void_p]
    L.fm_mesh_info.argtypes = [vp, i32p]
    L.fm_mesh_level.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.fm_mesh_coords.argtypes = [vp, f64p]
    L.fm_mesh_conn.argtypes = [vp, C.c_int, vp, vp, vp]
    L.fm_gencase_matrix_block.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, C.POINTER(C.c_int64)]
    L.fm_gencase_transfer_block.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.POINTER(C.c_int64)]
    L.fm_gencase_boundary.argtypes = [vp, C.c_int, vp, C.POINTER(C.c_int)]
    L.fm_system_create.argtypes = [vp, vp, C.c_int, C.c_int, pp]
    L.fm_system_destroy.argtypes = [vp]
    L.fm_system_ndof.restype = C.c_int64
    L.fm_system_ndof.argtypes = [vp, C.c_int]
    L.fm_system_node_dof.argtypes = [vp, C.c_int, i32p]
    L.fm_system_genbc.argtypes = [vp, BCFN, vp]
    L.fm_system_get_bc.argtypes = [vp, vp, vp]
    L.fm_system_init.argtypes = [vp]
    L.fm_system_csr_nnz.restype = C.c_int64
    L.fm_system_csr_nnz.argtypes = [vp, C.c_int, C.c_int]
    L.fm_system_csr_get.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
    for fn in ("fm_system_set_xold",):
        getattr(L, fn).argtypes = [vp, f64p]
    for fn in ("fm_system_get_x", "fm_system_set_x", "fm_system_get_rhs", "fm_system_set_rhs", "fm_get_res"):
        getattr(L, fn).argtypes = [vp, C.c_int, f64p]
    L.fm_system_set_elem_coef.argtypes = [vp, C.c_int, vp]
    L.fm_system_set_velocity.argtypes = [vp, vp]
    L.fm_heat_genmatrhs.argtypes = [vp, C.c_int, C.c_int, C.POINTER(HeatParams)]
    L.fm_ns_genmatrhs.argtypes = [vp, C.c_int, C.c_int, C.POINTER(NSParams)]
    L.fm_spmv.argtypes = [vp, C.c_int, f64p, f64p]
    for fn in ("fm_residual", "fm_restrict", "fm_prolong"):
        getattr(L, fn).argtypes = [vp, C.c_int]
    L.fm_mg_setup.argtypes = [vp, C.POINTER(MGOpts)]
    L.fm_vanka_patch.argtypes = [vp, C.c_int, C.c_int, vp, vp]
    L.fm_mg_solve.argtypes = [vp, C.POINTER(MGOpts), C.POINTER(MGReport)]
    L.fm_adjoint_solve.argtypes = [vp, C.POINTER(MGOpts), f64p, f64p, C.POINTER(MGReport)]
    L.fm_heat_grad_coef.argtypes = [vp, C.POINTER(HeatParams), f64p, f64p, f64p]
    L.fm_heat_genmat_transpose.argtypes = [vp, C.c_int, C.POINTER(HeatParams)]
    L.fm_ns_genmat_transpose.argtypes = [vp, C.c_int, C.POINTER(NSParams)]
    L.fm_ns_genjac.argtypes = [vp, C.c_int, C.c_int, C.POINTER(NSParams)]
    L.fm_get_adjoint_csr.argtypes = [vp, C.c_int, f64p]
    L.fm_heat_dissipation.argtypes = [vp, C.POINTER(HeatParams), f64p, C.POINTER(C.c_double), vp, vp]
    L.fm_bench_spmv.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.fm_bench_stream.argtypes = [vp, C.c_int64, C.c_int, C.POINTER(C.c_double)]
    L.fm_ns_dissipation.argtypes = [vp, C.POINTER(NSParams), f64p, C.POINTER(C.c_double), vp, vp]
    L.fm_ns_grad_coef.argtypes = [vp, C.POINTER(NSParams), f64p, f64p, f64p]
    L.fm_system_set_bcdata.argtypes = [vp, vp]
    L.fm_heat_bc_rhs.argtypes = [vp, C.POINTER(HeatParams), f64p, f64p]
    L.fm_ns_bc_rhs.argtypes = [vp, C.POINTER(NSParams), f64p, f64p]
    L.fm_ns_pressure_drop.argtypes = [vp, f64p, C.POINTER(C.c_double), vp]
    L.fm_ns_set_buoyancy.argtypes = [vp, vp, vp]
    L.fm_heat_nusselt.argtypes = [vp, C.POINTER(HeatParams), C.c_int, f64p, C.POINTER(C.c_double), vp, vp]
    L.fm_heat_nusselt_noalloc.argtypes = [vp, C.POINTER(HeatParams), C.c_int, f64p, C.POINTER(C.c_double), vp, vp]
    L.fm_heat_coupling.argtypes = [vp, C.POINTER(HeatParams), C.c_int, C.c_double, vp, vp, vp, vp]
    L.fm_adjoint_spmv.argtypes = [vp, C.c_int, f64p, f64p]
    _LIB = L
    return L


def _vec(a, n, what):
    """a as a contiguous float64 array of exactly n entries: the library reads n values through the pointer, so a
    shorter array would be read past its end."""
    v = np.ascontiguousarray(a, np.float64).reshape(-1)
    if v.size != n:
        raise ValueError(f"{what}: expected {n} values, got {v.size}")
    return v


def _ck(rc, ok=(0,)):
    if rc not in ok:
        raise FemusError(rc, lib().fm_last_error().decode(errors="replace"))
    return rc


class Context:
    """One HIP device + one stream (fm_ctx).  Systems created on it are closed first when it closes."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        self._systems = weakref.WeakSet()
        _ck(lib().fm_ctx_create(device, C.byref(self.h)))

    def sync(self):
        _ck(lib().fm_sync(self.h))

    def stream_copy(self, nbytes, reps=20):
        """Average seconds of a device copy of nbytes (reads nbytes, writes nbytes)."""
        sec = C.c_double()
        _ck(lib().fm_bench_stream(self.h, int(nbytes), reps, C.byref(sec)))
        return sec.value

    def close(self):
        for s in list(self._systems):
            s.close()
        if self.h:
            _ck(lib().fm_ctx_destroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mesh:
    """gencase hierarchy of a box: build_square/build_cube + uniformly_refine + printMesh ordering."""

    def __init__(self, dim, nx0, lo, hi, nlev):
        nx = np.ones(3, np.int32); nx[:dim] = nx0
        l = np.zeros(3); l[:dim] = lo
        h = np.zeros(3); h[:dim] = hi
        self.h = C.c_void_p()
        _ck(lib().fm_mesh_create_box(dim, nx, l, h, nlev, C.byref(self.h)))
        self._info()

    def _info(self):
        info = np.zeros(6, np.int32)
        _ck(lib().fm_mesh_info(self.h, info))
        self.dim, self.nlev, self.n2, self.n1, self.nsides, self.N = (int(v) for v in info)

    @classmethod
    def from_arrays(cls, dim, xyz, off_el, el_map, el_neighbor, nodes_q2, nodes_q1, parent=None, child=None):
        """A caller's hierarchy in MGMesh's layout (fm_mesh_create): xyz [dim][N], off_el [nlev+1], el_map
        [elements][9|27] top-level node ids, el_neighbor [elements][2 dim] global element ids of the same level
        (gencase's EL_NEIG numbering, -1 on the boundary), node counts per level, and for levels >= 1 each element's
        parent (level L-1 index) and child index."""
        self = cls.__new__(cls)
        self.h = C.c_void_p()
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        off = np.ascontiguousarray(off_el, np.int32)
        nlev = len(off) - 1
        i32 = lambda a: np.ascontiguousarray(a, np.int32).reshape(-1)
        par = None if parent is None else i32(parent)
        ch = None if child is None else i32(child)
        self._keep = (par, ch)
        ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        _ck(lib().fm_mesh_create(dim, nlev, len(xyz) // dim, xyz, off, i32(el_map), i32(el_neighbor), i32(nodes_q2),
                                 i32(nodes_q1), ptr(par), ptr(ch), C.byref(self.h)))
        self._info()
        return self

    def level(self, L):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _ck(lib().fm_mesh_level(self.h, L, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def coords(self):
        x = np.zeros(self.dim * self.N)
        _ck(lib().fm_mesh_coords(self.h, x))
        return x.reshape(self.dim, self.N)

    def conn(self, L):
        ne = self.level(L)[0]
        c = np.zeros(ne * self.n2, np.int32); nb = np.zeros(ne * self.nsides, np.int32); p = np.zeros(ne, np.int32)
        _ck(lib().fm_mesh_conn(self.h, L, c.ctypes.data_as(C.c_void_p), nb.ctypes.data_as(C.c_void_p),
                               p.ctypes.data_as(C.c_void_p)))
        return c.reshape(ne, self.n2), nb.reshape(ne, self.nsides), p

    def _rows(self, L, kind):
        ne, nq2, nq1 = self.level(L)
        return {2: nq2, 1: nq1, 0: ne}[kind]

    def gencase_matrix_block(self, L, rkind, ckind):
        """Block (rkind, ckind) of level L's matrix pattern in gencase's order (fm_gencase_matrix_block): the
        LEN (row offsets) and POS (columns) datasets.  Kinds: 2 Q2 nodes, 1 Q1 nodes, 0 elements."""
        ln = np.zeros(self._rows(L, rkind) + 1, np.int32)
        nnz = C.c_int64()
        vp = lambda a: a.ctypes.data_as(C.c_void_p)
        _ck(lib().fm_gencase_matrix_block(self.h, L, rkind, ckind, vp(ln), None, C.byref(nnz)))
        pos = np.zeros(nnz.value, np.int32)
        _ck(lib().fm_gencase_matrix_block(self.h, L, rkind, ckind, vp(ln), vp(pos), C.byref(nnz)))
        return ln, pos

    def gencase_transfer_block(self, L, kind, rest):
        """Prolongation (rest False: fine L <- coarse L-1) or restriction (rest True: coarse L <- fine L+1) block
        of one kind in gencase's order (fm_gencase_transfer_block): LEN, POS, VAL."""
        ln = np.zeros(self._rows(L, kind) + 1, np.int32)
        nnz = C.c_int64()
        vp = lambda a: a.ctypes.data_as(C.c_void_p)
        _ck(lib().fm_gencase_transfer_block(self.h, L, kind, int(bool(rest)), vp(ln), None, None, C.byref(nnz)))
        pos = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value, np.float64)
        _ck(lib().fm_gencase_transfer_block(self.h, L, kind, int(bool(rest)), vp(ln), vp(pos), vp(val),
                                            C.byref(nnz)))
        return ln, pos, val

    def boundary(self, L):
        """Boundary mesh of level L (fm_gencase_boundary): [sides][3 | 9] top-level node ids."""
        n = C.c_int()
        _ck(lib().fm_gencase_boundary(self.h, L, None, C.byref(n)))
        nb2 = 3 if self.dim == 2 else 9
        nodes = np.zeros(n.value * nb2, np.int32)
        _ck(lib().fm_gencase_boundary(self.h, L, nodes.ctypes.data_as(C.c_void_p), C.byref(n)))
        return nodes.reshape(n.value, nb2)

    def __del__(self):
        try:
            if self.h:
                lib().fm_mesh_destroy(self.h)
                self.h = C.c_void_p()
        except Exception:
            pass


class MGSolDA:
    """One multigrid equation system (MGSolBase + MGSolDA state and operators)."""

    def __init__(self, ctx, mesh, nq2, nq1):
        self.ctx, self.mesh, self.nq2, self.nq1 = ctx, mesh, nq2, nq1
        self.h = C.c_void_p()
        _ck(lib().fm_system_create(ctx.h, mesh.h, nq2, nq1, C.byref(self.h)))
        ctx._systems.add(self)
        self._NoLevels = mesh.nlev
        self._Dim = [int(lib().fm_system_ndof(self.h, L)) for L in range(mesh.nlev)]
        self._cb = None

    def close(self):
        if self.h:
            lib().fm_system_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- DoF / BC / operators -------------------------------------------------
    def node_dof(self, L):
        nd = np.zeros((self.nq2 + self.nq1) * self.mesh.N, np.int32)
        _ck(lib().fm_system_node_dof(self.h, L, nd))
        return nd

    def GenBc(self, bc_read):
        """bc_read(xp, bc_Neum, bc_flag) fills lists over the variables (MGSolDA::GenBc)."""
        nv = self.nq2 + self.nq1
        dim = self.mesh.dim

        def cb(xp, neum, flag, user):
            x = [xp[i] for i in range(dim)]
            nn = [neum[i] for i in range(nv)]
            ff = [flag[i] for i in range(nv)]
            bc_read(x, nn, ff)
            for i in range(nv):
                neum[i] = int(nn[i]); flag[i] = int(ff[i])

        self._cb = BCFN(cb)
        _ck(lib().fm_system_genbc(self.h, self._cb, None))

    def bc(self):
        n = self._Dim[-1]
        b0 = np.zeros(n, np.int32); b1 = np.zeros(n, np.int32)
        _ck(lib().fm_system_get_bc(self.h, b0.ctypes.data_as(C.c_void_p), b1.ctypes.data_as(C.c_void_p)))
        return b0, b1

    def init(self):
        """MGSolDA::init for every level: A pattern, Rst (masked), Prl on the device."""
        _ck(lib().fm_system_init(self.h))

    def GenIc(self, ic_read):
        """ic_read(xp) -> list of values per variable; sets x_old[top] (MGSolDA::GenIc)."""
        xyz = self.mesh.coords()
        nd = self.node_dof(self._NoLevels - 1)
        N = self.mesh.N
        x = np.zeros(self._Dim[-1])
        for k in range(N):
            vals = ic_read(xyz[:, k])
            for iv in range(self.nq2 + self.nq1):
                d = nd[k + iv * N]
                if d >= 0:
                    x[d] = vals[iv]
        self.set_xold(x)
        return x

    def csr(self, which, L):
        nnz = int(lib().fm_system_csr_nnz(self.h, which, L))
        nr = self._Dim[L]
        rp = np.zeros(nr + 1, np.int64); ci = np.zeros(nnz, np.int32); v = np.zeros(nnz)
        _ck(lib().fm_system_csr_get(self.h, which, L, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                                    v.ctypes.data_as(C.c_void_p)))
        return rp, ci, v

    def set_matrix(self, L, val):
        """The caller's values of A on level L in the CSR order of csr(0, L) (SparseMatrixM::set over the pattern)."""
        nnz = int(lib().fm_system_csr_nnz(self.h, 0, L))
        v = np.ascontiguousarray(val, np.float64)
        if v.size != nnz:
            raise ValueError(f"set_matrix: {v.size} values for a pattern of {nnz}")
        _ck(lib().fm_system_set_matrix(self.h, L, v.ctypes.data_as(C.c_void_p)))

    def set_xold(self, x):
        _ck(lib().fm_system_set_xold(self.h, _vec(x, self._Dim[-1], "x_old")))

    def get_x(self, L=None):
        L = self._NoLevels - 1 if L is None else L
        x = np.zeros(self._Dim[L]); _ck(lib().fm_system_get_x(self.h, L, x)); return x

    def set_x(self, x, L=None):
        L = self._NoLevels - 1 if L is None else L
        _ck(lib().fm_system_set_x(self.h, L, _vec(x, self._Dim[L], "x")))

    def get_rhs(self, L=None):
        L = self._NoLevels - 1 if L is None else L
        b = np.zeros(self._Dim[L]); _ck(lib().fm_system_get_rhs(self.h, L, b)); return b

    def set_rhs(self, b, L=None):
        L = self._NoLevels - 1 if L is None else L
        _ck(lib().fm_system_set_rhs(self.h, L, _vec(b, self._Dim[L], "rhs")))

    def set_pattern(self, L, rowptr, colind):
        """The caller's pattern of A on level L (ReadMatrix), before init(); None drops it."""
        if rowptr is None:
            _ck(lib().fm_system_set_pattern(self.h, L, None, None))
            return
        rp = np.ascontiguousarray(rowptr, np.int64); ci = np.ascontiguousarray(colind, np.int32)
        if rp.size != self._Dim[L] + 1 or ci.size != rp[-1]:
            raise ValueError("pattern: row pointers / columns of the wrong length")
        _ck(lib().fm_system_set_pattern(self.h, L, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p)))

    def set_transfer(self, which, L, rowptr, colind, val):
        """The caller's restriction (which 1, Rst[L]) or prolongation (2, Prl[L]) in the reference numbering."""
        rp = np.ascontiguousarray(rowptr, np.int64); ci = np.ascontiguousarray(colind, np.int32)
        v = np.ascontiguousarray(val, np.float64)
        if rp.size != self._Dim[L] + 1 or ci.size != rp[-1] or v.size != rp[-1]:
            raise ValueError("transfer: arrays of the wrong length")
        _ck(lib().fm_system_set_transfer(self.h, which, L, rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
                                         v.ctypes.data_as(C.c_void_p)))

    def dev_vector(self, fn, L, ptr):
        """fm_system_{get,set}_{x,rhs}_dev with a device address (e.g. a torch tensor's data_ptr())."""
        L = self._NoLevels - 1 if L is None else L
        _ck(getattr(lib(), fn)(self.h, L, C.c_void_p(int(ptr))))

    def get_res(self, L):
        r = np.zeros(self._Dim[L]); _ck(lib().fm_get_res(self.h, L, r)); return r

    def node_values(self, x, var=0):
        """Nodal field of variable `var` on the top-level nodes from a top-level DoF vector (0 where absent)."""
        N = self.mesh.N
        nd = self.node_dof(self._NoLevels - 1)[var * N:(var + 1) * N]
        out = np.zeros(N)
        ok = nd >= 0
        out[ok] = np.asarray(x)[nd[ok]]
        return out

    def set_bcdata(self, g):
        """Dirichlet data of the penalty RHS (top-level DoFs); None = x_old (the reference's source)."""
        if g is None:
            _ck(lib().fm_system_set_bcdata(self.h, None))
        else:
            c = _vec(g, self._Dim[-1], "bcdata")
            _ck(lib().fm_system_set_bcdata(self.h, c.ctypes.data_as(C.c_void_p)))

    def bc_rhs(self, g):
        """Boundary-face penalty RHS of the top level for Dirichlet data g (linear in g)."""
        out = np.zeros(self._Dim[-1])
        fn = lib().fm_heat_bc_rhs if self.nq1 == 0 else lib().fm_ns_bc_rhs
        _ck(fn(self.h, C.byref(self.params), _vec(g, self._Dim[-1], "g"), out))
        return out

    def set_elem_coef(self, L, coef):
        c = None if coef is None else _vec(coef, self.mesh.level(L)[0], "element coefficients")
        _ck(lib().fm_system_set_elem_coef(self.h, L, None if c is None else c.ctypes.data_as(C.c_void_p)))
        self._keep_coef = c

    # -- LA operators (NumericVectorM::resid / matrix_mult) -----------------------
    def spmv(self, L, x):
        y = np.zeros(self._Dim[L]); _ck(lib().fm_spmv(self.h, L, _vec(x, self._Dim[L], "x"), y)); return y

    def couple(self, src, which, coef=None):
        """This system's coupling field from src's top-level solution, on the device: which 0 = src's temperature
        as the buoyancy field (coef per direction), 1 = src's velocity as the advection field."""
        c = np.zeros(3)
        if coef is not None:
            c[:len(coef)] = coef
        _ck(lib().fm_system_couple(self.h, src.h, int(which), c.ctypes.data_as(C.c_void_p)))

    def advance(self):
        """x_old = x on the top level (on the device); returns |x - x_old| / |x| before the copy."""
        r = C.c_double()
        _ck(lib().fm_system_advance(self.h, C.byref(r)))
        return r.value

    def set_vanka_store(self, mask):
        """Precisions in which the Vanka patch inverses are kept (bit 0 FP64, bit 1 FP32); returns the previous mask."""
        rc = lib().fm_system_set_vanka_store(self.h, int(mask))
        if rc < 0:
            _ck(rc)
        return rc

    def release_mg(self, adjoint=False):
        """Free the patch and coarse inverses of the forward (or transposed) hierarchy."""
        _ck(lib().fm_mg_release(self.h, 1 if adjoint else 0))

    def sell_info(self, which, L):
        """SELL storage of matrix `which` (0 A, 1 R, 2 P) on level L: slices, slots, 8/16/32-bit offset slots, nnz,
        2-bit lane-predicted slots."""
        info = np.zeros(7, np.int64)
        _ck(lib().fm_system_sell_info(self.h, which, L, info.ctypes.data_as(C.c_void_p)))
        return dict(zip(("slices", "slots", "slots_c8", "slots_c16", "slots_c32", "nnz", "slots_c2"),
                        (int(v) for v in info)))

    def resid(self, L):
        _ck(lib().fm_residual(self.h, L))

    def restrict(self, L):
        _ck(lib().fm_restrict(self.h, L))

    def prolong(self, L):
        _ck(lib().fm_prolong(self.h, L))

    # -- MGSolve (src/MGSolverBase.C:163-208) --------------------------------------
    @staticmethod
    def opts(precision=FM_PREC_FP64, smoother=FM_SMOOTH_CHEBYSHEV, nu_pre=2, nu_post=2, correction=1,
             max_iter=100, fgmres_restart=30, tol=1e-8, cheb_lo=0.2, cheb_hi=1.05, omega=None, vanka_colored=0,
             gamma=1, use_graph=1, coarse_mult=0):
        if omega is None:
            omega = 0.5 if smoother == FM_SMOOTH_VANKA else 0.6
        return MGOpts(precision, smoother, nu_pre, nu_post, correction, max_iter, fgmres_restart, tol,
                      cheb_lo, cheb_hi, omega, vanka_colored, gamma, use_graph, coarse_mult)

    def MGSetup(self, opts):
        _ck(lib().fm_mg_setup(self.h, C.byref(opts)))

    def print_u(self, base, names, L=None, x=None, refvalue=None):
        """Solution file of this system on level L (MGSolDA::print_u): nodal datasets named after the
        variables; x defaults to the current solution of the top level."""
        from . import solio
        L = self._NoLevels - 1 if L is None else L
        x = self.get_x(L) if x is None else x
        return solio.print_u(self.mesh, L, base, [(names, self.node_dof(L), x, self.nq2, self.nq1, refvalue)])

    def read_u(self, base, names, L=None, refvalue=None):
        """Restart (MGSolDA::read_u): x_old and x of level L from a solution file written on that level."""
        from . import solio
        L = self._NoLevels - 1 if L is None else L
        x = solio.read_u(self.mesh, L, base, names, self.node_dof(L), self.nq2, self.nq1, self._Dim[L], refvalue)
        if L == self._NoLevels - 1:
            self.set_xold(x)
        self.set_x(x, L)
        return x

    def vanka_patch(self, L, e):
        """(dofs, K_e^-1): the FP64 local inverse the Vanka smoother applies for element e on level L."""
        nl = (self.nq2 * 27 if self.mesh.dim == 3 else self.nq2 * 9) + (self.nq1 * (8 if self.mesh.dim == 3 else 4))
        dofs = np.zeros(nl, np.int32)
        K = np.zeros((nl, nl))
        _ck(lib().fm_vanka_patch(self.h, L, e, dofs.ctypes.data_as(C.c_void_p), K.ctypes.data_as(C.c_void_p)))
        return dofs, K

    def MGSolve(self, Eps=1e-8, MaxIter=100, Gamma=1, Nc_pre=2, Nc_coarse=40, Nc_post=2, opts=None,
                raise_on_noconv=False):
        """V-cycles until ||b-Ax||/||b|| <= Eps (FP64 residual) or MaxIter; returns MGReport."""
        o = opts if opts is not None else self.opts(nu_pre=Nc_pre, nu_post=Nc_post, max_iter=MaxIter, tol=Eps)
        rep = MGReport()
        rc = lib().fm_mg_solve(self.h, C.byref(o), C.byref(rep))
        _ck(rc, ok=(0,) if raise_on_noconv else (0, FM_ENOCONV))
        return rep

    def adjoint_csr(self, L):
        nnz = int(lib().fm_system_csr_nnz(self.h, 0, L))
        v = np.zeros(nnz)
        _ck(lib().fm_get_adjoint_csr(self.h, L, v))
        return v

    def AdjointSolve(self, g, opts):
        """Transposed multigrid solve A^T lambda = g (returns lambda, report)."""
        lam = np.zeros(self._Dim[-1])
        rep = MGReport()
        rc = lib().fm_adjoint_solve(self.h, C.byref(opts), _vec(g, self._Dim[-1], "g"), lam, C.byref(rep))
        _ck(rc, ok=(0, FM_ENOCONV))
        return lam, rep

    def adjoint_spmv(self, L, x):
        """A^T x with the transposed matrix of level L (after GenMatTranspose / GenJac(L, 1))."""
        y = np.zeros(self._Dim[L])
        _ck(lib().fm_adjoint_spmv(self.h, L, _vec(x, self._Dim[L], "x"), y))
        return y

    def bench_spmv(self, L, prec, reps):
        s = C.c_double()
        _ck(lib().fm_bench_spmv(self.h, L, prec, reps, C.byref(s)))
        return s.value


class MGSolT(MGSolDA):
    """Heat system (template_appl/nse/SRC/MGSolverT.C); one Q2 variable T."""

    def __init__(self, ctx, mesh, alpha0=1.0, dt=0.1, rhocp=1.0, time_coeff=1.0, source=0.0, upwind=1.0,
                 vel=(0.0, 0.0, 0.0)):
        super().__init__(ctx, mesh, 1, 0)
        self.params = HeatParams(alpha0, rhocp, dt, time_coeff, source, upwind, (C.c_double * 3)(*vel))

    def GenMatRhs(self, time, Level, mode):
        _ck(lib().fm_heat_genmatrhs(self.h, Level, mode, C.byref(self.params)))

    # -- discrete adjoint ---------------------------------------------------------
    def GenMatTranspose(self, Level):
        _ck(lib().fm_heat_genmat_transpose(self.h, Level, C.byref(self.params)))

    def coupling(self, mode, scale, T=None, lam=None, m=None):
        """Node-space coupling products of the Boussinesq system (fm_heat_coupling)."""
        N = self.mesh.N
        out = np.zeros((self.mesh.dim if mode == 1 else 1) * N)
        n = self._Dim[-1]
        T = None if T is None else _vec(T, n, "T")
        lam = None if lam is None else _vec(lam, n, "lam")
        m = None if m is None else _vec(m, N, "m")
        ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        _ck(lib().fm_heat_coupling(self.h, C.byref(self.params), mode, scale, ptr(T), ptr(lam), ptr(m),
                                   out.ctypes.data_as(C.c_void_p)))
        return out

    def nusselt(self, T, axis=0):
        """Nusselt number along `axis` of a differentially heated box with the current advection field:
        (J, dJ/dT, dJ/dk_e partials)."""
        J = C.c_double()
        g = np.zeros(self._Dim[-1])
        dk = np.zeros(self.mesh.level(self._NoLevels - 1)[0])
        _ck(lib().fm_heat_nusselt(self.h, C.byref(self.params), axis, _vec(T, self._Dim[-1], "T"),
                                  C.byref(J), g.ctypes.data_as(C.c_void_p), dk.ctypes.data_as(C.c_void_p)))
        return J.value, g, dk

    def dissipation(self, T):
        """J = alpha0 sum_e k_e int |grad T|^2, dJ/dT, dJ/dk_e (partial) on the top level."""
        J = C.c_double()
        g = np.zeros(self._Dim[-1])
        dk = np.zeros(self.mesh.level(self._NoLevels - 1)[0])
        _ck(lib().fm_heat_dissipation(self.h, C.byref(self.params), _vec(T, self._Dim[-1], "T"), C.byref(J),
                                      g.ctypes.data_as(C.c_void_p), dk.ctypes.data_as(C.c_void_p)))
        return J.value, g, dk

    def grad_coef(self, lam, T):
        """-lambda^T (dA/dk_e) T for every top element."""
        out = np.zeros(self.mesh.level(self._NoLevels - 1)[0])
        _ck(lib().fm_heat_grad_coef(self.h, C.byref(self.params), _vec(lam, self._Dim[-1], "lam"),
                                    _vec(T, self._Dim[-1], "T"), out))
        return out

    def set_velocity(self, vel_top):
        v = None if vel_top is None else _vec(vel_top, self.mesh.dim * self.mesh.N, "velocity")
        self._keep_vel = v
        _ck(lib().fm_system_set_velocity(self.h, None if v is None else v.ctypes.data_as(C.c_void_p)))

    def MGTimeStep(self, time, opts=None):
        """Assemble all levels (top with rhs), MGSolve, x_old <- x (MGSolT::MGTimeStep)."""
        top = self._NoLevels - 1
        self.GenMatRhs(time, top, 1)
        for L in range(top):
            self.GenMatRhs(time, L, 0)
        rep = self.MGSolve(opts=opts)
        self.set_xold(self.get_x())
        return rep


class MGSolNS(MGSolDA):
    """Coupled Q2-Q1 Navier-Stokes (template_appl/nse/SRC/MGSolverNS.C, NS_EQUATIONS==1)."""

    def __init__(self, ctx, mesh, IRe=0.01, dt=0.1, time_coeff=1.0, upwind=0.1, IFr=0.0, dirg=(0, 0, 0), src=(0, 0, 0)):
        super().__init__(ctx, mesh, mesh.dim, 1)
        self.params = NSParams(IRe, dt, time_coeff, upwind, IFr, (C.c_double * 3)(*dirg), (C.c_double * 3)(*src))

    def GenMatRhs(self, time, Level, mode):
        _ck(lib().fm_ns_genmatrhs(self.h, Level, mode, C.byref(self.params)))

    # -- discrete adjoint of the Oseen (Picard) system ----------------------------
    def GenMatTranspose(self, Level):
        _ck(lib().fm_ns_genmat_transpose(self.h, Level, C.byref(self.params)))

    def GenJac(self, Level, trans=0):
        """Newton Jacobian of the steady residual at x_old (trans = 1: its transpose, for the adjoint solve)."""
        _ck(lib().fm_ns_genjac(self.h, Level, trans, C.byref(self.params)))

    def dissipation(self, U):
        """J = sum_e nu_e int_e |grad u|^2, dJ/dU, dJ/dnu_e (partial) on the top level."""
        J = C.c_double()
        g = np.zeros(self._Dim[-1])
        dn = np.zeros(self.mesh.level(self._NoLevels - 1)[0])
        _ck(lib().fm_ns_dissipation(self.h, C.byref(self.params), _vec(U, self._Dim[-1], "U"), C.byref(J),
                                    g.ctypes.data_as(C.c_void_p), dn.ctypes.data_as(C.c_void_p)))
        return J.value, g, dn

    def set_buoyancy(self, T_node, coef):
        """Boussinesq body force coef * T (nodal temperature on the top-level nodes); None removes it."""
        if T_node is None:
            _ck(lib().fm_ns_set_buoyancy(self.h, None, None))
            return
        t = _vec(T_node, self.mesh.N, "nodal temperature")
        c = np.zeros(3); c[:len(coef)] = coef
        _ck(lib().fm_ns_set_buoyancy(self.h, t.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p)))
        self._keep_T = t

    def velocity_nodes(self, x=None):
        """Nodal velocity [dim * Ntop] (component-major) from a top-level DoF vector (default: x)."""
        x = self.get_x() if x is None else x
        return np.concatenate([self.node_values(x, d) for d in range(self.mesh.dim)])

    def pressure_drop(self, U):
        """J = mean p(inlet x = lo) - mean p(outlet x = hi) and dJ/dU."""
        J = C.c_double()
        g = np.zeros(self._Dim[-1])
        _ck(lib().fm_ns_pressure_drop(self.h, _vec(U, self._Dim[-1], "U"), C.byref(J),
                                      g.ctypes.data_as(C.c_void_p)))
        return J.value, g

    def grad_coef(self, lam, U):
        """-lambda^T (dA/dnu_e) U for every top element."""
        out = np.zeros(self.mesh.level(self._NoLevels - 1)[0])
        _ck(lib().fm_ns_grad_coef(self.h, C.byref(self.params), _vec(lam, self._Dim[-1], "lam"),
                                  _vec(U, self._Dim[-1], "U"), out))
        return out

    def MGTimeStep(self, time, opts=None):
        top = self._NoLevels - 1
        self.GenMatRhs(time, top, 1)
        for L in range(top):
            self.GenMatRhs(time, L, 0)
        rep = self.MGSolve(opts=opts)
        self.set_xold(self.get_x())
        return rep
