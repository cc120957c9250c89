"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- never by the product package exfemus_amd.
Each wrapper names the reference function its C body restates.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
BCFN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p)


class HeatParams(C.Structure):
    _fields_ = [("alpha0", C.c_double), ("rhocp", C.c_double), ("dt", C.c_double),
                ("time_coeff", C.c_double), ("source", C.c_double), ("upwind", C.c_double),
                ("vel_const", C.c_double * 3)]


class NSParams(C.Structure):
    """om_ns_params; buoy / Tn (Boussinesq body force from a nodal temperature field) default to none.  Keep
    the array passed as Tn alive while the structure is in use (set_buoyancy stores a reference)."""
    _fields_ = [("IRe", C.c_double), ("dt", C.c_double), ("time_coeff", C.c_double),
                ("upwind", C.c_double), ("IFr", C.c_double), ("dirg", C.c_double * 3),
                ("src", C.c_double * 3), ("buoy", C.c_double * 3), ("Tn", C.c_void_p)]

    def set_buoyancy(self, T_node, coef):
        self._Tn = None if T_node is None else np.ascontiguousarray(T_node, np.float64)
        self.Tn = None if self._Tn is None else self._Tn.ctypes.data
        c = [0.0, 0.0, 0.0]
        c[:len(coef)] = coef
        self.buoy = (C.c_double * 3)(*c)
        return self


class MGOpts(C.Structure):
    _fields_ = [("smoother", C.c_int), ("coarse", C.c_int), ("correction", C.c_int),
                ("nu_pre", C.c_int), ("nu_post", C.c_int), ("gamma", C.c_int),
                ("max_cycles", C.c_int), ("eps", C.c_double), ("cheb_lo", C.c_double),
                ("cheb_hi", C.c_double), ("jacobi_omega", C.c_double), ("ref_exit", C.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.om_mesh_build.restype = vp
        L.om_mesh_build.argtypes = [C.c_int, i32p, f64p, f64p, C.c_int]
        L.om_mesh_free.argtypes = [vp]
        L.om_mesh_info.argtypes = [vp, i32p]
        L.om_mesh_level.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.om_mesh_coords.argtypes = [vp, f64p]
        L.om_mesh_set_coords.argtypes = [vp, f64p]
        L.om_mesh_conn.argtypes = [vp, C.c_int, vp, vp, vp]
        L.om_mesh_node_keys.argtypes = [vp, i32p, i32p, i32p]
        L.om_fe_tables.argtypes = [C.c_int, C.c_int, vp, vp, vp, C.POINTER(C.c_int)]
        L.om_embedding.argtypes = [C.c_int, C.c_int, f64p]
        L.om_init_dof.argtypes = [vp, C.c_int, C.c_int, C.c_int, i32p]
        L.om_genbc.argtypes = [vp, C.c_int, C.c_int, i32p, BCFN, vp, i32p, i32p]
        L.om_mat_block.restype = C.c_int64
        L.om_mat_block.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp]
        for fn in (L.om_prol_block, L.om_rest_block):
            fn.restype = C.c_int64
            fn.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
        L.om_system_pattern.restype = C.c_int64
        L.om_system_pattern.argtypes = [vp, C.c_int, C.c_int, C.c_int, i32p, vp, vp]
        L.om_system_rest.restype = C.c_int64
        L.om_system_rest.argtypes = [vp, C.c_int, C.c_int, C.c_int, i32p, i32p, i32p, i32p, vp, vp, vp]
        L.om_system_prol.restype = C.c_int64
        L.om_system_prol.argtypes = [vp, C.c_int, C.c_int, C.c_int, i32p, i32p, vp, vp, vp]
        L.om_heat_genmatrhs.argtypes = [vp, C.c_int, C.c_int, C.POINTER(HeatParams), i32p, i32p, i32p, i32p,
                                        f64p, vp, vp, i64p, i32p, f64p, vp]
        L.om_ns_genmatrhs.argtypes = [vp, C.c_int, C.c_int, C.POINTER(NSParams), i32p, i32p, i32p, i32p,
                                      f64p, vp, i64p, i32p, f64p, vp]
        L.om_heat_element.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(HeatParams), i32p, i32p, i32p,
                                      f64p, C.c_double, vp, f64p, f64p, i32p, vp]
        L.om_ns_element.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(NSParams), i32p, i32p, i32p,
                                    f64p, C.c_double, f64p, f64p, i32p, vp]
        L.om_spmv.argtypes = [C.c_int64, i64p, i32p, f64p, f64p, f64p]
        L.om_resid.argtypes = [C.c_int64, i64p, i32p, f64p, f64p, f64p, f64p]
        L.om_gmres_ilu.restype = C.c_int
        L.om_gmres_ilu.argtypes = [C.c_int64, i64p, i32p, f64p, f64p, f64p, C.c_int, C.c_int, C.c_double,
                                   C.c_int, C.POINTER(C.c_double)]
        L.om_hier_new.restype = vp
        L.om_hier_new.argtypes = [C.c_int]
        L.om_hier_free.argtypes = [vp]
        L.om_hier_set_A.argtypes = [vp, C.c_int, C.c_int64, i64p, i32p, f64p]
        L.om_hier_set_R.argtypes = [vp, C.c_int, C.c_int64, C.c_int64, i64p, i32p, f64p]
        L.om_hier_set_P.argtypes = [vp, C.c_int, C.c_int64, C.c_int64, i64p, i32p, f64p]
        L.om_mg_solve.restype = C.c_int
        L.om_mg_solve.argtypes = [vp, C.POINTER(MGOpts), f64p, f64p, f64p]
        L.om_lambda_max.restype = C.c_double
        L.om_lambda_max.argtypes = [C.c_int64, i64p, i32p, f64p, C.c_int]
        L.om_dense_lu_solve.argtypes = [C.c_int, f64p, f64p]
        _LIB = L
    return _LIB


def ref_lib():
    """LASPACK compiled from the reference's own sources (oracle/_ref), or None."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "liblaspack_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.lr_spmv.argtypes = [C.c_int64, i64p, i32p, f64p, f64p, f64p]
        R.lr_ilu_apply.argtypes = [C.c_int64, i64p, i32p, f64p, f64p, f64p]
        R.lr_gmres.argtypes = [C.c_int64, i64p, i32p, f64p, f64p, f64p, C.c_int, C.c_int, C.c_int, C.c_double,
                               C.POINTER(C.c_double), C.POINTER(C.c_int)]
        _REF = R
    return _REF


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Mesh:
    """Structured Quad9/Hex27 hierarchy (gencase GenCase + printMesh ordering)."""

    def __init__(self, dim, nx0, lo, hi, nlev):
        nx = np.zeros(3, np.int32); nx[:dim] = nx0
        l = np.zeros(3); l[:dim] = lo
        h = np.zeros(3); h[:dim] = hi
        self.h = lib().om_mesh_build(dim, nx, l, h, nlev)
        info = np.zeros(6, np.int32)
        lib().om_mesh_info(self.h, info)
        self.dim, self.nlev, self.n2, self.n1, self.nsides, self.N = (int(v) for v in info)

    def __del__(self):
        if getattr(self, "h", None):
            lib().om_mesh_free(self.h)
            self.h = None

    def level(self, L):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        lib().om_mesh_level(self.h, L, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def coords(self):
        x = np.zeros(self.dim * self.N)
        lib().om_mesh_coords(self.h, x)
        return x.reshape(self.dim, self.N)

    def set_coords(self, xyz):
        """Move the nodes (same connectivity): xyz[d][node]."""
        lib().om_mesh_set_coords(self.h, np.ascontiguousarray(xyz, np.float64).reshape(-1))

    def conn(self, L):
        ne = self.level(L)[0]
        c = np.zeros(ne * self.n2, np.int32)
        nb = np.zeros(ne * self.nsides, np.int32)
        p = np.zeros(ne, np.int32)
        lib().om_mesh_conn(self.h, L, _ptr(c), _ptr(nb), _ptr(p))
        return c.reshape(ne, self.n2), nb.reshape(ne, self.nsides), p

    def node_keys(self):
        a = np.zeros(self.N, np.int32); b = np.zeros(self.N, np.int32); c = np.zeros(self.N, np.int32)
        lib().om_mesh_node_keys(self.h, a, b, c)
        return a, b, c


def fe_tables(kind, dim):
    ns = {(2, 1): 3, (2, 2): 9, (2, 3): 27, (1, 1): 2, (1, 2): 4, (1, 3): 8}[(kind, dim)]
    ng = 3 ** dim
    phi = np.zeros(ns * ng); dphi = np.zeros(dim * ns * ng); w = np.zeros(ng)
    n = C.c_int()
    lib().om_fe_tables(kind, dim, _ptr(phi), _ptr(dphi), _ptr(w), C.byref(n))
    return phi.reshape(ns, ng), dphi.reshape(dim, ns, ng), w


def embedding(kind, dim):
    ns = {2: 3 ** dim, 1: 2 ** dim}[kind]
    E = np.zeros((2 ** dim) * ns * ns)
    lib().om_embedding(kind, dim, E)
    return E.reshape(2 ** dim, ns, ns)


def init_dof(m, L, nq2, nq1):
    d = np.zeros(m.N * (nq2 + nq1), np.int32)
    n = lib().om_init_dof(m.h, L, nq2, nq1, d)
    return d, n


def genbc(m, nq2, nq1, dof_top, fn):
    """fn(xp) -> (neum list, flag list) over the nq2+nq1 variables (bc_read)."""
    nv = nq2 + nq1
    ntop = int(dof_top.max()) + 1

    def cb(xp, neum, flag, user):
        x = [xp[i] for i in range(m.dim)]
        nn = [neum[i] for i in range(nv)]
        ff = [flag[i] for i in range(nv)]
        fn(x, nn, ff)
        for i in range(nv):
            neum[i] = nn[i]; flag[i] = ff[i]

    cfn = BCFN(cb)
    bc0 = np.zeros(ntop, np.int32); bc1 = np.zeros(ntop, np.int32)
    lib().om_genbc(m.h, nq2, nq1, dof_top, cfn, None, bc0, bc1)
    return bc0, bc1


def mat_block(m, L, rkind, ckind):
    nr = m.level(L)[1] if rkind == 2 else m.level(L)[2]
    ln = np.zeros(nr + 1, np.int32)
    nnz = lib().om_mat_block(m.h, L, rkind, ckind, _ptr(ln), None)
    pos = np.zeros(max(nnz, 1), np.int32)
    lib().om_mat_block(m.h, L, rkind, ckind, _ptr(ln), _ptr(pos))
    return ln, pos[:nnz]


def _transfer_block(fn, m, L, kind):
    nr = m.level(L)[1] if kind == 2 else m.level(L)[2]
    ln = np.zeros(nr + 1, np.int32)
    nnz = fn(m.h, L, kind, _ptr(ln), None, None)
    pos = np.zeros(max(nnz, 1), np.int32); val = np.zeros(max(nnz, 1))
    fn(m.h, L, kind, _ptr(ln), _ptr(pos), _ptr(val))
    return ln, pos[:nnz], val[:nnz]


def prol_block(m, L, kind):
    return _transfer_block(lib().om_prol_block, m, L, kind)


def rest_block(m, L, kind):
    return _transfer_block(lib().om_rest_block, m, L, kind)


def system_pattern(m, L, nq2, nq1, dofL):
    n = nq2 * m.level(L)[1] + nq1 * m.level(L)[2]
    rp = np.zeros(n + 1, np.int64)
    nnz = lib().om_system_pattern(m.h, L, nq2, nq1, dofL, _ptr(rp), None)
    ci = np.zeros(nnz, np.int32)
    lib().om_system_pattern(m.h, L, nq2, nq1, dofL, _ptr(rp), _ptr(ci))
    return rp, ci


def system_rest(m, L, nq2, nq1, dof_c, dof_f, dof_top, bc0):
    n = nq2 * m.level(L)[1] + nq1 * m.level(L)[2]
    rp = np.zeros(n + 1, np.int64)
    nnz = lib().om_system_rest(m.h, L, nq2, nq1, dof_c, dof_f, dof_top, bc0, _ptr(rp), None, None)
    ci = np.zeros(nnz, np.int32); v = np.zeros(nnz)
    lib().om_system_rest(m.h, L, nq2, nq1, dof_c, dof_f, dof_top, bc0, _ptr(rp), _ptr(ci), _ptr(v))
    return rp, ci, v


def system_prol(m, L, nq2, nq1, dof_f, dof_c):
    n = nq2 * m.level(L)[1] + nq1 * m.level(L)[2]
    rp = np.zeros(n + 1, np.int64)
    nnz = lib().om_system_prol(m.h, L, nq2, nq1, dof_f, dof_c, _ptr(rp), None, None)
    ci = np.zeros(nnz, np.int32); v = np.zeros(nnz)
    lib().om_system_prol(m.h, L, nq2, nq1, dof_f, dof_c, _ptr(rp), _ptr(ci), _ptr(v))
    return rp, ci, v


def heat_genmatrhs(m, L, mode, p, dofL, doftop, bc0, bc1, xold, rp, ci, k_elem=None, vel_top=None):
    val = np.zeros(len(ci)); rhs = np.zeros(len(rp) - 1)
    k = None if k_elem is None else np.ascontiguousarray(k_elem, np.float64)
    v = None if vel_top is None else np.ascontiguousarray(vel_top, np.float64)
    rc = lib().om_heat_genmatrhs(m.h, L, mode, C.byref(p), dofL, doftop, bc0, bc1, xold, _ptr(k), _ptr(v),
                                 rp, ci, val, _ptr(rhs))
    if rc != 0:
        raise RuntimeError("heat assembly: entry outside pattern")
    return val, rhs


def ns_genmatrhs(m, L, mode, p, dofL, doftop, bc0, bc1, xold, rp, ci, nu_elem=None):
    val = np.zeros(len(ci)); rhs = np.zeros(len(rp) - 1)
    nu = None if nu_elem is None else np.ascontiguousarray(nu_elem, np.float64)
    rc = lib().om_ns_genmatrhs(m.h, L, mode, C.byref(p), dofL, doftop, bc0, bc1, xold, _ptr(nu),
                               rp, ci, val, _ptr(rhs))
    if rc != 0:
        raise RuntimeError("NS assembly: entry outside pattern")
    return val, rhs


def spmv(rp, ci, v, x):
    y = np.zeros(len(rp) - 1)
    lib().om_spmv(len(rp) - 1, rp, ci, v, np.ascontiguousarray(x, np.float64), y)
    return y


def resid(rp, ci, v, b, x):
    r = np.zeros(len(rp) - 1)
    lib().om_resid(len(rp) - 1, rp, ci, v, b, x, r)
    return r


def gmres_ilu(rp, ci, v, b, x0, maxit, restart=30, tol=1e-20, use_ilu=True):
    x = np.array(x0, np.float64)
    acc = C.c_double()
    it = lib().om_gmres_ilu(len(rp) - 1, rp, ci, v, b, x, maxit, restart, tol, int(use_ilu), C.byref(acc))
    return x, it, acc.value


def lambda_max(rp, ci, v, iters=20):
    return lib().om_lambda_max(len(rp) - 1, rp, ci, v, iters)


class Hier:
    """MGSolBase level structures + MGSolve/MGStep (src/MGSolverBase.C:163-289)."""

    def __init__(self, nlev):
        self.h = lib().om_hier_new(nlev)
        self.nlev = nlev

    def __del__(self):
        if getattr(self, "h", None):
            lib().om_hier_free(self.h)
            self.h = None

    def set_A(self, L, rp, ci, v):
        lib().om_hier_set_A(self.h, L, len(rp) - 1, rp, ci, v)

    def set_R(self, L, rp, ci, v, ncols):
        lib().om_hier_set_R(self.h, L, len(rp) - 1, ncols, rp, ci, v)

    def set_P(self, L, rp, ci, v, ncols):
        lib().om_hier_set_P(self.h, L, len(rp) - 1, ncols, rp, ci, v)

    def solve(self, b, x0=None, smoother=0, coarse=0, correction=0, nu_pre=8, nu_post=8, gamma=1,
              max_cycles=40, eps=1e-6, cheb_lo=0.1, cheb_hi=1.05, jacobi_omega=0.6, ref_exit=1):
        o = MGOpts(smoother, coarse, correction, nu_pre, nu_post, gamma, max_cycles, eps, cheb_lo, cheb_hi,
                   jacobi_omega, ref_exit)
        x = np.zeros(len(b)) if x0 is None else np.array(x0, np.float64)
        hist = np.zeros(max_cycles)
        nc = lib().om_mg_solve(self.h, C.byref(o), np.ascontiguousarray(b, np.float64), x, hist)
        return x, hist[:nc]
