"""A hierarchy handed in by the caller (fm_mesh_create, MGMesh's arrays) instead of the library's own boxes:
curved elements (the isoparametric map (x, y + 0.2 x^2, z + 0.3 x y): full inverse Jacobians), the elements of
every level and the nodes of every level block renumbered at random (as another mesh generator would number
them), and the operators optionally handed in too (fm_system_set_pattern / fm_system_set_transfer, ReadMatrix /
ReadRest / ReadProl).  Parity with the oracle on the same curved mesh in gencase numbering, mapped through the
renumbering: the assembled matrices and right-hand sides of every level (1e-12), the transfers (exact), and the
multigrid solutions against direct solves.  The device-pointer vector exchanges round-trip through device
buffers of the caller."""
import functools

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from exfemus_amd import configs

pytestmark = pytest.mark.gpu


def curve(x):
    out = np.array(x, dtype=np.float64, copy=True)
    out[1] = x[1] + 0.2 * x[0] ** 2
    if x.shape[0] == 3:
        out[2] = x[2] + 0.3 * x[0] * x[1]
    return out


def uncurve(xp):
    x = list(xp)
    y = x[1] - 0.2 * x[0] ** 2
    out = [x[0], y]
    if len(x) == 3:
        out.append(x[2] - 0.3 * x[0] * y)
    return out


def caller_mesh(femus, dim, nx, nlev, seed, curved=True, permuted=True):
    """The library's box hierarchy, curved, with the elements of every level and the nodes of every level
    block (the Q1 / Q2 prefixes) renumbered at random; returns the caller mesh, the old -> new node map and,
    per level, the old -> new element map."""
    box = femus.Mesh(dim, nx, [0.0] * dim, [1.0] * dim, nlev)
    rng = np.random.default_rng(seed)
    N, nch = box.N, 2 ** dim
    counts = [box.level(L) for L in range(nlev)]
    # node blocks that keep every prefix: [0, nq1[0]), [nq1[0], nq2[0]), [nq2[L-1], nq2[L])
    cuts = [0, counts[0][2], counts[0][1]] + [counts[L][1] for L in range(1, nlev)]
    new_id = np.arange(N)
    if permuted:
        for a, b in zip(cuts[:-1], cuts[1:]):
            new_id[a:b] = a + rng.permutation(b - a)
    xyz = np.zeros((dim, N))
    xyz[:, new_id] = curve(box.coords()) if curved else box.coords()
    off, emap, eneigh, par, chd, pe = [0], [], [], [], [], []
    for L in range(nlev):
        conn, neigh, parent = box.conn(L)
        ne = conn.shape[0]
        p = rng.permutation(ne) if permuted else np.arange(ne)   # new position of old element e: p[e]
        inv = np.argsort(p)
        pe.append(p)
        emap.append(new_id[conn[inv]])
        nb = neigh[inv]
        eneigh.append(np.where(nb >= 0, p[np.maximum(nb, 0)] + off[-1], -1))
        if L > 0:
            old_par = parent[inv]
            par.append(pe[L - 1][old_par])
            chd.append(inv - old_par * nch)     # the box stores the children of a parent in embedding order
        else:
            par.append(np.full(ne, -1)); chd.append(np.zeros(ne, int))
        off.append(off[-1] + ne)
    m = femus.Mesh.from_arrays(dim, xyz, off, np.concatenate(emap), np.concatenate(eneigh),
                               [c[1] for c in counts], [c[2] for c in counts],
                               np.concatenate(par) if nlev > 1 else None, np.concatenate(chd) if nlev > 1 else None)
    return m, new_id, pe


def oracle_mesh(oracle, dim, nx, nlev, curved=True):
    om = oracle.Mesh(dim, nx, [0.0] * dim, [1.0] * dim, nlev)
    if curved:
        om.set_coords(curve(om.coords()))
    return om


def dof_map(sysg, odofs, new_id, L, nv):
    """g[oracle DoF] = GPU DoF of the same node and variable on level L."""
    N = len(new_id)
    gd = sysg.node_dof(L)
    g = np.zeros(sysg._Dim[L], np.int64)
    for iv in range(nv):
        od = odofs[iv * N:(iv + 1) * N]
        ok = od >= 0
        g[od[ok]] = gd[new_id[np.where(ok)[0]] + iv * N]
    return g


def mapped(rp, ci, v, g_rows, g_cols, shape):
    coo = sp.csr_matrix((v, ci, rp), shape=shape).tocoo()
    return sp.csr_matrix((coo.data, (g_rows[coo.row], g_cols[coo.col])), shape=shape)


def assert_same(A, B, rtol):
    D = (A - B).tocoo()
    scale = abs(B).max()
    assert (abs(D.data).max() if D.nnz else 0.0) <= rtol * scale


@pytest.mark.parametrize("dim,nx,nlev", [(2, [2, 3], 3), (3, [2, 2, 2], 3)])
@pytest.mark.parametrize("user_ops", [False, True])
def test_caller_mesh_heat_curved(femus, oracle, ctx, dim, nx, nlev, user_ops):
    m, new_id, pe = caller_mesh(femus, dim, nx, nlev, seed=dim)
    om = oracle_mesh(oracle, dim, nx, nlev)
    top = nlev - 1
    s = femus.MGSolT(ctx, m, alpha0=1.0, dt=0.1, time_coeff=0.0, source=1.0, upwind=1.0)
    s.GenBc(configs.heat_dirichlet_all)
    odofs = [oracle.init_dof(om, L, 1, 0)[0] for L in range(nlev)]
    g = [dof_map(s, odofs[L], new_id, L, 1) for L in range(nlev)]
    ob0, ob1 = oracle.genbc(om, 1, 0, odofs[top], configs.heat_dirichlet_all)
    pb0, pb1 = s.bc()
    np.testing.assert_array_equal(pb0[g[top]], ob0)
    np.testing.assert_array_equal(pb1[g[top]], ob1)
    if user_ops:
        for L in range(nlev):
            rp, ci = oracle.system_pattern(om, L, 1, 0, odofs[L])
            A = mapped(rp, ci, np.ones(len(ci)), g[L], g[L], (s._Dim[L],) * 2)
            s.set_pattern(L, A.indptr, A.indices)
            if L < top:
                r = oracle.system_rest(om, L, 1, 0, odofs[L], odofs[L + 1], odofs[top], np.ones_like(ob0))
                R = mapped(*r, g[L], g[L + 1], (s._Dim[L], s._Dim[L + 1]))
                s.set_transfer(1, L, R.indptr, R.indices, R.data)     # unmasked: the library applies the mask
            if L > 0:
                q = oracle.system_prol(om, L, 1, 0, odofs[L], odofs[L - 1])
                P = mapped(*q, g[L], g[L - 1], (s._Dim[L], s._Dim[L - 1]))
                s.set_transfer(2, L, P.indptr, P.indices, P.data)
    s.init()
    k_old = configs.coarsen_coef(configs.lognormal_k(om.level(top)[0], sigma=0.7), nlev, 2 ** dim)
    for L in range(nlev):
        k_new = np.zeros_like(k_old[L]); k_new[pe[L]] = k_old[L]
        s.set_elem_coef(L, k_new)
    s.set_xold(np.zeros(s._Dim[top]))
    p = oracle.HeatParams(1.0, 1.0, 0.1, 0.0, 1.0, 1.0, (0.0, 0.0, 0.0))
    for L in range(nlev):
        mode = 1 if L == top else 0
        s.GenMatRhs(0.0, L, mode)
        rp, ci = oracle.system_pattern(om, L, 1, 0, odofs[L])
        ov, orhs = oracle.heat_genmatrhs(om, L, mode, p, odofs[L], odofs[top], ob0, ob1, np.zeros(len(ob0)), rp, ci,
                                         k_elem=k_old[L])
        n = s._Dim[L]
        grp, gci, gv = s.csr(0, L)
        Ag = sp.csr_matrix((gv, gci, grp), shape=(n, n))
        assert_same(Ag, mapped(rp, ci, ov, g[L], g[L], (n, n)), 1e-12)
        if mode == 1:
            np.testing.assert_allclose(s.get_rhs(L)[g[L]], orhs, rtol=1e-12, atol=1e-13 * np.abs(orhs).max())
        if L < top:
            r = oracle.system_rest(om, L, 1, 0, odofs[L], odofs[L + 1], odofs[top], ob0)
            R = mapped(*r, g[L], g[L + 1], (n, s._Dim[L + 1]))
            q = s.csr(1, L)
            assert_same(sp.csr_matrix((q[2], q[1], q[0]), shape=R.shape), R, 0.0)
        if L > 0:
            r = oracle.system_prol(om, L, 1, 0, odofs[L], odofs[L - 1])
            P = mapped(*r, g[L], g[L - 1], (n, s._Dim[L - 1]))
            q = s.csr(2, L)
            assert_same(sp.csr_matrix((q[2], q[1], q[0]), shape=P.shape), P, 0.0)
    rp, ci, v = s.csr(0, top)
    n = s._Dim[top]
    xref = spla.spsolve(sp.csr_matrix((v, ci, rp), shape=(n, n)).tocsc(), s.get_rhs())
    for prec in (femus.FM_PREC_FP32_FGMRES, femus.FM_PREC_FP64_FGMRES):
        s.set_x(np.zeros(n))
        rep = s.MGSolve(opts=s.opts(precision=prec, nu_pre=3, nu_post=3, cheb_lo=0.1, tol=1e-10, max_iter=100))
        assert rep.converged == 1, (rep.iters, rep.rel_resid)
        assert np.linalg.norm(s.get_x() - xref) <= 1e-8 * np.linalg.norm(xref)
    s.close()


@pytest.mark.parametrize("dim,nx,nlev,curved,permuted", [(2, [2, 2], 3, False, False), (2, [2, 2], 3, False, True),
                                                        (2, [2, 2], 3, True, False), (2, [2, 2], 3, True, True),
                                                        (3, [2, 2, 2], 2, True, True)])
def test_caller_mesh_ns_curved_cavity(femus, oracle, ctx, dim, nx, nlev, curved, permuted):
    """The lid-driven cavity on the curved, renumbered caller mesh (the lid and walls found through the inverse
    map): assembly of every level against the oracle, and the Vanka-smoothed multigrid solution against a direct
    solve (velocity 1e-7, pressure up to its constant)."""
    m, new_id, pe = caller_mesh(femus, dim, nx, nlev, seed=10 + dim, curved=curved, permuted=permuted)
    om = oracle_mesh(oracle, dim, nx, nlev, curved=curved)
    top = nlev - 1
    inv = uncurve if curved else (lambda xp: list(xp))

    def bc(xp, neum, flag):
        # the cavity's walls; on the lid every velocity component carries data (the lid moves along the curved
        # top face: a lid velocity with a normal part would make the enclosed problem inconsistent)
        configs.cavity_bc(inv(xp), neum, flag, dim)
        if flag[0] == 4:
            for iv in range(dim):
                flag[iv] = 4

    def ic(xp):
        v = configs.cavity_ic(inv(xp), dim)
        if v[0] != 0.0 and curved:        # tangent of the mapped lid: d/dx (x, 1 + 0.2 x^2, z + 0.3 x)
            x0 = inv(xp)[0]
            v[1] = 0.4 * x0
            if dim == 3:
                v[2] = 0.3
        return v
    ns = femus.MGSolNS(ctx, m, IRe=0.01, dt=0.1, time_coeff=1.0, upwind=0.1)
    ns.GenBc(bc)
    ns.init()
    x0 = ns.GenIc(ic)
    odofs = [oracle.init_dof(om, L, dim, 1)[0] for L in range(nlev)]
    g = [dof_map(ns, odofs[L], new_id, L, dim + 1) for L in range(nlev)]
    ob0, ob1 = oracle.genbc(om, dim, 1, odofs[top], bc)
    np.testing.assert_array_equal(ns.bc()[0][g[top]], ob0)
    np.testing.assert_array_equal(ns.bc()[1][g[top]], ob1)
    assert (ob1 == 4).sum() > 0                  # the lid is found on the curved top face
    xold = x0[g[top]]
    p = oracle.NSParams(0.01, 0.1, 1.0, 0.1, 0.0, (0, 0, 0), (0, 0, 0))
    for L in range(nlev):
        mode = 1 if L == top else 0
        ns.GenMatRhs(0.0, L, mode)
        rp, ci = oracle.system_pattern(om, L, dim, 1, odofs[L])
        ov, orhs = oracle.ns_genmatrhs(om, L, mode, p, odofs[L], odofs[top], ob0, ob1, xold, rp, ci)
        n = ns._Dim[L]
        grp, gci, gv = ns.csr(0, L)
        assert_same(sp.csr_matrix((gv, gci, grp), shape=(n, n)), mapped(rp, ci, ov, g[L], g[L], (n, n)), 1e-12)
        if mode == 1:
            np.testing.assert_allclose(ns.get_rhs(L)[g[L]], orhs, rtol=1e-12, atol=1e-13 * np.abs(orhs).max())
        if L < top:
            r = oracle.system_rest(om, L, dim, 1, odofs[L], odofs[L + 1], odofs[top], ob0)
            R = mapped(*r, g[L], g[L + 1], (n, ns._Dim[L + 1]))
            q = ns.csr(1, L)
            assert_same(sp.csr_matrix((q[2], q[1], q[0]), shape=R.shape), R, 0.0)
        if L > 0:
            r = oracle.system_prol(om, L, dim, 1, odofs[L], odofs[L - 1])
            P = mapped(*r, g[L], g[L - 1], (n, ns._Dim[L - 1]))
            q = ns.csr(2, L)
            assert_same(sp.csr_matrix((q[2], q[1], q[0]), shape=P.shape), P, 0.0)
    o = ns.opts(precision=femus.FM_PREC_FP32_FGMRES, smoother=femus.FM_SMOOTH_VANKA, nu_pre=2, nu_post=2, tol=1e-9,
                max_iter=100, omega=0.5)
    ns.MGSetup(o)
    for L in range(1, nlev):
        # every sampled patch inverse times the patch matrix gathered from the level's CSR is the identity
        grp, gci, gv = ns.csr(0, L)
        AL = sp.csr_matrix((gv, gci, grp), shape=(ns._Dim[L],) * 2).toarray()
        for e in range(0, m.level(L)[0], max(1, m.level(L)[0] // 5)):
            dofs, Kinv = ns.vanka_patch(L, e)
            np.testing.assert_allclose(Kinv @ AL[np.ix_(dofs, dofs)], np.eye(len(dofs)), atol=1e-9)
    rp, ci, v = ns.csr(0, top)
    n = ns._Dim[top]
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    Ap = A.tolil(); b = ns.get_rhs().copy()
    Ap[n - 1, :] = 0.0; Ap[n - 1, n - 1] = 1.0; b[n - 1] = 0.0
    xref = spla.spsolve(Ap.tocsc(), b)
    ns.set_x(x0)
    rep = ns.MGSolve(opts=o)
    assert rep.converged == 1, (rep.iters, rep.rel_resid, list(rep.hist[:12]))
    x = ns.get_x()
    nd = ns.node_dof(top)
    N = m.N
    pd = nd[dim * N:][nd[dim * N:] >= 0]
    vd = np.setdiff1d(np.arange(n), pd)
    assert np.linalg.norm(x[vd] - xref[vd]) <= 1e-7 * np.linalg.norm(xref[vd])
    pm, pr = x[pd] - x[pd].mean(), xref[pd] - xref[pd].mean()
    assert np.linalg.norm(pm - pr) <= 1e-6 * np.linalg.norm(pr)
    ns.close()


def test_caller_pattern_missing_entry_and_bad_mesh(femus, oracle, ctx):
    """A caller's pattern without one element coupling is refused at init with the entry named; inconsistent
    hierarchies are refused by fm_mesh_create with a message; the system stays usable with the full pattern."""
    m, new_id, pe = caller_mesh(femus, 2, [2, 2], 2, seed=5)
    om = oracle_mesh(oracle, 2, [2, 2], 2)
    s = femus.MGSolT(ctx, m, alpha0=1.0, dt=0.1, time_coeff=0.0, source=1.0)
    s.GenBc(configs.heat_dirichlet_all)
    od = oracle.init_dof(om, 1, 1, 0)[0]
    g = dof_map(s, od, new_id, 1, 1)
    rp, ci = oracle.system_pattern(om, 1, 1, 0, od)
    A = mapped(rp, ci, np.ones(len(ci)), g, g, (s._Dim[1],) * 2).tolil()
    i = 7
    A[i, A.rows[i][-1]] = 0.0
    Ab = A.tocsr(); Ab.eliminate_zeros()
    s.set_pattern(1, Ab.indptr, Ab.indices)
    with pytest.raises(femus.FemusError) as e:
        s.init()
    assert e.value.code == -1 and "lacks entry" in str(e.value)
    s.set_pattern(1, None, None)
    s.init()
    s.set_xold(np.zeros(s._Dim[1]))
    for L in range(2):
        s.GenMatRhs(0.0, L, 1 if L == 1 else 0)
    assert s.MGSolve(opts=s.opts(precision=femus.FM_PREC_FP64_FGMRES, tol=1e-10, max_iter=50)).converged == 1
    s.close()
    # a node of the top level used by no element
    xyz = np.zeros((2, 10))
    with pytest.raises(femus.FemusError) as e:
        femus.Mesh.from_arrays(2, xyz, [0, 1], np.arange(9), -np.ones(4, int), [10], [4])
    assert "belongs to no element" in str(e.value)


def _hip():
    """The HIP runtime the library itself is bound to, for device buffers of the caller.

    The symbols are looked up through the library's own handle (dlsym on a handle searches the object and its
    dependencies), so they come from the runtime copy the library uses.  A process can hold two copies (torch
    bundles one with the same soname, and which copy the library binds to depends on the import order); a
    runtime opened by file name can then be the other copy, which does not share the library's device state."""
    import ctypes as C
    from exfemus_amd import femus
    h = C.CDLL(femus.LIBPATH)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def test_device_pointer_vectors(femus, ctx):
    """fm_system_{set,get}_{x,rhs}_dev with device buffers of the caller (hipMalloc) equal the host exchanges on
    every level; a host address is refused."""
    import ctypes as C
    hip = _hip()
    m = femus.Mesh(3, [2, 2, 2], [0, 0, 0], [1, 1, 1], 2)
    s = femus.MGSolT(ctx, m, alpha0=1.0, dt=0.1, time_coeff=0.0, source=1.0)
    s.GenBc(configs.heat_dirichlet_all)
    s.init()
    rng = np.random.default_rng(3)
    H2D, D2H = 1, 2
    for L in (0, 1):
        nL = s._Dim[L]
        buf = C.c_void_p()
        assert hip.hipMalloc(C.byref(buf), nL * 8) == 0
        try:
            a = rng.standard_normal(nL)
            assert hip.hipMemcpy(buf, a.ctypes.data_as(C.c_void_p), nL * 8, H2D) == 0
            s.dev_vector("fm_system_set_x_dev", L, buf.value)
            np.testing.assert_array_equal(s.get_x(L), a)
            s.dev_vector("fm_system_set_rhs_dev", L, buf.value)
            np.testing.assert_array_equal(s.get_rhs(L), a)
            b = rng.standard_normal(nL)
            out = np.zeros(nL)
            s.set_x(b, L)
            s.dev_vector("fm_system_get_x_dev", L, buf.value)
            assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), buf, nL * 8, D2H) == 0
            np.testing.assert_array_equal(out, b)
            s.set_rhs(2 * b, L)
            s.dev_vector("fm_system_get_rhs_dev", L, buf.value)
            assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), buf, nL * 8, D2H) == 0
            np.testing.assert_array_equal(out, 2 * b)
        finally:
            hip.hipFree(buf)
    host = np.zeros(s._Dim[1])
    with pytest.raises(femus.FemusError) as e:
        s.dev_vector("fm_system_set_x_dev", 1, host.ctypes.data)
    assert e.value.code == -1 and "device pointer" in str(e.value)
    s.close()
