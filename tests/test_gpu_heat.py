"""GPU parity of the heat (Q2) path through the C ABI against the CPU oracle.

Mesh hierarchy, DoF map, BC map, sparsity, restriction (with Dirichlet mask)
and prolongation must match bit-exactly; assembled FP64 matrices/RHS to
rtol 1e-12; multigrid solutions to the FP64 direct solve within the stated
tolerances.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from exfemus_amd import configs

pytestmark = pytest.mark.gpu

CASES = [
    (2, [2, 2], 3),
    (3, [2, 2, 2], 3),
    (3, [1, 2, 3], 2),
]


def _build(femus, oracle, ctx, dim, nx, nlev, k_sigma=1.0, source=1.0, time_coeff=0.0):
    lo, hi = [0.0] * dim, [1.0] * dim
    pm = femus.Mesh(dim, nx, lo, hi, nlev)
    om = oracle.Mesh(dim, nx, lo, hi, nlev)
    sysT = femus.MGSolT(ctx, pm, alpha0=1.0, dt=0.1, time_coeff=time_coeff, source=source, upwind=1.0)
    sysT.GenBc(configs.heat_dirichlet_all)
    sysT.init()
    ne_top = pm.level(nlev - 1)[0]
    k_top = configs.lognormal_k(ne_top, sigma=k_sigma)
    ks = configs.coarsen_coef(k_top, nlev, 2 ** dim)
    for L in range(nlev):
        sysT.set_elem_coef(L, ks[L])
    return pm, om, sysT, ks


@pytest.mark.parametrize("dim,nx,nlev", CASES)
def test_mesh_and_maps_bitexact(femus, oracle, ctx, dim, nx, nlev):
    pm, om, sysT, _ = _build(femus, oracle, ctx, dim, nx, nlev)
    assert pm.N == om.N
    np.testing.assert_array_equal(pm.coords(), om.coords())
    for L in range(nlev):
        assert pm.level(L) == om.level(L)
        pc, pn, pp = pm.conn(L)
        oc, on, op = om.conn(L)
        np.testing.assert_array_equal(pc, oc)
        np.testing.assert_array_equal(pn, on)
        np.testing.assert_array_equal(pp, op)
        od, ndof = oracle.init_dof(om, L, 1, 0)
        np.testing.assert_array_equal(sysT.node_dof(L), od)
        assert sysT._Dim[L] == ndof
    od_top, _ = oracle.init_dof(om, nlev - 1, 1, 0)
    ob0, ob1 = oracle.genbc(om, 1, 0, od_top, configs.heat_dirichlet_all)
    pb0, pb1 = sysT.bc()
    np.testing.assert_array_equal(pb0, ob0)
    np.testing.assert_array_equal(pb1, ob1)


@pytest.mark.parametrize("dim,nx,nlev", CASES)
def test_operators_bitexact(femus, oracle, ctx, dim, nx, nlev):
    pm, om, sysT, _ = _build(femus, oracle, ctx, dim, nx, nlev)
    dofs = [oracle.init_dof(om, L, 1, 0)[0] for L in range(nlev)]
    top = nlev - 1
    ob0, _ = oracle.genbc(om, 1, 0, dofs[top], configs.heat_dirichlet_all)
    for L in range(nlev):
        rp, ci = oracle.system_pattern(om, L, 1, 0, dofs[L])
        prp, pci, _ = sysT.csr(0, L)
        np.testing.assert_array_equal(prp, rp)
        np.testing.assert_array_equal(pci, ci)
        if L < top:
            rrp, rci, rv = oracle.system_rest(om, L, 1, 0, dofs[L], dofs[L + 1], dofs[top], ob0)
            q = sysT.csr(1, L)
            np.testing.assert_array_equal(q[0], rrp)
            np.testing.assert_array_equal(q[1], rci)
            np.testing.assert_array_equal(q[2], rv)
        if L > 0:
            prp2, pci2, pv2 = oracle.system_prol(om, L, 1, 0, dofs[L], dofs[L - 1])
            q = sysT.csr(2, L)
            np.testing.assert_array_equal(q[0], prp2)
            np.testing.assert_array_equal(q[1], pci2)
            np.testing.assert_array_equal(q[2], pv2)


@pytest.mark.parametrize("dim,nx,nlev", CASES)
@pytest.mark.parametrize("time_coeff", [0.0, 1.0])
def test_assembly_parity(femus, oracle, ctx, dim, nx, nlev, time_coeff):
    pm, om, sysT, ks = _build(femus, oracle, ctx, dim, nx, nlev, time_coeff=time_coeff)
    dofs = [oracle.init_dof(om, L, 1, 0)[0] for L in range(nlev)]
    top = nlev - 1
    ob0, ob1 = oracle.genbc(om, 1, 0, dofs[top], configs.heat_dirichlet_all)
    rng = np.random.default_rng(7)
    xold = rng.standard_normal(sysT._Dim[top])
    sysT.set_xold(xold)
    p = oracle.HeatParams(1.0, 1.0, 0.1, time_coeff, 1.0, 1.0, (0.0, 0.0, 0.0))
    for L in range(nlev):
        mode = 1 if L == top else 0
        sysT.GenMatRhs(0.0, L, mode)
        rp, ci = oracle.system_pattern(om, L, 1, 0, dofs[L])
        ov, orhs = oracle.heat_genmatrhs(om, L, mode, p, dofs[L], dofs[top], ob0, ob1, xold, rp, ci, k_elem=ks[L])
        _, _, pv = sysT.csr(0, L)
        scale = np.abs(ov).max()
        np.testing.assert_allclose(pv, ov, rtol=1e-12, atol=1e-13 * scale)
        if mode == 1:
            np.testing.assert_allclose(sysT.get_rhs(L), orhs, rtol=1e-12, atol=1e-13 * np.abs(orhs).max())


def _direct(sysT, L):
    rp, ci, v = sysT.csr(0, L)
    n = len(rp) - 1
    A = sp.csr_matrix((v, ci, rp), shape=(n, n))
    return spla.spsolve(A.tocsc(), sysT.get_rhs(L)), A


@pytest.mark.parametrize("dim,nx,nlev", [(2, [2, 2], 4), (3, [2, 2, 2], 3)])
@pytest.mark.parametrize("prec", [0, 1, 2])
def test_mg_solution(femus, oracle, ctx, dim, nx, nlev, prec):
    pm, om, sysT, ks = _build(femus, oracle, ctx, dim, nx, nlev)
    top = nlev - 1
    sysT.set_xold(np.zeros(sysT._Dim[top]))
    for L in range(nlev):
        sysT.GenMatRhs(0.0, L, 1 if L == top else 0)
    xref, A = _direct(sysT, top)
    sysT.set_x(np.zeros(sysT._Dim[top]))
    rep = sysT.MGSolve(opts=sysT.opts(precision=prec, tol=1e-10, max_iter=60))
    x = sysT.get_x()
    b = sysT.get_rhs()
    rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert rep.converged == 1, (rep.iters, rep.rel_resid)
    assert rel <= 1.5e-10
    err = np.linalg.norm(x - xref) / np.linalg.norm(xref)
    assert err <= 1e-8


@pytest.mark.parametrize("dim,nx,nlev", [(2, [2, 2], 3), (3, [2, 2, 2], 3)])
def test_mg_matches_oracle_cycles(femus, oracle, ctx, dim, nx, nlev):
    """Same V-cycle (Chebyshev smoother, reference correction rule, LU coarse) on GPU and oracle:
    the FP64 residual histories agree to fp tolerance."""
    pm, om, sysT, ks = _build(femus, oracle, ctx, dim, nx, nlev)
    top = nlev - 1
    sysT.set_xold(np.zeros(sysT._Dim[top]))
    for L in range(nlev):
        sysT.GenMatRhs(0.0, L, 1 if L == top else 0)
    H = oracle.Hier(nlev)
    for L in range(nlev):
        rp, ci, v = sysT.csr(0, L)
        H.set_A(L, rp, ci, v)
        if L < top:
            r = sysT.csr(1, L)
            H.set_R(L, r[0], r[1], r[2], sysT._Dim[L + 1])
        if L > 0:
            q = sysT.csr(2, L)
            H.set_P(L, q[0], q[1], q[2], sysT._Dim[L - 1])
    b = sysT.get_rhs()
    for corr in (0, 1):
        xo, hist_o = H.solve(b, smoother=1, coarse=1, correction=corr, nu_pre=2, nu_post=2, max_cycles=6,
                             eps=1e-30, cheb_lo=0.2, cheb_hi=1.05, ref_exit=0)
        sysT.set_x(np.zeros(sysT._Dim[top]))
        rep = sysT.MGSolve(opts=sysT.opts(precision=0, correction=corr, tol=1e-30, max_iter=6, cheb_lo=0.2,
                                          cheb_hi=1.05))
        hist_g = np.array(rep.hist[:rep.iters])
        assert rep.iters == len(hist_o)
        np.testing.assert_allclose(hist_g, hist_o, rtol=1e-6)
        np.testing.assert_allclose(sysT.get_x(), xo, rtol=1e-7, atol=1e-12 * np.abs(xo).max())


@pytest.mark.parametrize("smoother", [0, 1])
@pytest.mark.parametrize("prec", [0, 1, 2, 3])
def test_graph_replay_matches_eager(femus, oracle, ctx, prec, smoother):
    """V-cycles replayed from the captured HIP graph give bitwise the same iterates as eager launches,
    over repeated solves with eager outer-loop work and host synchronisation between the replays
    (Chebyshev and damped Jacobi smoothing; the Jacobi sweep's copy runs as a kernel inside the graph)."""
    pm, om, sysT, ks = _build(femus, oracle, ctx, 3, [2, 2, 2], 3)
    top = 2
    sysT.set_xold(np.zeros(sysT._Dim[top]))
    for L in range(3):
        sysT.GenMatRhs(0.0, L, 1 if L == top else 0)
    res = {}
    for g in (0, 1, 1, 0, 1):
        sysT.set_x(np.zeros(sysT._Dim[top]))
        rep = sysT.MGSolve(opts=sysT.opts(precision=prec, smoother=smoother, tol=1e-10, max_iter=200,
                                          use_graph=g))
        x = sysT.get_x()
        key = (rep.iters, tuple(rep.hist[:rep.iters]))
        res.setdefault(g, []).append((key, x))
    ref_key, ref_x = res[0][0]
    for g in (0, 1):
        for key, x in res[g]:
            assert key == ref_key
            np.testing.assert_array_equal(x, ref_x)


@pytest.mark.parametrize("dim,nx,nlev", CASES)
def test_level_operators_vs_csr(femus, oracle, ctx, dim, nx, nlev):
    """NumericVectorM::resid / matrix_mult(R), matrix_mult(P) (contrib/matrix/laspack_vectorM.C:147-189) through
    the C ABI: the device runs them on SELL copies in its internal DoF order; vectors cross the ABI in the
    reference numbering and must equal the same products formed with the reference-order CSR matrices."""
    pm, om, sysT, ks = _build(femus, oracle, ctx, dim, nx, nlev)
    top = nlev - 1
    sysT.set_xold(np.zeros(sysT._Dim[top]))
    for L in range(nlev):
        sysT.GenMatRhs(0.0, L, 1 if L == top else 0)
    rng = np.random.default_rng(11)
    for L in range(nlev):
        n = sysT._Dim[L]
        rp, ci, v = sysT.csr(0, L)
        A = sp.csr_matrix((v, ci, rp), shape=(n, n))
        x = rng.standard_normal(n)
        b = rng.standard_normal(n)
        # y = A x through the CSR kernel (reference order) against the oracle's LASPACK-order product
        np.testing.assert_allclose(sysT.spmv(L, x), oracle.spmv(rp, ci, v, x), rtol=1e-13, atol=1e-13 * np.abs(v).sum())
        sysT.set_x(x, L)
        sysT.set_rhs(b, L)
        np.testing.assert_array_equal(sysT.get_x(L), x)
        np.testing.assert_array_equal(sysT.get_rhs(L), b)
        sysT.resid(L)
        res = sysT.get_res(L)
        ref = b - A @ x
        np.testing.assert_allclose(res, ref, rtol=1e-12, atol=1e-13 * (np.abs(b).max() + np.abs(A).sum(axis=1).max()))
        if L > 0:
            sysT.restrict(L)
            rr, rc, rv = sysT.csr(1, L - 1)
            R = sp.csr_matrix((rv, rc, rr), shape=(sysT._Dim[L - 1], n))
            np.testing.assert_allclose(sysT.get_rhs(L - 1), R @ res, rtol=1e-12, atol=1e-12 * np.abs(res).max())
            xc = rng.standard_normal(sysT._Dim[L - 1])
            sysT.set_x(xc, L - 1)
            sysT.prolong(L)
            pr, pc, pv = sysT.csr(2, L)
            P = sp.csr_matrix((pv, pc, pr), shape=(n, sysT._Dim[L - 1]))
            np.testing.assert_allclose(sysT.get_res(L), P @ xc, rtol=1e-12, atol=1e-12 * np.abs(xc).max())


def test_split_slice_products_match(femus, oracle, ctx):
    """The level products with 2, 4 or 8 wavefronts sharing a slice (the launch used for levels with few
    slices) against the one-wave-per-slice launch and the reference-order CSR products: the threshold W is
    swept over powers of two so that every level runs with S = 1, 2, 4 and 8 in turn."""
    pm, om, sysT, ks = _build(femus, oracle, ctx, 3, [2, 2, 2], 3)
    top = 2
    sysT.set_xold(np.zeros(sysT._Dim[top]))
    for L in range(3):
        sysT.GenMatRhs(0.0, L, 1 if L == top else 0)
    rng = np.random.default_rng(5)
    lib = femus.lib()
    W0 = lib.fm_set_sell_split(-1)
    try:
        for L in range(3):
            n, nc = sysT._Dim[L], sysT._Dim[L - 1] if L > 0 else 0
            x, b = rng.standard_normal(n), rng.standard_normal(n)
            xc = rng.standard_normal(nc) if L > 0 else None
            rp, ci, v = sysT.csr(0, L)
            ref_res = b - sp.csr_matrix((v, ci, rp), shape=(n, n)) @ x
            outs = {}
            for W in [0] + [2 ** k for k in range(0, 17)] + [2 ** 30]:
                lib.fm_set_sell_split(W)
                sysT.set_x(x, L)
                sysT.set_rhs(b, L)
                sysT.resid(L)
                res = sysT.get_res(L)
                got = [res]
                if L > 0:
                    sysT.restrict(L)
                    got.append(sysT.get_rhs(L - 1))
                    sysT.set_x(xc, L - 1)
                    sysT.prolong(L)
                    got.append(sysT.get_res(L))
                outs[W] = got
                np.testing.assert_allclose(res, ref_res, rtol=1e-12,
                                           atol=1e-13 * (np.abs(b).max() + np.abs(v).sum()))
            for W, got in outs.items():
                for a, r in zip(got, outs[0]):
                    np.testing.assert_allclose(a, r, rtol=1e-13, atol=1e-13 * np.abs(r).max(), err_msg=f"L={L} W={W}")
    finally:
        lib.fm_set_sell_split(W0)
