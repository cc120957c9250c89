"""GPU parity of the coupled Q2-Q1 Navier-Stokes path (MGSolNS) against the oracle.

Bit-exact: mesh, DoF, BC maps, sparsity, restriction/prolongation.  Assembled
Oseen matrices and RHS (from a non-trivial x_old) to rtol 1e-12.  Multigrid
with the Vanka smoother: relative residual <= 1e-8 and the solution equal to
the FP64 direct solve (velocity to 1e-7 relative, pressure up to its constant
for the enclosed cavity).  Picard / pseudo-time steps (MGTimeStep) match the
oracle's steps with direct solves.
"""
import functools

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from exfemus_amd import configs

pytestmark = pytest.mark.gpu

CASES = [
    ("cavity2d", 2, [2, 2], 3),
    ("cavity3d", 3, [1, 1, 1], 3),
    ("channel3d", 3, [4, 2, 2], 2),
]


def _setup(femus, oracle, ctx, case, dim, nx, nlev, IRe=0.01):
    if case == "channel3d":
        lo, hi = [0, 0, 0], [4, 1, 1]
        bc = configs.channel_bc
        ic = configs.channel_ic
    else:
        lo, hi = [0] * dim, [1] * dim
        bc = functools.partial(configs.cavity_bc, dim=dim)
        ic = functools.partial(configs.cavity_ic, dim=dim)
    pm = femus.Mesh(dim, nx, lo, hi, nlev)
    om = oracle.Mesh(dim, nx, lo, hi, nlev)
    ns = femus.MGSolNS(ctx, pm, IRe=IRe, dt=0.1, time_coeff=1.0, upwind=0.1)
    ns.GenBc(bc)
    ns.init()
    x0 = ns.GenIc(ic)
    return pm, om, ns, bc, x0


def _oracle_maps(oracle, om, dim, nlev, bc):
    dofs = [oracle.init_dof(om, L, dim, 1)[0] for L in range(nlev)]
    b0, b1 = oracle.genbc(om, dim, 1, dofs[nlev - 1], bc)
    return dofs, b0, b1


@pytest.mark.parametrize("case,dim,nx,nlev", CASES)
def test_ns_maps_and_operators_bitexact(femus, oracle, ctx, case, dim, nx, nlev):
    pm, om, ns, bc, _ = _setup(femus, oracle, ctx, case, dim, nx, nlev)
    top = nlev - 1
    dofs, b0, b1 = _oracle_maps(oracle, om, dim, nlev, bc)
    pb0, pb1 = ns.bc()
    np.testing.assert_array_equal(pb0, b0)
    np.testing.assert_array_equal(pb1, b1)
    for L in range(nlev):
        np.testing.assert_array_equal(ns.node_dof(L), dofs[L])
        rp, ci = oracle.system_pattern(om, L, dim, 1, dofs[L])
        q = ns.csr(0, L)
        np.testing.assert_array_equal(q[0], rp)
        np.testing.assert_array_equal(q[1], ci)
        if L < top:
            r = oracle.system_rest(om, L, dim, 1, dofs[L], dofs[L + 1], dofs[top], b0)
            q = ns.csr(1, L)
            for a, b in zip(q, r):
                np.testing.assert_array_equal(a, b)
        if L > 0:
            r = oracle.system_prol(om, L, dim, 1, dofs[L], dofs[L - 1])
            q = ns.csr(2, L)
            for a, b in zip(q, r):
                np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("case,dim,nx,nlev", CASES)
def test_ns_assembly_parity(femus, oracle, ctx, case, dim, nx, nlev):
    pm, om, ns, bc, x0 = _setup(femus, oracle, ctx, case, dim, nx, nlev)
    top = nlev - 1
    dofs, b0, b1 = _oracle_maps(oracle, om, dim, nlev, bc)
    # a non-trivial old velocity (keeps the Dirichlet values of the IC)
    rng = np.random.default_rng(11)
    xold = x0 + 0.3 * rng.standard_normal(len(x0)) * (b0 == 1)
    ns.set_xold(xold)
    nu = None
    if case == "cavity2d":   # per-element viscosity field on every level
        nu = configs.coarsen_coef(0.01 * configs.lognormal_k(pm.level(top)[0], sigma=0.5), nlev, 2 ** dim)
        for L in range(nlev):
            ns.set_elem_coef(L, nu[L])
    p = oracle.NSParams(0.01, 0.1, 1.0, 0.1, 0.0, (0, 0, 0), (0, 0, 0))
    for L in range(nlev):
        mode = 1 if L == top else 0
        ns.GenMatRhs(0.0, L, mode)
        rp, ci = oracle.system_pattern(om, L, dim, 1, dofs[L])
        ov, orhs = oracle.ns_genmatrhs(om, L, mode, p, dofs[L], dofs[top], b0, b1, xold, rp, ci,
                                       nu_elem=None if nu is None else nu[L])
        _, _, pv = ns.csr(0, L)
        np.testing.assert_allclose(pv, ov, rtol=1e-12, atol=1e-13 * np.abs(ov).max())
        if mode == 1:
            np.testing.assert_allclose(ns.get_rhs(L), orhs, rtol=1e-12, atol=1e-13 * np.abs(orhs).max())


def _ref_solve(ns, L, enclosed, dim):
    rp, ci, v = ns.csr(0, L)
    n = len(rp) - 1
    A = sp.csr_matrix((v, ci, rp), shape=(n, n)).tolil()
    b = ns.get_rhs(L).copy()
    if enclosed:     # pin the last pressure DoF for the reference direct solve
        A[n - 1, :] = 0.0
        A[n - 1, n - 1] = 1.0
        b[n - 1] = 0.0
    return spla.spsolve(A.tocsc(), b), sp.csr_matrix((v, ci, rp), shape=(n, n))


def _compare(ns, x, xref, dim, enclosed):
    nvel = dim * ns.mesh.level(ns._NoLevels - 1)[1]
    ev = np.linalg.norm(x[:nvel] - xref[:nvel]) / np.linalg.norm(xref[:nvel])
    p, pr = x[nvel:], xref[nvel:]
    if enclosed:
        p = p - p.mean(); pr = pr - pr.mean()
    ep = np.linalg.norm(p - pr) / max(np.linalg.norm(pr), 1e-300)
    return ev, ep


@pytest.mark.parametrize("case,dim,nx,nlev", [("cavity2d", 2, [2, 2], 3), ("channel3d", 3, [4, 2, 2], 2)])
@pytest.mark.parametrize("prec", [3, 2])
def test_ns_mg_vanka_solution(femus, oracle, ctx, case, dim, nx, nlev, prec):
    pm, om, ns, bc, x0 = _setup(femus, oracle, ctx, case, dim, nx, nlev)
    top = nlev - 1
    for L in range(nlev):
        ns.GenMatRhs(0.0, L, 1 if L == top else 0)
    enclosed = case.startswith("cavity")
    xref, A = _ref_solve(ns, top, enclosed, dim)
    ns.set_x(x0)
    o = ns.opts(precision=prec, smoother=femus.FM_SMOOTH_VANKA, nu_pre=2, nu_post=2, tol=1e-9, max_iter=80,
                omega=0.5)
    rep = ns.MGSolve(opts=o)
    x = ns.get_x()
    b = ns.get_rhs()
    assert rep.converged == 1, (rep.iters, rep.rel_resid, list(rep.hist[:rep.iters + 1]))
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) <= 2e-9
    ev, ep = _compare(ns, x, xref, dim, enclosed)
    assert ev < 1e-7 and ep < 1e-6, (ev, ep)


def test_vanka_store_mask_and_release(femus, oracle, ctx):
    """The Vanka patch inverses kept in one precision only: a solve whose cycles need the other precision fails
    with FM_ESTATE and leaves the system usable; FP32-only inverses give the mixed solve the same iterations as
    both stores; after fm_mg_release the next solve sets the hierarchy up again."""
    pm, om, ns, bc, x0 = _setup(femus, oracle, ctx, "channel3d", 3, [4, 2, 2], 2)
    for L in range(2):
        ns.GenMatRhs(0.0, L, 1 if L == 1 else 0)
    kw = dict(smoother=femus.FM_SMOOTH_VANKA, nu_pre=2, nu_post=2, tol=1e-9, max_iter=80, omega=0.5)
    o32, o64 = ns.opts(precision=femus.FM_PREC_FP32_FGMRES, **kw), ns.opts(precision=femus.FM_PREC_FP64_FGMRES, **kw)

    def solve(o):
        ns.set_x(x0)
        return ns.MGSolve(opts=o), ns.get_x()

    r32, x32 = solve(o32)
    r64, x64 = solve(o64)
    assert ns.set_vanka_store(2) == 3
    ns.release_mg()
    r, x = solve(o32)
    assert r.iters == r32.iters and np.array_equal(x, x32)
    with pytest.raises(femus.FemusError) as e:
        solve(o64)
    assert e.value.code == -5 and "FP64" in str(e.value)
    assert ns.set_vanka_store(3) == 2
    r, x = solve(o64)
    assert r.iters == r64.iters and np.array_equal(x, x64)
    ns.close()


def test_ns_picard_steps_match_oracle(femus, oracle, ctx):
    """Three MGTimeStep (assemble all levels, MGSolve, x_old <- x) on the cavity vs the oracle with
    direct solves (MGSolverNS.C:722-756)."""
    dim, nx, nlev = 2, [2, 2], 3
    pm, om, ns, bc, x0 = _setup(femus, oracle, ctx, "cavity2d", dim, nx, nlev, IRe=0.01)
    top = nlev - 1
    dofs, b0, b1 = _oracle_maps(oracle, om, dim, nlev, bc)
    p = oracle.NSParams(0.01, 0.1, 1.0, 0.1, 0.0, (0, 0, 0), (0, 0, 0))
    o = ns.opts(precision=3, smoother=femus.FM_SMOOTH_VANKA, nu_pre=2, nu_post=2, tol=1e-11, max_iter=80,
                omega=0.5)
    xo = x0.copy()
    rp, ci = oracle.system_pattern(om, top, dim, 1, dofs[top])
    n = len(rp) - 1
    for step in range(3):
        ns.set_x(xo)
        ns.set_xold(xo)
        rep = ns.MGTimeStep(0.1 * step, opts=o)
        assert rep.converged == 1
        xg = ns.get_x()
        v, rhs = oracle.ns_genmatrhs(om, top, 1, p, dofs[top], dofs[top], b0, b1, xo, rp, ci)
        A = sp.csr_matrix((v, ci, rp), shape=(n, n)).tolil()
        A[n - 1, :] = 0.0; A[n - 1, n - 1] = 1.0
        rhs[n - 1] = 0.0
        xo_new = spla.spsolve(A.tocsc(), rhs)
        ev, ep = _compare(ns, xg, xo_new, dim, True)
        assert ev < 1e-8 and ep < 1e-6, (step, ev, ep)
        # continue both from the oracle's iterate (pressure constant of the GPU one is arbitrary)
        xo = xo_new


def test_ns_graph_replay_matches_eager(femus, oracle, ctx):
    pm, om, ns, bc, x0 = _setup(femus, oracle, ctx, "cavity2d", 2, [2, 2], 3)
    for L in range(3):
        ns.GenMatRhs(0.0, L, 1 if L == 2 else 0)
    out = []
    for g in (0, 1, 1):
        ns.set_x(x0)
        rep = ns.MGSolve(opts=ns.opts(precision=2, smoother=femus.FM_SMOOTH_VANKA, tol=1e-9, max_iter=60, use_graph=g))
        out.append((rep.iters, tuple(rep.hist[:rep.iters]), ns.get_x()))
    for it, h, x in out[1:]:
        assert (it, h) == out[0][:2]
        np.testing.assert_array_equal(x, out[0][2])


@pytest.mark.parametrize("case,dim,nx,nlev", CASES)
def test_ns_level_operators_vs_csr(femus, oracle, ctx, case, dim, nx, nlev):
    """Residual, restriction and prolongation of the coupled system through the C ABI (SELL copies in the
    device's internal node-interleaved order) against the reference-order CSR products."""
    pm, om, ns, bc, x0 = _setup(femus, oracle, ctx, case, dim, nx, nlev)
    top = nlev - 1
    for L in range(nlev):
        ns.GenMatRhs(0.0, L, 1 if L == top else 0)
    rng = np.random.default_rng(12)
    for L in range(nlev):
        n = ns._Dim[L]
        rp, ci, v = ns.csr(0, L)
        A = sp.csr_matrix((v, ci, rp), shape=(n, n))
        x = rng.standard_normal(n)
        b = rng.standard_normal(n)
        ns.set_x(x, L)
        ns.set_rhs(b, L)
        ns.resid(L)
        res = ns.get_res(L)
        np.testing.assert_allclose(res, b - A @ x, rtol=1e-12,
                                   atol=1e-13 * (np.abs(b).max() + np.abs(A).sum(axis=1).max()))
        if L > 0:
            ns.restrict(L)
            rr, rc, rv = ns.csr(1, L - 1)
            R = sp.csr_matrix((rv, rc, rr), shape=(ns._Dim[L - 1], n))
            np.testing.assert_allclose(ns.get_rhs(L - 1), R @ res, rtol=1e-12, atol=1e-12 * np.abs(res).max())
            xc = rng.standard_normal(ns._Dim[L - 1])
            ns.set_x(xc, L - 1)
            ns.prolong(L)
            pr, pc, pv = ns.csr(2, L)
            P = sp.csr_matrix((pv, pc, pr), shape=(n, ns._Dim[L - 1]))
            np.testing.assert_allclose(ns.get_res(L), P @ xc, rtol=1e-12, atol=1e-12 * np.abs(xc).max())


def test_ns_channel_re400_picard_steps_match_oracle(femus, oracle, ctx):
    """The channel at C3's Reynolds number (IRe = 1/400, the reference's upwinding 0.1), 8 x 4 x 4 top: the
    assembled Oseen system of every level against the oracle (1e-12), and three MGTimeStep pseudo-time steps
    (mixed-precision multigrid with the coloured Vanka smoother) against the oracle's steps with direct solves."""
    IRe = 1.0 / 400.0
    dim, nx, nlev = 3, [4, 2, 2], 2
    pm, om, ns, bc, x0 = _setup(femus, oracle, ctx, "channel3d", dim, nx, nlev, IRe=IRe)
    top = nlev - 1
    dofs, b0, b1 = _oracle_maps(oracle, om, dim, nlev, bc)
    p = oracle.NSParams(IRe, 0.1, 1.0, 0.1, 0.0, (0, 0, 0), (0, 0, 0))
    rng = np.random.default_rng(13)
    xr = x0 + 0.3 * rng.standard_normal(len(x0)) * (b0 == 1)
    ns.set_xold(xr)
    for L in range(nlev):
        mode = 1 if L == top else 0
        ns.GenMatRhs(0.0, L, mode)
        rp, ci = oracle.system_pattern(om, L, dim, 1, dofs[L])
        ov, orhs = oracle.ns_genmatrhs(om, L, mode, p, dofs[L], dofs[top], b0, b1, xr, rp, ci)
        np.testing.assert_allclose(ns.csr(0, L)[2], ov, rtol=1e-12, atol=1e-13 * np.abs(ov).max())
        if mode == 1:
            np.testing.assert_allclose(ns.get_rhs(L), orhs, rtol=1e-12, atol=1e-13 * np.abs(orhs).max())
    o = ns.opts(precision=femus.FM_PREC_FP32_FGMRES, smoother=femus.FM_SMOOTH_VANKA, vanka_colored=1, nu_pre=3,
                nu_post=3, coarse_mult=1, tol=1e-11, max_iter=100)
    rp, ci = oracle.system_pattern(om, top, dim, 1, dofs[top])
    n = len(rp) - 1
    xo = x0.copy()
    for step in range(3):
        ns.set_x(xo)
        ns.set_xold(xo)
        rep = ns.MGTimeStep(0.1 * step, opts=o)
        assert rep.converged == 1, (step, rep.iters, rep.rel_resid)
        xg = ns.get_x()
        v, rhs = oracle.ns_genmatrhs(om, top, 1, p, dofs[top], dofs[top], b0, b1, xo, rp, ci)
        xo_new = spla.spsolve(sp.csr_matrix((v, ci, rp), shape=(n, n)).tocsc(), rhs)
        ev, ep = _compare(ns, xg, xo_new, dim, False)
        assert ev < 1e-9 and ep < 1e-8, (step, ev, ep)
        xo = xo_new
    ns.close()
