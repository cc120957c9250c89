"""Dense inverse of the coarsest level (blocked exchange inversion with partial pivoting, coarse.hip).

On a one-level hierarchy one V-cycle is exactly x = A^-1 b, so the solve is compared with a sparse direct
solve.  Sizes cover a ragged last panel (729 = 11 * 64 + 25), an open flow (nonsingular saddle point) and an
enclosed flow whose pressure is defined up to a constant: there the last pressure DoF is pinned to 0 and the
remaining system must be solved exactly."""
import functools

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

pytestmark = pytest.mark.gpu


def _csr(s):
    rp, ci, v = s.csr(0, 0)
    n = len(rp) - 1
    return sp.csr_matrix((v, ci, rp), shape=(n, n))


def test_coarse_inverse_heat_ragged_panel(femus, ctx):
    from exfemus_amd import configs
    m = femus.Mesh(3, [4, 4, 4], [0, 0, 0], [1, 1, 1], 1)
    s = femus.MGSolT(ctx, m, alpha0=1.0, dt=0.1, time_coeff=0.0, source=1.0, upwind=1.0)
    s.GenBc(configs.heat_dirichlet_all)
    s.init()
    s.set_elem_coef(0, configs.lognormal_k(m.level(0)[0], seed=3))
    s.set_xold(np.zeros(s._Dim[0]))
    s.GenMatRhs(0.0, 0, 1)
    assert s._Dim[0] == 729
    A, b = _csr(s), s.get_rhs()
    xr = spla.spsolve(A.tocsc(), b)
    s.set_x(np.zeros(729))
    rep = s.MGSolve(opts=s.opts(precision=femus.FM_PREC_FP64, tol=1e-30, max_iter=1))
    assert rep.iters == 1
    x = s.get_x()
    assert np.linalg.norm(x - xr) <= 1e-12 * np.linalg.norm(xr)
    # FP32 copy of the inverse inside FP64 refinement converges to the same solution
    s.set_x(np.zeros(729))
    rep = s.MGSolve(opts=s.opts(precision=femus.FM_PREC_FP32_IR, tol=1e-12, max_iter=20))
    assert rep.converged == 1
    assert np.linalg.norm(s.get_x() - xr) <= 1e-11 * np.linalg.norm(xr)
    s.close()


def test_coarse_inverse_open_channel(femus, ctx):
    from exfemus_amd import configs
    # two elements across: with one, Q2-Q1 on this box has spurious pressure modes (three zero singular
    # values) and the inverse correctly reports a singular matrix
    m = femus.Mesh(3, [4, 2, 2], [0, 0, 0], [4, 1, 1], 1)
    s = femus.MGSolNS(ctx, m, IRe=1.0 / 400, dt=0.1, time_coeff=0.0, upwind=0.1)
    s.GenBc(configs.channel_bc)
    s.init()
    s.GenIc(configs.channel_ic)
    s.GenMatRhs(0.0, 0, 1)
    n = s._Dim[0]
    A, b = _csr(s), s.get_rhs()
    xr = spla.spsolve(A.tocsc(), b)
    s.set_x(np.zeros(n))
    rep = s.MGSolve(opts=s.opts(precision=femus.FM_PREC_FP64, smoother=femus.FM_SMOOTH_VANKA, tol=1e-30, max_iter=1))
    assert rep.iters == 1
    assert np.linalg.norm(s.get_x() - xr) <= 1e-10 * np.linalg.norm(xr)
    s.close()


def test_coarse_inverse_enclosed_cavity_pinned(femus, ctx):
    from exfemus_amd import configs
    m = femus.Mesh(2, [8, 8], [0, 0], [1, 1], 1)
    s = femus.MGSolNS(ctx, m, IRe=0.01, dt=0.1, time_coeff=0.0, upwind=0.1)
    s.GenBc(functools.partial(configs.cavity_bc, dim=2))
    s.init()
    s.GenIc(functools.partial(configs.cavity_ic, dim=2))
    s.GenMatRhs(0.0, 0, 1)
    n = s._Dim[0]
    A, b = _csr(s), s.get_rhs()
    s.set_x(np.zeros(n))
    rep = s.MGSolve(opts=s.opts(precision=femus.FM_PREC_FP64, smoother=femus.FM_SMOOTH_VANKA, tol=1e-30, max_iter=1))
    assert rep.iters == 1
    x = s.get_x()
    assert x[n - 1] == 0.0                                   # the pinned pressure DoF (last reference DoF)
    assert np.linalg.norm(A @ x - b) <= 1e-10 * np.linalg.norm(b)
    # the same solution from a direct solve with that DoF removed and one continuity row dropped
    keep = np.arange(n - 1)
    xr = spla.spsolve(A[keep][:, keep].tocsc(), b[keep])
    np.testing.assert_allclose(x[:n - 1], xr, rtol=0, atol=1e-9 * np.abs(xr).max())
    s.close()


@pytest.mark.parametrize("dim,nx,nlev", [(2, [2, 2], 3), (3, [2, 2, 2], 2)])
def test_vanka_patch_inverses(femus, ctx, dim, nx, nlev):
    """Every stored Vanka patch inverse (register-tiled exchange inversion with partial pivoting) inverts the
    patch matrix gathered from the assembled level matrix, on every level."""
    from exfemus_amd import configs
    m = femus.Mesh(dim, nx, [0] * dim, [1] * dim, nlev)
    s = femus.MGSolNS(ctx, m, IRe=0.01, dt=0.1, time_coeff=1.0, upwind=0.1)
    s.GenBc(functools.partial(configs.cavity_bc, dim=dim))
    s.init()
    s.GenIc(functools.partial(configs.cavity_ic, dim=dim))
    top = nlev - 1
    for L in range(nlev):
        s.GenMatRhs(0.0, L, 1 if L == top else 0)
    s.MGSetup(s.opts(smoother=femus.FM_SMOOTH_VANKA))
    rng = np.random.default_rng(0)
    for L in range(1, nlev):
        A = _csr_level(s, L)
        ne = m.level(L)[0]
        for e in sorted(set(rng.integers(0, ne, 6).tolist()) | {0, ne - 1}):
            dofs, Ki = s.vanka_patch(L, e)
            K = A[dofs][:, dofs].toarray()
            err = np.abs(Ki @ K - np.eye(len(dofs))).max()
            assert err < 1e-10 * max(1.0, np.linalg.cond(K) * 1e-6), (L, e, err)
    s.close()


def _csr_level(s, L):
    rp, ci, v = s.csr(0, L)
    n = len(rp) - 1
    return sp.csr_matrix((v, ci, rp), shape=(n, n))
