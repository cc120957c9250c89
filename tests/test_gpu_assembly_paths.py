"""The two assembly routes give the same matrices and right-hand sides: the element-buffer gather (element
matrices of all elements to a buffer, summed into the SELL slots in element order) against the coloured
scatter (element matrices added colour by colour), through the C ABI, on every level, for

* heat with a nodal velocity field, upwinding, transient terms, per-element conductivity and the Dirichlet /
  Robin / Neumann faces (GenMatRhs and the transposed matrix of the adjoint);
* Navier-Stokes with buoyancy, per-element viscosity, every penalty bit and the Neumann face terms (GenMatRhs,
  the Newton Jacobian and its transpose, the transposed Oseen matrix).

The two routes sum the same terms in different orders, so they agree to rounding (1e-13 relative); the gather
route repeats bit for bit.  Both routes are separately in parity with the oracle (the other GPU tests run the
default route, the gather)."""
import functools

import numpy as np
import pytest

from exfemus_amd import configs

from test_gpu_bc_branches import heat_mixed_bc, ns_mixed_bc

pytestmark = pytest.mark.gpu


def close(a, b, what):
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= 1e-13 * scale, what


def both_routes(femus, fn):
    """fn() under the gather route, then under the coloured scatter (the knob restored afterwards)."""
    lib = femus.lib()
    old = lib.fm_set_assembly_gather(1)
    try:
        g = fn()
        g2 = fn()
        lib.fm_set_assembly_gather(0)
        c = fn()
    finally:
        lib.fm_set_assembly_gather(old)
    return g, g2, c


@pytest.mark.parametrize("dim,nx,nlev", [(2, [3, 2], 3), (3, [2, 2, 1], 3)])
def test_heat_routes_agree(femus, ctx, dim, nx, nlev):
    pm = femus.Mesh(dim, nx, [0.0] * dim, [1.0] * dim, nlev)
    s = femus.MGSolT(ctx, pm, alpha0=0.02, dt=0.1, time_coeff=1.0, source=1.0, upwind=1.0, vel=(0.7, -0.4, 0.5))
    s.GenBc(functools.partial(heat_mixed_bc, dim=dim))
    s.init()
    top = nlev - 1
    ks = configs.coarsen_coef(configs.lognormal_k(pm.level(top)[0], seed=5, sigma=0.5), nlev, 2 ** dim)
    for L in range(nlev):
        s.set_elem_coef(L, ks[L])
    rng = np.random.default_rng(41)
    s.set_xold(rng.standard_normal(s._Dim[top]))
    s.set_velocity(rng.uniform(-1.0, 1.0, dim * pm.N))

    def run():
        out = []
        for L in range(nlev):
            s.GenMatRhs(0.0, L, 1)
            out.append((s.csr(0, L)[2].copy(), s.get_rhs(L).copy()))
            s.GenMatTranspose(L)
            out.append((s.adjoint_csr(L).copy(), None))
        return out

    g, g2, c = both_routes(femus, run)
    for k, ((va, ra), (vb, rb), (vc, rc)) in enumerate(zip(g, g2, c)):
        np.testing.assert_array_equal(va, vb)
        close(va, vc, f"matrix {k}")
        if ra is not None:
            np.testing.assert_array_equal(ra, rb)
            close(ra, rc, f"rhs {k}")
    s.close()


@pytest.mark.parametrize("dim,nx,nlev", [(2, [2, 2], 3), (3, [2, 1, 1], 3)])
def test_ns_routes_agree(femus, ctx, dim, nx, nlev):
    pm = femus.Mesh(dim, nx, [0.0] * dim, [1.0] * dim, nlev)
    IRe = 0.004
    s = femus.MGSolNS(ctx, pm, IRe=IRe, dt=0.1, time_coeff=1.0, upwind=0.1, IFr=0.3, dirg=(0.0, -1.0, 0.0),
                      src=(0.2, 0.1, -0.3))
    s.GenBc(functools.partial(ns_mixed_bc, dim=dim))
    s.init()
    top = nlev - 1
    nus = configs.coarsen_coef(IRe * configs.lognormal_k(pm.level(top)[0], seed=7, sigma=0.5), nlev, 2 ** dim)
    for L in range(nlev):
        s.set_elem_coef(L, nus[L])
    rng = np.random.default_rng(43)
    s.set_xold(rng.standard_normal(s._Dim[top]))
    s.set_buoyancy(rng.uniform(-1.0, 1.0, pm.N), [0.0, 2.5, 0.0][:dim])

    def run():
        out = []
        for L in range(nlev):
            s.GenMatRhs(0.0, L, 1)
            out.append((s.csr(0, L)[2].copy(), s.get_rhs(L).copy()))
            s.GenMatTranspose(L)
            out.append((s.adjoint_csr(L).copy(), None))
            s.GenJac(L, 0)
            out.append((s.csr(0, L)[2].copy(), None))
            s.GenJac(L, 1)
            out.append((s.adjoint_csr(L).copy(), None))
        return out

    g, g2, c = both_routes(femus, run)
    for k, ((va, ra), (vb, rb), (vc, rc)) in enumerate(zip(g, g2, c)):
        np.testing.assert_array_equal(va, vb)
        close(va, vc, f"matrix {k}")
        if ra is not None:
            np.testing.assert_array_equal(ra, rb)
            close(ra, rc, f"rhs {k}")
    s.close()


def test_route_knob(femus):
    lib = femus.lib()
    old = lib.fm_set_assembly_gather(0)
    assert lib.fm_set_assembly_gather(old) == 0
    assert lib.fm_set_assembly_gather(2) == -1  # FM_EINVAL
