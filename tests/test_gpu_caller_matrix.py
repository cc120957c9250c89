"""A caller-assembled matrix handed to the solver (SparseMatrixM::set over the pattern, then LinearSolverM::solve,
contrib/matrix/sparse_matrixM.h:113-116, contrib/matrix/linear_solverM.h:150-154) through fm_system_set_matrix:

* the values come back bit for bit through fm_system_csr_get, and the level's products use them (fm_spmv against
  a host product of the same CSR matrix, 1e-12 relative);
* multigrid on matrices the library did not assemble (every level row- and column-scaled by a random diagonal,
  D A D, so the hierarchy is no longer the one GenMatRhs would build) solves them: heat with the Chebyshev
  smoother, Navier-Stokes with the Vanka smoother, against sparse direct solves;
* the device-pointer form gives the same SELL values as the host form; wrong lengths and host pointers are
  refused."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from exfemus_amd import configs

from test_gpu_caller_mesh import _hip
from test_gpu_ns import _setup as ns_setup

pytestmark = pytest.mark.gpu


def _scaled(s, L, rng):
    rp, ci, v = s.csr(0, L)
    n = len(rp) - 1
    d = 1.0 + 0.5 * rng.random(n)
    rows = np.repeat(np.arange(n), np.diff(rp))
    return rp, ci, d[rows] * v * d[ci]


def _csr(rp, ci, v):
    n = len(rp) - 1
    return sp.csr_matrix((v, ci, rp), shape=(n, n))


@pytest.mark.parametrize("prec", [0, 3])
def test_caller_matrix_heat(femus, ctx, prec):
    m = femus.Mesh(3, [2, 2, 2], [0, 0, 0], [1, 1, 1], 3)
    s = femus.MGSolT(ctx, m, alpha0=1.0, dt=0.1, time_coeff=0.0, source=1.0)
    s.GenBc(configs.heat_dirichlet_all)
    s.init()
    s.set_xold(np.zeros(s._Dim[-1]))
    rng = np.random.default_rng(11)
    mats = []
    for L in range(3):
        s.GenMatRhs(0.0, L, 1 if L == 2 else 0)
        rp, ci, v = _scaled(s, L, rng)
        s.set_matrix(L, v)
        rp2, ci2, v2 = s.csr(0, L)
        np.testing.assert_array_equal(rp2, rp)
        np.testing.assert_array_equal(ci2, ci)
        np.testing.assert_array_equal(v2, v)
        A = _csr(rp, ci, v)
        x = rng.standard_normal(s._Dim[L])
        y = s.spmv(L, x)
        ref = A @ x
        assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()
        mats.append(A)
    b = rng.standard_normal(s._Dim[2])
    s.set_rhs(b)
    s.set_x(np.zeros(s._Dim[2]))
    rep = s.MGSolve(opts=s.opts(precision=prec, tol=1e-10, max_iter=80))
    assert rep.converged == 1, (rep.iters, rep.rel_resid)
    x = s.get_x()
    xref = spla.spsolve(mats[2].tocsc(), b)
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) <= 1e-8
    s.close()


def test_caller_matrix_ns_vanka(femus, oracle, ctx):
    pm, om, ns, bc, x0 = ns_setup(femus, oracle, ctx, "channel3d", 3, [4, 2, 2], 2)
    rng = np.random.default_rng(12)
    mats = []
    for L in range(2):
        ns.GenMatRhs(0.0, L, 1 if L == 1 else 0)
        rp, ci, v = _scaled(ns, L, rng)
        ns.set_matrix(L, v)
        np.testing.assert_array_equal(ns.csr(0, L)[2], v)
        mats.append(_csr(rp, ci, v))
    b = ns.get_rhs()
    ns.set_x(np.zeros(ns._Dim[1]))
    o = ns.opts(precision=femus.FM_PREC_FP32_FGMRES, smoother=femus.FM_SMOOTH_VANKA, nu_pre=2, nu_post=2, tol=1e-9,
                max_iter=80, omega=0.5)
    rep = ns.MGSolve(opts=o)
    assert rep.converged == 1, (rep.iters, rep.rel_resid)
    x = ns.get_x()
    assert np.linalg.norm(b - mats[1] @ x) / np.linalg.norm(b) <= 2e-9
    xref = spla.spsolve(mats[1].tocsc(), b)
    assert np.linalg.norm(x - xref) / np.linalg.norm(xref) <= 1e-7
    ns.close()


def test_caller_matrix_device_pointer_and_errors(femus, ctx):
    hip = _hip()
    m = femus.Mesh(2, [2, 2], [0, 0], [1, 1], 2)
    s = femus.MGSolT(ctx, m, alpha0=1.0, dt=0.1, time_coeff=0.0, source=1.0)
    s.GenBc(configs.heat_dirichlet_all)
    s.init()
    s.set_xold(np.zeros(s._Dim[-1]))
    s.GenMatRhs(0.0, 1, 1)
    rng = np.random.default_rng(13)
    rp, ci, v = _scaled(s, 1, rng)
    nnz = v.size
    buf = C.c_void_p()
    assert hip.hipMalloc(C.byref(buf), nnz * 8) == 0
    try:
        assert hip.hipMemcpy(buf, v.ctypes.data_as(C.c_void_p), nnz * 8, 1) == 0
        femus._ck(femus.lib().fm_system_set_matrix_dev(s.h, 1, buf))
        np.testing.assert_array_equal(s.csr(0, 1)[2], v)
        x = rng.standard_normal(s._Dim[1])
        y_dev = s.spmv(1, x)
        s.set_matrix(1, 2.0 * v)
        np.testing.assert_array_equal(s.spmv(1, x), 2.0 * y_dev)
    finally:
        hip.hipFree(buf)
    with pytest.raises(ValueError):
        s.set_matrix(1, v[:-1])
    host = np.zeros(nnz)
    with pytest.raises(femus.FemusError) as e:
        femus._ck(femus.lib().fm_system_set_matrix_dev(s.h, 1, host.ctypes.data_as(C.c_void_p)))
    assert e.value.code == -1 and "device pointer" in str(e.value)
    with pytest.raises(femus.FemusError) as e:
        femus._ck(femus.lib().fm_system_set_matrix(s.h, 7, v.ctypes.data_as(C.c_void_p)))
    s.close()


@pytest.mark.parametrize("case,dim,nx,nlev", [("channel3d", 3, [4, 2, 2], 2), ("cavity", 2, [2, 2], 3)])
def test_singular_vanka_patch_reported(femus, oracle, ctx, case, dim, nx, nlev):
    """A caller matrix with the row and the column of one DoF zeroed makes every element patch holding that DoF
    singular: the patch inversion stops pivoting in such a patch, writes it as zeros and the set-up reports
    FM_ESINGULAR with the level; once the assembled matrix is set again the same solve sets up and converges."""
    pm, om, ns, bc, x0 = ns_setup(femus, oracle, ctx, case, dim, nx, nlev)
    top = nlev - 1
    for L in range(nlev):
        ns.GenMatRhs(0.0, L, 1 if L == top else 0)
    rp, ci, v = ns.csr(0, top)
    n = len(rp) - 1
    d = n // 2
    rows = np.repeat(np.arange(n), np.diff(rp))
    vs = v.copy()
    vs[(rows == d) | (ci == d)] = 0.0
    ns.set_matrix(top, vs)
    ns.set_x(np.zeros(n))
    o = ns.opts(precision=femus.FM_PREC_FP64_FGMRES, smoother=femus.FM_SMOOTH_VANKA, tol=1e-9, max_iter=80)
    with pytest.raises(femus.FemusError) as e:
        ns.MGSolve(opts=o)
    assert e.value.code == -4
    assert "singular element patch" in str(e.value) and f"on level {top}" in str(e.value)
    ns.set_matrix(top, v)
    ns.set_x(np.zeros(n))
    rep = ns.MGSolve(opts=o)
    assert rep.converged == 1, (rep.iters, rep.rel_resid)
    ns.close()
