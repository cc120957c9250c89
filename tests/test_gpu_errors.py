"""Error paths of the C ABI (the reference aborts through std::cerr / abort(), laspack_linear_solverM.C:118-122;
here every failure is a status code plus fm_last_error, and the system stays usable afterwards)."""
import ctypes as C
import functools

import numpy as np
import pytest

from exfemus_amd import configs

pytestmark = pytest.mark.gpu

FM_EINVAL, FM_ENOCONV, FM_ESINGULAR, FM_ESTATE = -1, -3, -4, -5


def test_invalid_arguments(femus, ctx):
    L = femus.lib()
    h = C.c_void_p()
    nx = np.array([2, 2, 1], np.int32)
    lo, hi = np.zeros(3), np.ones(3)
    assert L.fm_mesh_create_box(4, nx, lo, hi, 2, C.byref(h)) == FM_EINVAL          # dim 4
    assert L.fm_mesh_create_box(2, nx, lo, hi, 0, C.byref(h)) == FM_EINVAL          # no level
    m = femus.Mesh(2, [2, 2], [0, 0], [1, 1], 2)
    s = C.c_void_p()
    assert L.fm_system_create(ctx.h, m.h, -1, 0, C.byref(s)) == FM_EINVAL          # negative variable count
    sysT = femus.MGSolT(ctx, m, alpha0=1.0, dt=0.1, time_coeff=0.0, source=1.0)
    # calls out of order: assembly and solve before init
    with pytest.raises(femus.FemusError) as e:
        sysT.GenMatRhs(0.0, 1, 1)
    assert e.value.code == FM_ESTATE
    sysT.GenBc(configs.heat_dirichlet_all)
    sysT.init()
    with pytest.raises(femus.FemusError) as e:
        sysT.GenMatRhs(0.0, 7, 1)                                                   # no level 7
    assert e.value.code == FM_ESTATE
    bad = femus.HeatParams()
    C.pointer(bad)[0] = sysT.params
    bad.dt = 0.0
    assert L.fm_heat_genmatrhs(sysT.h, 1, 1, C.byref(bad)) == FM_EINVAL             # dt = 0
    assert L.fm_vanka_patch(sysT.h, 1, 0, np.zeros(9, np.int32).ctypes.data_as(C.c_void_p),
                            np.zeros(81).ctypes.data_as(C.c_void_p)) == FM_ESTATE   # no Vanka inverses
    assert b"Vanka" in L.fm_last_error()
    # null host pointers are invalid arguments that name the call (a second handle on the library takes raw
    # pointers, the mirror's argument types would refuse None before the call)
    raw = C.CDLL(femus.LIBPATH)
    raw.fm_system_set_xold.argtypes = [C.c_void_p, C.c_void_p]
    raw.fm_spmv.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    buf = np.zeros(sysT._Dim[1])
    assert raw.fm_system_set_xold(sysT.h, None) == FM_EINVAL
    assert b"fm_system_set_xold" in L.fm_last_error()
    assert raw.fm_spmv(sysT.h, 1, None, buf.ctypes.data_as(C.c_void_p)) == FM_EINVAL
    assert raw.fm_spmv(sysT.h, 1, buf.ctypes.data_as(C.c_void_p), None) == FM_EINVAL
    assert b"fm_spmv" in L.fm_last_error()
    assert L.fm_set_sell_offsets(12) == FM_EINVAL
    # the system still works: assemble and solve
    sysT.set_xold(np.zeros(sysT._Dim[1]))
    for lev in range(2):
        sysT.GenMatRhs(0.0, lev, 1 if lev == 1 else 0)
    rep = sysT.MGSolve(opts=sysT.opts(precision=femus.FM_PREC_FP64_FGMRES, tol=1e-10, max_iter=50))
    assert rep.converged == 1
    sysT.close()


def test_singular_coarse_level_and_noconv(femus, ctx):
    """A one-element enclosed cavity as the coarse level (every velocity DoF Dirichlet: the whole pressure block
    is decoupled) is reported singular; a solve limited to 1 cycle reports FM_ENOCONV with its residual."""
    m = femus.Mesh(3, [1, 1, 1], [0, 0, 0], [1, 1, 1], 2)
    ns = femus.MGSolNS(ctx, m, IRe=0.01, dt=0.1, time_coeff=1.0, upwind=0.1)
    ns.GenBc(functools.partial(configs.cavity_bc, dim=3))
    ns.init()
    ns.GenIc(functools.partial(configs.cavity_ic, dim=3))
    for lev in range(2):
        ns.GenMatRhs(0.0, lev, 1 if lev == 1 else 0)
    with pytest.raises(femus.FemusError) as e:
        ns.MGSolve(opts=ns.opts(precision=femus.FM_PREC_FP64_FGMRES, smoother=femus.FM_SMOOTH_VANKA, tol=1e-10))
    assert e.value.code == FM_ESINGULAR
    assert "singular" in str(e.value)
    ns.close()
    m2 = femus.Mesh(2, [2, 2], [0, 0], [1, 1], 3)
    s2 = femus.MGSolT(ctx, m2, alpha0=1.0, dt=0.1, time_coeff=0.0, source=1.0)
    s2.GenBc(configs.heat_dirichlet_all)
    s2.init()
    s2.set_xold(np.zeros(s2._Dim[2]))
    for lev in range(3):
        s2.GenMatRhs(0.0, lev, 1 if lev == 2 else 0)
    rep = s2.MGSolve(opts=s2.opts(precision=femus.FM_PREC_FP64, tol=1e-14, max_iter=1))
    assert rep.converged == 0 and rep.iters == 1 and 0.0 < rep.rel_resid < 1.0
    rc = femus.lib().fm_mg_solve(s2.h, C.byref(s2.opts(precision=femus.FM_PREC_FP64, tol=1e-14, max_iter=1)),
                                 C.byref(femus.MGReport()))
    assert rc == FM_ENOCONV
    s2.close()


def test_memory_and_coupling_controls_reject_bad_arguments(femus, ctx):
    """fm_system_couple needs two systems of one mesh and a known field; the Vanka store mask is 1..3 (0 reads);
    fm_system_sell_info and fm_system_advance need an initialised system; fm_mg_release on a hierarchy that was
    never set up is a no-op."""
    L = femus.lib()
    m1 = femus.Mesh(2, [2, 2], [0, 0], [1, 1], 2)
    m2 = femus.Mesh(2, [2, 2], [0, 0], [1, 1], 2)
    a = femus.MGSolT(ctx, m1, alpha0=1.0, dt=0.1, time_coeff=0.0, source=1.0)
    b = femus.MGSolT(ctx, m2, alpha0=1.0, dt=0.1, time_coeff=0.0, source=1.0)
    info = np.zeros(7, np.int64)
    assert L.fm_system_sell_info(a.h, 0, 0, info.ctypes.data_as(C.c_void_p)) == FM_EINVAL   # before init
    r = C.c_double()
    assert L.fm_system_advance(a.h, C.byref(r)) == FM_ESTATE                                  # before init
    for s_ in (a, b):
        s_.GenBc(configs.heat_dirichlet_all)
        s_.init()
    zero = np.zeros(3)
    assert L.fm_system_couple(a.h, b.h, 1, zero.ctypes.data_as(C.c_void_p)) == FM_EINVAL     # different meshes
    assert L.fm_system_couple(a.h, a.h, 5, zero.ctypes.data_as(C.c_void_p)) == FM_EINVAL     # unknown field
    assert L.fm_system_couple(a.h, a.h, 0, zero.ctypes.data_as(C.c_void_p)) == FM_EINVAL     # heat has no buoyancy
    assert L.fm_system_set_vanka_store(a.h, 7) == FM_EINVAL
    assert L.fm_system_set_vanka_store(a.h, 0) == 3                                           # read only
    assert L.fm_mg_release(a.h, 0) == 0 and L.fm_mg_release(a.h, 1) == 0
    assert L.fm_system_sell_info(a.h, 3, 1, info.ctypes.data_as(C.c_void_p)) == FM_EINVAL    # no matrix 3
    assert L.fm_system_sell_info(a.h, 0, 1, info.ctypes.data_as(C.c_void_p)) == 0
    assert info[5] == a.csr(0, 1)[1].size and info[1] >= info[5] and info[0] * 64 >= a._Dim[1]
    a.close()
    b.close()


def test_wrapper_rejects_arrays_of_the_wrong_length(femus, ctx):
    """The Python mirror checks every host array it hands to the library against the length the library reads
    (level DoFs, elements, nodes) and raises before any copy; the system stays usable."""
    m = femus.Mesh(2, [2, 2], [0, 0], [1, 1], 2)
    s = femus.MGSolT(ctx, m, alpha0=1.0, dt=0.1, time_coeff=0.0, source=1.0)
    s.GenBc(configs.heat_dirichlet_all)
    s.init()
    n, nel, N = s._Dim[1], m.level(1)[0], m.N
    for call in (lambda: s.set_x(np.zeros(n - 1)), lambda: s.set_rhs(np.zeros(n + 1), 0),
                 lambda: s.set_xold(np.zeros(n - 1)), lambda: s.spmv(1, np.zeros(s._Dim[0])),
                 lambda: s.set_elem_coef(1, np.ones(nel - 1)), lambda: s.set_velocity(np.zeros(N)),
                 lambda: s.nusselt(np.zeros(n - 1)), lambda: s.coupling(2, 1.0, m=np.zeros(N - 1))):
        with pytest.raises(ValueError):
            call()
    s.set_x(np.ones((1, n)))                                    # any shape with the right number of entries
    assert np.array_equal(s.get_x(), np.ones(n))
    s.set_xold(np.zeros(n))
    for lev in range(2):
        s.GenMatRhs(0.0, lev, 1 if lev == 1 else 0)
    rep = s.MGSolve(opts=s.opts(precision=femus.FM_PREC_FP64_FGMRES, tol=1e-10, max_iter=50))
    assert rep.converged == 1
    s.close()
