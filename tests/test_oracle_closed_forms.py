"""CPU pins of the oracle's element routines (om_heat_element, om_ns_element) against closed forms that the
mathematics fixes, computed here independently of the oracle's Gauss tables: on an affine (axis-aligned,
anisotropic) element every volume term of MGSolverT.C:300-365 / MGSolverNS.C:285-489 is a Kronecker product of
exact 1D Q2 / Q1 integrals (mass, stiffness, the advection matrix int phi_i phi_j', the mixed Q2-Q1 matrices),
the boundary penalty rows of MGSolverT.C:401-438 / MGSolverNS.C:532-601 are nodal penalties whose bits project
the data on the face's normal or tangent space, the Neumann-face terms of MGSolverT.C:441-506 /
MGSolverNS.C:604-681 are face mass matrices (the NS ones lumped on the diagonal), and the upwind diffusion
coefficient 0.5 (coth Pe - 1/Pe) h / |u| (MGSolverT.C:276-279, MGSolverNS.C:330-335) has the two limits
h^2 / (12 kappa) (Pe -> 0) and h / (2 |u|) (Pe -> infinity).  A dropped term, a wrong sign, a transposed
operand or a wrong axis fails one of them (the element is anisotropic and the velocity has distinct
components).  A manufactured Stokes / Oseen solution pins the assembled system: velocity L2 errors fall at
rate 3, pressure at rate 2, Poiseuille flow is reproduced exactly."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
from numpy.polynomial import polynomial as Pn

Q9 = [(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0), (0, 0)]
H27 = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1),
       (0, -1, -1), (1, 0, -1), (0, 1, -1), (-1, 0, -1), (-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0),
       (0, -1, 1), (1, 0, 1), (0, 1, 1), (-1, 0, 1), (0, 0, -1), (0, -1, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0),
       (0, 0, 1), (0, 0, 0)]
# the element side on the x = lo face and its local nodes (libMesh Quad9 / Hex27 side numbering)
XSIDE = {2: (3, [3, 0, 7]), 3: (4, [3, 0, 4, 7, 11, 12, 19, 15, 24])}

# 1D Lagrange bases on the reference interval [-1, 1]: Q2 nodes (-1, 0, 1), Q1 nodes (-1, 1)
PHI2 = [np.array([0.0, -0.5, 0.5]), np.array([1.0, 0.0, -1.0]), np.array([0.0, 0.5, 0.5])]
PHI1 = [np.array([0.5, -0.5]), np.array([0.5, 0.5])]


def _int(c):
    a = Pn.polyint(c)
    return Pn.polyval(1.0, a) - Pn.polyval(-1.0, a)


def one_d(h):
    """Exact 1D integrals on an interval of length h (dx = h/2 dxi, d/dx = 2/h d/dxi)."""
    d2 = [Pn.polyder(p) for p in PHI2]
    M = np.array([[h / 2 * _int(Pn.polymul(a, b)) for b in PHI2] for a in PHI2])
    S = np.array([[2 / h * _int(Pn.polymul(a, b)) for b in d2] for a in d2])
    Cm = np.array([[_int(Pn.polymul(a, db)) for db in d2] for a in PHI2])            # int phi_i phi_j'
    M21 = np.array([[h / 2 * _int(Pn.polymul(a, b)) for b in PHI1] for a in PHI2])    # int phi_i psi_j
    D21 = np.array([[_int(Pn.polymul(da, b)) for b in PHI1] for da in d2])           # int phi_i' psi_j
    m = np.array([h / 2 * _int(a) for a in PHI2])                                    # int phi_i
    return dict(M=M, S=S, C=Cm, M21=M21, D21=D21, m=m)


def kron(dim, rows, cols, facs):
    """Matrix over local nodes: entry (a, b) = prod_t facs[t][rows[a][t], cols[b][t]]."""
    out = np.ones((len(rows), len(cols)))
    for t in range(dim):
        F = facs[t]
        out *= F[np.ix_([r[t] for r in rows], [c[t] for c in cols])]
    return out


def q2_idx(dim):
    return [tuple(c + 1 for c in t) for t in (Q9 if dim == 2 else H27)]


def q1_idx(dim):
    return [tuple((c + 1) // 2 for c in t) for t in (Q9 if dim == 2 else H27)[:2 ** dim]]


def _box(oracle, dim, L_len):
    """A 3 x 3 (x 3) single-level box with edge lengths L_len: elements of size L/3, one interior element and
    one element whose only boundary side is the x = lo face."""
    m = oracle.Mesh(dim, [3] * dim, [0.0] * dim, list(L_len), 1)
    conn, neigh, _ = m.conn(0)
    interior = [e for e in range(conn.shape[0]) if np.all(neigh[e] >= 0)][0]
    side = XSIDE[dim][0]
    xface = [e for e in range(conn.shape[0]) if neigh[e, side] == -1 and np.sum(neigh[e] == -1) == 1][0]
    return m, conn, interior, xface


LEN = {2: [0.9, 1.5], 3: [0.9, 1.5, 1.2]}


def _heat_elem(oracle, m, e, p, bc0, bc1, xold, k_e, vel=None, mode=1):
    d, n = oracle.init_dof(m, 0, 1, 0)
    KeM = np.zeros(27 * 27); FeM = np.zeros(27); dofs = np.zeros(27, np.int32)
    v = None if vel is None else np.ascontiguousarray(vel, np.float64)
    oracle.lib().om_heat_element(m.h, 0, e, mode, C.byref(p), d, bc0, bc1, xold, k_e,
                                 None if v is None else v.ctypes.data_as(C.c_void_p), KeM, FeM, dofs, None)
    n2 = m.n2
    return KeM[:n2 * n2].reshape(n2, n2), FeM[:n2]


def _ns_elem(oracle, m, e, p, bc0, bc1, xold, nu_e, mode=1):
    d, n = oracle.init_dof(m, 0, m.dim, 1)
    KeM = np.zeros(89 * 89); FeM = np.zeros(89); dofs = np.zeros(89, np.int32)
    oracle.lib().om_ns_element(m.h, 0, e, mode, C.byref(p), d, bc0, bc1, xold, nu_e, KeM, FeM, dofs, None)
    nl = m.dim * m.n2 + m.n1
    return KeM[:nl * nl].reshape(nl, nl), FeM[:nl]


@pytest.mark.parametrize("dim", [2, 3])
def test_heat_element_advection_kronecker(oracle, dim):
    """Constant velocity a, no upwinding, k_e = 1.3: K = alpha0 k_e sum_d S_d + rhocp sum_d a_d C_d + rhocp M / dt
    (each a Kronecker product over the axes), F = (rhocp T_old / dt + f) int phi for T_old constant."""
    m, conn, e, _ = _box(oracle, dim, LEN[dim])
    h = [L / 3 for L in LEN[dim]]
    D = [one_d(hh) for hh in h]
    a = (0.7, -0.4, 0.5)[:dim] + (0.0,) * (3 - dim)
    dt, alpha0, rhocp, src, k_e, T0 = 0.37, 0.8, 1.7, 0.6, 1.3, 2.5
    p = oracle.HeatParams(alpha0, rhocp, dt, 1.0, src, 0.0, a)
    d, n = oracle.init_dof(m, 0, 1, 0)
    ones = np.ones(n, np.int32)
    K, F = _heat_elem(oracle, m, e, p, ones, ones, np.full(n, T0), k_e)
    r = q2_idx(dim)
    Mk = kron(dim, r, r, [D[t]["M"] for t in range(dim)])
    Kref = rhocp * Mk / dt
    for dd in range(dim):
        Kref += alpha0 * k_e * kron(dim, r, r, [D[t]["S" if t == dd else "M"] for t in range(dim)])
        Kref += rhocp * a[dd] * kron(dim, r, r, [D[t]["C" if t == dd else "M"] for t in range(dim)])
    np.testing.assert_allclose(K, Kref, rtol=1e-12, atol=1e-13 * np.abs(Kref).max())
    mvec = np.prod([[D[t]["m"][ri[t]] for t in range(dim)] for ri in r], axis=1)
    np.testing.assert_allclose(F, (rhocp * T0 / dt + src) * mvec, rtol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("limit", ["small", "large"])
def test_heat_upwind_coefficient_limits(oracle, dim, limit):
    """The upwind diffusion UP_WIND_T f_d a_d^2 int d_d phi_j d_d phi_i (MGSolverT.C:276-300) with
    f_d = 0.5 (coth Pe_d - 1/Pe_d) h_d / |a|, Pe_d = |a| h_d / (2 alpha0): the difference of the element matrices
    with and without upwinding is sum_d c_d S_d with c_d -> a_d^2 h_d^2 / (12 alpha0) as Pe -> 0 and
    c_d -> a_d^2 h_d / (2 |a|) as Pe -> infinity (h_d the element's extent along d)."""
    m, conn, e, _ = _box(oracle, dim, LEN[dim])
    h = [L / 3 for L in LEN[dim]]
    D = [one_d(hh) for hh in h]
    a = np.array((0.7, -0.4, 0.5)[:dim])
    # small: Pe ~ 2e-3 (the series coth x - 1/x = x/3 - x^3/45 + ... is then exact to 3e-7; much smaller Pe
    # lose digits to the cancellation of coth x and 1/x); large: Pe ~ 4e11
    alpha0 = 100.0 if limit == "small" else 1e-12
    d, n = oracle.init_dof(m, 0, 1, 0)
    ones = np.ones(n, np.int32)
    av = tuple(a) + (0.0,) * (3 - dim)
    K1, _ = _heat_elem(oracle, m, e, oracle.HeatParams(alpha0, 1.0, 0.1, 0.0, 0.0, 1.0, av), ones, ones, np.zeros(n), 1.0)
    K0, _ = _heat_elem(oracle, m, e, oracle.HeatParams(alpha0, 1.0, 0.1, 0.0, 0.0, 0.0, av), ones, ones, np.zeros(n), 1.0)
    r = q2_idx(dim)
    amod = np.linalg.norm(a)
    Kref = np.zeros_like(K1)
    for dd in range(dim):
        c = a[dd] ** 2 * h[dd] ** 2 / (12 * alpha0) if limit == "small" else a[dd] ** 2 * h[dd] / (2 * amod)
        Kref += c * kron(dim, r, r, [D[t]["S" if t == dd else "M"] for t in range(dim)])
    np.testing.assert_allclose(K1 - K0, Kref, rtol=2e-6, atol=2e-6 * np.abs(Kref).max())


@pytest.mark.parametrize("dim", [2, 3])
def test_ns_element_oseen_kronecker(oracle, dim):
    """Constant advection field a (x_old), no upwinding, per-element nu: the velocity blocks are
    nu (sum_k S_k + S_iv) + sum_k a_k C_k + M / dt on the diagonal and nu kron(C on axis iv, C^T on axis jv) off it
    (the viscous tensor d_iv phi_j d_jv phi_i), B^T = -kron(D21 on axis iv, M21 elsewhere),
    B = kron(D21^T on axis jv, M21^T elsewhere); the RHS of a constant x_old, gravity, source and a linear nodal
    temperature with the buoyancy coefficient is (a_iv / dt + IFr g_iv + s_iv) int phi_i + b_iv (M T)_i."""
    m, conn, e, _ = _box(oracle, dim, LEN[dim])
    h = [L / 3 for L in LEN[dim]]
    D = [one_d(hh) for hh in h]
    a = np.array((0.7, -0.4, 0.5)[:dim])
    nu, dt, IFr = 0.013, 0.37, 0.4
    g = (0.0, -1.0, 0.3)
    s = (0.2, 0.1, -0.3)
    buoy = [0.9, -1.1, 2.5][:dim]
    d, n = oracle.init_dof(m, 0, dim, 1)
    N = m.N
    xyz = m.coords()
    xold = np.zeros(n)
    for iv in range(dim):
        xold[d[iv * N:(iv + 1) * N][d[iv * N:(iv + 1) * N] >= 0]] = a[iv]
    Tn = 0.4 + 0.7 * xyz[0] - 0.3 * xyz[1] + (0.5 * xyz[2] if dim == 3 else 0.0)
    p = oracle.NSParams(0.02, dt, 1.0, 0.0, IFr, g, s).set_buoyancy(Tn, buoy)
    ones = np.ones(n, np.int32)
    K, F = _ns_elem(oracle, m, e, p, ones, ones, xold, nu)
    r2, r1 = q2_idx(dim), q1_idx(dim)
    n2 = len(r2)
    facs = lambda kind: [D[t][kind] for t in range(dim)]
    M = kron(dim, r2, r2, facs("M"))
    for iv in range(dim):
        for jv in range(dim):
            blk = K[iv * n2:(iv + 1) * n2, jv * n2:(jv + 1) * n2]
            if iv == jv:
                ref = M / dt
                for k in range(dim):
                    ref = ref + nu * kron(dim, r2, r2, [D[t]["S" if t == k else "M"] for t in range(dim)])
                    ref = ref + a[k] * kron(dim, r2, r2, [D[t]["C" if t == k else "M"] for t in range(dim)])
                ref = ref + nu * kron(dim, r2, r2, [D[t]["S" if t == iv else "M"] for t in range(dim)])
            else:
                ref = nu * kron(dim, r2, r2, [D[t]["C"] if t == iv else (D[t]["C"].T if t == jv else D[t]["M"])
                                              for t in range(dim)])
            np.testing.assert_allclose(blk, ref, rtol=1e-12, atol=1e-13 * np.abs(K).max(), err_msg=f"block {iv},{jv}")
        BT = K[iv * n2:(iv + 1) * n2, dim * n2:]
        refBT = -kron(dim, r2, r1, [D[t]["D21" if t == iv else "M21"] for t in range(dim)])
        np.testing.assert_allclose(BT, refBT, rtol=1e-12, atol=1e-14)
        B = K[dim * n2:, iv * n2:(iv + 1) * n2]
        refB = kron(dim, r1, r2, [D[t]["D21"].T if t == iv else D[t]["M21"].T for t in range(dim)])
        np.testing.assert_allclose(B, refB, rtol=1e-12, atol=1e-14)
    assert np.all(K[dim * n2:, dim * n2:] == 0.0)
    mvec = np.prod([[D[t]["m"][ri[t]] for t in range(dim)] for ri in r2], axis=1)
    Te = Tn[conn[e]]
    for iv in range(dim):
        ref = (a[iv] / dt + IFr * g[iv] + s[iv]) * mvec + buoy[iv] * (M @ Te)
        np.testing.assert_allclose(F[iv * n2:(iv + 1) * n2], ref, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("dim", [2, 3])
def test_ns_upwind_large_peclet_limit(oracle, dim):
    """UP_WIND_NS at Pe -> infinity (MGSolverNS.C:330-435): the upwind part of the velocity blocks is
    sum_k f_k a_k^2 S_k + f_iv a_iv^2 S_iv on the diagonal and f_iv a_iv a_jv kron(C on iv, C^T on jv) off it,
    with f_d = h_d / (2 |a|)."""
    m, conn, e, _ = _box(oracle, dim, LEN[dim])
    h = [L / 3 for L in LEN[dim]]
    D = [one_d(hh) for hh in h]
    a = np.array((0.7, -0.4, 0.5)[:dim])
    d, n = oracle.init_dof(m, 0, dim, 1)
    N = m.N
    xold = np.zeros(n)
    for iv in range(dim):
        xold[d[iv * N:(iv + 1) * N][d[iv * N:(iv + 1) * N] >= 0]] = a[iv]
    ones = np.ones(n, np.int32)
    IRe = 1e-12
    K1, _ = _ns_elem(oracle, m, e, oracle.NSParams(IRe, 0.1, 0.0, 0.7, 0.0, (0, 0, 0), (0, 0, 0)), ones, ones, xold, 0.0)
    K0, _ = _ns_elem(oracle, m, e, oracle.NSParams(IRe, 0.1, 0.0, 0.0, 0.0, (0, 0, 0), (0, 0, 0)), ones, ones, xold, 0.0)
    r2 = q2_idx(dim)
    n2 = len(r2)
    f = 0.7 * np.array(h) / (2 * np.linalg.norm(a))
    dK = K1 - K0
    for iv in range(dim):
        for jv in range(dim):
            if iv == jv:
                ref = sum(f[k] * a[k] ** 2 * kron(dim, r2, r2, [D[t]["S" if t == k else "M"] for t in range(dim)])
                          for k in range(dim))
                ref = ref + f[iv] * a[iv] ** 2 * kron(dim, r2, r2, [D[t]["S" if t == iv else "M"] for t in range(dim)])
            else:
                ref = f[iv] * a[iv] * a[jv] * kron(dim, r2, r2, [D[t]["C"] if t == iv else
                                                                 (D[t]["C"].T if t == jv else D[t]["M"]) for t in range(dim)])
            np.testing.assert_allclose(dK[iv * n2:(iv + 1) * n2, jv * n2:(jv + 1) * n2], ref, rtol=1e-7,
                                       atol=1e-9 * np.abs(dK).max(), err_msg=f"block {iv},{jv}")


def _face_penalty(dim, L_len, dt):
    """Nodal penalty of a face: det of the face map at a Gauss point / dt (affine face: area / 2^(dim-1))."""
    area = np.prod(L_len[1:]) / 3 ** (dim - 1)
    return area / 2 ** (dim - 1) / dt


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("bits", [1, 0])
def test_heat_dirichlet_face_penalty(oracle, dim, bits):
    """Dirichlet rows of the x = lo face: volume terms dropped, the nodal penalty area / 2^(dim-1) / dt on the
    diagonal, and with the non-homogeneous bit the data T_old times the penalty on the RHS."""
    m, conn, _, e = _box(oracle, dim, LEN[dim])
    dt = 0.37
    d, n = oracle.init_dof(m, 0, 1, 0)
    face = XSIDE[dim][1]
    bc0 = np.ones(n, np.int32); bc1 = np.ones(n, np.int32)
    fd = d[conn[e][face]]
    bc0[fd] = 0
    bc1[fd] = bits
    xold = np.random.default_rng(2).standard_normal(n)
    p = oracle.HeatParams(0.8, 1.0, dt, 1.0, 0.6, 0.0, (0.3, 0.1, 0.2))
    K, F = _heat_elem(oracle, m, e, p, bc0, bc1, xold, 1.0)
    Kv, Fv = _heat_elem(oracle, m, e, p, np.ones(n, np.int32), np.ones(n, np.int32), xold, 1.0)
    pen = _face_penalty(dim, LEN[dim], dt)
    rest = [i for i in range(m.n2) if i not in face]
    for i in face:
        ref = np.zeros(m.n2); ref[i] = pen
        np.testing.assert_allclose(K[i], ref, rtol=1e-13, atol=1e-15)
        assert abs(F[i] - bits * pen * xold[d[conn[e][i]]]) <= 1e-13 * pen * (1 + abs(xold).max())
    np.testing.assert_allclose(K[rest], Kv[rest], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(F[rest], Fv[rest], rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("dim", [2, 3])
def test_heat_robin_face_mass(oracle, dim):
    """Neumann face with the Robin bit (alpha): the face rows gain alpha0 int_face phi_i phi_j, the face mass
    matrix (a Kronecker product over the face's axes); without the bit nothing changes."""
    m, conn, _, e = _box(oracle, dim, LEN[dim])
    h = [L / 3 for L in LEN[dim]]
    D = [one_d(hh) for hh in h]
    d, n = oracle.init_dof(m, 0, 1, 0)
    face = XSIDE[dim][1]
    alpha0 = 0.8
    p = oracle.HeatParams(alpha0, 1.0, 0.37, 0.0, 0.0, 0.0, (0.0, 0.0, 0.0))
    ones = np.ones(n, np.int32)
    K0, _ = _heat_elem(oracle, m, e, p, ones, ones, np.zeros(n), 1.0)
    bc1 = np.ones(n, np.int32)
    bc1[d[conn[e][face]]] = 2
    K2, _ = _heat_elem(oracle, m, e, p, ones, bc1, np.zeros(n), 1.0)
    bc1[d[conn[e][face]]] = 0
    Kz, _ = _heat_elem(oracle, m, e, p, ones, bc1, np.zeros(n), 1.0)
    r = q2_idx(dim)
    fr = [r[i][1:] for i in face]
    Mf = kron(dim - 1, fr, fr, [D[t]["M"] for t in range(1, dim)])
    dK = K2 - K0
    np.testing.assert_allclose(dK[np.ix_(face, face)], alpha0 * Mf, rtol=1e-12, atol=1e-15)
    mask = np.ones_like(dK, bool); mask[np.ix_(face, face)] = False
    assert np.all(dK[mask] == 0.0)
    np.testing.assert_array_equal(Kz, K0)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("bits", [4, 5, 6, 1, 2, 0])
def test_ns_dirichlet_face_penalty_bits(oracle, dim, bits):
    """Dirichlet velocity rows of the x = lo face (normal -e_x) with bits 4 (data) / 2 (tangential) / 1 (normal):
    the rows hold only the nodal penalty on the diagonal, on the normal component alone with bit 1, on the
    tangential components alone with bit 2, on all components otherwise; with bit 4 the RHS is the penalty times
    the data projected the same way (normal part v_x, tangential parts v_y, v_z, or all of v)."""
    m, conn, _, e = _box(oracle, dim, LEN[dim])
    dt = 0.37
    d, n = oracle.init_dof(m, 0, dim, 1)
    N = m.N
    face = XSIDE[dim][1]
    n2 = m.n2
    bc0 = np.ones(n, np.int32); bc1 = np.ones(n, np.int32)
    for iv in range(dim):
        fd = d[conn[e][face] + iv * N]
        bc0[fd] = 0
        bc1[fd] = bits
    xold = np.random.default_rng(3).standard_normal(n)
    p = oracle.NSParams(0.02, dt, 1.0, 0.0, 0.0, (0, 0, 0), (0, 0, 0))
    K, F = _ns_elem(oracle, m, e, p, bc0, bc1, xold, 0.02)
    pen = _face_penalty(dim, LEN[dim], dt)
    nbit, tbit, rbit = bits & 1, (bits >> 1) & 1, (bits >> 2) & 1
    for iv in range(dim):
        normal = iv == 0
        on = (1 if normal else 0) if nbit else ((0 if normal else 1) if tbit else 1)
        for i in face:
            row = iv * n2 + i
            ref = np.zeros(K.shape[1]); ref[row] = on * pen
            np.testing.assert_allclose(K[row], ref, rtol=1e-13, atol=1e-15, err_msg=f"bits {bits} comp {iv}")
            v = xold[d[conn[e][i] + iv * N]]
            assert abs(F[row] - rbit * on * pen * v) <= 1e-13 * pen * (1 + abs(v))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("bits", [1, 2, 3])
def test_ns_neumann_face_lumped_mass(oracle, dim, bits):
    """Neumann velocity rows of the x = lo face with the normal (1) / tangential (2) bits (MGSolverNS.C:604-681):
    the diagonal gains the lumped face mass int_face phi_i on the normal component with bit 1, on the tangential
    components with bit 2 (both: on every component)."""
    m, conn, _, e = _box(oracle, dim, LEN[dim])
    h = [L / 3 for L in LEN[dim]]
    D = [one_d(hh) for hh in h]
    d, n = oracle.init_dof(m, 0, dim, 1)
    N = m.N
    face = XSIDE[dim][1]
    n2 = m.n2
    p = oracle.NSParams(0.02, 0.37, 1.0, 0.0, 0.0, (0, 0, 0), (0, 0, 0))
    ones = np.ones(n, np.int32)
    bc1 = np.ones(n, np.int32)
    for iv in range(dim):
        bc1[d[conn[e][face] + iv * N]] = bits
    K1, _ = _ns_elem(oracle, m, e, p, ones, bc1, np.zeros(n), 0.02)
    # bc1 = 1 everywhere is itself the normal bit: compare against a face without bits
    bcz = np.ones(n, np.int32)
    for iv in range(dim):
        bcz[d[conn[e][face] + iv * N]] = 0
    Kz, _ = _ns_elem(oracle, m, e, p, ones, bcz, np.zeros(n), 0.02)
    r = q2_idx(dim)
    lump = np.array([np.prod([D[t]["m"][r[i][t]] for t in range(1, dim)]) for i in face])
    dK = K1 - Kz
    expect = np.zeros_like(dK)
    for iv in range(dim):
        normal = iv == 0
        on = ((bits & 1) if normal else 0) + ((bits >> 1 & 1) if not normal else 0)
        for k, i in enumerate(face):
            expect[iv * n2 + i, iv * n2 + i] = on * lump[k]
    np.testing.assert_allclose(dK, expect, rtol=1e-12, atol=1e-15)


# ----------------------------------------------------------------------------- manufactured solutions
def _mms_fields(x, y, nu, oseen):
    """u = curl(sin^2(pi x) sin^2(pi y)) (divergence-free, zero on the unit square's boundary), p = cos(pi x) cos(pi y);
    f = -nu lap u + (u . grad) u [Oseen] + grad p."""
    pi = np.pi
    s2x, s2y, c2x, c2y = np.sin(2 * pi * x), np.sin(2 * pi * y), np.cos(2 * pi * x), np.cos(2 * pi * y)
    ux = pi / 2 * (1 - c2x) * s2y
    uy = -pi / 2 * s2x * (1 - c2y)
    lux = 2 * pi ** 3 * s2y * (2 * c2x - 1)
    luy = -2 * pi ** 3 * s2x * (2 * c2y - 1)
    fx = -nu * lux - pi * np.sin(pi * x) * np.cos(pi * y)
    fy = -nu * luy - pi * np.cos(pi * x) * np.sin(pi * y)
    if oseen:
        dxux, dyux = pi ** 2 * s2x * s2y, pi ** 2 * (1 - c2x) * c2y
        dxuy, dyuy = -pi ** 2 * c2x * (1 - c2y), -pi ** 2 * s2x * s2y
        fx = fx + ux * dxux + uy * dyux
        fy = fy + ux * dxuy + uy * dyuy
    return ux, uy, np.cos(pi * x) * np.cos(pi * y), fx, fy


def test_mms_fields_satisfy_the_equations():
    """The hand-derived forcing of the manufactured solution against central differences of u and p."""
    rng = np.random.default_rng(4)
    x, y = rng.uniform(0.1, 0.9, 20), rng.uniform(0.1, 0.9, 20)
    nu, e = 0.1, 1e-4
    for oseen in (False, True):
        ux, uy, p, fx, fy = _mms_fields(x, y, nu, oseen)
        F = lambda xx, yy: _mms_fields(xx, yy, nu, oseen)
        lap = lambda k: (F(x + e, y)[k] + F(x - e, y)[k] + F(x, y + e)[k] + F(x, y - e)[k] - 4 * F(x, y)[k]) / e ** 2
        dx = lambda k: (F(x + e, y)[k] - F(x - e, y)[k]) / (2 * e)
        dy = lambda k: (F(x, y + e)[k] - F(x, y - e)[k]) / (2 * e)
        rx = -nu * lap(0) + dx(2) + (ux * dx(0) + uy * dy(0) if oseen else 0)
        ry = -nu * lap(1) + dy(2) + (ux * dx(1) + uy * dy(1) if oseen else 0)
        np.testing.assert_allclose(rx, fx, rtol=1e-5, atol=1e-4)
        np.testing.assert_allclose(ry, fy, rtol=1e-5, atol=1e-4)
        np.testing.assert_allclose(dx(0) + dy(1), 0.0, atol=1e-6)


def _solve_ns_oracle(oracle, nx, nu, oseen, exact_poiseuille=False):
    """Steady Stokes / Oseen on the unit square with the oracle's assembly: no-slip (penalty Dirichlet with the
    data of x_old), body force through the nodal buoyancy field (one assembly per component: the system is
    linear), advection field = the interpolated exact velocity, last pressure DoF pinned to its exact value."""
    m = oracle.Mesh(2, [nx, nx], [0, 0], [1, 1], 1)
    d, n = oracle.init_dof(m, 0, 2, 1)
    N = m.N
    xyz = m.coords()

    def bc(xp, neum, flag):
        on = min(xp[0], xp[1], 1 - xp[0], 1 - xp[1]) < 1e-12
        for iv in range(2):
            neum[iv], flag[iv] = (0, 4) if on else (1, 0)
        neum[2], flag[2] = 1, 0

    b0, b1 = oracle.genbc(m, 2, 1, d, bc)
    rp, ci = oracle.system_pattern(m, 0, 2, 1, d)
    if exact_poiseuille:
        ux, uy = 4 * xyz[1] * (1 - xyz[1]), 0 * xyz[0]
        pex = -8 * nu * xyz[0]
        fx = fy = 0 * xyz[0]
    else:
        ux, uy, pex, fx, fy = _mms_fields(xyz[0], xyz[1], nu, oseen)
    xold = np.zeros(n)
    for iv, f in enumerate((ux, uy)):
        dd = d[iv * N:(iv + 1) * N]
        xold[dd[dd >= 0]] = f[dd >= 0]
    if not oseen:
        # advection from x_old is switched off by a zero interior field (x_old still carries the boundary data)
        inner = (b0 == 1)
        xold[inner] = 0.0
    A = None
    rhs = np.zeros(n)
    for comp, f in ((0, fx), (1, fy)):
        coef = [0.0, 0.0]; coef[comp] = 1.0
        p = oracle.NSParams(nu, 1.0, 0.0, 0.0, 0.0, (0, 0, 0), (0, 0, 0)).set_buoyancy(f, coef)
        v, r = oracle.ns_genmatrhs(m, 0, 1, p, d, d, b0, b1, xold, rp, ci)
        rhs += r if comp == 0 else r - _bd_rhs(oracle, m, d, b0, b1, xold, rp, ci, nu)
        A = v
    # pin the last pressure DoF
    q0, q1 = rp[n - 1], rp[n]
    A = A.copy()
    A[q0:q1] = np.where(ci[q0:q1] == n - 1, 1.0, 0.0)
    pd = d[2 * N:3 * N]
    node_last = np.where(pd == n - 1)[0][0]
    rhs[n - 1] = pex[node_last]
    x = spla.spsolve(sp.csr_matrix((A, ci, rp), shape=(n, n)).tocsc(), rhs)
    return m, d, x


def _bd_rhs(oracle, m, d, b0, b1, xold, rp, ci, nu):
    """The penalty (boundary data) part of an assembly's RHS: assembled twice it would count twice."""
    p = oracle.NSParams(nu, 1.0, 0.0, 0.0, 0.0, (0, 0, 0), (0, 0, 0))
    _, r = oracle.ns_genmatrhs(m, 0, 1, p, d, d, b0, b1, xold, rp, ci)
    return r


def _l2_errors(oracle, m, d, x, nu, oseen):
    """L2 errors of velocity and (mean-free) pressure by the 3-point Gauss rule on every element."""
    phi2, _, w = oracle.fe_tables(2, 2)
    phi1, _, _ = oracle.fe_tables(1, 2)
    conn, _, _ = m.conn(0)
    N = m.N
    xyz = m.coords()
    gp = np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)])
    eu = ep = 0.0
    psum = area = 0.0
    pv = []
    for e in range(conn.shape[0]):
        c = conn[e]
        lo, hi = xyz[:, c[0]], xyz[:, c[2]]
        J = np.prod(hi - lo) / 4
        for qp in range(9):
            xi, eta = gp[qp // 3], gp[qp % 3]
            X = lo[0] + (xi + 1) / 2 * (hi[0] - lo[0])
            Y = lo[1] + (eta + 1) / 2 * (hi[1] - lo[1])
            ux, uy, pe, _, _ = _mms_fields(np.array(X), np.array(Y), nu, oseen)
            uh = [sum(phi2[s, qp] * x[d[c[s] + iv * N]] for s in range(9)) for iv in range(2)]
            ph = sum(phi1[s, qp] * x[d[c[s] + 2 * N]] for s in range(4))
            eu += w[qp] * J * ((uh[0] - ux) ** 2 + (uh[1] - uy) ** 2)
            pv.append((w[qp] * J, ph - pe))
    wsum = sum(a for a, _ in pv)
    mean = sum(a * b for a, b in pv) / wsum
    ep = sum(a * (b - mean) ** 2 for a, b in pv)
    return np.sqrt(eu), np.sqrt(ep)


@pytest.mark.parametrize("oseen", [False, True])
def test_mms_stokes_oseen_convergence_rates(oracle, oseen):
    """Manufactured Stokes (and Oseen, advected by the interpolated exact velocity) on 4^2, 8^2, 16^2 Q2-Q1:
    velocity L2 error rate 3 (>= 2.8 on the last refinement), pressure rate 2 (>= 1.8)."""
    nu = 0.1
    errs = []
    for nx in (4, 8, 16):
        m, d, x = _solve_ns_oracle(oracle, nx, nu, oseen)
        errs.append(_l2_errors(oracle, m, d, x, nu, oseen))
    eu = np.array([e[0] for e in errs]); ep = np.array([e[1] for e in errs])
    ru, rpr = np.log2(eu[:-1] / eu[1:]), np.log2(ep[:-1] / ep[1:])
    assert ru[-1] >= 2.8, (eu, ru)
    assert rpr[-1] >= 1.8, (ep, rpr)


def test_poiseuille_reproduced_exactly(oracle):
    """u = (4 y (1 - y), 0), p = -8 nu x solve Stokes with no body force: Q2-Q1 holds them exactly, so the
    assembled system (viscous terms, B^T with its sign, penalty Dirichlet data) returns them to rounding."""
    nu = 0.1
    m, d, x = _solve_ns_oracle(oracle, 4, nu, False, exact_poiseuille=True)
    N = m.N
    xyz = m.coords()
    for iv, ex in enumerate((4 * xyz[1] * (1 - xyz[1]), 0 * xyz[0])):
        dd = d[iv * N:(iv + 1) * N]
        np.testing.assert_allclose(x[dd[dd >= 0]], ex[dd >= 0], atol=1e-11)
    pd = d[2 * N:3 * N]
    np.testing.assert_allclose(x[pd[pd >= 0]], -8 * nu * xyz[0][pd >= 0], atol=1e-10)


# ----------------------------------------------------------------------------- curved elements
def curved_map(xyz):
    """(x, y + 0.2 x^2, z + 0.3 x y): curved, volume preserving (det DF = 1), and in the Q2 space, so the Q2
    isoparametric elements hold it exactly: the discrete domain has volume 1 and linear fields are exact."""
    out = np.array(xyz, dtype=np.float64, copy=True)
    out[1] = xyz[1] + 0.2 * xyz[0] ** 2
    if xyz.shape[0] == 3:
        out[2] = xyz[2] + 0.3 * xyz[0] * xyz[1]
    return out


def _natural_bc(nv):
    def bc(xp, neum, flag):
        for iv in range(nv):
            neum[iv], flag[iv] = 1, 0
    return bc


@pytest.mark.parametrize("dim", [2, 3])
def test_curved_heat_stiffness_on_linear_fields(oracle, dim):
    """On a curved mesh the element Jacobians are full (off-diagonal inverse terms): the assembled stiffness
    (alpha0 = 1, no advection, steady, natural boundaries) satisfies u^T K v = (c . d) |Omega| = c . d for the
    linear fields u = c . x, v = d . x; a transposed inverse Jacobian fails it."""
    m = oracle.Mesh(dim, [2] * dim, [0.0] * dim, [1.0] * dim, 1)
    m.set_coords(curved_map(m.coords()))
    xyz = m.coords()
    d, n = oracle.init_dof(m, 0, 1, 0)
    b0, b1 = oracle.genbc(m, 1, 0, d, _natural_bc(1))
    rp, ci = oracle.system_pattern(m, 0, 1, 0, d)
    p = oracle.HeatParams(1.0, 1.0, 0.1, 0.0, 0.0, 0.0, (0.0, 0.0, 0.0))
    v, _ = oracle.heat_genmatrhs(m, 0, 1, p, d, d, b0, b1, np.zeros(n), rp, ci)
    K = sp.csr_matrix((v, ci, rp), shape=(n, n))
    rng = np.random.default_rng(8)
    for _ in range(3):
        c, dd = rng.standard_normal(dim), rng.standard_normal(dim)
        u, w = np.zeros(n), np.zeros(n)
        u[d[:m.N]], w[d[:m.N]] = c @ xyz, dd @ xyz
        assert abs(w @ (K @ u) - c @ dd) <= 1e-12 * (1 + abs(c @ dd))
    # constants are in the kernel
    one = np.ones(n)
    np.testing.assert_allclose(K @ one, 0.0, atol=1e-13)


@pytest.mark.parametrize("dim", [2, 3])
def test_curved_viscous_tensor_and_divergence_on_linear_fields(oracle, dim):
    """The NS viscous block on the curved mesh with linear velocity fields u = G x, v = H x (natural boundaries,
    no advection, steady): v^T K u = nu (H : G + tr(H G)) |Omega| (the weak form of -div(nu (grad u + grad u^T)));
    the pressure rows give 1^T B u = tr(G) |Omega| (partition of unity of the Q1 shapes)."""
    m = oracle.Mesh(dim, [2] * dim, [0.0] * dim, [1.0] * dim, 1)
    m.set_coords(curved_map(m.coords()))
    xyz = m.coords()
    N = m.N
    d, n = oracle.init_dof(m, 0, dim, 1)
    b0, b1 = oracle.genbc(m, dim, 1, d, _natural_bc(dim + 1))
    rp, ci = oracle.system_pattern(m, 0, dim, 1, d)
    nu = 0.37
    p = oracle.NSParams(nu, 0.1, 0.0, 0.0, 0.0, (0, 0, 0), (0, 0, 0))
    v, _ = oracle.ns_genmatrhs(m, 0, 1, p, d, d, b0, b1, np.zeros(n), rp, ci)
    K = sp.csr_matrix((v, ci, rp), shape=(n, n))
    nq2 = m.level(0)[1]
    vel = np.concatenate([d[iv * N:iv * N + nq2] for iv in range(dim)])
    pre = d[dim * N:dim * N + m.level(0)[2]]
    rng = np.random.default_rng(9)
    for _ in range(3):
        G, H = rng.standard_normal((dim, dim)), rng.standard_normal((dim, dim))
        u, w = np.zeros(n), np.zeros(n)
        for iv in range(dim):
            u[d[iv * N:iv * N + nq2]] = G[iv] @ xyz[:, :nq2]
            w[d[iv * N:iv * N + nq2]] = H[iv] @ xyz[:, :nq2]
        Ku = K @ u
        ref = nu * (np.sum(H * G) + np.trace(H @ G))
        assert abs(w[vel] @ Ku[vel] - ref) <= 1e-12 * (1 + abs(ref))
        assert abs(Ku[pre].sum() - np.trace(G)) <= 1e-12 * (1 + abs(np.trace(G)))
