"""Problem set-ups (BC callbacks, initial conditions, coefficient fields).

bc_read / ic_read callbacks follow the reference's conventions
(template_appl/nse/SRC/User_NS.C, User_T.C): bc_Neum 0 = Dirichlet row,
bc_flag bits for NS: 4 non-homogeneous, 2 tangential, 1 normal; for T:
bit 1 = beta (non-homogeneous), bit 2 = alpha (Robin).
"""
import numpy as np

TOL = 1e-9


def heat_dirichlet_all(xp, neum, flag):
    """T = 0 on every face (homogeneous Dirichlet)."""
    neum[0] = 0
    flag[0] = 0


def cavity_bc(xp, neum, flag, dim=2):
    """Lid-driven cavity: u = (1,0[,0]) on the lid y = 1, no-slip elsewhere; p free."""
    for iv in range(dim):
        neum[iv] = 0
        flag[iv] = 0
    if xp[1] > 1.0 - TOL:
        flag[0] = 4
    # side walls and bottom override the lid on shared corners
    if xp[0] < TOL or xp[0] > 1.0 - TOL or xp[1] < TOL or (dim == 3 and (xp[2] < TOL or xp[2] > 1.0 - TOL)):
        for iv in range(dim):
            flag[iv] = 0
    neum[dim] = 1
    flag[dim] = 0


def cavity_ic(xp, dim=2, lid=None):
    """x_old: lid velocity on y = 1 (interior of the lid), zero elsewhere."""
    v = [0.0] * (dim + 1)
    if xp[1] > 1.0 - TOL and xp[0] > TOL and xp[0] < 1.0 - TOL:
        v[0] = 1.0 if lid is None else lid(xp[0])
    return v


def channel_bc(xp, neum, flag, Lx=4.0):
    """3D channel [0,Lx]x[0,1]x[0,1]: parabolic inflow at x=0, zero traction at x=Lx, no-slip walls."""
    for iv in range(3):
        neum[iv] = 0
        flag[iv] = 0
    if xp[0] < TOL:
        flag[0] = 4
    if xp[0] > Lx - TOL:
        for iv in range(3):
            neum[iv] = 1
            flag[iv] = 0
    if xp[1] < TOL or xp[1] > 1 - TOL or xp[2] < TOL or xp[2] > 1 - TOL:
        for iv in range(3):
            neum[iv] = 0
            flag[iv] = 0
    neum[3] = 1
    flag[3] = 0


def channel_ic(xp):
    v = [0.0, 0.0, 0.0, 0.0]
    if xp[0] < TOL:
        y, z = xp[1], xp[2]
        v[0] = 16.0 * y * (1 - y) * z * (1 - z)
    return v


def lognormal_k(n_top_elem, seed=20261203, mu=0.0, sigma=1.0):
    """Per-element conductivity of the top level: exp(N(mu, sigma)), one draw per element in element order."""
    rng = np.random.Generator(np.random.Philox(seed))
    return np.exp(mu + sigma * rng.standard_normal(n_top_elem))


def coarsen_coef(k_top, nlev, nch, mean="arithmetic"):
    """Per-level coefficients: each coarse element takes the mean (arithmetic, geometric or harmonic) of its
    children (children are stored parent by parent, so level L+1 element e has parent e // nch).  Coarse
    coefficients shape only the multigrid preconditioner, not the solution."""
    ks = [None] * nlev
    ks[nlev - 1] = np.asarray(k_top, np.float64)
    for L in range(nlev - 2, -1, -1):
        c = ks[L + 1].reshape(-1, nch)
        if mean == "geometric":
            ks[L] = np.exp(np.log(c).mean(axis=1))
        elif mean == "harmonic":
            ks[L] = 1.0 / (1.0 / c).mean(axis=1)
        else:
            ks[L] = c.mean(axis=1)
    return ks
