"""The GPU discrete adjoint as a torch.autograd.Function, for gradient checks.

One Oseen (Picard) step of the coupled Navier-Stokes system, A(nu) U = b(g),
with the advection field frozen at x_old (MGSolNS::GenMatRhs,
template_appl/nse/SRC/MGSolverNS.C:108-717), per-element viscosity nu_e on
the top level (coarse levels take the mean of their children; they shape only
the preconditioner) and Dirichlet data g = g0 + sum_k a_k G_k on the boundary
(the lid profile of the cavity, the inflow profile of the channel).

Forward: assemble every level, A^T on every level, one multigrid solve, and
the scalar functional J(U): the viscous dissipation sum_e nu_e int_e |grad u|^2
or the inlet-outlet pressure drop.  Backward: one transposed multigrid solve
A^T lambda = dJ/dU, then
    dJ/dnu_e = partial_nu_e J - lambda^T (dA/dnu_e) U,
    dJ/da_k  = lambda^T b_bd(G_k)      (b is linear in the boundary data).
Everything runs through the C ABI; torch only carries the tensors.
"""
import numpy as np
import torch

from . import configs


class OseenProblem:
    """State shared by forward and backward: the system, the parameter maps and the solver options."""

    def __init__(self, ns, G, g0=None, functional="dissipation", opts=None):
        self.ns = ns
        self.G = np.asarray(G, np.float64)          # (K, ndof_top) boundary-data basis
        self.g0 = np.zeros(self.G.shape[1]) if g0 is None else np.asarray(g0, np.float64)
        self.functional = functional
        self.opts = opts
        self.nlev = ns._NoLevels
        self.nch = 2 ** ns.mesh.dim
        self.last = None

    def _set(self, nu, a):
        ks = configs.coarsen_coef(nu, self.nlev, self.nch)
        for L in range(self.nlev):
            self.ns.set_elem_coef(L, ks[L])
        self.ns.set_bcdata(self.g0 + a @ self.G)

    def solve(self, nu, a):
        ns = self.ns
        self._set(nu, a)
        top = self.nlev - 1
        for L in range(self.nlev):
            ns.GenMatRhs(0.0, L, 1 if L == top else 0)
            ns.GenMatTranspose(L)
        ns.set_x(np.zeros(ns._Dim[top]))
        rep = ns.MGSolve(opts=self.opts)
        if rep.converged != 1:
            raise RuntimeError(f"forward multigrid did not converge: {rep.iters} its, {rep.rel_resid:.3e}")
        U = ns.get_x()
        if self.functional == "dissipation":
            J, dJdU, dJdnu = ns.dissipation(U)
        else:
            J, dJdU = ns.pressure_drop(U)
            dJdnu = np.zeros(len(nu))
        self.last = (U, dJdU, dJdnu)
        return J

    def gradient(self):
        ns = self.ns
        U, dJdU, dJdnu = self.last
        lam, rep = ns.AdjointSolve(dJdU, self.opts)
        if rep.converged != 1:
            raise RuntimeError(f"adjoint multigrid did not converge: {rep.iters} its, {rep.rel_resid:.3e}")
        g_nu = dJdnu + ns.grad_coef(lam, U)
        g_a = np.array([lam @ ns.bc_rhs(Gk) for Gk in self.G])
        return g_nu, g_a


class OseenFunctional(torch.autograd.Function):
    """J(nu, a) with the GPU adjoint as its backward (float64 tensors on the CPU)."""

    @staticmethod
    def forward(ctx, nu, a, problem):
        J = problem.solve(nu.detach().cpu().numpy().astype(np.float64), a.detach().cpu().numpy().astype(np.float64))
        g_nu, g_a = problem.gradient()
        ctx.save_for_backward(torch.from_numpy(g_nu), torch.from_numpy(g_a))
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        g_nu, g_a = ctx.saved_tensors
        return gJ * g_nu, gJ * g_a, None


def lid_fourier_basis(ns, K, dim=2):
    """G_k: sin(k pi x) on the u-DoFs of the lid y = 1 whose boundary flag is non-homogeneous (bit 4)."""
    top = ns._NoLevels - 1
    N = ns.mesh.N
    xyz = ns.mesh.coords()
    nd = ns.node_dof(top)
    _, bc1 = ns.bc()
    G = np.zeros((K, ns._Dim[top]))
    for node in range(N):
        d = nd[node]
        if d < 0 or xyz[1, node] < 1.0 - configs.TOL or not (bc1[d] & 4):
            continue
        for k in range(K):
            G[k, d] = np.sin((k + 1) * np.pi * xyz[0, node])
    return G


def control_loop(problem, nu, a, eta=1.0, max_func_iter=50, toll_conv=1e-6, which="lid", bounds=None):
    """Steepest-descent control driven by the discrete gradient, with the step rule of the reference's
    optimal-control loop (MGEquationsSystem::eqnmap_timestep_loop_control, src/MGEquationsSystem.C:435-561):
    solve the state, solve the adjoint, try a step of size eta along the negative gradient; a step that lowers
    the functional is kept and eta grows by 1.5, a step that raises it is undone and eta shrinks by 0.66; stop
    when two consecutive functionals differ by less than toll_conv (relative) or after max_func_iter trials.

    which = "lid" moves the boundary-data coefficients a, "nu" the element viscosities nu (kept inside bounds,
    a (lo, hi) pair, when given).  Returns (nu, a, history) with history = [(J, eta, accepted), ...]."""
    nu = np.asarray(nu, np.float64).copy()
    a = np.asarray(a, np.float64).copy()
    J0 = problem.solve(nu, a)
    g_nu, g_a = problem.gradient()
    history = [(J0, eta, True)]
    for _ in range(max_func_iter):
        if which == "lid":
            step = -eta * g_a
            a1, nu1 = a + step, nu
        else:
            nu1 = nu - eta * g_nu * nu          # relative step on the (positive) viscosities
            if bounds is not None:
                nu1 = np.clip(nu1, bounds[0], bounds[1])
            a1 = a
        J1 = problem.solve(nu1, a1)
        if abs(J0 - J1) / abs(J0) < toll_conv:
            history.append((J1, eta, J1 < J0))
            if J1 < J0:
                nu, a, J0 = nu1, a1, J1
            break
        if J1 < J0:
            nu, a, J0 = nu1, a1, J1
            g_nu, g_a = problem.gradient()
            history.append((J1, eta, True))
            eta *= 1.5
        else:
            history.append((J1, eta, False))
            eta *= 0.66
    problem.solve(nu, a)                        # leave the state at the accepted control
    return nu, a, history
