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
import ctypes as C

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


class SteadyNSProblem(OseenProblem):
    """The steady state of the same discretisation (no time term; upwinding as set in the system's parameters):
        R(U; nu, g) = A(U; nu) U - b(g) = 0,
    the fixed point of the reference's Picard step (MGSolNS::MGTimeStep, MGSolverNS.C:722-756, iterated with
    x_old <- x).  Solved by Picard steps until the relative update is below picard_tol, then Newton steps
    J_N(U) dU = b - A(U) U with the Jacobian of fm_ns_genjac, multigrid-solved, until ||R|| <= tol ||b||.
    The adjoint of J(U) at the steady state solves J_N(U)^T lambda = dJ/dU (transposed multigrid); then
    dJ/dnu_e = partial - lambda^T (dA/dnu_e) U and dJ/da_k = lambda^T b(G_k).
    Only the top level carries the Newton Jacobian (FGMRES measures its residual there); the coarse levels of
    the preconditioner carry the Picard matrix at the same state with upwinding coarse_upwind (the Galerkin
    Newton Jacobian of a coarse grid at high mesh Peclet number makes a poor coarse operator)."""

    def __init__(self, ns, G, g0=None, functional="dissipation", opts=None, tol=1e-12, picard_tol=1e-3,
                 max_picard=200, max_newton=30, coarse_upwind=0.1, newton_rtol=1e-6, picard_rtol=1e-6):
        super().__init__(ns, G, g0, functional, opts)
        self.tol, self.picard_tol = tol, picard_tol
        self.max_picard, self.max_newton = max_picard, max_newton
        self.coarse_upwind = coarse_upwind
        self.newton_rtol, self.picard_rtol = newton_rtol, picard_rtol
        self.U = None
        self.stats = {}

    def _residual(self, U):
        ns = self.ns
        top = self.nlev - 1
        ns.set_xold(U)
        ns.GenMatRhs(0.0, top, 1)
        b = ns.get_rhs()
        return b - ns.spmv(top, U), np.linalg.norm(b)

    def _step_opts(self, tol):
        """The solver options with another relative tolerance (Picard and Newton steps need not be exact)."""
        o = type(self.opts)()
        C.pointer(o)[0] = self.opts
        o.tol = tol
        return o

    def _coarse(self, transpose):
        """Picard matrices (or their transposes) with upwinding on the coarse levels, at the current x_old."""
        ns = self.ns
        up = ns.params.upwind
        ns.params.upwind = self.coarse_upwind
        try:
            for L in range(self.nlev - 1):
                if transpose:
                    ns.GenMatTranspose(L)
                else:
                    ns.GenMatRhs(0.0, L, 0)
        finally:
            ns.params.upwind = up

    def steady_state(self, U0=None):
        """Picard, then Newton, from U0 (default: the last steady state, else zero).  Returns U."""
        ns = self.ns
        top = self.nlev - 1
        N = ns._Dim[top]
        tc = ns.params.time_coeff
        ns.params.time_coeff = 0.0
        try:
            U = np.zeros(N) if U0 is None else np.array(U0, np.float64)
            npic = 0
            if U0 is None:
                for npic in range(1, self.max_picard + 1):
                    ns.set_xold(U)
                    for L in range(self.nlev):
                        ns.GenMatRhs(0.0, L, 1 if L == top else 0)
                    ns.set_x(U)
                    rep = ns.MGSolve(opts=self._step_opts(self.picard_rtol))
                    if rep.converged != 1:
                        raise RuntimeError(f"Picard step: {rep.iters} its, {rep.rel_resid:.2e}")
                    U1 = ns.get_x()
                    du = np.linalg.norm(U1 - U) / max(np.linalg.norm(U1), 1e-300)
                    U = U1
                    if du < self.picard_tol:
                        break
            rprev = np.inf
            its_newton = []
            for nnew in range(self.max_newton + 1):
                r, bn = self._residual(U)
                rn = np.linalg.norm(r)
                # converged, or at the rounding floor of the residual (no longer halving below 1e-10)
                if rn <= self.tol * bn or (rn <= 1e-10 * bn and rn > 0.5 * rprev):
                    break
                rprev = rn
                if nnew == self.max_newton:
                    raise RuntimeError(f"Newton: ||R||/||b|| = {np.linalg.norm(r) / bn:.2e} after {nnew} steps")
                ns.GenJac(top, 0)
                self._coarse(False)
                ns.set_rhs(r)
                ns.set_x(np.zeros(N))
                rep = ns.MGSolve(opts=self._step_opts(self.newton_rtol))
                its_newton.append(rep.iters)
                if rep.converged != 1 and not rep.rel_resid < 1e-3:      # inexact Newton: a reduced residual will do
                    raise RuntimeError(f"Newton step: {rep.iters} its, {rep.rel_resid:.2e}")
                U = U + ns.get_x()
            self.stats = {"picard": npic, "newton": nnew, "newton_mg_its": its_newton,
                          "rel_residual": float(np.linalg.norm(r) / bn)}
        finally:
            ns.params.time_coeff = tc
        return U

    def solve(self, nu, a, U0=None):
        self._set(nu, a)
        U = self.steady_state(U0)
        ns = self.ns
        ns.set_xold(U)
        ns.set_x(U)
        if self.functional == "dissipation":
            J, dJdU, dJdnu = ns.dissipation(U)
        else:
            J, dJdU = ns.pressure_drop(U)
            dJdnu = np.zeros(len(nu))
        self.U = U
        self.last = (U, dJdU, dJdnu)
        return J

    def gradient(self):
        ns = self.ns
        U, dJdU, dJdnu = self.last
        ns.set_xold(U)
        tc = ns.params.time_coeff
        ns.params.time_coeff = 0.0
        try:
            ns.GenJac(self.nlev - 1, 1)
            self._coarse(True)
        finally:
            ns.params.time_coeff = tc
        lam, rep = ns.AdjointSolve(dJdU, self.opts)
        self.stats["adjoint_mg_its"] = rep.iters
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
