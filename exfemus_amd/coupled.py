"""Coupled heat + Navier-Stokes (Boussinesq) driver: the differentially heated box.

The reference couples its equations through external fields (MGSolDA::_data_eq,
src/MGSolverDA.C:101-543) and steps them in turn (MGEquationsSystem::
eqnmap_timestep_loop, src/MGEquationsSystem.C:422-433); its nse application keeps
the temperature dependence of the density commented out (MGSolverNS.C:344).
Here the two systems share the mesh and are stepped in turn in pseudo-time:
  Navier-Stokes (Oseen step) with the body force T e_up,
  heat (implicit step) advected by the new velocity,
in convective units: nu = sqrt(Pr / Ra), kappa = 1 / sqrt(Ra Pr).  Walls:
no-slip; T = 1 at x = 0 (hot), T = 0 at x = 1 (cold), adiabatic elsewhere.
The Nusselt number is the box-averaged heat flux along x in units of the
conductive one (fm_heat_nusselt).
"""
import os

import numpy as np

from . import configs, femus as F

TOL = configs.TOL


def dhc_ns_bc(xp, neum, flag, dim=3):
    for iv in range(dim):
        neum[iv] = 0
        flag[iv] = 0
    neum[dim] = 1
    flag[dim] = 0


def dhc_T_bc(xp, neum, flag):
    if xp[0] < TOL or xp[0] > 1.0 - TOL:
        neum[0] = 0          # Dirichlet, value from x_old (T = 1 - x at start)
        flag[0] = 1
    else:
        neum[0] = 1          # adiabatic
        flag[0] = 0


class Boussinesq:
    def __init__(self, ctx, dim, nx0, nlev, Ra, Pr=0.71, dt=1.0, upwind_ns=0.1, upwind_T=1.0):
        self.dim = dim
        self.mesh = F.Mesh(dim, [nx0] * dim, [0.0] * dim, [1.0] * dim, nlev)
        self.nu = np.sqrt(Pr / Ra)
        self.kappa = 1.0 / np.sqrt(Ra * Pr)
        self.ns = F.MGSolNS(ctx, self.mesh, IRe=self.nu, dt=dt, time_coeff=1.0, upwind=upwind_ns)
        self.ns.GenBc(lambda xp, n, f: dhc_ns_bc(xp, n, f, dim))
        self.ns.init()
        self.ns.GenIc(lambda xp: [0.0] * (dim + 1))
        self.T = F.MGSolT(ctx, self.mesh, alpha0=self.kappa, dt=dt, rhocp=1.0, time_coeff=1.0, source=0.0,
                          upwind=upwind_T)
        self.T.GenBc(dhc_T_bc)
        self.T.init()
        self.xT = self.T.GenIc(lambda xp: [1.0 - xp[0]])
        self.buoy = [0.0] * 3
        self.buoy[dim - 1] = 1.0                 # gravity along -y (2D) / -z (3D)
        self.o_ns = self.ns.opts(precision=F.FM_PREC_FP32_FGMRES, smoother=F.FM_SMOOTH_VANKA, vanka_colored=1,
                                 nu_pre=3, nu_post=3, coarse_mult=1, tol=1e-9, max_iter=200)
        self.o_T = self.T.opts(precision=F.FM_PREC_FP32_FGMRES, nu_pre=3, nu_post=3, cheb_lo=0.1, tol=1e-10,
                               max_iter=200)

    def state(self):
        """(velocity-pressure vector, temperature vector) of the current step, top level, reference order."""
        return self.ns.get_x().copy(), self.xT.copy()

    def set_state(self, st):
        U, T = st
        self.ns.set_x(U)
        self.ns.set_xold(U)
        self.T.set_x(T)
        self.T.set_xold(T)
        self.T.set_velocity(self.ns.velocity_nodes(U))
        self.xT = T.copy()

    def step(self, t=0.0):
        """One pseudo-time step of each system; returns the relative changes (NS, T).  The two multigrid
        reports are kept in self.last_reps."""
        T_old = self.xT
        u_prev = self.ns.get_x()
        self.ns.set_buoyancy(self.T.node_values(T_old), self.buoy)
        rep_ns = self.ns.MGTimeStep(t, opts=self.o_ns)
        if rep_ns.converged != 1:
            raise RuntimeError(f"NS step: {rep_ns.iters} its, {rep_ns.rel_resid:.2e}")
        u = self.ns.get_x()
        self.T.set_velocity(self.ns.velocity_nodes(u))
        rep_T = self.T.MGTimeStep(t, opts=self.o_T)
        if rep_T.converged != 1:
            raise RuntimeError(f"T step: {rep_T.iters} its, {rep_T.rel_resid:.2e}")
        self.last_reps = (rep_ns, rep_T)
        T_new = self.T.get_x()
        self.xT = T_new
        du = np.linalg.norm(u - u_prev) / max(np.linalg.norm(u), 1e-300)
        dT = np.linalg.norm(T_new - T_old) / max(np.linalg.norm(T_new), 1e-300)
        return du, dT

    def names(self):
        return ["u", "v", "w"][:self.dim] + ["p"], ["T"]

    def print_step(self, outdir, step, time):
        """sol.<step>.bin/.json with u, v, (w), p and T on the top level, its XDMF file, and the mesh file
        once (print_u of both systems + print_soln_xmf)."""
        from . import solio
        os.makedirs(outdir, exist_ok=True)
        top = self.mesh.nlev - 1
        mbase = os.path.join(outdir, "mesh")
        if not os.path.exists(mbase + ".json"):
            solio.print_mesh(self.mesh, top, mbase)
        base = os.path.join(outdir, f"sol.{step:04d}")
        ns_names, T_names = self.names()
        solio.print_u(self.mesh, top, base,
                      [(ns_names, self.ns.node_dof(top), self.ns.get_x(), self.dim, 1, None),
                       (T_names, self.T.node_dof(top), self.xT, 1, 0, None)])
        solio.print_xmf(base + ".xmf", time, self.mesh, top, mbase, base, ns_names + T_names)
        return base + ".xmf"

    def read_step(self, outdir, step):
        """Restart from sol.<step> (read_u of both systems); the next step continues from that state."""
        base = os.path.join(outdir, f"sol.{step:04d}")
        ns_names, T_names = self.names()
        U = self.ns.read_u(base, ns_names)
        self.xT = self.T.read_u(base, T_names)
        self.T.set_velocity(self.ns.velocity_nodes(U))

    def run(self, max_steps=400, tol=1e-8, callback=None):
        """Pseudo-time steps until both relative changes fall below tol; returns the number of steps."""
        for k in range(max_steps):
            du, dT = self.step(float(k + 1))
            if callback:
                callback(k, du, dT)
            if du < tol and dT < tol:
                return k + 1
        return max_steps

    def nusselt(self):
        """(Nu, dNu/dT, dNu/dk_e partials) at the current state, advection field frozen."""
        return self.T.nusselt(self.xT, axis=0)

    def nusselt_gradient_k(self, opts=None):
        """Steady heat problem with the current velocity frozen (time term off): the temperature, its Nusselt
        number and dNu/dk_e from one transposed multigrid solve: dNu/dk_e = partial - lambda^T (dA/dk_e) T.
        Returns (Nu, gradient, T)."""
        top = self.mesh.nlev - 1
        tc = self.T.params.time_coeff
        self.T.params.time_coeff = 0.0
        try:
            self.T.set_xold(self.xT)             # Dirichlet data
            for L in range(self.mesh.nlev):
                self.T.GenMatRhs(0.0, L, 1 if L == top else 0)
                self.T.GenMatTranspose(L)
            o = opts or self.T.opts(precision=F.FM_PREC_FP64_FGMRES, nu_pre=3, nu_post=3, cheb_lo=0.1, tol=1e-12,
                                    max_iter=200)
            self.T.set_x(self.xT)
            rep = self.T.MGSolve(opts=o)
            if rep.converged != 1:
                raise RuntimeError(f"steady heat: {rep.iters} its, {rep.rel_resid:.2e}")
            Ts = self.T.get_x()
            J, dJdT, dJdk = self.T.nusselt(Ts, axis=0)
            lam, rep = self.T.AdjointSolve(dJdT, o)
            if rep.converged != 1:
                raise RuntimeError(f"adjoint: {rep.iters} its, {rep.rel_resid:.2e}")
            return J, dJdk + self.T.grad_coef(lam, Ts), Ts
        finally:
            self.T.params.time_coeff = tc
