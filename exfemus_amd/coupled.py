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


def fgmres(matvec, precond, b, tol=1e-10, maxit=100, restart=50, callback=None):
    """Right-preconditioned flexible GMRES on host vectors (Saad 1993): the operator and the preconditioner
    (inexact multigrid solves) run on the GPU.  Returns (x, iterations, relative residual)."""
    n = len(b)
    x = np.zeros(n)
    bn = np.linalg.norm(b)
    if bn == 0.0:
        return x, 0, 0.0
    r = b.copy()
    its = 0
    while its < maxit:
        beta = np.linalg.norm(r)
        if beta <= tol * bn:
            break
        m = min(restart, maxit - its)
        V = np.zeros((m + 1, n))
        Z = np.zeros((m, n))
        H = np.zeros((m + 1, m))
        V[0] = r / beta
        k = 0
        for k in range(m):
            Z[k] = precond(V[k])
            w = matvec(Z[k])
            # classical Gram-Schmidt; a second pass only when the first lost more than half of |w|^2 (Daniel,
            # Gragg, Kaufman and Stewart, Math. Comp. 30 (1976))
            w0 = np.linalg.norm(w)
            for p in range(2):
                h = V[:k + 1] @ w
                w -= h @ V[:k + 1]
                H[:k + 1, k] += h
                wn = np.linalg.norm(w)
                if p == 0 and wn >= w0 / np.sqrt(2.0):
                    break
            H[k + 1, k] = wn
            its += 1
            e1 = np.zeros(k + 2)
            e1[0] = beta
            y = np.linalg.lstsq(H[:k + 2, :k + 1], e1, rcond=None)[0]
            res = np.linalg.norm(H[:k + 2, :k + 1] @ y - e1)
            if callback:
                callback(its, res / bn)
            if H[k + 1, k] == 0.0 or res <= tol * bn:
                break
            V[k + 1] = w / H[k + 1, k]
        x += y @ Z[:k + 1]
        r = b - matvec(x)
    return x, its, float(np.linalg.norm(r) / bn)


class Boussinesq:
    def __init__(self, ctx, dim, nx0, nlev, Ra, Pr=0.71, dt=1.0, upwind_ns=0.1, upwind_T=1.0, coarse_upwind_ns=None,
                 coarse_upwind_T=None):
        """coarse_upwind_*: upwinding of the coarse levels of each system's multigrid (the preconditioner only;
        None = the system's own).  With Galerkin top levels (upwind 0) an upwinded coarse operator keeps the
        V-cycle effective where the coarse grids' mesh Peclet numbers exceed 1."""
        self.dim = dim
        self.cu_ns = upwind_ns if coarse_upwind_ns is None else coarse_upwind_ns
        self.cu_T = upwind_T if coarse_upwind_T is None else coarse_upwind_T
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
        self.T.set_x(self.xT)                    # the device iterate is the coupling source
        self.buoy = [0.0] * 3
        self.buoy[dim - 1] = 1.0                 # gravity along -y (2D) / -z (3D)
        self.o_ns = self.ns.opts(precision=F.FM_PREC_FP32_FGMRES, smoother=F.FM_SMOOTH_VANKA, vanka_colored=1,
                                 nu_pre=3, nu_post=3, coarse_mult=1, tol=1e-9, max_iter=200)
        self.o_T = self.T.opts(precision=F.FM_PREC_FP32_FGMRES, nu_pre=3, nu_post=3, cheb_lo=0.1, tol=1e-10,
                               max_iter=200)

    @property
    def xT(self):
        """Temperature of the current step (top level, reference order), fetched from the device on demand."""
        if self._xT is None:
            self._xT = self.T.get_x()
        return self._xT

    @xT.setter
    def xT(self, v):
        self._xT = v

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

    def _assemble(self, sysm, t, coarse_upwind, transpose=False):
        """All levels of one system: the top with its rhs and own upwinding, coarse levels with coarse_upwind
        (MGTimeStep's assembly, src/MGSolverBase.C:426-454); transpose = 1 adds the transposed matrices."""
        top = self.mesh.nlev - 1
        sysm.GenMatRhs(t, top, 1)
        if transpose:
            sysm.GenMatTranspose(top)
        up = sysm.params.upwind
        sysm.params.upwind = coarse_upwind
        try:
            for L in range(top):
                sysm.GenMatRhs(t, L, 0)
                if transpose:
                    sysm.GenMatTranspose(L)
        finally:
            sysm.params.upwind = up

    def _timestep(self, sysm, t, opts, coarse_upwind):
        """Assembly and solve of one system; x_old <- x and the relative change on the device (MGTimeStep)."""
        self._assemble(sysm, t, coarse_upwind)
        rep = sysm.MGSolve(opts=opts)
        return rep, sysm.advance()

    def step(self, t=0.0):
        """One pseudo-time step of each system; returns the relative changes (NS, T).  The two multigrid
        reports are kept in self.last_reps.  The coupling fields (the temperature driving the buoyancy, the
        velocity advecting the heat) pass from one system to the other on the device."""
        self.ns.couple(self.T, 0, self.buoy)
        rep_ns, du = self._timestep(self.ns, t, self.o_ns, self.cu_ns)
        if rep_ns.converged != 1:
            raise RuntimeError(f"NS step: {rep_ns.iters} its, {rep_ns.rel_resid:.2e}")
        self.T.couple(self.ns, 1)
        rep_T, dT = self._timestep(self.T, t, self.o_T, self.cu_T)
        if rep_T.converged != 1:
            raise RuntimeError(f"T step: {rep_T.iters} its, {rep_T.rel_resid:.2e}")
        self.last_reps = (rep_ns, rep_T)
        self._xT = None
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

    def nusselt_gradient_coupled(self, tol=1e-10, maxit=200, inner_tol=1e-2, inner_maxit=4,
                                 inner_precision=F.FM_PREC_FP32_FGMRES, callback=None, lean=False):
        """dNu/dk_e of the steady coupled state, through the adjoint of the whole coupled system (velocity not
        frozen): with R_NS(U, T) = 0, R_T(T, U) = 0 at the state,
            [ J_NS^T     (dR_T/dU)^T ] [lam_U]   [dNu/dU]
            [ (dR_NS/dT)^T    A_T^T  ] [lam_T] = [dNu/dT],   dNu/dk_e = partial - lam_T^T (dA_T/dk_e) T,
        J_NS the Newton Jacobian of the NS residual, A_T the heat matrix at the state's velocity, dR_NS/dT = -coef M
        (buoyancy), (dR_T/dU) the derivative of the advection.  Solved by flexible GMRES with the two transposed
        multigrid solves as a block-diagonal preconditioner.  The upwind terms of both systems are differentiated
        (fm_ns_genjac, fm_heat_coupling); the state must be steady (the time terms are left out of the Jacobian).
        lean: release the forward NS multigrid data first and keep the transposed hierarchy's Vanka inverses in
        the inner solves' precision only (the memory of the largest meshes).  Returns (Nu, gradient, info)."""
        ns, T = self.ns, self.T
        if lean:
            ns.release_mg()
            fp32 = inner_precision in (F.FM_PREC_FP32_IR, F.FM_PREC_FP32_FGMRES)
            store0 = ns.set_vanka_store(2 if fp32 else 1)
            try:
                return self.nusselt_gradient_coupled(tol, maxit, inner_tol, inner_maxit, inner_precision, callback)
            finally:
                ns.release_mg(adjoint=True)
                ns.set_vanka_store(store0)
        top = self.mesh.nlev - 1
        N, dim = self.mesh.N, self.dim
        U = ns.get_x()
        Tv = self.xT.copy()
        tcn, tct = ns.params.time_coeff, T.params.time_coeff
        ns.params.time_coeff = 0.0
        T.params.time_coeff = 0.0
        try:
            ns.set_xold(U)
            ns.GenJac(top, 1)
            up = ns.params.upwind
            ns.params.upwind = self.cu_ns if self.cu_ns != 0.0 else 0.1   # coarse levels: preconditioner only
            try:
                for L in range(top):
                    ns.GenMatTranspose(L)
            finally:
                ns.params.upwind = up
            T.set_velocity(ns.velocity_nodes(U))
            T.set_xold(Tv)
            self._assemble(T, 0.0, self.cu_T, transpose=True)
        finally:
            ns.params.time_coeff = tcn
            T.params.time_coeff = tct
        Nu, dJdT, dJdk = T.nusselt(Tv, axis=0)
        ndU = ns.node_dof(top)
        ndT = T.node_dof(top)[:N]
        bc0U = ns.bc()[0]
        nU, nT = len(U), len(Tv)
        kappa, rhocp = T.params.alpha0, T.params.rhocp
        ib = dim - 1                                  # the buoyant velocity component
        coef = self.buoy[ib]
        okT = ndT >= 0

        def to_U(nodal, comp, out):
            d = ndU[comp * N:(comp + 1) * N]
            ok = d >= 0
            out[d[ok]] += nodal[ok]

        gU = np.zeros(nU)
        to_U(T.coupling(0, 1.0 / kappa, T=Tv), 0, gU)
        g = np.concatenate([gU, dJdT])

        def matvec(v):
            lU, lT = v[:nU], v[nU:]
            yU = ns.adjoint_spmv(top, lU)
            cT = T.coupling(1, rhocp, T=Tv, lam=lT)
            for d in range(dim):
                to_U(cT[d * N:(d + 1) * N], d, yU)
            db = ndU[ib * N:(ib + 1) * N]
            m = np.zeros(N)
            okb = db >= 0
            m[okb] = lU[db[okb]] * (bc0U[db[okb]] != 0)
            Mm = T.coupling(2, -coef, m=m)
            yT = T.adjoint_spmv(top, lT)
            yT[ndT[okT]] += Mm[okT]
            return np.concatenate([yU, yT])

        # inexact block solves, a bounded number of mixed-precision iterations each (flexible GMRES absorbs the
        # variation; the outer residual is FP64).  Measured at 48^3 (Ra 1e5, to 1e-8): at most 15 inner
        # iterations to 1e-4 take 42 outer iterations and 22.2 s, at most 4 to 1e-2 take 48 and 9.4 s, one takes
        # 65 and 6.9 s; the outer Krylov vectors live on the host, so the outer count is kept moderate.
        o_ns = ns.opts(precision=inner_precision, smoother=F.FM_SMOOTH_VANKA, vanka_colored=1, nu_pre=3,
                       nu_post=3, coarse_mult=1, tol=inner_tol, max_iter=inner_maxit)
        o_T = T.opts(precision=inner_precision, nu_pre=3, nu_post=3, cheb_lo=0.1, tol=inner_tol, max_iter=inner_maxit)
        inner = []

        def precond(r):
            zU, rU = ns.AdjointSolve(r[:nU], o_ns)
            zT, rT = T.AdjointSolve(r[nU:], o_T)
            inner.append((rU.iters, rT.iters))
            return np.concatenate([zU, zT])

        lam, its, rel = fgmres(matvec, precond, g, tol=tol, maxit=maxit, callback=callback)
        if not rel <= tol:
            raise RuntimeError(f"coupled adjoint: {its} its, {rel:.2e}")
        grad = dJdk + T.grad_coef(lam[nU:], Tv)
        return Nu, grad, {"fgmres_its": its, "rel_resid": rel, "inner_its": inner}

    def nusselt_gradient_k(self, opts=None):
        """Steady heat problem with the current velocity frozen (time term off): the temperature, its Nusselt
        number and dNu/dk_e from one transposed multigrid solve: dNu/dk_e = partial - lambda^T (dA/dk_e) T.
        Returns (Nu, gradient, T)."""
        top = self.mesh.nlev - 1
        tc = self.T.params.time_coeff
        self.T.params.time_coeff = 0.0
        try:
            self.T.set_xold(self.xT)             # Dirichlet data
            self._assemble(self.T, 0.0, self.cu_T, transpose=True)
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
