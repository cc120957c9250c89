#!/usr/bin/env python3
"""Benchmark: MDoF/s to 1e-8 relative residual on one MI355X.

Workload (default): 3D scalar heat -div(k grad T) = 1 on the unit cube, Q2
hex27, 4^3 coarse -> 64^3 top (5 multigrid levels, 2,146,689 top DoF),
per-element k = exp(N(0,1)) (Philox seed 20261203), T = 0 on every face.
A "step" is one full solve from x = 0 with the hierarchy assembled:
FP32 V-cycle (FP32 values/accumulation) inside FP64 iterative refinement /
FGMRES until ||b - A x||_2 / ||b||_2 <= 1e-8 with the residual in FP64.
The all-FP64 multigrid solve of the same system is timed beside it, and the
top-level SpMV is timed against the HBM roofline.  The CPU baseline is the
oracle's restatement of the reference CPU path (GMRES(30)+ILU(0) smoothed
V-cycle, src/MGSolverBase.C:163-289) on a bounded sample, one host core.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_heat(F, ctx, nx0, nlev, k_sigma=1.0):
    from exfemus_amd import configs
    m = F.Mesh(3, nx0, [0, 0, 0], [1, 1, 1], nlev)
    s = F.MGSolT(ctx, m, alpha0=1.0, dt=0.1, time_coeff=0.0, source=1.0, upwind=1.0)
    s.GenBc(configs.heat_dirichlet_all)
    s.init()
    top = nlev - 1
    k_top = configs.lognormal_k(m.level(top)[0], sigma=k_sigma)
    # coarse levels: geometric mean of the children (log-normal field); it shapes only the preconditioner
    # (11 instead of 16 FGMRES iterations against the arithmetic mean at 64^3)
    ks = configs.coarsen_coef(k_top, nlev, 8, mean="geometric")
    for L in range(nlev):
        s.set_elem_coef(L, ks[L])
    s.set_xold(np.zeros(s._Dim[top]))
    t0 = time.perf_counter()
    for L in range(nlev):
        s.GenMatRhs(0.0, L, 1 if L == top else 0)
    t_asm = time.perf_counter() - t0
    return m, s, t_asm


def assembly_rate(F, ctx, s, heat, reps=5):
    """Top-level GenMatRhs (matrix + rhs) timed alone: elements/s and two byte rates.  GBs_out: the bytes an
    assembly has to produce -- every SELL slot of the level matrix in FP64 and FP32 (12 B) and the rhs (8 B per
    DoF) -- plus the element inputs (per element its Q2 nodes' coordinates and nodal data, 8 B each).  GBs_route:
    the bytes the element-buffer route moves -- the element matrices written and read back (2 x NL^2 x 8 B per
    element), the gather's index (per slot a count and a start, 3 B; per element-matrix entry a 2-byte LDS
    position), the outputs and the inputs."""
    m = s.mesh
    top = m.nlev - 1
    nel = m.level(top)[0]
    n2 = 27 if m.dim == 3 else 9
    nl = n2 if heat else m.dim * n2 + (8 if m.dim == 3 else 4)
    info = s.sell_info(0, top)
    ts = []
    for _ in range(reps):
        ctx.sync()
        t0 = time.perf_counter()
        s.GenMatRhs(0.0, top, 1)
        ctx.sync()
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    n = s._Dim[top]
    inputs = nel * n2 * (m.dim + (1 if heat else m.dim)) * 8
    out = info["slots"] * 12 + n * 8
    route = 2 * nel * nl * nl * 8 + info["slots"] * 3 + nel * nl * nl * 2 + out + inputs
    return {"elements_top": nel, "ms_top": round(t * 1e3, 3), "Melements_per_s": round(nel / t / 1e6, 2),
            "GBs_out": round((out + inputs) / t / 1e9, 1), "GBs_route": round(route / t / 1e9, 1),
            "bytes_out": out + inputs, "bytes_route": route}


def layout_key(kernel, info):
    """Key of a SELL layout for the committed PMC summary: the kernel and every count that sets its bytes."""
    return "|".join([kernel] + [str(info[k]) for k in ("slices", "slots", "slots_c8", "slots_c16", "slots_c32",
                                                        "nnz", "slots_c2")])


def pmc_traffic(kernel, info):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary (FETCH_SIZE + WRITE_SIZE,
    scaled by the copy kernel's readings), when it was measured on this exact SELL layout; otherwise None."""
    path = os.path.join(ROOT, "profiles", "sell_heat64_pmc_traffic.json")
    try:
        d = json.load(open(path))
    except (OSError, ValueError):
        return None
    return d.get("layouts", {}).get(layout_key(kernel, info), {}).get("traffic_bytes")


def spmv_bytes(n, nnz, v, nvec=2):
    """Algorithmic bytes of one CSR product with int32 columns: values and columns once, row pointers, and
    nvec vector streams of n entries (x gathered once + y written once = 2)."""
    p = 8 if nnz > 2 ** 31 else 4
    return nnz * (v + 4) + (n + 1) * p + nvec * n * v


def sell_bytes(info, n, v, nvec=2):
    """Algorithmic bytes of one product in the SELL-64 format as stored: every slot's value (padding
    included), its 2-bit lane-predicted residual or 8-, 16- or 32-bit column offset, one int32 base per 64
    slots, the slice headers (offsets into values and the four column arrays, width: 44 B) and the slice's 64
    row indices, plus nvec vector streams of n entries."""
    return (info["slots"] * v + info.get("slots_c2", 0) // 4 + info["slots_c8"] + 2 * info["slots_c16"] +
            4 * info["slots_c32"] + info["slots"] // 64 * 4 + info["slices"] * (44 + 64 * 4) + nvec * n * v)


def roofline_measure(s, top, N, reps):
    """Top-level products timed with HIP events on the library stream: y = A x (FP64, FP32) and the Chebyshev
    smoothing step (FP32: the kernel the mixed solve spends most time in, and FP64), with their format and
    CSR-equivalent byte counts."""
    info = s.sell_info(0, top)
    nnz = info["nnz"]
    out = {}
    for name, prec, v, nvec in (("y_fp64", 0, 8, 2), ("y_fp32", 1, 4, 2), ("cheb_fp64", 4, 8, 7),
                                ("cheb_fp32", 5, 4, 7)):
        sec = s.bench_spmv(top, prec, reps)
        bf, bc = sell_bytes(info, N, v, nvec), spmv_bytes(N, nnz, v, nvec)
        out[name] = {"us": round(sec * 1e6, 2), "bytes": bf, "GBs": round(bf / sec / 1e9, 1),
                     "frac": round(bf / sec / 1e9 / HBM_PEAK_GBS, 4), "csr_equiv_bytes": bc,
                     "csr_equiv_frac": round(bc / sec / 1e9 / HBM_PEAK_GBS, 4)}
    return info, out


KERNELS = {"y_fp64": "k_sell<double, 0, 8>", "y_fp32": "k_sell<float, 0, 8>", "cheb_fp64": "k_sell<double, 2, 8>",
           "cheb_fp32": "k_sell<float, 2, 8>"}


# Cycles the oracle's V-cycle (GMRES(30)+ILU(0) 8/8 smoothing, the reference path) needs on the benchmark's
# hierarchy to reach ||b - A x|| / ||b|| <= 1e-8, from one full run of the same solve (deterministic: the same
# on any host; 9.13e-9 after cycle 40, 13.5 s per cycle on a 2.7 GHz Xeon core); the bench times a bounded
# number of its cycles and scales.
CPU_CYCLES_TO_TOL = 40


def cpu_baseline(cycles=1):
    """Oracle (restated reference CPU path) on the benchmark's own workload (3D heat hex27 64^3, 5 levels, the
    same log-normal k and geometric-mean coarse coefficients), one core: GMRES(30)+ILU(0) 8/8 smoothing (x2 below
    the top), 60-iteration GMRES coarse solve, the reference's 0.95 conditional correction.  The assembly, the
    transfers and the ILU factors (the smoother's set-up, kept by LASPACK across smoother calls, as the GPU's
    multigrid set-up is kept across solves) are built first and not timed; `cycles` V-cycles from x = 0 are
    timed, and the throughput to 1e-8 uses the cycle count of a full run."""
    from oracle import oracle as O
    from exfemus_amd import configs
    nx0, nlev = [4, 4, 4], 5
    t0 = time.perf_counter()
    m = O.Mesh(3, nx0, [0, 0, 0], [1, 1, 1], nlev)
    top = nlev - 1
    dofs = [O.init_dof(m, L, 1, 0)[0] for L in range(nlev)]
    b0, b1 = O.genbc(m, 1, 0, dofs[top], configs.heat_dirichlet_all)
    k_top = configs.lognormal_k(m.level(top)[0])
    ks = configs.coarsen_coef(k_top, nlev, 8, mean="geometric")
    p = O.HeatParams(1.0, 1.0, 0.1, 0.0, 1.0, 1.0, (0.0, 0.0, 0.0))
    H = O.Hier(nlev)
    xold = np.zeros(len(b0))
    b = None
    for L in range(nlev):
        rp, ci = O.system_pattern(m, L, 1, 0, dofs[L])
        v, rhs = O.heat_genmatrhs(m, L, 1 if L == top else 0, p, dofs[L], dofs[top], b0, b1, xold, rp, ci, k_elem=ks[L])
        H.set_A(L, rp, ci, v)
        if L == top:
            b = rhs
        if L < top:
            r = O.system_rest(m, L, 1, 0, dofs[L], dofs[L + 1], dofs[top], b0)
            H.set_R(L, r[0], r[1], r[2], len(dofs[L + 1][dofs[L + 1] >= 0]))
        if L > 0:
            q = O.system_prol(m, L, 1, 0, dofs[L], dofs[L - 1])
            H.set_P(L, q[0], q[1], q[2], len(dofs[L - 1][dofs[L - 1] >= 0]))
        del rp, ci, v
        log(f"cpu leg: level {L} assembled ({time.perf_counter() - t0:.0f} s)")
    n = len(b)
    t_setup = time.perf_counter() - t0
    # the ILU factors of every level are formed by the first smoother call
    t0 = time.perf_counter()
    H.solve(b, smoother=0, coarse=0, correction=0, nu_pre=8, nu_post=8, max_cycles=1, eps=1e-30, ref_exit=0)
    t_ilu = time.perf_counter() - t0
    log(f"cpu leg: ILU factors and a first cycle {t_ilu:.0f} s")
    t0 = time.perf_counter()
    x, hist = H.solve(b, smoother=0, coarse=0, correction=0, nu_pre=8, nu_post=8, max_cycles=cycles, eps=1e-8,
                      ref_exit=0)
    dt = time.perf_counter() - t0
    t_cycle = dt / len(hist)
    ncyc = len(hist) if hist[-1] <= 1e-8 else CPU_CYCLES_TO_TOL
    import platform
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": None if ncyc is None else round(n / (t_cycle * ncyc) / 1e6, 6), "unit": "MDoF/s", "cores": 1,
            "kind": "port", "s_per_cycle": round(t_cycle, 3), "cycles_timed": len(hist),
            "rel_resid_after": [float(h) for h in hist], "cycles_to_1e-8": ncyc,
            "cycles_to_1e-8_source": "this run" if hist[-1] <= 1e-8 else "a full run of the same solve",
            "untimed_setup_s": round(t_setup, 1), "untimed_ilu_s": round(t_ilu, 1), "cpu": model,
            "sample": f"the metric's workload itself: heat hex27 64^3 top ({n} DoF), 5 levels, oracle V-cycle with "
                      f"GMRES(30)+ILU(0) 8/8 sweeps (x2 below top), 60-it GMRES coarse, 0.95 conditional correction; "
                      f"{len(hist)} timed cycle(s) of {t_cycle:.2f} s from x = 0, throughput = N / (s per cycle x "
                      f"cycles to 1e-8)"}


def _cpu_baseline_child(cycles):
    """The CPU leg in a child process started before the GPU is touched (one host core, no thread pools), so
    that it runs while the GPU legs are timed."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    return subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", "--cpu-cycles",
                             str(cycles)], stdout=subprocess.PIPE, stderr=None, text=True, env=env)


def build_ns(F, ctx, workload):
    """C3: channel [0,4]x[0,1]^2, 96x24x24 hex27 top (12x3x3 coarse, 4 levels), Re = 400, parabolic inflow,
    zero traction at x = 4; C1: lid-driven cavity 32^2 quad9 (8^2 coarse, 3 levels), Re = 100.  One Oseen
    (Picard) system with x_old = the initial condition, assembled on every level."""
    import functools
    from exfemus_amd import configs
    if workload == "channel":
        m = F.Mesh(3, [12, 3, 3], [0, 0, 0], [4, 1, 1], 4)
        s = F.MGSolNS(ctx, m, IRe=1.0 / 400, dt=0.1, time_coeff=0.0, upwind=0.1)
        s.GenBc(configs.channel_bc)
        s.init()
        s.GenIc(configs.channel_ic)
    else:
        m = F.Mesh(2, [8, 8], [0, 0], [1, 1], 3)
        s = F.MGSolNS(ctx, m, IRe=0.01, dt=0.1, time_coeff=0.0, upwind=0.1)
        s.GenBc(functools.partial(configs.cavity_bc, dim=2))
        s.init()
        s.GenIc(functools.partial(configs.cavity_ic, dim=2))
    top = m.nlev - 1
    t0 = time.perf_counter()
    for L in range(m.nlev):
        s.GenMatRhs(0.0, L, 1 if L == top else 0)
    ctx.sync()
    return m, s, time.perf_counter() - t0


def adjoint_workload(F, ctx, args):
    """C4: 2D lid-driven cavity 128^2 quad9 Q2-Q1 (16^2 coarse, 4 levels), Re 400: per-element viscosity
    log-normal around 0.0025 (sigma 1, seeded) and 16 Fourier coefficients of the lid profile
    u(x) = sum_k a_k sin(k pi x) (a_1 = 1 + noise, a_k ~ N(0,1)/k); J = viscous dissipation.
    Upwinding 0.1 (the reference's UP_WIND_NS).  --adjoint-problem steady (default): J of the steady state
    R(U) = A(U) U - b = 0, solved by Picard then Newton; the adjoint uses the transposed Newton Jacobian (upwind
    coefficients differentiated).  oseen: J after one Oseen step with the
    advection frozen at the lid profile.  Times the forward solve, the transposed multigrid solve and the
    gradient assembly; checks the most sensitive components against central differences (h = 1e-6)."""
    import functools
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from exfemus_amd import adjoint, configs
    steady = args.adjoint_problem == "steady"
    m = F.Mesh(2, [16, 16], [0, 0], [1, 1], 4)
    ns = F.MGSolNS(ctx, m, IRe=0.0025, dt=0.1, time_coeff=0.0, upwind=0.1)
    ns.GenBc(functools.partial(configs.cavity_bc, dim=2))
    ns.init()
    K = 16
    G = adjoint.lid_fourier_basis(ns, K)
    rng = np.random.default_rng(2026)
    a0 = rng.standard_normal(K) / np.arange(1, K + 1)
    if steady:
        a0[0] += 1.0
    top = 3
    ne = m.level(top)[0]
    nu = 0.0025 * configs.lognormal_k(ne, seed=7, sigma=1.0)
    o = ns.opts(precision=F.FM_PREC_FP64_FGMRES, smoother=F.FM_SMOOTH_VANKA, vanka_colored=1, nu_pre=4, nu_post=4,
                tol=1e-12, max_iter=1000)
    if steady:
        prob = adjoint.SteadyNSProblem(ns, G, functional="dissipation", opts=o)
    else:
        ns.set_xold(a0 @ G)
        prob = adjoint.OseenProblem(ns, G, functional="dissipation", opts=o)
    t0 = time.perf_counter()
    J = prob.solve(nu, a0)
    t_first = time.perf_counter() - t0
    U0 = ns.get_x().copy()
    stats = dict(getattr(prob, "stats", {}))
    log(f"forward {t_first * 1e3:.1f} ms, J = {J:.10e}, {stats}")
    t_fwd, t_adj = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        if steady:
            prob.solve(nu, a0)                      # Picard + Newton from zero
        else:
            ns.set_x(np.zeros(ns._Dim[top]))
            r = ns.MGSolve(opts=o)
        ctx.sync()
        t_fwd.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        g_nu, g_a = prob.gradient()                 # transposed assembly + transposed MG + gradient assembly
        ctx.sync()
        t_adj.append(time.perf_counter() - t0)
    adj_its = getattr(prob, "stats", {}).get("adjoint_mg_its")

    if steady:
        def J_of(nu_t, a):
            return prob.solve(nu_t, a, U0=U0)
    else:
        def J_of(nu_t, a):
            ns.set_elem_coef(top, nu_t)
            ns.set_bcdata(a @ G)
            ns.GenMatRhs(0.0, top, 1)
            rp, ci, v = ns.csr(0, top)
            n = len(rp) - 1
            Uk = spla.spsolve(sp.csr_matrix((v, ci, rp), shape=(n, n)).tocsc(), ns.get_rhs())
            return ns.dissipation(Uk)[0]

    h = 1e-6
    errs_nu, errs_a = [], []
    for e in np.argsort(-np.abs(g_nu * nu))[:4]:
        up = nu.copy(); up[e] *= 1 + h
        um = nu.copy(); um[e] *= 1 - h
        fd = (J_of(up, a0) - J_of(um, a0)) / (2 * h * nu[e])
        errs_nu.append(abs(fd - g_nu[e]) / abs(g_nu[e]))
    for k in np.argsort(-np.abs(g_a))[:4]:
        ap = a0.copy(); ap[k] += h
        am = a0.copy(); am[k] -= h
        fd = (J_of(nu, ap) - J_of(nu, am)) / (2 * h)
        errs_a.append(abs(fd - g_a[k]) / abs(g_a[k]))
    prob.solve(nu, a0, U0=U0) if steady else None
    t_total = float(np.median(t_fwd) + np.median(t_adj))
    N = ns._Dim[top]
    what = ("the steady state (upwind 0.1, Picard + Newton; adjoint with the transposed Newton Jacobian)" if steady
            else "one Oseen step (advection frozen at the lid profile)")
    out = {"metric": "adjoint gradient of the dissipation w.r.t. 16384 element viscosities + 16 lid coefficients",
           "value": round(t_total * 1e3, 3), "unit": "ms per gradient (forward solve + adjoint solve + gradient)",
           "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": round(t_total * 1e3, 3),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded log-normal viscosity, seeded lid coefficients)",
           "config": {"workload": "2D NS cavity quad9 Q2-Q1 128^2, 4 levels, Re 400 (nu 0.0025 log-normal, sigma 1), "
                                  f"16 lid Fourier coefficients, J = viscous dissipation of {what}",
                      "n_dof_top": N, "elements_top": ne,
                      "smoother": "coloured element Vanka 4/4, FP64 FGMRES to 1e-12"},
           "J": J, "forward_ms": round(float(np.median(t_fwd)) * 1e3, 3),
           "adjoint_and_gradient_ms": round(float(np.median(t_adj)) * 1e3, 3),
           "forward_stats": stats, "adjoint_mg_its": adj_its,
           "fd_rel_err_nu_max": float(max(errs_nu)), "fd_rel_err_lid_max": float(max(errs_a)), "fd_h": h}
    return out


def dhc_workload(F, ctx, args):
    """Coupled heat + Navier-Stokes (Boussinesq) in the differentially heated cube, Ra 1e5, Pr 0.71, hex27
    Q2-Q1 / Q2 on (6 * 2^(nlev-1))^3 (default 48^3, 4 levels).  A "step" is one coupled pseudo-time step
    (dt = 1 in convective units): Navier-Stokes assembled on every level with the buoyancy of the current
    temperature and solved, then the heat equation assembled with the new velocity and solved, each to a
    relative residual of 1e-8.  After spin-up steps the state is saved; K steps run with the mixed-precision
    multigrid (FP32 V-cycle in FP64 FGMRES), the state is restored and the same K steps run all-FP64; the
    final fields are compared per field.  Upwinding as in the reference's nse application (0.1 NS, 1 heat).
    Then the run continues to the steady state and the Nusselt number's gradient with respect to the 110,592
    (at 48^3) per-element conductivities is computed twice: through the adjoint of the whole coupled system
    (flexible GMRES preconditioned by the two transposed multigrid solves; the upwind coefficients are
    differentiated) and, for comparison, with the velocity frozen (steady heat adjoint only).  --dhc-nlev 5 is
    the 96^3 configuration (the coupled adjoint then runs memory-lean: forward NS multigrid data released,
    transposed Vanka inverses in FP32 only); --dhc-steady-steps 0 stops after the timed steps."""
    from exfemus_amd import coupled
    nx0, nlev = args.dhc_nx0, args.dhc_nlev
    t0 = time.perf_counter()
    b = coupled.Boussinesq(ctx, 3, nx0, nlev, 1e5, Pr=0.71, dt=1.0, upwind_ns=0.1, upwind_T=1.0)
    ns, T = b.ns, b.T
    top = nlev - 1
    Nns, NT = ns._Dim[top], T._Dim[top]
    nu = args.nu or 3
    kw_ns = dict(smoother=F.FM_SMOOTH_VANKA, vanka_colored=args.vanka_colored, nu_pre=nu, nu_post=nu, coarse_mult=1, tol=args.tol,
                 max_iter=200, use_graph=0 if args.no_graph else 1, omega=args.omega)
    kw_T = dict(nu_pre=3, nu_post=3, cheb_lo=0.1, tol=args.tol, max_iter=200, use_graph=0 if args.no_graph else 1)
    mixed = (ns.opts(precision=F.FM_PREC_FP32_FGMRES, **kw_ns), T.opts(precision=F.FM_PREC_FP32_FGMRES, **kw_T))
    fp64 = (ns.opts(precision=F.FM_PREC_FP64_FGMRES, **kw_ns), T.opts(precision=F.FM_PREC_FP64_FGMRES, **kw_T))
    b.o_ns, b.o_T = mixed
    mem_used = device_memory_used_gb()
    log(f"dhc setup {time.perf_counter() - t0:.1f}s  N_ns={Nns} N_T={NT}  device memory in use {mem_used} GB")
    k = 0
    for _ in range(args.dhc_spinup):
        k += 1
        du, dT = b.step(float(k))
        log(f"spin-up step {k}: du {du:.2e} dT {dT:.2e} its {b.last_reps[0].iters}/{b.last_reps[1].iters}")
    st0 = b.state()

    def timed(opts):
        b.set_state(st0)
        b.o_ns, b.o_T = opts
        ctx.sync()
        wall, sns, sT, its = [], [], [], []
        for j in range(args.steps):
            t1 = time.perf_counter()
            b.step(float(k + 1 + j))
            ctx.sync()
            wall.append(time.perf_counter() - t1)
            sns.append(b.last_reps[0].seconds)
            sT.append(b.last_reps[1].seconds)
            its.append((b.last_reps[0].iters, b.last_reps[1].iters))
        return wall, sns, sT, its, b.state()

    w64, s64n, s64T, its64, st64 = timed(fp64)
    wm, smn, smT, itsm, stm = timed(mixed)
    N = Nns + NT
    var = field_ids(ns)
    errs = {nm: field_err(stm[0][var == iv], st64[0][var == iv], nm == "p")   # enclosed box: p up to a constant
            for iv, nm in enumerate(["u", "v", "w", "p"])}
    errs["T"] = float(np.linalg.norm(stm[1] - st64[1]) / np.linalg.norm(st64[1]))
    t_step = float(np.median(wm))
    value = N / t_step / 1e6
    log(f"mixed step {t_step * 1e3:.1f} ms (NS MG {np.median(smn) * 1e3:.1f}, T MG {np.median(smT) * 1e3:.1f}) its {itsm}; "
        f"fp64 step {np.median(w64) * 1e3:.1f} ms (NS MG {np.median(s64n) * 1e3:.1f}, T MG {np.median(s64T) * 1e3:.1f}) "
        f"its {its64}; err {errs}")
    adj = {}
    nsteady = None
    if args.dhc_steady_steps > 0:   # 0: the timed steps only (no steady state, no adjoint)
        t1 = time.perf_counter()
        nsteady = b.run(max_steps=args.dhc_steady_steps, tol=1e-8,
                        callback=lambda j, du, dT: j % 5 == 4 and log(f"steady step {j + 1}: du {du:.2e} dT {dT:.2e}"))
        t_steady = time.perf_counter() - t1
        log(f"steady after {nsteady} more steps ({t_steady:.1f} s)")
        t1 = time.perf_counter()
        Nu, grad, cinfo = b.nusselt_gradient_coupled(
            tol=1e-8, callback=lambda it, r: log(f"coupled adjoint FGMRES {it}: {r:.2e}"), lean=True)
        ctx.sync()
        t_grad = time.perf_counter() - t1
        t1 = time.perf_counter()
        Nu_f, grad_f, _ = b.nusselt_gradient_k()
        ctx.sync()
        t_grad_f = time.perf_counter() - t1
        rel = float(np.linalg.norm(grad_f - grad) / np.linalg.norm(grad))
        log(f"Nu {Nu:.5f}; coupled adjoint {t_grad * 1e3:.1f} ms ({cinfo['fgmres_its']} its); frozen-velocity adjoint "
            f"{t_grad_f * 1e3:.1f} ms; |g_frozen - g| / |g| = {rel:.3f}")
        adj = {"nusselt": Nu, "nusselt_grad_coupled_ms": round(t_grad * 1e3, 2),
               "nusselt_grad_coupled_fgmres_its": cinfo["fgmres_its"],
               "nusselt_grad_coupled_inner_its": [list(x) for x in cinfo["inner_its"]],
               "nusselt_grad_norm": float(np.linalg.norm(grad)),
               "nusselt_grad_frozen_velocity_ms": round(t_grad_f * 1e3, 2), "frozen_vs_coupled_rel_diff": rel}
    nnz = int(F.lib().fm_system_csr_nnz(ns.h, 0, top))
    sec32 = ns.bench_spmv(top, 1, args.spmv_reps)
    # the bytes the SELL format moves (as for the heat workload), the CSR-equivalent count beside them
    b32 = sell_bytes(ns.sell_info(0, top), Nns, 4)
    b32_csr = spmv_bytes(Nns, nnz, 4)
    gbs32 = b32 / sec32 / 1e9
    n = nx0 * 2 ** (nlev - 1)
    out = {"metric": "MDoF/s to 1e-8 resid: FP32-MG+FP64 refine vs FP64 MG, 1 GPU; SpMV %HBM peak",
           "value": round(value, 3), "unit": "MDoF/s (NS + T DoF per coupled step)", "n_gpus": 1,
           "steps": args.steps, "warmup": args.dhc_spinup, "ms_per_step": round(t_step * 1e3, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32 V-cycle + fp64 residual",
           "data": "synthetic (analytic boundary data, conduction profile start)",
           "config": {"workload": f"3D Boussinesq differentially heated cube hex27 Q2-Q1/Q2 {n}^3, {nlev} levels, "
                                  f"Ra 1e5, Pr 0.71, upwind 0.1 / 1, pseudo-time dt 1, each solve to {args.tol:g}",
                      "n_dof_top": N, "n_dof_ns": Nns, "n_dof_T": NT, "nnz_ns_top": nnz,
                      "elements_top": b.mesh.level(top)[0],
                      "smoother": "NS: coloured element Vanka 3/3; T: Chebyshev-Jacobi 3/3", "parallelism": "single GPU"},
           "fp64_mg": {"value": round(N / float(np.median(w64)) / 1e6, 3), "unit": "MDoF/s",
                       "ms_per_step": round(float(np.median(w64)) * 1e3, 3),
                       "ns_solve_ms": round(float(np.median(s64n)) * 1e3, 3), "T_solve_ms": round(float(np.median(s64T)) * 1e3, 3),
                       "iters": its64},
           "mixed": {"ns_solve_ms": round(float(np.median(smn)) * 1e3, 3), "T_solve_ms": round(float(np.median(smT)) * 1e3, 3),
                     "iters": itsm},
           "speedup_vs_fp64": round(float(np.median(w64)) / t_step, 3),
           "ns_solve_speedup_vs_fp64": round(float(np.median(s64n)) / float(np.median(smn)), 3),
           "rel_l2_err_vs_fp64": errs,
           "steady_after_steps": nsteady, **adj,
           "device_memory_used_GB": mem_used,
           "roofline": {"bound": "hbm", "achieved": round(gbs32, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs32 / HBM_PEAK_GBS, 4), "traffic": None, "bytes_per_launch": b32,
                        "us_per_launch": round(sec32 * 1e6, 2),
                        "csr_equiv_frac": round(b32_csr / sec32 / 1e9 / HBM_PEAK_GBS, 4),
                        "kernel": f"k_sell<float, 0, 8> y = A x ({F.SPMV_FORMAT}) on the top Navier-Stokes matrix"}}
    ns.close()
    T.close()
    return out


def device_memory_used_gb():
    """Device memory in use on the current device (hipMemGetInfo), in GB."""
    import ctypes as C
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        return None
    free, total = C.c_size_t(), C.c_size_t()
    if hip.hipMemGetInfo(C.byref(free), C.byref(total)) != 0:
        return None
    return round((total.value - free.value) / 1e9, 1)


def field_err(a, ref, mean_free=False):
    """Relative L2 difference of a field from its reference (mean-free for a pressure defined up to a constant)."""
    if mean_free:
        a, ref = a - a.mean(), ref - ref.mean()
    return float(np.linalg.norm(a - ref) / max(np.linalg.norm(ref), 1e-300))


def field_ids(s):
    """Variable index of every top-level DoF (0..nq2-1 velocity components / T, nq2 = pressure)."""
    top = s._NoLevels - 1
    nd = s.node_dof(top)
    N = s.mesh.N
    var = np.full(s._Dim[top], -1)
    for iv in range(s.nq2 + s.nq1):
        d = nd[iv * N:(iv + 1) * N]
        var[d[d >= 0]] = iv
    return var


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="heat64", choices=["heat64", "channel", "cavity", "cavity-adjoint", "dhc"])
    ap.add_argument("--nx0", type=int, default=4)
    ap.add_argument("--nlev", type=int, default=5)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--outer", default="fgmres", choices=["ir", "fgmres"])
    ap.add_argument("--nu", type=int, default=None)
    ap.add_argument("--omega", type=float, default=None, help="Vanka damping (default 0.7 channel, 0.5 otherwise)")
    ap.add_argument("--gamma", type=int, default=1, help="cycle index below the top (1 V-cycle, 2 W-cycle)")
    ap.add_argument("--vanka-colored", type=int, default=1, help="1: multiplicative over element colours, 0: additive")
    ap.add_argument("--cheb-lo", type=float, default=0.1)
    ap.add_argument("--coarse-mult", type=int, default=None,
                    help="sweep multiplier below the top level (0: the reference's doubling; default 1 for heat64)")
    ap.add_argument("--smoother", type=int, default=0, help="heat64 smoother: 0 Chebyshev (first kind), 3 fourth kind")
    ap.add_argument("--spmv-reps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--adjoint-problem", default="steady", choices=["steady", "oseen"])
    ap.add_argument("--dhc-nx0", type=int, default=6)
    ap.add_argument("--dhc-nlev", type=int, default=4)
    ap.add_argument("--dhc-spinup", type=int, default=8)
    ap.add_argument("--dhc-steady-steps", type=int, default=200)
    ap.add_argument("--no-graph", action="store_true", help="launch V-cycles eagerly instead of replaying HIP graphs")
    ap.add_argument("--cpu-baseline-only", action="store_true", help="only the CPU leg (run as a child process)")
    ap.add_argument("--cpu-cycles", type=int, default=1, help="V-cycles the CPU leg times")
    ap.add_argument("--assembly-only", type=int, default=0, metavar="REPS",
                    help="only the top-level GenMatRhs, timed REPS times (heat64 or channel; for the PMC runs)")
    ap.add_argument("--roofline-only", action="store_true",
                    help="heat64: only the top-level product timings and their layout keys (for the PMC runs)")
    args = ap.parse_args()
    if args.cpu_baseline_only:
        print(json.dumps(cpu_baseline(args.cpu_cycles)))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "exfemus_amd runs on exactly one GPU: multi-GPU sharding (subdomain halos "
                                   "over RCCL) is not implemented"}))
        sys.exit(2)
    cpu_child = None
    if (not args.no_cpu and args.workload == "heat64" and not args.roofline_only and not args.assembly_only
            and args.nx0 == 4 and args.nlev == 5):
        cpu_child = _cpu_baseline_child(args.cpu_cycles)
    from exfemus_amd import femus as F
    ctx = F.Context(0)
    if args.workload in ("cavity-adjoint", "dhc"):
        out = adjoint_workload(F, ctx, args) if args.workload == "cavity-adjoint" else dhc_workload(F, ctx, args)
        for sysm in list(ctx._systems):
            sysm.close()
        ctx.close()
        print(json.dumps(out))
        return
    t0 = time.perf_counter()
    heat = args.workload == "heat64"
    if heat:
        nu = args.nu or 4
        m, s, t_asm = build_heat(F, ctx, [args.nx0] * 3, args.nlev)
        # the same sweep count on every level (the reference's MGStep doubles them below the top; here the
        # doubled coarse sweeps cost more than they save: 35.5 vs 29.0 ms, 11 iterations either way)
        cm = 1 if args.coarse_mult is None else args.coarse_mult
        kw = dict(nu_pre=nu, nu_post=nu, cheb_lo=args.cheb_lo, coarse_mult=cm, smoother=args.smoother, gamma=args.gamma)
        smoother = (f"Chebyshev-Jacobi {nu}/{nu}" + (" (4th kind)" if args.smoother == F.FM_SMOOTH_CHEBYSHEV4 else "") +
                    (", same sweeps on every level" if cm == 1 else ", sweeps doubled below the top" if cm in (0, 2) else ""))
        workload = (f"3D heat hex27 {args.nx0 * 16}^3, {args.nlev} levels, lognormal k, "
                    f"FP32-MG+FP64-{args.outer.upper()} to {args.tol:g}")
        data = "synthetic (seeded log-normal k, f = 1; coarse-level k = geometric mean of the children)"
    else:
        # coloured Vanka with the same sweep count on every level (the reference's MGStep doubles the sweeps
        # below the top; here that costs more than the iterations it saves: 409 vs 333 ms at C3).  Channel: 2/2
        # sweeps damped by 0.7 (23 iterations, 233 ms; 3/3 by 0.5: 18 iterations, 262 ms; 2/2 by 0.5: 36; one
        # sweep does not converge); cavity: 3/3 by 0.5
        chan = args.workload == "channel"
        nu = args.nu or (2 if chan else 3)
        omega = args.omega if args.omega is not None else (0.7 if chan else 0.5)
        m, s, t_asm = build_ns(F, ctx, args.workload)
        kw = dict(nu_pre=nu, nu_post=nu, smoother=F.FM_SMOOTH_VANKA, vanka_colored=args.vanka_colored,
                  coarse_mult=1 if args.coarse_mult is None else args.coarse_mult, omega=omega, gamma=args.gamma)
        smoother = f"coloured element Vanka {nu}/{nu} (damping {omega:g}), same sweeps on every level"
        workload = ("3D NS channel hex27 Q2-Q1 96x24x24, 4 levels, Re 400, Oseen step from the inflow profile"
                    if args.workload == "channel" else
                    "2D NS lid-driven cavity quad9 Q2-Q1 32^2, 3 levels, Re 100, Oseen step from the lid data")
        workload += f", FP32-MG+FP64-{args.outer.upper()} to {args.tol:g}"
        data = "synthetic (analytic boundary data)"
    top = m.nlev - 1
    N = s._Dim[top]
    nel_top = m.level(top)[0]
    log(f"setup {time.perf_counter() - t0:.1f}s  N_top={N}  assembly(all levels)={t_asm * 1e3:.1f} ms")
    if args.assembly_only:
        asm = assembly_rate(F, ctx, s, heat, reps=args.assembly_only)
        # the 256 MiB copy (k_copy16): the yardstick that scales the PMC byte counters
        cp = {"GBs": round(2 * (1 << 28) / ctx.stream_copy(1 << 28, reps=3) / 1e9, 1), "bytes_each_way": 1 << 28}
        s.close()
        ctx.close()
        print(json.dumps({"workload": args.workload, "assembly_top": asm, "copy": cp}))
        return
    if args.roofline_only:
        info, rf = roofline_measure(s, top, N, args.spmv_reps)
        out = {k: dict(r, kernel=KERNELS[k], layout_key=layout_key(KERNELS[k], info)) for k, r in rf.items()}
        # the 1 GiB copy (k_copy16): the yardstick that scales the PMC byte counters
        out["copy"] = {"GBs": round(2 * (1 << 30) / ctx.stream_copy(1 << 30, reps=3) / 1e9, 1), "bytes_each_way": 1 << 30}
        s.close()
        ctx.close()
        print(json.dumps(out))
        return
    prec_mixed = F.FM_PREC_FP32_IR if args.outer == "ir" else F.FM_PREC_FP32_FGMRES
    prec_fp64 = F.FM_PREC_FP64 if args.outer == "ir" else F.FM_PREC_FP64_FGMRES
    kw["use_graph"] = 0 if args.no_graph else 1
    om = s.opts(precision=prec_mixed, tol=args.tol, max_iter=200, **kw)
    o64 = s.opts(precision=prec_fp64, tol=args.tol, max_iter=200, **kw)
    t0 = time.perf_counter()
    s.MGSetup(o64)
    ctx.sync()
    t_mgsetup = time.perf_counter() - t0
    asm = assembly_rate(F, ctx, s, heat)
    s.MGSetup(o64)
    log(f"top-level assembly {asm}")
    zero = np.zeros(N)

    def run(o):
        s.set_x(zero)
        rep = s.MGSolve(opts=o)
        return rep

    for _ in range(args.warmup):
        run(om); run(o64)
    # reference FP64 solution for the per-field error
    rep64 = run(o64)
    x64 = s.get_x()
    t64 = []
    for _ in range(args.steps):
        r = run(o64)
        t64.append(r.seconds)
    ctx.sync()
    tm = []
    reps = []
    t_wall0 = time.perf_counter()
    for _ in range(args.steps):
        r = run(om)
        tm.append(r.seconds)
        reps.append(r)
    ctx.sync()
    wall = time.perf_counter() - t_wall0
    xm = s.get_x()
    var = field_ids(s)
    names = ["T"] if heat else (["u", "v", "w"][:m.dim] + ["p"])
    # the cavity is enclosed: its pressure is defined up to a constant, compared mean-free
    errs = {nm: field_err(xm[var == iv], x64[var == iv], nm == "p" and args.workload == "cavity")
            for iv, nm in enumerate(names)}
    t_mixed = float(np.sum(tm))
    value = N * args.steps / t_mixed / 1e6
    mdofs_fp64 = N * args.steps / float(np.sum(t64)) / 1e6
    log(f"mixed: iters={reps[-1].iters} rel={reps[-1].rel_resid:.3e} t={np.median(tm) * 1e3:.2f} ms; "
        f"fp64: iters={rep64.iters} t={np.median(t64) * 1e3:.2f} ms; err={errs}")
    # PCIe-inclusive rate beside the metric (never the metric): around each mixed solve the rhs and the zero
    # start vector come in from host arrays and the solution goes back to one, timed on the host clock
    try:
        b_host = s.get_rhs()
        tp = []
        for _ in range(args.steps):
            ctx.sync()
            t0 = time.perf_counter()
            s.set_rhs(b_host)
            s.set_x(zero)
            s.MGSolve(opts=om)
            s.get_x()
            tp.append(time.perf_counter() - t0)
        pcie = {"value": round(N * args.steps / float(np.sum(tp)) / 1e6, 3), "unit": "MDoF/s",
                "ms_per_solve": round(float(np.median(tp)) * 1e3, 3), "host_bytes_per_solve": 3 * 8 * N,
                "what": "rhs and zero start vector copied from host arrays, solution copied back, host wall clock"}
        log(f"pcie-inclusive: {pcie['value']} MDoF/s ({pcie['ms_per_solve']} ms per solve)")
    except Exception as e:  # the metric above stands without it
        pcie = {"error": str(e)}
    # roofline of the top-level products: the FP32 Chebyshev step dominates the mixed solve (8 of the 9
    # top-level products of an iteration); y = A x in both precisions beside it
    info, rf = roofline_measure(s, top, N, args.spmv_reps)
    nnz = info["nnz"]
    for k, r in rf.items():
        log(f"{k}: {r['us']} us, {r['GBs']} GB/s format bytes ({r['frac']}), CSR-equivalent {r['csr_equiv_frac']}")
    # STREAM copy yardstick: 1 GiB read + 1 GiB written per launch
    copy_bytes = 1 << 30
    sec_copy = ctx.stream_copy(copy_bytes, reps=10)
    gbs_copy = 2 * copy_bytes / sec_copy / 1e9
    log(f"stream copy {gbs_copy:.0f} GB/s")
    dom = rf["cheb_fp32"]
    traffic = pmc_traffic(KERNELS["cheb_fp32"], info)
    out = {
        "metric": "MDoF/s to 1e-8 resid: FP32-MG+FP64 refine vs FP64 MG, 1 GPU; SpMV %HBM peak",
        "value": round(value, 3),
        "unit": "MDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 V-cycle + fp64 residual",
        "data": data,
        "config": {"workload": workload, "n_dof_top": N, "nnz_top": nnz, "elements_top": nel_top,
                   "outer": args.outer, "smoother": smoother, "parallelism": "single GPU"},
        "fp64_mg": {"value": round(mdofs_fp64, 3), "unit": "MDoF/s", "ms_per_solve": round(float(np.median(t64)) * 1e3, 3),
                    "iters": rep64.iters},
        "mixed_iters": reps[-1].iters,
        "speedup_vs_fp64": round(value / mdofs_fp64, 3),
        "rel_l2_err_vs_fp64": errs,
        "final_rel_resid": reps[-1].rel_resid,
        "pcie_inclusive": pcie,
        "assembly_ms_all_levels": round(t_asm * 1e3, 2),
        "assembly_top": asm,
        "mg_setup_ms": round(t_mgsetup * 1e3, 2),
        "spmv": {**{k: r for k, r in rf.items()}, "stream_copy_GBs": round(gbs_copy, 1), "storage": info,
                 "kernels": KERNELS},
        # achieved = the bytes the SELL format moves per launch (values, 8/16/32-bit offsets, bases, slice
        # headers, the seven vector streams of the step) / the launch time; csr_equiv_frac counts a CSR product
        # with int32 columns instead (what the metric's SpMV definition assumes)
        "roofline": {"bound": "hbm", "achieved": dom["GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["frac"], "traffic": traffic,
                     "traffic_frac": None if traffic is None else round(traffic / (dom["us"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                     "bytes_per_launch": dom["bytes"], "us_per_launch": dom["us"],
                     "csr_equiv_frac": dom["csr_equiv_frac"],
                     "kernel": f"{KERNELS['cheb_fp32']}: Chebyshev step r -= A d, d' = c1 d + c2 D^-1 r, x += d' "
                               f"({F.SPMV_FORMAT}) on the top matrix"},
    }
    if cpu_child is not None:
        log("waiting for the CPU leg")
        try:
            so, _ = cpu_child.communicate(timeout=900)
            lines = [l for l in so.splitlines() if l.startswith("{")]
            out["cpu_baseline"] = json.loads(lines[-1]) if lines else {"error": f"CPU leg exited {cpu_child.returncode}"}
        except Exception as e:  # the oracle may be unbuilt on a bare box
            cpu_child.kill()
            out["cpu_baseline"] = {"error": str(e)}
    s.close()
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
