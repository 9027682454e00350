#!/usr/bin/env python3
"""Headline benchmark: solved DOF/s at 1e-10 relative residual, FP32-inner /
FP64-refined GMRES vs plain FP64 GMRES, on the steady convection-diffusion
workload (100^3 jittered lattice, 1,000,000 vertices, cellular flow at cell
Peclet 5, D = 1e-2, c = 1 at x = 0, zero diffusive flux elsewhere).

A step is one linear solve of the assembled system from x0 = 0 to
||b - A x||_2 / ||b||_2 <= 1e-10 (FP64 residual), preconditioner set-up
included, with the matrix, b and x resident in HBM.  Mesh, dual and assembly
are built before the timed region (assembly time is reported beside it).

Beside it, the adjoint line of the transient heat workload (126^3 jittered lattice,
2,000,376 vertices, lognormal conductivity from a seeded Gaussian random field with
correlation length 0.1, implicit Euler dt = 1e-3 for 200 steps, J = mean T on the x = 1
face): one forward pass with a checkpoint every 10 steps and one backward pass (segment
replays + transposed GMRES + assembly VJP) for dJ/dk at every vertex, FP32-inner /
FP64-refined solves, reported as the backward/forward time ratio.

And the flow line (lid-driven cavity at Re = 100 on an 80^3 jittered lattice, 2,048,000
DOF): the coupled velocity-pressure system after a few Picard iterations, solved FP64 and
FP32-inner / FP64-refined with GMRES(100) + 4x4 block Jacobi.

Prints one JSON line.  --gpus N with N > 1 exits: the solver runs on one GPU.
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vfvchem-vfvm3d_amd"))

HBM_PEAK_GBS = 8000.0


def progress(msg):
    """Stage line on stderr (stdout carries only the JSON line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def cellular_velocity(pos, U0):
    import numpy as np
    x, y = pos[:, 0], pos[:, 1]
    return np.stack([U0 * np.sin(np.pi * x) * np.cos(np.pi * y), -U0 * np.cos(np.pi * x) * np.sin(np.pi * y), 0.0 * x])


def build_problem(vfvm, n, D=1e-2):
    """Fields and operator list of the steady mass transport on a prescribed flow, in
    the operator order of steady_massflow.rs:455-545 (mass part)."""
    import numpy as np
    mesh = vfvm.Mesh.lattice(n, 0, 0.3, device=0)  # Voronoi dual built on the GPU
    ctx = vfvm.Context(0)
    ctx.upload(mesh)
    pos = mesh.get("pos")
    h = 1.0 / (n - 1)
    U0 = 5.0 * D / h  # cell Peclet number 5
    c = ctx.scldom_add(vfvm.FIELD_UNKNOWN, 0.0)
    Dd = ctx.scldom_add(vfvm.FIELD_CONSTANT, D)
    mp = ctx.scldom_add(vfvm.FIELD_CONSTANT, 0.0)
    ms = ctx.scldom_add(vfvm.FIELD_CONSTANT, 0.0)
    den = ctx.scldom_add(vfvm.FIELD_CONSTANT, 1.0)
    visc = ctx.scldom_add(vfvm.FIELD_CONSTANT, 1.0)
    pres = ctx.scldom_add(vfvm.FIELD_CONSTANT, 0.0)
    vel = ctx.vecdom_add(vfvm.FIELD_FUNCTION, values=cellular_velocity(pos, U0))
    fce = ctx.vecdom_add(vfvm.FIELD_CONSTANT, cnst3=np.zeros(3))
    f = ctx.sclflw_add(vel, pres, den, visc, fce)
    cin = ctx.sclbnd_add(0, 1.0)
    flux_b = [1, 2, 3, 4, 5]
    q = {b: ctx.sclbnd_add(b, 0.0) for b in flux_b}
    ops = [(vfvm.OP_SCL_DOM_DIFFUSION, 0, c, Dd, c), (vfvm.OP_SCL_DOM_PRODUCTION, 0, c, mp, -1),
           (vfvm.OP_SCL_DOM_SOURCE, 0, c, ms, -1), (vfvm.OP_SCL_DOM_ADVECTION_UNITY, 0, c, Dd, f),
           (vfvm.OP_SCL_BND_DIRICHLET, 0, c, cin, -1)]
    for b in flux_b:
        ops += [(vfvm.OP_SCL_BND_DIFFUSION, b, c, Dd, c), (vfvm.OP_SCL_BND_PRODUCTION, b, c, mp, -1),
                (vfvm.OP_SCL_BND_SOURCE, b, c, ms, -1), (vfvm.OP_SCL_BND_NEUMANN, b, c, q[b], -1)]
    for b in flux_b:
        ops.append((vfvm.OP_SCL_BND_ADVECTION_UNITY, b, c, Dd, f))
    return mesh, ctx, ops


def build_heat(vfvm, n, seed=0):
    """Transient heat on the jittered lattice (TransientHeat operator order,
    transient_heat.rs:152-171): rho c_p = 1, lognormal k, no source, T = 1 at x = 0, zero
    heat flux on the other faces, T0 = 0."""
    from vfvm.fields import lognormal_conductivity
    mesh = vfvm.Mesh.lattice(n, seed, 0.3, device=0)  # Voronoi dual built on the GPU
    k = lognormal_conductivity(mesh.get("pos"), ell=0.1, modes=1024, seed=seed)
    ctx = vfvm.Context(0)
    ctx.upload(mesh)
    T = ctx.scldom_add(vfvm.FIELD_UNKNOWN, 0.0)
    w = ctx.scldom_add(vfvm.FIELD_CONSTANT, 1.0)
    kk = ctx.scldom_add(vfvm.FIELD_FUNCTION, 0.0, k)
    hp = ctx.scldom_add(vfvm.FIELD_CONSTANT, 0.0)
    hs = ctx.scldom_add(vfvm.FIELD_CONSTANT, 0.0)
    tin = ctx.sclbnd_add(0, 1.0)
    ops = [(vfvm.OP_SCL_DOM_TIME, 0, T, w, -1), (vfvm.OP_SCL_DOM_DIFFUSION, 0, T, kk, T),
           (vfvm.OP_SCL_DOM_PRODUCTION, 0, T, hp, -1), (vfvm.OP_SCL_DOM_SOURCE, 0, T, hs, -1),
           (vfvm.OP_SCL_BND_DIRICHLET, 0, T, tin, -1)]
    for b in (1, 2, 3, 4, 5):
        q = ctx.sclbnd_add(b, 0.0)
        ops += [(vfvm.OP_SCL_BND_TIME, b, T, w, -1), (vfvm.OP_SCL_BND_DIFFUSION, b, T, kk, T),
                (vfvm.OP_SCL_BND_PRODUCTION, b, T, hp, -1), (vfvm.OP_SCL_BND_SOURCE, b, T, hs, -1),
                (vfvm.OP_SCL_BND_NEUMANN, b, T, q, -1)]
    return mesh, ctx, ops, T, kk


def run_adjoint(vfvm, args):
    """Forward + backward of the transient heat workload; the adjoint/forward time ratio."""
    import numpy as np
    t0 = time.time()
    mesh, ctx, ops, T, kk = build_heat(vfvm, args.adjoint_n)
    t_setup = time.time() - t0
    N = mesh.num_tot
    lin = dict(tol=args.tol, restart=args.restart, max_outer=2000, precision=vfvm.FP32_REFINED,
               precond=vfvm.PRECOND_ILU0, inner_tol=args.inner_tol, ortho=1)
    p = mesh.get("bnd_ptr")
    x1 = mesh.num_pts + mesh.get("bnd_nid")[p[1]:p[2]].astype(np.int64)
    seed = np.zeros(N)
    seed[x1] = 1.0 / x1.size
    fst = ctx.adjoint_forward(ops, dt=1e-3, num_ts=args.adjoint_steps, checkpoint_every=args.adjoint_every, lin=lin)
    J = float(seed @ ctx.scldom_get(T))
    (gk,), st = ctx.adjoint_backward(seed, [(vfvm.WRT_SCLDOM, kk)], lin=lin)
    ctx.close()
    return {
        "workload": f"transient heat, {args.adjoint_n}^3 jittered lattice ({N} vertices), lognormal k (RFF, l=0.1, "
                    f"1024 modes, seed 0), implicit Euler dt=1e-3 x {args.adjoint_steps} steps, J = mean T on x=1, "
                    f"dJ/dk per vertex, checkpoint every {args.adjoint_every}, GMRES({args.restart}) + ilu0 "
                    f"FP32-inner/FP64-refined, tol {args.tol}",
        "forward_s": fst.forward_seconds, "backward_s": st.backward_seconds,
        "adjoint_over_forward": st.backward_seconds / fst.forward_seconds,
        "recompute_s": st.recompute_seconds, "transposed_solve_s": st.adjoint_solve_seconds,
        "setup_s": st.setup_seconds, "forward_iters": fst.forward_iterations,
        "recompute_iters": st.recompute_iterations, "adjoint_iters": st.adjoint_iterations,
        "recomputed_steps": st.recomputed_steps, "checkpoints": st.checkpoints,
        "replay_mismatches": st.replay_mismatches, "J": J, "grad_abs_max": float(np.abs(gk).max()),
        "grad_sum": float(gk.sum()), "mesh_and_field_s": t_setup,
        "forward_dof_steps_per_s": N * args.adjoint_steps / fst.forward_seconds,
    }


def build_lid_cavity(vfvm, n, rho=1.0, mu=0.01, lid=1.0):
    """Steady lid-driven cavity (SteadyFlow composition of flow_lid.rs): no-slip walls, lid
    (lid, 0, 0) on z = 1, zero body force, reference pressure 0 at node 0."""
    import numpy as np
    mesh = vfvm.Mesh.lattice(n, 0, 0.3, device=0)  # Voronoi dual built on the GPU
    mesh.build_grad_coeff(device=0)
    ctx = vfvm.Context(0)
    ctx.upload(mesh)
    vel = ctx.vecdom_add(vfvm.FIELD_UNKNOWN, cnst3=np.zeros(3))
    pres = ctx.scldom_add(vfvm.FIELD_UNKNOWN, 0.0)
    den = ctx.scldom_add(vfvm.FIELD_CONSTANT, rho)
    visc = ctx.scldom_add(vfvm.FIELD_CONSTANT, mu)
    fce = ctx.vecdom_add(vfvm.FIELD_CONSTANT, cnst3=np.zeros(3))
    flw = ctx.sclflw_add(vel, pres, den, visc, fce)
    ops = [(vfvm.OP_VEC_DOM_ADVECTION, 0, vel, den, visc, flw), (vfvm.OP_VEC_DOM_GRADIENT, 0, vel, pres, -1),
           (vfvm.OP_VEC_DOM_DIFFUSION, 0, vel, visc, -1), (vfvm.OP_VEC_DOM_TRANSPOSE, 0, vel, visc, -1),
           (vfvm.OP_VEC_DOM_DILATATORY, 0, vel, visc, -1), (vfvm.OP_VEC_DOM_SOURCE, 0, vel, fce, -1),
           (vfvm.OP_SCL_DOM_DIVERGENCE, 0, pres, vel, den), (vfvm.OP_SCL_DOM_PRESSURE_STABILIZER, 0, pres, den, fce, flw)]
    for b in range(6):
        vb = ctx.vecbnd_add(b, cnst3=np.array([lid, 0.0, 0.0]) if b == 5 else np.zeros(3))
        ops += [(vfvm.OP_VEC_BND_DIRICHLET, b, vel, vb, -1), (vfvm.OP_SCL_BND_DIVERGENCE, b, pres, vel, den),
                (vfvm.OP_SCL_BND_PRESSURE_STABILIZER, b, pres, den, fce, flw)]
    ops.append((vfvm.OP_SCL_DOM_PRESSURE_REFERENCE, 0, pres, 0, -1, -1, 0.0))
    return mesh, ctx, ops, vel, pres


def timed_block_solves(vfvm, ctx, ndof, tol, restart, inner_tol, steps, label, max_outer=2000, events_first=False):
    """The assembled coupled system solved from x0 = 0 to tol with GMRES(restart) + block
    Jacobi, FP64 and FP32-inner/FP64-refined: one untimed solve, `steps` timed solves, then
    one solve with per-kernel HIP events (per-class averages).  events_first: the solve with
    per-kernel events is the untimed first solve (one solve fewer per variant)."""
    import torch
    x = torch.zeros(ndof, dtype=torch.float64, device="cuda")
    xp = x.data_ptr()
    out = {}
    sols = {}

    def kernel_pass(kw):
        ctx.timing_enable(-1)
        ctx.timing_reset()
        x.zero_()
        torch.cuda.synchronize()
        ctx.solve_device(xp, **kw)
        torch.cuda.synchronize()
        ctx.timing_enable(0)
        kern = {}
        for name, k in (("spmv", vfvm.K_SPMV), ("mdot", vfvm.K_MDOT), ("update", vfvm.K_UPDATE), ("bjacobi", vfvm.K_PRECOND)):
            nl, ms, by = ctx.timing_get(k)
            if nl:
                kern[name] = dict(launches=nl, avg_us=1e3 * ms / nl, total_ms=ms, bytes_per_launch=by / nl,
                                  gbs=by / (ms * 1e-3) / 1e9)
        return kern

    for var in ("fp64", "mixed"):
        prec = vfvm.FP64 if var == "fp64" else vfvm.FP32_REFINED
        kw = dict(tol=tol, restart=restart, max_outer=max_outer, precision=prec, precond=vfvm.PRECOND_BJACOBI,
                  inner_tol=inner_tol, ortho=1)
        if events_first:
            kern = kernel_pass(kw)
        else:
            ctx.timing_enable(0)
            x.zero_()
            torch.cuda.synchronize()
            ctx.solve_device(xp, **kw)
        ts = []
        for _ in range(steps):
            x.zero_()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            st = ctx.solve_device(xp, **kw)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t1)
        if not events_first:
            kern = kernel_pass(kw)
        ms = 1e3 * sum(ts) / len(ts)
        progress(f"{label} {var}: {ms:.1f} ms per solve ({st.iterations} steps)")
        out[var] = dict(ms_per_step=ms, dof_per_s=ndof / (ms * 1e-3), iters=st.iterations, refinements=st.refinements,
                        rel_residual=st.rel_residual, setup_s=st.setup_seconds, kernels=kern)
        sols[var] = x.clone()
    # error of the refined mixed-precision result against the FP64 solve of the same system
    out["mixed"]["rel_l2_vs_fp64"] = float(torch.linalg.vector_norm(sols["mixed"] - sols["fp64"]).item() /
                                           torch.linalg.vector_norm(sols["fp64"]).item())
    return out


def build_heat_mass_flow(vfvm, n, Ra, Pr, Sc):
    """Natural convection in the unit cube (TransientHeatMassFlow composition,
    transient_heatmassflow.rs:248-600) in free-fall units (velocity sqrt(g beta dT L), time
    L / that velocity): rho = 1, mu = sqrt(Pr / Ra), rho c_p = 1, k = 1 / sqrt(Ra Pr), species
    diffusivity mu / Sc, body force (0, 0, T - 1/2) evaluated on the device from T before every
    assembly (Boussinesq, an affine closure of the temperature), T = C = 1 on x = 0 and 0 on
    x = 1, zero heat / species flux and no-slip on every wall, reference pressure 0 at node 0;
    fluid at rest, T = C = 0.  Returns mesh, ctx, ops and the field ids."""
    import numpy as np
    mesh = vfvm.Mesh.lattice(n, 0, 0.3, device=0)  # Voronoi dual built on the GPU
    mesh.build_grad_coeff(device=0)
    ctx = vfvm.Context(0)
    ctx.upload(mesh)
    N = mesh.num_tot
    vel = ctx.vecdom_add(vfvm.FIELD_UNKNOWN, cnst3=np.zeros(3))
    pres = ctx.scldom_add(vfvm.FIELD_UNKNOWN, 0.0)
    den = ctx.scldom_add(vfvm.FIELD_CONSTANT, 1.0)
    mu = (Pr / Ra) ** 0.5
    visc = ctx.scldom_add(vfvm.FIELD_CONSTANT, mu)
    fce = ctx.vecdom_add(vfvm.FIELD_FUNCTION, cnst3=np.zeros(3))
    flw = ctx.sclflw_add(vel, pres, den, visc, fce)
    T = ctx.scldom_add(vfvm.FIELD_UNKNOWN, 0.0)
    rcp = ctx.scldom_add(vfvm.FIELD_CONSTANT, 1.0)
    k = ctx.scldom_add(vfvm.FIELD_CONSTANT, (Ra * Pr) ** -0.5)
    q = ctx.scldom_add(vfvm.FIELD_CONSTANT, 0.0)
    C = ctx.scldom_add(vfvm.FIELD_UNKNOWN, 0.0)
    D = ctx.scldom_add(vfvm.FIELD_CONSTANT, mu / Sc)
    ops = [(vfvm.OP_VEC_DOM_TIME, 0, vel, den, -1),
           (vfvm.OP_VEC_DOM_ADVECTION, 0, vel, den, visc, flw), (vfvm.OP_VEC_DOM_GRADIENT, 0, vel, pres, -1),
           (vfvm.OP_VEC_DOM_DIFFUSION, 0, vel, visc, -1), (vfvm.OP_VEC_DOM_TRANSPOSE, 0, vel, visc, -1),
           (vfvm.OP_VEC_DOM_DILATATORY, 0, vel, visc, -1), (vfvm.OP_VEC_DOM_SOURCE, 0, vel, fce, -1),
           (vfvm.OP_SCL_DOM_DIVERGENCE, 0, pres, vel, den), (vfvm.OP_SCL_DOM_PRESSURE_STABILIZER, 0, pres, den, fce, flw)]
    for b in range(6):
        vb = ctx.vecbnd_add(b, cnst3=np.zeros(3))
        ops += [(vfvm.OP_VEC_BND_DIRICHLET, b, vel, vb, -1), (vfvm.OP_SCL_BND_DIVERGENCE, b, pres, vel, den),
                (vfvm.OP_SCL_BND_PRESSURE_STABILIZER, b, pres, den, fce, flw)]
    # heat, then species (the reference's unknown order [vel, pres, temp, conc])
    hot = {}
    for u, w, dif, src, adv in ((T, rcp, k, q, vfvm.OP_SCL_DOM_ADVECTION), (C, den, D, -1, vfvm.OP_SCL_DOM_ADVECTION_UNITY)):
        ops += [(vfvm.OP_SCL_DOM_TIME_PREVIOUS, 0, u, w, -1), (vfvm.OP_SCL_DOM_DIFFUSION, 0, u, dif, u)]
        if src >= 0:
            ops.append((vfvm.OP_SCL_DOM_SOURCE, 0, u, src, -1))
        ops.append((adv, 0, u, w, dif, flw) if u == T else (adv, 0, u, dif, flw))
        hot[u] = ctx.sclbnd_add(0, 1.0)
        ops += [(vfvm.OP_SCL_BND_DIRICHLET, 0, u, hot[u], -1),
                (vfvm.OP_SCL_BND_DIRICHLET, 1, u, ctx.sclbnd_add(1, 0.0), -1)]
        for b in range(2, 6):
            ops += [(vfvm.OP_SCL_BND_TIME_PREVIOUS, b, u, w, -1), (vfvm.OP_SCL_BND_DIFFUSION, b, u, dif, u),
                    (vfvm.OP_SCL_BND_NEUMANN, b, u, ctx.sclbnd_add(b, 0.0), -1)]
            ops.append((vfvm.OP_SCL_BND_ADVECTION, b, u, w, dif, flw) if u == T else
                       (vfvm.OP_SCL_BND_ADVECTION_UNITY, b, u, dif, flw))
    ops.append((vfvm.OP_SCL_DOM_PRESSURE_REFERENCE, 0, pres, 0, -1, -1, 0.0))
    zero = np.zeros(N)
    ctx.vecdom_func_affine(fce, T, np.stack([zero, zero, zero - 0.5]), np.stack([zero, zero, zero + 1.0]))
    ids = dict(vel=vel, pres=pres, T=T, C=C, fce=fce, hot_T=hot[T])
    return mesh, ctx, ops, ids


def hmf_rest(ctx, ids, N):
    """The heat-mass-flow fields back to the initial state (rest, T = C = 0)."""
    import numpy as np
    ctx.vecdom_set(ids["vel"], np.zeros((3, N)))
    for f in ("pres", "T", "C"):
        ctx.scldom_set(ids[f], np.zeros(N))


def run_heat_mass_flow(vfvm, args):
    """Config-5 system (natural convection with a passive species, six unknowns per vertex):
    implicit-Euler steps from rest with Picard per step, then the linear solve of the
    transient system at the last state timed FP64 vs FP32-inner/FP64-refined with GMRES(100)
    + 6x6 block Jacobi (solved DOF/s at 1e-10); then the adjoint of the run: J = time-averaged
    hot-wall Nusselt number over hmf_ts_total steps, dJ/dtheta of a 32 x 32 bilinear profile of
    the hot wall's temperature, with every-10-step checkpoints and segment replays, and a
    central-difference check of the gradient on a short horizon of the same mesh."""
    import torch
    t0 = time.time()
    mesh, ctx, ops, ids = build_heat_mass_flow(vfvm, args.hmf_n, 1e4, 0.71, 1.0)
    t_setup = time.time() - t0
    N = mesh.num_tot
    ndof = 6 * N
    dt = 1e-2
    lin = dict(tol=1e-8, restart=args.hmf_restart, max_outer=100, precision=vfvm.FP32_REFINED,
               precond=vfvm.PRECOND_BJACOBI, inner_tol=1e-4, ortho=1)
    tp = time.time()
    note = None
    st = None
    try:
        st = ctx.transient_solve(ops, dt, args.hmf_ts, 8, 1e-6, 1.0, lin=lin)
    except vfvm.VFVChemError as e:
        note = str(e)
    t_steps = time.time() - tp
    progress(f"heat-mass-flow {args.hmf_n}^3: {args.hmf_ts} time steps in {t_steps:.1f} s")
    _, nnz = ctx.system_info()
    ta = time.time()
    ctx.assemble(ops, dt)
    torch.cuda.synchronize()
    t_asm = time.time() - ta
    out = timed_block_solves(vfvm, ctx, ndof, args.tol, args.hmf_restart, args.hmf_inner_tol, args.hmf_steps,
                             "heat-mass-flow", max_outer=100, events_first=True)
    adj = None
    if st is not None and args.hmf_ts_total > 0:
        adj = hmf_adjoint(vfvm, args, mesh, ctx, ops, ids, lin, dt)
    ctx.close()
    return {
        "workload": f"natural convection Ra=1e4, Pr=0.71, Sc=1 (free-fall units) with a passive species (u, v, w, p, T, C), {args.hmf_n}^3 "
                    f"jittered lattice ({N} vertices, {ndof} DOF, {nnz} nnz), implicit Euler dt=1e-2, transient system "
                    f"after {args.hmf_ts} time steps from rest, GMRES({args.hmf_restart}) + 6x6 block Jacobi, tol {args.tol}",
        "dof": ndof, "nnz": nnz, "fp64": out["fp64"], "mixed": out["mixed"],
        "mixed_over_fp64": out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"],
        "assembly_wall_ms": 1e3 * t_asm, "time_steps_s": t_steps,
        "picard_iterations": st.picard_iterations if st is not None else None, "steps_note": note,
        "mesh_dual_gc_s": t_setup,
        "adjoint": adj,
    }


def hmf_adjoint(vfvm, args, mesh, ctx, ops, ids, lin, dt):
    """Forward pass with checkpoints, backward pass and the profile gradient of J = mean
    hot-wall Nusselt number; a central-difference check on args.hmf_fd_steps steps."""
    import numpy as np
    from vfvm.objectives import WallProfile, wall_flux_seed
    N = mesh.num_tot
    n = 6 * N
    M, K = args.hmf_ts_total, args.hmf_every
    # Nu of the x = 0 wall: sum over its nodes' faces of (A / d)(T_v - T_n), the conductive flux
    # over k dT / L (dT = L = wall area = 1); 1 for the conduction profile T = 1 - x
    seed = wall_flux_seed(mesh, 0, 4 * N, n)
    prof = WallProfile(mesh, 0, (32, 32))
    # the transposed solves precondition the outer GCR: one FP32 cycle of GMRES(30) each (the
    # outer iteration corrects what it leaves)
    inner = dict(tol=args.hmf_adj_inner_tol, restart=args.hmf_adj_inner_restart, max_outer=1, precision=vfvm.FP32_REFINED,
                 precond=vfvm.PRECOND_BJACOBI, inner_tol=1e-4, ortho=1)

    def forward(steps, theta, tol, ln):
        hmf_rest(ctx, ids, N)
        ctx.sclbnd_set(ids["hot_T"], prof.values(theta, base=1.0))
        return ctx.flow_adjoint_forward(ops, dt, steps, max_iter=20, tol=tol, damp=1.0, checkpoint_every=K, lin=ln,
                                        seed=seed, weights=np.full(steps, 1.0 / steps))

    def on_state(step, outer, inner_steps, rel):
        if step % 10 == 0 or step <= 2:
            progress(f"heat-mass-flow adjoint: state {step}, {outer} outer iterations, {inner_steps} inner steps, "
                     f"residual {rel:.2e}")

    def backward(steps, tol):
        g, bst = ctx.flow_adjoint_backward(seed, [(vfvm.WRT_SCLBND, ids["hot_T"])], weights=np.full(steps, 1.0 / steps),
                                           inner=inner, tol=tol, max_iter=200, restart=20, on_state=on_state)
        return prof.grad(g[0]), bst

    theta0 = np.zeros((32, 32))
    fst = forward(M, theta0, 1e-6, lin)
    progress(f"heat-mass-flow adjoint: forward {M} steps in {fst.base.forward_seconds:.1f} s, J = {fst.objective:.6f}")
    grad, bst = backward(M, args.hmf_adj_tol)
    progress(f"heat-mass-flow adjoint: backward in {bst.base.backward_seconds:.1f} s "
             f"(replay {bst.base.recompute_seconds:.1f} s, {bst.outer_iterations} outer iterations)")
    res = {
        "workload": f"J = time-averaged hot-wall Nusselt number over {M} implicit-Euler steps (dt 1e-2) from rest, "
                    f"dJ/dtheta of a 32 x 32 bilinear hot-wall temperature profile, checkpoint every {K}, forward "
                    f"Picard tol 1e-6 with GMRES({args.hmf_restart}) + 6x6 block Jacobi FP32-inner/FP64-refined at 1e-8; "
                    f"backward: GCR on J^T (tol {args.hmf_adj_tol}) preconditioned by one FP32 cycle of transposed "
                    f"GMRES({args.hmf_adj_inner_restart}) + 6x6 block Jacobi on A^T with the reference-pressure mode deflated",
        "J": fst.objective, "forward_s": fst.base.forward_seconds, "backward_s": bst.base.backward_seconds,
        "adjoint_over_forward": bst.base.backward_seconds / fst.base.forward_seconds,
        "recompute_s": bst.base.recompute_seconds, "transposed_solve_s": bst.base.adjoint_solve_seconds,
        "setup_s": fst.base.setup_seconds + bst.base.setup_seconds,
        "forward_picard_iterations": fst.picard_iterations, "forward_linear_iterations": fst.base.forward_iterations,
        "recompute_picard_iterations": bst.recompute_picard_iterations, "recomputed_steps": bst.base.recomputed_steps,
        "replay_mismatches": bst.base.replay_mismatches, "outer_iterations": bst.outer_iterations,
        "inner_arnoldi_steps": bst.base.adjoint_iterations, "max_outer_residual": bst.max_outer_residual,
        "checkpoints": fst.base.checkpoints, "grad_abs_max": float(np.abs(grad).max()),
        "grad_l2": float(np.linalg.norm(grad)), "grad_sum": float(grad.sum()),
    }
    if args.hmf_fd_steps > 0:
        # central differences of the device forward pass on a short horizon of the same mesh,
        # Picard and linear solves tightened (1e-10, 1e-11) so that J carries ~1e-11 of relative
        # error, against the adjoint gradient of the same horizon (outer tolerance 1e-12)
        S = args.hmf_fd_steps
        tight = dict(lin, tol=1e-11, max_outer=400)
        forward(S, theta0, 1e-10, tight)
        g_s, bs = backward(S, 1e-12)
        yy, zz = np.meshgrid(np.linspace(0, 1, 32), np.linspace(0, 1, 32), indexing="ij")
        delta = np.sin(np.pi * yy) * np.sin(2 * np.pi * zz) + 0.5 * np.cos(3 * np.pi * zz)
        eps = 1e-3
        jp = forward(S, eps * delta, 1e-10, tight).objective
        jm = forward(S, -eps * delta, 1e-10, tight).objective
        fd = (jp - jm) / (2 * eps)
        ad = float(np.sum(g_s * delta))
        res["fd_check"] = {"steps": S, "direction": "sin(pi y) sin(2 pi z) + 0.5 cos(3 pi z) on the 32 x 32 grid",
                           "eps": eps, "adjoint": ad, "central_difference": fd, "rel_err": abs(ad - fd) / abs(fd),
                           "max_outer_residual": bs.max_outer_residual}
        progress(f"heat-mass-flow adjoint: {S}-step gradient check, adjoint {ad:.8e}, fd {fd:.8e}")
    return res


def run_flow(vfvm, args):
    """Lid-driven cavity at Re = 100: Picard iterations of the coupled velocity-pressure system
    with GMRES(100) + 4x4 block Jacobi; the linear solve of the system after a few Picard
    iterations timed FP64 vs FP32-inner/FP64-refined (solved DOF/s at 1e-10)."""
    import torch
    t0 = time.time()
    mesh, ctx, ops, vel, pres = build_lid_cavity(vfvm, args.flow_n)
    t_setup = time.time() - t0
    N = mesh.num_tot
    ndof = 4 * N
    lin = dict(tol=args.tol, restart=args.flow_restart, max_outer=400, precision=vfvm.FP32_REFINED,
               precond=vfvm.PRECOND_BJACOBI, inner_tol=1e-4, ortho=1)
    # SteadyBase::solve (steady_base.rs:74-117) to steady state: damped Picard (0.8) until the
    # scaled residual is below 1e-6, each linearised system solved to the metric's tolerance
    tp = time.time()
    picard = None
    pst = None
    try:
        pst = ctx.steady_solve(ops, args.flow_picard, 1e-6, 0.8, lin=lin)
    except vfvm.VFVChemError as e:
        picard = str(e)
    t_picard = time.time() - tp
    progress(f"lid cavity: Picard {'converged' if pst is not None else 'stopped'} in {t_picard:.1f} s "
             f"({pst.picard_iterations if pst is not None else '?'} iterations)")
    _, nnz = ctx.system_info()
    ctx.timing_enable(-1)
    ctx.timing_reset()
    ta = time.time()
    ctx.assemble(ops)
    torch.cuda.synchronize()
    t_asm = time.time() - ta
    out = timed_block_solves(vfvm, ctx, ndof, args.tol, args.flow_restart, args.inner_tol, args.flow_steps, "lid cavity")
    ctx.close()
    return {
        "workload": f"steady lid-driven cavity Re=100 (rho=1, mu=0.01, lid (1,0,0)), {args.flow_n}^3 jittered lattice "
                    f"({N} vertices, {ndof} DOF, {nnz} nnz), Picard to steady state (damping 0.8, scaled residual 1e-6), "
                    f"then the converged system's linear solve timed, GMRES({args.flow_restart}) + 4x4 block Jacobi, tol {args.tol}",
        "dof": ndof, "nnz": nnz, "fp64": out["fp64"], "mixed": out["mixed"],
        "mixed_over_fp64": out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"],
        "assembly_wall_ms": 1e3 * t_asm, "mesh_dual_gc_s": t_setup,
        "steady_state": {"converged": pst is not None, "picard_iterations": pst.picard_iterations if pst else None,
                         "linear_iterations": pst.linear_iterations if pst else None,
                         "scaled_residual": pst.residual if pst else None, "time_to_steady_s": t_picard,
                         "assembly_s": pst.time_assembly if pst else None, "solve_s": pst.time_solve if pst else None,
                         "note": picard},
    }


def kernel_source_hash():
    """sha256 of the Krylov kernel source whose launches the traffic figure describes."""
    import hashlib
    with open(os.path.join(ROOT, "vfvchem-vfvm3d_amd", "csrc", "krylov.hip"), "rb") as fh:
        return hashlib.sha256(fh.read()).hexdigest()


def measured_traffic(kernel, variant):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload (scaled by a copy kernel of known size), or None; only a summary
    measured on the current kernel source (its recorded sha256 equal to krylov.hip's) counts.
    Returns (bytes or None, note)."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_convdiff_100cubed.md")
    if variant != "mixed" or not os.path.exists(path):
        return None, "no counter summary for this variant"
    value, src = None, None
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                value = float(parts[2])
            if len(parts) == 3 and parts[0] == "source_sha256" and parts[1] == "krylov.hip":
                src = parts[2]
    if value is None:
        return None, "the counter summary has no figure for this kernel"
    if src != kernel_source_hash():
        return None, "the counter summary was measured on another kernel source"
    return value, "rocprofv3 FETCH_SIZE + WRITE_SIZE of the current kernel source, " + os.path.relpath(path, ROOT)


def kernel_class(name):
    """Kernel class of a rocprofv3 kernel name (the classes bench.py times with events)."""
    if "k_ilu_tile" in name or "k_ilu_lower" in name or "k_ilu_upper" in name:
        return "ilu_sweep"
    if "k_spmv" in name:
        return "spmv"
    if "copyBuffer" in name or "elementwise" in name or "copy_kernel" in name:
        return "copy"
    return None


def pmc_summary(fetch_dir, write_dir, copy_mb):
    """Markdown summary of two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE, KiB per dispatch) of
    the same command: per kernel the average counters, scaled to bytes by what the same counters
    read for the plain device copy of copy_mb MiB in that command; the per-launch HBM traffic of
    the ILU(0) sweep and SpMV classes in the lines measured_traffic() reads, with the sha256 of the
    kernel source they were measured on."""
    import csv
    import glob

    def load(d, counter):
        per = {}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name") != counter:
                        continue
                    name = row["Kernel_Name"]
                    per.setdefault(name, []).append(float(row["Counter_Value"]))
        return per

    fetch, write = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    copy_bytes = copy_mb * float(1 << 20)
    cf = [max(v) for k, v in fetch.items() if kernel_class(k) == "copy"]
    cw = [max(v) for k, v in write.items() if kernel_class(k) == "copy"]
    f_scale = copy_bytes / (max(cf) * 1024.0) if cf else 1.0
    w_scale = copy_bytes / (max(cw) * 1024.0) if cw else 1.0
    out = [f"Calibration: the {copy_mb} MiB device copy reads {max(cf) if cf else float('nan'):.0f} KiB by FETCH_SIZE and "
           f"writes {max(cw) if cw else float('nan'):.0f} KiB by WRITE_SIZE: scale factors {f_scale:.3f} (fetch) and "
           f"{w_scale:.3f} (write) to bytes.", "",
           "| kernel | dispatches | FETCH_SIZE avg KiB | WRITE_SIZE avg KiB | scaled HBM bytes per launch |",
           "|---|---:|---:|---:|---:|"]
    cls_tot = {}
    for name in sorted(fetch, key=lambda k: -sum(fetch[k])):
        c = kernel_class(name)
        if c not in ("ilu_sweep", "spmv"):
            continue
        fv = fetch[name]
        wv = write.get(name, [0.0])
        fa, wa = sum(fv) / len(fv), sum(wv) / len(wv)
        b = fa * 1024.0 * f_scale + wa * 1024.0 * w_scale
        out.append(f"| `{name[:90]}` | {len(fv)} | {fa:.1f} | {wa:.1f} | {b:.4e} |")
        t = cls_tot.setdefault(c, [0, 0.0])
        t[0] += len(fv)
        t[1] += b * len(fv)
    out.append("")
    for c, (cnt, tot) in sorted(cls_tot.items()):
        out.append(f"traffic_bytes_per_launch {c} {tot / cnt:.6e}")
    out.append(f"source_sha256 krylov.hip {kernel_source_hash()}")
    return "\n".join(out)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100, help="lattice points per axis")
    ap.add_argument("--precond", default="ilu0", choices=["jacobi", "ilu0", "none"])
    ap.add_argument("--restart", type=int, default=50)
    ap.add_argument("--inner-tol", type=float, default=1e-6)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--cpu-sample-steps", type=int, default=50, help="Arnoldi steps timed on the CPU (0: skip)")
    ap.add_argument("--adjoint-n", type=int, default=126, help="lattice points per axis of the adjoint workload")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-every", type=int, default=10)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--flow-n", type=int, default=80, help="lattice points per axis of the lid-driven cavity")
    ap.add_argument("--flow-picard", type=int, default=100, help="Picard iterations allowed to reach the steady state")
    ap.add_argument("--flow-restart", type=int, default=100)
    ap.add_argument("--flow-steps", type=int, default=2)
    ap.add_argument("--no-flow", action="store_true")
    ap.add_argument("--hmf-n", type=int, default=100, help="lattice points per axis of the heat-mass-flow system")
    ap.add_argument("--hmf-ts", type=int, default=10, help="implicit-Euler steps before the timed heat-mass-flow solves")
    ap.add_argument("--hmf-ts-total", type=int, default=100, help="implicit-Euler steps of the heat-mass-flow adjoint run (objective J; 0: skip)")
    ap.add_argument("--hmf-steps", type=int, default=1)
    ap.add_argument("--hmf-restart", type=int, default=100, help="GMRES restart of the heat-mass-flow solves")
    ap.add_argument("--hmf-inner-tol", type=float, default=1e-6, help="inner tolerance of the mixed heat-mass-flow solve")
    ap.add_argument("--hmf-every", type=int, default=2, help="checkpoint interval of the heat-mass-flow adjoint "
                    "(51 states of 48 MB at 100 steps)")
    ap.add_argument("--hmf-adj-tol", type=float, default=1e-8, help="outer tolerance of the heat-mass-flow adjoint solves")
    ap.add_argument("--hmf-adj-inner-tol", type=float, default=1e-2, help="tolerance of the transposed solves (outer preconditioner)")
    ap.add_argument("--hmf-adj-inner-restart", type=int, default=30, help="GMRES restart of the transposed solves")
    ap.add_argument("--hmf-fd-steps", type=int, default=3, help="steps of the central-difference gradient check (0: skip)")
    ap.add_argument("--no-hmf", action="store_true")
    ap.add_argument("--hmf-only", action="store_true",
                    help="run only the heat-mass-flow line and print it (profiling aid; not the headline line)")
    ap.add_argument("--flow-only", action="store_true",
                    help="run only the lid-driven cavity line and print it (profiling aid; not the headline line)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="first copy MB megabytes device-to-device 5 times (a plain copy kernel of known size, "
                         "used to scale rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="print the traffic summary of two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of a command "
                         "with --copy-calibration MB, and exit (no GPU work)")
    ap.add_argument("--variants", default="fp64,mixed")
    ap.add_argument("--ortho", default="cgs", choices=["auto", "cgs", "cgs2"])
    args = ap.parse_args()
    if args.pmc_summary:
        print(pmc_summary(args.pmc_summary[0], args.pmc_summary[1], args.copy_calibration or 256))
        return
    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding is a design only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import vfvm

    if args.copy_calibration > 0:
        src = torch.ones(args.copy_calibration * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.flow_only:
        print(json.dumps({"flow": run_flow(vfvm, args)}))
        return
    if args.hmf_only:
        print(json.dumps({"heat_mass_flow": run_heat_mass_flow(vfvm, args)}))
        return
    pc = {"jacobi": vfvm.PRECOND_JACOBI, "ilu0": vfvm.PRECOND_ILU0, "none": vfvm.PRECOND_NONE}[args.precond]
    t0 = time.time()
    mesh, ctx, ops = build_problem(vfvm, args.n)
    t_mesh = time.time() - t0
    progress(f"convection-diffusion {args.n}^3: mesh and fields in {t_mesh:.1f} s")
    nnz = ctx.pattern_build(ops)
    ctx.timing_enable(-1)
    ctx.timing_reset()
    ta = time.time()
    ctx.assemble(ops)
    t_asm_wall = time.time() - ta
    _, asm_ms, asm_bytes = ctx.timing_get(vfvm.K_ASSEMBLE)
    N = mesh.num_tot
    x = torch.zeros(N, dtype=torch.float64, device="cuda")
    xp = x.data_ptr()

    results = {}
    solutions = {}
    variants = [v for v in args.variants.split(",") if v]
    for var in variants:
        prec = vfvm.FP64 if var == "fp64" else vfvm.FP32_REFINED
        kw = dict(tol=args.tol, restart=args.restart, max_outer=2000, precision=prec, precond=pc, inner_tol=args.inner_tol,
                  ortho={"auto": 0, "cgs": 1, "cgs2": 2}[args.ortho])
        ctx.timing_enable(0)
        for _ in range(args.warmup):
            x.zero_()
            torch.cuda.synchronize()
            ctx.solve_device(xp, **kw)
        # timed region: K solves, no per-kernel events (they perturb the launch stream)
        stats = []
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            x.zero_()
            torch.cuda.synchronize()
            st = ctx.solve_device(xp, **kw)
            stats.append((st.iterations, st.refinements, st.rel_residual, st.seconds, st.setup_seconds))
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        # error of the result against the exact solution c = 1 (the inlet value carried by a
        # divergence-free flow with zero diffusive flux on the other walls)
        err = float((x - 1.0).abs().max().item())
        err_l2 = float(torch.linalg.vector_norm(x - 1.0).item() / N ** 0.5)
        solutions[var] = x.clone()
        # kernel pass: one more solve with HIP start/stop events on every Krylov kernel launch
        ctx.timing_enable(-1)
        ctx.timing_reset()
        x.zero_()
        torch.cuda.synchronize()
        ctx.solve_device(xp, **kw)
        torch.cuda.synchronize()
        ctx.timing_enable(0)
        kern = {}
        for name, k in (("spmv", vfvm.K_SPMV), ("mdot", vfvm.K_MDOT), ("update", vfvm.K_UPDATE), ("ilu_sweep", vfvm.K_PRECOND)):
            nl, ms, by = ctx.timing_get(k)
            if nl:
                kern[name] = dict(launches=nl, avg_us=1e3 * ms / nl, total_ms=ms, bytes_per_launch=by / nl,
                                  gbs=by / (ms * 1e-3) / 1e9)
        progress(f"convection-diffusion {var}: {1e3 * (t2 - t1) / args.steps:.1f} ms per solve")
        results[var] = dict(ms_per_step=1e3 * (t2 - t1) / args.steps, iters=stats[-1][0], refinements=stats[-1][1],
                            rel_residual=stats[-1][2], device_s=stats[-1][3], setup_s=stats[-1][4], kernels=kern,
                            max_abs_err_vs_exact=err, rel_l2_err_vs_exact=err_l2)
    if "fp64" in solutions and "mixed" in solutions:
        a, b = solutions["fp64"], solutions["mixed"]
        results["mixed"]["rel_l2_vs_fp64"] = float(torch.linalg.vector_norm(b - a).item() / torch.linalg.vector_norm(a).item())
    solutions.clear()

    head = "mixed" if "mixed" in results else variants[0]
    r = results[head]
    value = N / (r["ms_per_step"] * 1e-3)
    # dominant kernel of the headline variant: largest total time
    dom_name, dom = max(r["kernels"].items(), key=lambda kv: kv[1]["total_ms"])
    achieved = dom["bytes_per_launch"] / (dom["avg_us"] * 1e-6) / 1e9

    cpu = None
    if args.cpu_sample_steps > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import vfvm_oracle as O
        vals, b = ctx.matrix_get()
        rp, col = ctx.pattern_get()
        tc = time.perf_counter()
        _, _, steps_done, _ = O.gmres(rp, col, vals, b, None, args.restart, 1, args.tol, args.cpu_sample_steps)
        tcpu = time.perf_counter() - tc
        per_step = tcpu / max(steps_done, 1)
        # the same sample with the vector loops on every host thread the job may use
        nth = int(os.environ.get("OMP_NUM_THREADS") or 0) or (os.cpu_count() or 1)
        tc = time.perf_counter()
        _, _, steps_mt, _ = O.gmres(rp, col, vals, b, None, args.restart, 1, args.tol, args.cpu_sample_steps, threads=nth)
        tcpu_mt = time.perf_counter() - tc
        per_step_mt = tcpu_mt / max(steps_mt, 1)
        # Arnoldi steps the reference algorithm (Jacobi-GMRES(m), FP64) needs on this system,
        # counted by one untimed FP64 Jacobi solve on the GPU
        x.zero_()
        torch.cuda.synchronize()
        need = ctx.solve_device(xp, tol=args.tol, restart=args.restart, max_outer=5000, precision=vfvm.FP64,
                                precond=vfvm.PRECOND_JACOBI, ortho=1).iterations
        cpu = {"value": N / (per_step_mt * need), "unit": "DOF/s", "cores": nth, "kind": "port",
               "sample": f"{steps_mt} Arnoldi steps of Jacobi-GMRES({args.restart}) FP64 (oracle/gmres_ref.c, OpenMP vector loops on "
                         f"{nth} threads) on the same {N}-DOF system ({tcpu_mt:.2f} s), extrapolated to the {need} steps the "
                         f"FP64 solve needs",
               "single_thread": {"value": N / (per_step * need), "cores": 1,
                                 "sample": f"{steps_done} steps on 1 thread ({tcpu:.1f} s), as every reference example runs "
                                           f"(LinearGmres::new(1))"},
               "cpu": cpu_model(), "nproc": os.cpu_count()}

    ctx.close()
    def side_line(name, skip, fn):
        # a failing side workload is reported in its own field; the headline line still prints
        if skip:
            return None
        progress(f"{name} workload")
        try:
            return fn(vfvm, args)
        except Exception as e:  # noqa: BLE001 - the error text is the report
            progress(f"{name} workload failed: {e}")
            return {"error": f"{type(e).__name__}: {e}"}

    adjoint = side_line("adjoint", args.no_adjoint, run_adjoint)
    flow = side_line("lid-driven cavity", args.no_flow, run_flow)
    hmf = side_line("heat-mass-flow", args.no_hmf, run_heat_mass_flow)

    traffic = measured_traffic(dom_name, head)
    line = {
        "metric": "solved DOF/s at 1e-10 rel residual (FP32-inner/FP64-refined GMRES; FP64 GMRES beside it); "
                  "adjoint/forward time ratio in 'adjoint'",
        "value": value,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-refined" if head == "mixed" else "f64",
        "data": "synthetic (seeded jittered lattice, prescribed cellular velocity)",
        "config": {"workload": f"steady convection-diffusion, {args.n}^3 jittered lattice ({N} vertices, {N} DOF, {nnz} nnz), "
                               f"cell Peclet 5, D=1e-2, GMRES({args.restart}) + {args.precond}, tol {args.tol}",
                   "dof": N, "nnz": nnz, "precond": args.precond, "restart": args.restart},
        "fp64": results.get("fp64"),
        "mixed": results.get("mixed"),
        "mixed_over_fp64": (results["fp64"]["ms_per_step"] / results["mixed"]["ms_per_step"]) if ("fp64" in results and "mixed" in results) else None,
        "assembly": {"kernel_ms": asm_ms, "wall_ms": 1e3 * t_asm_wall, "bytes": asm_bytes,
                     "gbs": asm_bytes / (asm_ms * 1e-3) / 1e9 if asm_ms > 0 else None},
        "mesh_and_dual_s": t_mesh,
        "mesh_note": "jittered lattice and exact Delaunay on the host, Voronoi dual on the GPU (bitwise equal to the host build)",
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic[0], "traffic_note": traffic[1],
                     "algorithmic_bytes_per_launch": dom["bytes_per_launch"]},
        "cpu_baseline": cpu,
        "adjoint": adjoint,
        "flow": flow,
        "heat_mass_flow": hmf,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
