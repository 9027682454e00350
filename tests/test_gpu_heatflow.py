"""Parity of the coupled heat-flow system on the GPU (through the C ABI) with the oracle.

The solution vector is [u, v, w, p, T] (SteadyHeatFlow::initial_start): the flow rows of
the flow operators and the temperature rows of the scalar operators, whose advection reads
the flow's own Rhie-Chow flux; the Boussinesq body force is a function of T set by a host
update hook before every assembly.

Bars: pattern, values and b of the coupled system equal to the oracle's bit for bit; the
steady Boussinesq cavity (Picard with damping, block Jacobi or ILU(0), FP64 or mixed) and
three implicit-Euler steps of it within 1e-8 (FP64) / 1e-7 (mixed) of the oracle's Picard
iteration with SuperLU solves, with the same number of Picard iterations.
"""
import numpy as np
import pytest

import vfvm
import vfvm_oracle as O
from problems import HeatFlowProblem, flow_general, lid_cavity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vfvm.Context(0)
    yield c
    c.close()


def _mismatch(a, b):
    d = np.abs(a - b)
    return int(np.count_nonzero(a != b)), float(np.max(d) if a.size else 0.0)


@pytest.mark.parametrize("n,transient", [(5, False), (6, True)])
def test_coupled_assembly_vs_oracle(ctx, n, transient):
    print(f"[heat-flow assemble n={n} transient={transient}]", flush=True)
    mesh = vfvm.Mesh.lattice(n, 3, 0.3)
    pos = mesh.get("pos")
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    fp = flow_general(mesh)
    T0 = 0.4 + 0.3 * np.sin(2.0 * x) * np.cos(y) + 0.2 * z
    hp = HeatFlowProblem(fp, rcp=2.0 + 0.5 * y, k=0.05 * (1.0 + 0.3 * z), q=0.3 * np.cos(3.0 * y), gamma=5.0, T0=T0,
                         transient=transient)
    ops = hp.to_device(ctx)
    N = mesh.num_tot
    nnz = ctx.pattern_build(ops)
    assert ctx.system_info()[0] == 5 * N
    dt = 0.05 if transient else 0.0
    vprev, dprev, Tprev = fp.vel.copy(), fp.den.copy(), T0.copy()
    if transient:
        # previous state = the uploaded state (kept by the fields since they were added);
        # current state moved away from it
        fp.vel = 0.8 * vprev + 0.02
        ctx.vecdom_set(fp.ids[0], fp.vel)
        hp.T0 = 1.1 * T0
        ctx.scldom_set(hp.hids[0], hp.T0)
    ctx.assemble(ops, dt)
    rp, col = ctx.pattern_get()
    vals, b = ctx.matrix_get()
    g, gc, f, (orp, ocol, oval, ob) = fp.oracle(dt=dt, vel_prev=vprev if transient else None,
                                                 den_prev=dprev if transient else None)
    r, c, v, bh = hp.oracle_heat_rows(g, dt)(hp.T0, Tprev, f)
    hrp, hcol, hval = O.coo_to_csr(r, c, v, N)
    erp = np.concatenate([orp, orp[-1] + hrp[1:]])
    ecol = np.concatenate([ocol, hcol + 4 * N])
    eval_ = np.concatenate([oval, hval])
    eb = np.concatenate([ob, bh])
    assert nnz == ecol.size
    assert np.array_equal(rp, erp) and np.array_equal(col, ecol)
    assert np.array_equal(vals, eval_), _mismatch(vals, eval_)
    assert np.array_equal(b, eb), _mismatch(b, eb)


def _cavity(mesh, transient=False):
    fp = lid_cavity(mesh, rho=1.0, mu=0.1, lid=0.0)
    return HeatFlowProblem(fp, rcp=1.0, k=0.1, q=0.0, gamma=20.0, transient=transient)


@pytest.fixture(scope="module")
def cavity_oracle():
    mesh = vfvm.Mesh.lattice(5, 0, 0.3)
    hp = _cavity(mesh)
    hp.hids = (0, 1, 2, 3, 0, 0, 1, {2: 2, 3: 3, 4: 4, 5: 5})
    g = O.dual(mesh.get("pos"), mesh.get("tets"), mesh.num_pts)
    gc = O.grad_coeff(g)
    fp = hp.fp
    res = O.heatflow_solve(g, gc, fp.vel, fp.pres, hp.T0, fp.den, fp.visc, hp.fce_of, fp.bnd_vel,
                           [mesh.bnd_nodes(k) for k in range(6)], fp.pref, hp.oracle_heat_rows(g),
                           hp.oracle_t_dirichlet(), 60, 1e-9, 0.8)
    return mesh, res


@pytest.mark.parametrize("precond,precision,bar", [(vfvm.PRECOND_BJACOBI, vfvm.FP64, 1e-8),
                                                   (vfvm.PRECOND_ILU0, vfvm.FP64, 1e-8),
                                                   (vfvm.PRECOND_BJACOBI, vfvm.FP32_REFINED, 1e-7)])
def test_boussinesq_cavity_picard_vs_oracle(ctx, cavity_oracle, precond, precision, bar):
    print(f"[heat-flow cavity precond={precond} precision={precision}]", flush=True)
    mesh, (ovel, opres, oT, its) = cavity_oracle
    hp = _cavity(mesh)
    ops = hp.to_device(ctx)
    ctx.set_update_hook(hp.hook(ctx))
    try:
        st = ctx.steady_solve(ops, 60, 1e-9, 0.8, lin=dict(tol=1e-13, restart=100, max_outer=400, precision=precision,
                                                          precond=precond, inner_tol=1e-6))
    finally:
        ctx.set_update_hook(None)
    vel = ctx.vecdom_get(hp.fp.ids[0])
    pres = ctx.scldom_get(hp.fp.ids[1])
    T = ctx.scldom_get(hp.hids[0])
    assert st.picard_iterations == its, (st.picard_iterations, its)
    assert np.max(np.abs(vel - ovel)) <= bar * np.max(np.abs(ovel))
    assert np.max(np.abs(pres - opres)) <= bar * np.max(np.abs(opres))
    assert np.max(np.abs(T - oT)) <= bar * np.max(np.abs(oT))


def test_transient_boussinesq_vs_oracle(ctx):
    print("[heat-flow transient cavity]", flush=True)
    mesh = vfvm.Mesh.lattice(5, 0, 0.3)
    hp = _cavity(mesh, transient=True)
    ops = hp.to_device(ctx)
    dt, steps = 0.05, 3
    ctx.set_update_hook(hp.hook(ctx))
    try:
        st = ctx.transient_solve(ops, dt, steps, 40, 1e-9, 1.0, lin=dict(tol=1e-13, restart=100, max_outer=400,
                                                                         precision=vfvm.FP64, precond=vfvm.PRECOND_BJACOBI))
    finally:
        ctx.set_update_hook(None)
    vel = ctx.vecdom_get(hp.fp.ids[0])
    T = ctx.scldom_get(hp.hids[0])
    g = O.dual(mesh.get("pos"), mesh.get("tets"), mesh.num_pts)
    gc = O.grad_coeff(g)
    fp = hp.fp
    ovel, opres, oT, its = O.heatflow_solve(g, gc, fp.vel, fp.pres, hp.T0, fp.den, fp.visc, hp.fce_of, fp.bnd_vel,
                                            [mesh.bnd_nodes(k) for k in range(6)], fp.pref, hp.oracle_heat_rows(g, dt),
                                            hp.oracle_t_dirichlet(), 40, 1e-9, 1.0, dt=dt, num_ts=steps)
    assert st.picard_iterations == its, (st.picard_iterations, its)
    assert np.max(np.abs(vel - ovel)) <= 1e-8 * np.max(np.abs(ovel))
    assert np.max(np.abs(T - oT)) <= 1e-8 * np.max(np.abs(oT))
