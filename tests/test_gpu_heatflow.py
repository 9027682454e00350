"""Parity of the coupled heat(-mass)-flow system on the GPU (through the C ABI) with the oracle.

The solution vector is [u, v, w, p, T] or [u, v, w, p, T, C] (SteadyHeatFlow /
TransientHeatMassFlow::initial_start): the flow rows of the flow operators and the rows of
the scalar operators, whose advection reads the flow's own Rhie-Chow flux; the Boussinesq
body force is a function of T set by a host update hook before every assembly.

Bars: pattern, values and b of the coupled system equal to the oracle's bit for bit; the
steady Boussinesq cavity (Picard with damping; block Jacobi with 5 or 6 unknowns per vertex
or ILU(0); FP64 or mixed) and three implicit-Euler steps of it within 1e-8 (FP64) / 1e-7
(mixed) of the oracle's Picard iteration with SuperLU solves, with the same number of
Picard iterations.
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


@pytest.mark.parametrize("n,transient,species", [(5, False, False), (6, True, False), (5, False, True), (6, True, True)])
def test_coupled_assembly_vs_oracle(ctx, n, transient, species):
    print(f"[heat-flow assemble n={n} transient={transient} species={species}]", flush=True)
    mesh = vfvm.Mesh.lattice(n, 3, 0.3)
    pos = mesh.get("pos")
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    fp = flow_general(mesh)
    T0 = 0.4 + 0.3 * np.sin(2.0 * x) * np.cos(y) + 0.2 * z
    C0 = 0.5 + 0.2 * np.cos(3.0 * z) * x
    hp = HeatFlowProblem(fp, rcp=2.0 + 0.5 * y, k=0.05 * (1.0 + 0.3 * z), q=0.3 * np.cos(3.0 * y), gamma=5.0, T0=T0,
                         transient=transient, species=species, D=0.02 * (1.0 + x * y), C0=C0)
    ops = hp.to_device(ctx)
    N = mesh.num_tot
    K = 2 if species else 1
    nnz = ctx.pattern_build(ops)
    assert ctx.system_info()[0] == (4 + K) * N
    dt = 0.05 if transient else 0.0
    vprev, dprev, Sprev = fp.vel.copy(), fp.den.copy(), hp.S0()
    S = Sprev.copy()
    if transient:
        # previous state = the uploaded state (kept by the fields since they were added);
        # current state moved away from it
        fp.vel = 0.8 * vprev + 0.02
        ctx.vecdom_set(fp.ids[0], fp.vel)
        S = 1.1 * Sprev
        ctx.scldom_set(hp.hid["T"], S[:N])
        if species:
            ctx.scldom_set(hp.hid["C"], S[N:])
    ctx.assemble(ops, dt)
    rp, col = ctx.pattern_get()
    vals, b = ctx.matrix_get()
    g, gc, f, (orp, ocol, oval, ob) = fp.oracle(dt=dt, vel_prev=vprev if transient else None,
                                                 den_prev=dprev if transient else None)
    r, c, v, bh = hp.oracle_heat_rows(g, dt)(S, Sprev, f)
    hrp, hcol, hval = O.coo_to_csr(r, c, v, K * N)
    erp = np.concatenate([orp, orp[-1] + hrp[1:]])
    ecol = np.concatenate([ocol, hcol + 4 * N])
    eval_ = np.concatenate([oval, hval])
    eb = np.concatenate([ob, bh])
    assert nnz == ecol.size
    assert np.array_equal(rp, erp) and np.array_equal(col, ecol)
    assert np.array_equal(vals, eval_), _mismatch(vals, eval_)
    assert np.array_equal(b, eb), _mismatch(b, eb)


def _cavity(mesh, transient=False, species=False):
    fp = lid_cavity(mesh, rho=1.0, mu=0.1, lid=0.0)
    return HeatFlowProblem(fp, rcp=1.0, k=0.1, q=0.0, gamma=20.0, transient=transient, species=species, D=0.07)


def _oracle_solve(mesh, hp, g, gc, max_iter, tol, damp, dt=0.0, steps=0):
    hp.host_ids()
    fp = hp.fp
    return O.heatflow_solve(g, gc, fp.vel, fp.pres, hp.S0(), fp.den, fp.visc, hp.fce_of, fp.bnd_vel,
                            [mesh.bnd_nodes(k) for k in range(6)], fp.pref, hp.oracle_heat_rows(g, dt),
                            hp.oracle_t_dirichlet(), max_iter, tol, damp, dt=dt, num_ts=steps)


@pytest.fixture(scope="module")
def cavity_oracle():
    mesh = vfvm.Mesh.lattice(5, 0, 0.3)
    g = O.dual(mesh.get("pos"), mesh.get("tets"), mesh.num_pts)
    gc = O.grad_coeff(g)
    return mesh, {sp: _oracle_solve(mesh, _cavity(mesh, species=sp), g, gc, 60, 1e-9, 0.8) for sp in (False, True)}


@pytest.mark.parametrize("species,precond,precision,bar", [(False, vfvm.PRECOND_BJACOBI, vfvm.FP64, 1e-8),
                                                           (False, vfvm.PRECOND_ILU0, vfvm.FP64, 1e-8),
                                                           (False, vfvm.PRECOND_BJACOBI, vfvm.FP32_REFINED, 1e-7),
                                                           (True, vfvm.PRECOND_BJACOBI, vfvm.FP64, 1e-8),
                                                           (True, vfvm.PRECOND_BJACOBI, vfvm.FP32_REFINED, 1e-7),
                                                           (True, vfvm.PRECOND_JACOBI, vfvm.FP64, 1e-8)])
def test_boussinesq_cavity_picard_vs_oracle(ctx, cavity_oracle, species, precond, precision, bar):
    print(f"[heat-flow cavity species={species} precond={precond} precision={precision}]", flush=True)
    mesh, res = cavity_oracle
    ovel, opres, oS, its = res[species]
    hp = _cavity(mesh, species=species)
    ops = hp.to_device(ctx)
    ctx.set_update_hook(hp.hook(ctx))
    try:
        st = ctx.steady_solve(ops, 60, 1e-9, 0.8, lin=dict(tol=1e-13, restart=256, max_outer=400, precision=precision,
                                                          precond=precond, inner_tol=1e-6))
    finally:
        ctx.set_update_hook(None)
    vel = ctx.vecdom_get(hp.fp.ids[0])
    pres = ctx.scldom_get(hp.fp.ids[1])
    S = hp.scalar_values(ctx)
    assert st.picard_iterations == its, (st.picard_iterations, its)
    assert np.max(np.abs(vel - ovel)) <= bar * np.max(np.abs(ovel))
    assert np.max(np.abs(pres - opres)) <= bar * np.max(np.abs(opres))
    assert np.max(np.abs(S - oS)) <= bar * np.max(np.abs(oS))


@pytest.mark.parametrize("species", [False, True])
def test_transient_boussinesq_vs_oracle(ctx, species):
    print(f"[heat-flow transient cavity species={species}]", flush=True)
    mesh = vfvm.Mesh.lattice(5, 0, 0.3)
    hp = _cavity(mesh, transient=True, species=species)
    ops = hp.to_device(ctx)
    dt, steps = 0.05, 3
    ctx.set_update_hook(hp.hook(ctx))
    try:
        st = ctx.transient_solve(ops, dt, steps, 40, 1e-9, 1.0, lin=dict(tol=1e-13, restart=100, max_outer=400,
                                                                         precision=vfvm.FP64, precond=vfvm.PRECOND_BJACOBI))
    finally:
        ctx.set_update_hook(None)
    vel = ctx.vecdom_get(hp.fp.ids[0])
    S = hp.scalar_values(ctx)
    g = O.dual(mesh.get("pos"), mesh.get("tets"), mesh.num_pts)
    gc = O.grad_coeff(g)
    ovel, opres, oS, its = _oracle_solve(mesh, _cavity(mesh, transient=True, species=species), g, gc, 40, 1e-9, 1.0,
                                         dt=dt, steps=steps)
    assert st.picard_iterations == its, (st.picard_iterations, its)
    assert np.max(np.abs(vel - ovel)) <= 1e-8 * np.max(np.abs(ovel))
    assert np.max(np.abs(S - oS)) <= 1e-8 * np.max(np.abs(oS))


def _mirror_cavity(mesh, transient, species):
    """The Boussinesq cavity of _cavity built through the mirrored reference API
    (Variables / SteadyHeatFlow / TransientHeatMassFlow, a body-force closure of T)."""
    vars = vfvm.Variables(mesh)
    dom = vars.add_dom("cube")
    bnds = [vars.add_bnd(dom, name) for name in ("x0", "x1", "y0", "y1", "z0", "z1")]
    vel = vars.add_vecdom_unkn(dom, (0.0, 0.0, 0.0))
    pres = vars.add_scldom_unkn(dom, 0.0)
    den = vars.add_scldom_cnst(dom, 1.0)
    visc = vars.add_scldom_cnst(dom, 0.1)
    temp = vars.add_scldom_unkn(dom, 0.0)
    fce = vars.add_vecdom_func(dom, lambda t, pos, u: np.stack([0.0 * u[0], 0.0 * u[0], 20.0 * (u[0] - 0.5)]), [temp])
    vlcp = vars.add_scldom_cnst(dom, 1.0)
    cond = vars.add_scldom_cnst(dom, 0.1)
    zero = vars.add_scldom_cnst(dom, 0.0)
    phys = (vfvm.TransientHeatMassFlow.new(1) if species else vfvm.TransientHeatFlow.new()) if transient else vfvm.SteadyHeatFlow.new()
    phys.add_flow_dom(dom, vel, pres, den, visc, fce)
    for b in bnds:
        phys.add_vel_bnd(b, vars.add_vecbnd_cnst(b, (0.0, 0.0, 0.0)))
    phys.set_pres_ref(dom, 0, 0.0)
    phys.add_heat_dom(dom, temp, vlcp, cond, zero, zero)
    phys.add_temp_bnd(bnds[0], vars.add_sclbnd_cnst(bnds[0], 1.0))
    phys.add_temp_bnd(bnds[1], vars.add_sclbnd_cnst(bnds[1], 0.0))
    for b in bnds[2:]:
        phys.add_hflx_bnd(b, vars.add_sclbnd_cnst(b, 0.0))
    conc = None
    if species:
        conc = vars.add_scldom_unkn(dom, 0.0)
        dif = vars.add_scldom_cnst(dom, 0.07)
        phys.add_mass_dom(0, dom, conc, {0: dif}, zero, zero)
        phys.add_conc_bnd(0, bnds[0], vars.add_sclbnd_cnst(bnds[0], 1.0))
        phys.add_conc_bnd(0, bnds[1], vars.add_sclbnd_cnst(bnds[1], 0.0))
        for b in bnds[2:]:
            phys.add_mflx_bnd(0, b, vars.add_sclbnd_cnst(b, 0.0))
    return vars, phys, vel, pres, temp, conc


def test_reference_api_mirror_steady_heatflow(cavity_oracle):
    """SteadyHeatFlow through the mirrored API (the Boussinesq force a vecdom closure of T,
    refreshed by the update hook) gives the oracle cavity's solution."""
    print("[heat-flow api mirror steady]", flush=True)
    mesh, res = cavity_oracle
    ovel, opres, oS, its = res[False]
    mesh.build_grad_coeff()
    vars, phys, vel, pres, temp, _ = _mirror_cavity(mesh, False, False)
    lin = vfvm.LinearGmres.new_with_options(1, 1e-13, 256, 400, precond=vfvm.PRECOND_BJACOBI)
    st = phys.solve(vars, lin, 60, 1e-9, 0.8)
    assert st.picard_iterations == its
    assert np.max(np.abs(vars.value_vec(vel) - ovel)) <= 1e-8 * np.max(np.abs(ovel))
    assert np.max(np.abs(vars.value(temp) - oS)) <= 1e-8 * np.max(np.abs(oS))


def test_reference_api_mirror_transient_heatmassflow():
    """TransientHeatMassFlow through the mirrored API (reference operator order: the species
    time operator after diffusion / production / source) against the oracle's stepping of
    the same physics (species time operator first: equal up to rounding)."""
    print("[heat-mass-flow api mirror transient]", flush=True)
    mesh = vfvm.Mesh.lattice(5, 0, 0.3)
    mesh.build_grad_coeff()
    vars, phys, vel, pres, temp, conc = _mirror_cavity(mesh, True, True)
    lin = vfvm.LinearGmres.new_with_options(1, 1e-13, 256, 400, precond=vfvm.PRECOND_BJACOBI)
    dt, steps = 0.05, 3
    st = phys.solve(vars, lin, dt, steps, 1, 40, 1e-9, 1.0)
    g = O.dual(mesh.get("pos"), mesh.get("tets"), mesh.num_pts)
    gc = O.grad_coeff(g)
    ovel, opres, oS, its = _oracle_solve(mesh, _cavity(mesh, transient=True, species=True), g, gc, 40, 1e-9, 1.0,
                                         dt=dt, steps=steps)
    S = np.concatenate([vars.value(temp), vars.value(conc)])
    assert st.picard_iterations == its
    assert np.max(np.abs(vars.value_vec(vel) - ovel)) <= 1e-8 * np.max(np.abs(ovel))
    assert np.max(np.abs(S - oS)) <= 1e-8 * np.max(np.abs(oS))


def test_adaptive_heatmassflow(ctx):
    """The adaptive driver on the six-unknown system: fixed steps (no intermediate output
    times) equal the fixed-step driver's (same Picard iterations, fields within 1e-10), and an
    error-controlled run reaches t_end with every field finite and T, C within their wall
    values."""
    print("[heat-mass-flow adaptive]", flush=True)
    mesh = vfvm.Mesh.lattice(5, 0, 0.3)
    lin = dict(tol=1e-13, restart=256, max_outer=400, precond=vfvm.PRECOND_BJACOBI)
    dt, steps = 0.05, 3

    def run(adaptive):
        hp = _cavity(mesh, transient=True, species=True)
        ops = hp.to_device(ctx)
        ctx.set_update_hook(hp.hook(ctx))
        try:
            if adaptive:
                st, _ = ctx.transient_adaptive(ops, dt, dt * steps, lin=lin, fixed=1, max_iter=40, tol=1e-9, damp=1.0,
                                               out_times=[])
                its = st.picard_iterations
            else:
                its = ctx.transient_solve(ops, dt, steps, 40, 1e-9, 1.0, lin=lin).picard_iterations
        finally:
            ctx.set_update_hook(None)
        return its, ctx.vecdom_get(hp.fp.ids[0]), hp.scalar_values(ctx)

    its_a, vel_a, S_a = run(True)
    its_f, vel_f, S_f = run(False)
    assert its_a == its_f
    assert np.max(np.abs(vel_a - vel_f)) <= 1e-10 * np.max(np.abs(vel_f))
    assert np.max(np.abs(S_a - S_f)) <= 1e-10 * np.max(np.abs(S_f))

    hp = _cavity(mesh, transient=True, species=True)
    ops = hp.to_device(ctx)
    ctx.set_update_hook(hp.hook(ctx))
    try:
        # from rest the velocity's error estimate is 0.5 / rtol for any step (the predictor of
        # the first step is the zero state), so one fixed step first gives a state the
        # controller can follow, as for the impulsively started lid
        ctx.transient_adaptive(ops, 0.05, 0.05, lin=lin, fixed=1, max_iter=40, tol=1e-7, out_times=[])
        st, stp = ctx.transient_adaptive(ops, 0.05, 0.4, lin=dict(lin, tol=1e-10, max_outer=100), rtol=2e-2, max_iter=40,
                                         tol=1e-7, scl_exclude=[hp.fp.ids[1]])
    finally:
        ctx.set_update_hook(None)
    S = hp.scalar_values(ctx)
    assert st.t_final == 0.4 and st.accepted >= 3
    assert np.all(np.isfinite(S)) and np.min(S) >= -1e-6 and np.max(S) <= 1.0 + 1e-6
    assert np.all(np.isfinite(ctx.vecdom_get(hp.fp.ids[0])))


def test_coupled_list_errors(ctx):
    """Rejected coupled lists name the reason: a third scalar unknown (seven per vertex), a
    scalar operator on the flow's pressure, a scalar operator on a non-unknown field, and
    cross-diffusion into another unknown."""
    print("[heat-flow list errors]", flush=True)
    mesh = vfvm.Mesh.lattice(4, 0, 0.3)
    hp = _cavity(mesh, species=True)
    ops = hp.to_device(ctx)
    T, C = hp.hid["T"], hp.hid["C"]
    extra = ctx.scldom_add(vfvm.FIELD_UNKNOWN, 0.0)
    k = hp.hid["k"]
    cases = [
        (ops + [(vfvm.OP_SCL_DOM_DIFFUSION, 0, extra, k, extra)], vfvm.VFVChemError, "at most two scalar unknowns"),
        (ops + [(vfvm.OP_SCL_DOM_DIFFUSION, 0, k, k, k)], vfvm.VFVChemError, "unknown scalar field"),
        ([o if not (o[0] == vfvm.OP_SCL_DOM_DIFFUSION and o[2] == C) else (o[0], o[1], C, o[3], T) for o in ops],
         vfvm.VFVChemError, "cross-diffusion"),
    ]
    for bad, exc, msg in cases:
        with pytest.raises(exc, match=msg):
            ctx.pattern_build(bad)
    # the valid list still builds afterwards
    assert ctx.pattern_build(ops) > 0
    # a scalar operator on the pressure (with one scalar block, so the block count holds)
    hp = _cavity(mesh)
    ops = hp.to_device(ctx)
    pres, k = hp.fp.ids[1], hp.hid["k"]
    with pytest.raises(vfvm.VFVChemError, match="pressure"):
        ctx.pattern_build(ops + [(vfvm.OP_SCL_DOM_DIFFUSION, 0, pres, k, pres)])
