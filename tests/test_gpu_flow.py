"""Parity of the flow operators on the GPU (through the C ABI) with the oracle.

Bars: flow data (rc, momentum term, face mass flux) and the assembled coupled system
(pattern, values, b) equal to the oracle's bit for bit (the kernels sum every entry in the
reference's push order, as the oracle does); block-Jacobi application within 1e-12 of a dense
solve per block; steady lid-driven cavity through the device Picard driver within 1e-8
(FP64) / 1e-7 (mixed) of the oracle's Picard iteration with SuperLU solves, with the same
number of Picard iterations.
"""
import numpy as np
import pytest

import vfvm
import vfvm_oracle as O
from problems import flow_general, lid_cavity

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vfvm.Context(0)
    yield c
    c.close()


def _mismatch(a, b):
    d = np.abs(a - b)
    s = np.maximum(np.abs(b), 1e-300)
    return int(np.count_nonzero(a != b)), float(np.max(d / np.maximum(s, np.max(np.abs(b)) * 1e-300)) if a.size else 0.0)


@pytest.mark.parametrize("n", [5, 7])
def test_flow_data_vs_oracle(ctx, n):
    print(f"[flow data n={n}]", flush=True)
    mesh = vfvm.Mesh.lattice(n, 1, 0.3)
    p = flow_general(mesh)
    ops = p.to_device(ctx)
    g, gc, f, _ = p.oracle()
    flw = p.ids[5]
    ctx.update_flow(flw)
    N = mesh.num_tot
    rc = ctx.sclflw_get(flw, vfvm.FLW_RC)
    mom = ctx.sclflw_get(flw, vfvm.FLW_MOM).reshape(3, N)
    flux = ctx.sclflw_get(flw, vfvm.FLW_FLUX)
    assert np.array_equal(rc, f["rc"]), _mismatch(rc, f["rc"])
    assert np.array_equal(mom, f["mom"]), _mismatch(mom, f["mom"])
    assert np.array_equal(flux, f["flux"]), _mismatch(flux, f["flux"])


@pytest.mark.parametrize("n,pref", [(5, None), (6, (3, 0.25)), (8, (0, 0.0))])
def test_flow_assembly_vs_oracle(ctx, n, pref):
    print(f"[flow assemble n={n}]", flush=True)
    mesh = vfvm.Mesh.lattice(n, 2, 0.3)
    p = flow_general(mesh)
    p.pref = pref
    ops = p.to_device(ctx)
    nnz = ctx.pattern_build(ops)
    ctx.assemble(ops)
    rp, col = ctx.pattern_get()
    vals, b = ctx.matrix_get()
    _, _, _, (orp, ocol, oval, ob) = p.oracle()
    assert nnz == ocol.size
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.array_equal(vals, oval), _mismatch(vals, oval)
    assert np.array_equal(b, ob), _mismatch(b, ob)


def test_block_jacobi_apply(ctx):
    print("[flow block jacobi]", flush=True)
    mesh = vfvm.Mesh.lattice(6, 0, 0.3)
    p = flow_general(mesh)
    ops = p.to_device(ctx)
    ctx.pattern_build(ops)
    ctx.assemble(ops)
    rp, col = ctx.pattern_get()
    vals, _ = ctx.matrix_get()
    r = np.random.default_rng(3).standard_normal(4 * mesh.num_tot)
    z_ref = O.bjacobi_apply(rp, col, vals, r, mesh.num_tot)
    z = ctx.precond_apply(r, precond=vfvm.PRECOND_BJACOBI)
    assert np.max(np.abs(z - z_ref)) <= 1e-12 * np.max(np.abs(z_ref))
    z32 = ctx.precond_apply(r, precond=vfvm.PRECOND_BJACOBI, precision=vfvm.FP32_REFINED)
    assert np.max(np.abs(z32 - z_ref)) <= 1e-4 * np.max(np.abs(z_ref))


@pytest.fixture(scope="module")
def lid_oracle():
    mesh = vfvm.Mesh.lattice(6, 0, 0.3)
    p = lid_cavity(mesh, rho=1.0, mu=0.05, lid=1.0)
    g = O.dual(mesh.get("pos"), mesh.get("tets"), mesh.num_pts)
    gc = O.grad_coeff(g)
    return mesh, O.picard_flow(g, gc, p.vel, p.pres, p.den, p.visc, p.fce, p.bnd_vel,
                               [mesh.bnd_nodes(k) for k in range(6)], p.pref, 50, 1e-6, 0.8)


@pytest.mark.parametrize("precision,bar", [(vfvm.FP64, 1e-8), (vfvm.FP32_REFINED, 1e-7)])
def test_lid_cavity_picard_vs_oracle(ctx, lid_oracle, precision, bar):
    print(f"[flow lid cavity precision={precision}]", flush=True)
    mesh, (ovel, opres, its, hist) = lid_oracle
    p = lid_cavity(mesh, rho=1.0, mu=0.05, lid=1.0)
    ops = p.to_device(ctx)
    st = ctx.steady_solve(ops, 50, 1e-6, 0.8, lin=dict(tol=1e-13, restart=100, max_outer=400, precision=precision,
                                                      precond=vfvm.PRECOND_BJACOBI, inner_tol=1e-6))
    vel = ctx.vecdom_get(p.ids[0])
    pres = ctx.scldom_get(p.ids[1])
    assert st.picard_iterations == its, (st.picard_iterations, its, hist)
    assert np.max(np.abs(vel - ovel)) <= bar * np.max(np.abs(ovel))
    assert np.max(np.abs(pres - opres)) <= bar * max(np.max(np.abs(opres)), 1e-30)
    # the lid drives a recirculation: u < 0 somewhere in the lower half
    pos = mesh.get("pos")
    assert np.min(vel[0][:mesh.num_pts][pos[:mesh.num_pts, 2] < 0.5]) < 0.0


def test_reference_api_mirror_steady_flow(ctx, lid_oracle):
    """flow_lid.rs through the mirrored API (Variables / SteadyFlow / LinearGmres) gives the
    same solution as the operator list of the oracle's lid cavity."""
    print("[flow api mirror]", flush=True)
    mesh, (ovel, opres, its, hist) = lid_oracle
    mesh.build_grad_coeff()
    vars = vfvm.Variables(mesh)
    dom = vars.add_dom("cube")
    bnds = [vars.add_bnd(dom, name) for name in ("x0", "x1", "y0", "y1", "z0", "z1")]
    vel = vars.add_vecdom_unkn(dom, (0.0, 0.0, 0.0))
    pres = vars.add_scldom_unkn(dom, 0.0)
    den = vars.add_scldom_cnst(dom, 1.0)
    visc = vars.add_scldom_cnst(dom, 0.05)
    fce = vars.add_vecdom_cnst(dom, (0.0, 0.0, 0.0))
    phys = vfvm.SteadyFlow.new()
    phys.add_flow_dom(dom, vel, pres, den, visc, fce)
    for k, b in enumerate(bnds):
        phys.add_vel_bnd(b, vars.add_vecbnd_cnst(b, (1.0, 0.0, 0.0) if k == 5 else (0.0, 0.0, 0.0)))
    phys.set_pres_ref(dom, 0, 0.0)
    lin = vfvm.LinearGmres.new_with_options(1, 1e-13, 100, 400, precond=vfvm.PRECOND_BJACOBI)
    st = phys.solve(vars, lin, 50, 1e-6, 0.8)
    assert st.picard_iterations == its
    v = vars.value_vec(vel)
    assert np.max(np.abs(v - ovel)) <= 1e-8 * np.max(np.abs(ovel))


def test_transient_flow_assembly_vs_oracle(ctx):
    """TransientFlow system (velocity time operator, divergence time part) bitwise equal to the
    oracle, with previous velocity and density different from the current ones."""
    print("[transient flow assemble]", flush=True)
    mesh = vfvm.Mesh.lattice(6, 4, 0.3)
    p = flow_general(mesh)
    p.transient = True
    vprev, dprev = p.vel.copy(), p.den.copy()
    ops = p.to_device(ctx)
    vel_id, pres_id, den_id = p.ids[0], p.ids[1], p.ids[2]
    p.vel = 0.7 * vprev + 0.05
    p.den = 1.1 * dprev
    ctx.vecdom_set(vel_id, p.vel)
    ctx.scldom_set(den_id, p.den)
    ctx.pattern_build(ops)
    dt = 0.05
    ctx.assemble(ops, dt)
    rp, col = ctx.pattern_get()
    vals, b = ctx.matrix_get()
    _, _, _, (orp, ocol, oval, ob) = p.oracle(dt=dt, vel_prev=vprev, den_prev=dprev)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
    assert np.array_equal(vals, oval), _mismatch(vals, oval)
    assert np.array_equal(b, ob), _mismatch(b, ob)


def test_transient_lid_cavity_vs_oracle(ctx):
    print("[transient lid cavity]", flush=True)
    mesh = vfvm.Mesh.lattice(5, 0, 0.3)
    p = lid_cavity(mesh, rho=1.0, mu=0.05, lid=1.0)
    p.transient = True
    ops = p.to_device(ctx)
    dt, steps = 0.05, 3
    st = ctx.transient_solve(ops, dt, steps, 30, 1e-7, 1.0, lin=dict(tol=1e-13, restart=100, max_outer=400,
                                                                         precision=vfvm.FP64, precond=vfvm.PRECOND_BJACOBI))
    vel = ctx.vecdom_get(p.ids[0])
    pres = ctx.scldom_get(p.ids[1])
    g = O.dual(mesh.get("pos"), mesh.get("tets"), mesh.num_pts)
    gc = O.grad_coeff(g)
    ovel, opres, iters = O.transient_flow(g, gc, p.vel, p.pres, p.den, p.visc, p.fce, p.bnd_vel,
                                          [mesh.bnd_nodes(k) for k in range(6)], p.pref, dt, steps, 30, 1e-7, 1.0)
    assert st.picard_iterations == sum(iters), (st.picard_iterations, iters)
    assert np.max(np.abs(vel - ovel)) <= 1e-8 * np.max(np.abs(ovel))
    assert np.max(np.abs(pres - opres)) <= 1e-8 * np.max(np.abs(opres))


@pytest.fixture(scope="module")
def lid_oracle_re100():
    mesh = vfvm.Mesh.lattice(6, 0, 0.3)
    p = lid_cavity(mesh, rho=1.0, mu=0.01, lid=1.0)
    g = O.dual(mesh.get("pos"), mesh.get("tets"), mesh.num_pts)
    gc = O.grad_coeff(g)
    return mesh, O.picard_flow(g, gc, p.vel, p.pres, p.den, p.visc, p.fce, p.bnd_vel,
                               [mesh.bnd_nodes(k) for k in range(6)], p.pref, 100, 1e-6, 0.8)


@pytest.mark.parametrize("precision,bar", [(vfvm.FP64, 1e-8), (vfvm.FP32_REFINED, 1e-7)])
def test_lid_cavity_re100_picard_vs_oracle(ctx, lid_oracle_re100, precision, bar):
    """The lid-driven cavity at its configured Reynolds number (Re = 100: rho = 1, mu = 0.01, lid
    speed 1, unit cavity) through the device Picard driver (steady_base.rs:74-117, damping 0.8)
    to a scaled residual of 1e-6: the oracle's number of Picard iterations and its fields."""
    print(f"[flow lid cavity Re=100 precision={precision}]", flush=True)
    mesh, (ovel, opres, its, hist) = lid_oracle_re100
    p = lid_cavity(mesh, rho=1.0, mu=0.01, lid=1.0)
    ops = p.to_device(ctx)
    st = ctx.steady_solve(ops, 100, 1e-6, 0.8, lin=dict(tol=1e-13, restart=250, max_outer=400, precision=precision,
                                                       precond=vfvm.PRECOND_BJACOBI, inner_tol=1e-6))
    vel = ctx.vecdom_get(p.ids[0])
    pres = ctx.scldom_get(p.ids[1])
    assert st.picard_iterations == its, (st.picard_iterations, its, hist)
    assert np.max(np.abs(vel - ovel)) <= bar * np.max(np.abs(ovel))
    assert np.max(np.abs(pres - opres)) <= bar * max(np.max(np.abs(opres)), 1e-30)
