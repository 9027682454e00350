"""Convergence scaling on the device drivers against the oracle: user reference scales
(Variables::set_scldom_scale / set_vecdom_scale, vars.rs:203-213) and scale_off (vars.rs:31)
change the scaled Picard residual (scaling.rs:88-135) exactly as in the oracle: the lid cavity
takes the oracle's number of Picard iterations under each setting, with the same fields."""
import numpy as np
import pytest

import vfvm
import vfvm_oracle as O
from problems import lid_cavity

pytestmark = pytest.mark.gpu

LIN = dict(tol=1e-13, restart=100, max_outer=400, precond=vfvm.PRECOND_BJACOBI)


@pytest.fixture(scope="module")
def ctx():
    c = vfvm.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def cavity():
    mesh = vfvm.Mesh.lattice(6, 0, 0.3)
    g = O.dual(mesh.get("pos"), mesh.get("tets"), mesh.num_pts)
    return mesh, g, O.grad_coeff(g)


CASES = [("none", {}), ("vel scale 5", dict(scale_vel=5.0)), ("pres scale 40", dict(scale_pres=40.0)),
         ("scale off", dict(scale_off=True))]


@pytest.mark.parametrize("name,scales", CASES)
def test_picard_with_scales_vs_oracle(ctx, cavity, name, scales):
    print(f"[scaling {name}]", flush=True)
    mesh, g, gc = cavity
    p = lid_cavity(mesh, rho=1.0, mu=0.05, lid=1.0)
    ovel, opres, its, hist = O.picard_flow(g, gc, p.vel, p.pres, p.den, p.visc, p.fce, p.bnd_vel,
                                           [mesh.bnd_nodes(k) for k in range(6)], p.pref, 60, 1e-6, 0.8, scales=scales)
    ops = p.to_device(ctx)
    vel_id, pres_id = p.ids[0], p.ids[1]
    if "scale_vel" in scales:
        ctx.set_vecdom_scale(vel_id, scales["scale_vel"])
    if "scale_pres" in scales:
        ctx.set_scldom_scale(pres_id, scales["scale_pres"])
    ctx.set_scale_off(scales.get("scale_off", False))
    try:
        st = ctx.steady_solve(ops, 60, 1e-6, 0.8, lin=LIN)
    finally:
        ctx.set_scale_off(False)
    vel = ctx.vecdom_get(vel_id)
    assert st.picard_iterations == its, (st.picard_iterations, its, hist)
    assert abs(st.residual - hist[-1]) <= 1e-4 * hist[-1]
    assert np.max(np.abs(vel - ovel)) <= 1e-8 * np.max(np.abs(ovel))


def test_scale_errors(ctx, cavity):
    mesh = cavity[0]
    ctx.upload(mesh)
    fid = ctx.scldom_add(vfvm.FIELD_UNKNOWN, 0.0)
    vid = ctx.vecdom_add(vfvm.FIELD_UNKNOWN, cnst3=(0.0, 0.0, 0.0))
    for bad in (0.0, -1.0, float("nan")):
        with pytest.raises(vfvm.VFVChemError) as e:
            ctx.set_scldom_scale(fid, bad)
        assert e.value.code == 60
        with pytest.raises(vfvm.VFVChemError) as e:
            ctx.set_vecdom_scale(vid, bad)
        assert e.value.code == 60
    with pytest.raises(vfvm.VFVChemError) as e:
        ctx.set_scldom_scale(fid + 7, 1.0)
    assert e.value.code == 1


def test_mirror_scales(cavity):
    """The Variables mirror carries the scales to its context (flow_lid.rs with a velocity scale)."""
    print("[scaling api mirror]", flush=True)
    mesh, g, gc = cavity
    p = lid_cavity(mesh, rho=1.0, mu=0.05, lid=1.0)
    _, _, its, _ = O.picard_flow(g, gc, p.vel, p.pres, p.den, p.visc, p.fce, p.bnd_vel,
                                 [mesh.bnd_nodes(k) for k in range(6)], p.pref, 60, 1e-6, 0.8, scales=dict(scale_vel=5.0))
    mesh.build_grad_coeff()
    vars = vfvm.Variables(mesh)
    dom = vars.add_dom("cube")
    bnds = [vars.add_bnd(dom, name) for name in ("x0", "x1", "y0", "y1", "z0", "z1")]
    vel = vars.add_vecdom_unkn(dom, (0.0, 0.0, 0.0))
    pres = vars.add_scldom_unkn(dom, 0.0)
    den = vars.add_scldom_cnst(dom, 1.0)
    visc = vars.add_scldom_cnst(dom, 0.05)
    fce = vars.add_vecdom_cnst(dom, (0.0, 0.0, 0.0))
    vars.set_vecdom_scale(vel, 5.0)
    phys = vfvm.SteadyFlow.new()
    phys.add_flow_dom(dom, vel, pres, den, visc, fce)
    for k, b in enumerate(bnds):
        phys.add_vel_bnd(b, vars.add_vecbnd_cnst(b, (1.0, 0.0, 0.0) if k == 5 else (0.0, 0.0, 0.0)))
    phys.set_pres_ref(dom, 0, 0.0)
    lin = vfvm.LinearGmres.new_with_options(1, 1e-13, 100, 400, precond=vfvm.PRECOND_BJACOBI)
    st = phys.solve(vars, lin, 60, 1e-6, 0.8)
    assert st.picard_iterations == its
