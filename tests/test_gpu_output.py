"""Field output through the reference API mirror (the drivers' _write, steady_base.rs:213-233,
transient_base.rs:67, 141-143, writing CSV as write_csv.rs:5-100 does): file names and
numbering, contents against the solved fields, and transient runs written in chunks equal to
a run without output (time-dependent function fields included)."""
import os

import numpy as np
import pytest

import vfvm


def read_csv(path):
    with open(path) as fh:
        head = fh.readline().strip()
    return head, np.loadtxt(path, delimiter=",", skiprows=1, ndmin=2)


def heat_problem(n, out_dir=None, src_of_t=False):
    vars = vfvm.Variables.new_lattice(n)
    dom = vars.add_dom()
    b0, b1 = vars.add_bnd(dom, "x0"), vars.add_bnd(dom, "x1")
    path = (lambda name: os.path.join(out_dir, name)) if out_dir else (lambda name: "")
    T = vars.add_scldom_unkn(dom, 0.0, path("temp.csv"))
    c = vars.add_scldom_cnst(dom, 1.0)
    k = vars.add_scldom_cnst(dom, 1.0)
    zero = vars.add_scldom_cnst(dom, 0.0)
    if src_of_t:
        q = vars.add_scldom_func(dom, lambda t, pos, u: np.full(pos.shape[0], 1.0 + 10.0 * t), (), path("src.csv"))
    else:
        q = vars.add_scldom_cnst(dom, 1.0)
    t0 = vars.add_sclbnd_cnst(b0, 0.0, path("t_x0.csv"))
    t1 = vars.add_sclbnd_func(b1, lambda t, pos, u: 1.0 + pos[:, 1] * t, (), path("t_x1.csv"))
    phys = vfvm.TransientHeat.new()
    phys.add_heat_dom(dom, T, c, k, zero, q)
    phys.add_temp_bnd(b0, t0)
    phys.add_temp_bnd(b1, t1)
    for name in ("y0", "y1", "z0", "z1"):
        b = vars.add_bnd(dom, name)
        phys.add_hflx_bnd(b, vars.add_sclbnd_cnst(b, 0.0))
    return vars, phys, T


@pytest.mark.gpu
def test_steady_output_matches_solution(tmp_path):
    print("[output] steady heat", flush=True)
    vars = vfvm.Variables.new_lattice(8)
    dom = vars.add_dom()
    b0, b1 = vars.add_bnd(dom, "x0"), vars.add_bnd(dom, "x1")
    T = vars.add_scldom_unkn(dom, 0.0, str(tmp_path / "temp.csv"))
    k, zero, q = vars.add_scldom_cnst(dom, 1.0), vars.add_scldom_cnst(dom, 0.0), vars.add_scldom_cnst(dom, 1.0)
    phys = vfvm.SteadyHeat.new()
    phys.add_heat_dom(dom, T, k, zero, q)
    phys.add_temp_bnd(b0, vars.add_sclbnd_cnst(b0, 0.0))
    phys.add_temp_bnd(b1, vars.add_sclbnd_cnst(b1, 1.0, str(tmp_path / "tb.csv")))
    for name in ("y0", "y1", "z0", "z1"):
        b = vars.add_bnd(dom, name)
        phys.add_hflx_bnd(b, vars.add_sclbnd_cnst(b, 0.0))
    phys.solve(vars, vfvm.LinearGmres.new_with_options(1, 1e-12, 50, 200, precond=vfvm.PRECOND_ILU0), 10, 1e-8, 1.0)
    head, d = read_csv(str(tmp_path / "temp_0000.csv"))
    assert head == "x,y,z,value" and d.shape == (vars.mesh.num_tot, 4)
    assert np.abs(d[:, :3] - vars.pos).max() <= 5e-10 and np.abs(d[:, 3] - vars.value(T)).max() <= 5e-10
    _, db = read_csv(str(tmp_path / "tb_0000.csv"))
    nodes = vars.mesh.num_pts + vars.mesh.bnd_nodes(1)
    assert np.abs(db[:, :3] - vars.pos[nodes]).max() <= 5e-10 and np.all(db[:, 3] == 1.0)
    with pytest.raises(vfvm.VFVChemError):
        vars.add_scldom_cnst(dom, 1.0, str(tmp_path / "x.vtk"))


@pytest.mark.gpu
@pytest.mark.parametrize("src_of_t", [False, True])
def test_transient_output_in_chunks_equals_plain_run(tmp_path, src_of_t):
    print(f"[output] transient heat, time-dependent source {src_of_t}", flush=True)
    lin = vfvm.LinearGmres.new_with_options(1, 1e-12, 50, 200, precond=vfvm.PRECOND_ILU0)
    v_plain, p_plain, T_plain = heat_problem(7, None, src_of_t)
    p_plain.solve(v_plain, lin, 1e-2, 5, 1, 10, 1e-9, 1.0)
    v_out, p_out, T_out = heat_problem(7, str(tmp_path), src_of_t)
    p_out.solve(v_out, lin, 1e-2, 5, 2, 10, 1e-9, 1.0)
    # written at step 0 and every 2 steps: files 0, 1, 2
    for k in range(3):
        assert os.path.exists(tmp_path / f"temp_{k:04d}.csv") and os.path.exists(tmp_path / f"t_x1_{k:04d}.csv")
    assert not os.path.exists(tmp_path / "temp_0003.csv")
    Tp, To = v_plain.value(T_plain), v_out.value(T_out)
    assert np.abs(To - Tp).max() <= 1e-12 * max(1.0, np.abs(Tp).max())
    # file 2 holds the state after 4 steps: between the initial and the final state
    _, d2 = read_csv(str(tmp_path / "temp_0002.csv"))
    assert np.all(np.isfinite(d2[:, 3])) and np.abs(d2[:, 3]).max() > 0.0
    # the boundary function field is written at the output time t = 4 dt
    _, b2 = read_csv(str(tmp_path / "t_x1_0002.csv"))
    assert np.abs(b2[:, 3] - (1.0 + b2[:, 1] * 0.04)).max() <= 1e-9
