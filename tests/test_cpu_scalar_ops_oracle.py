"""CPU pins of the oracle's OpSclDomAdvection / OpSclBndAdvection, OpSclDomTimePrevious and
OpSclDomTimeUnity restatements against the operators they reduce to:
  advection with weight = density   == advection-unity (m_f / rho_f = 1 exactly)
  advection with weight = 2 density == advection-unity on the doubled mass flux
  time-unity                         == time with weight 1
  time-previous with prev = current  == time
and of OpSclDomProduction / OpSclDomTime against closed forms: the production of a linear
field integrates it exactly over the cube (first-moment identity of the dual), and time +
production on a uniform state is the implicit-Euler map of m dc/dt = s c."""
import numpy as np

import vfvm
import vfvm_oracle as O
from problems import heat_flow_c5like


def _system(p, g, ops, dt=0.0, prev=None):
    scl, sb, flux, bnd_nid, _ = p.oracle_inputs(g)
    ops = [O.Op(*o) for o in ops]
    return O.assemble(g, ops, scl, sb, flux, bnd_nid, dt=dt, value_prev=prev, flow_den=p.flow_den), scl, flux


def test_advection_reduces_to_unity():
    mesh = vfvm.Mesh.lattice(6, 0, 0.3)
    p = heat_flow_c5like(mesh)
    g = O.dual(mesh.get("pos"), mesh.get("tets"), mesh.num_pts)
    T, k, f = p.unknown, 2, 0
    den = p.flw[0][2]
    adv = [(vfvm.OP_SCL_DOM_ADVECTION, 0, T, den, k, f)] + [(vfvm.OP_SCL_BND_ADVECTION, b, T, den, k, f) for b in range(2, 6)]
    uni = [(vfvm.OP_SCL_DOM_ADVECTION_UNITY, 0, T, k, f)] + [(vfvm.OP_SCL_BND_ADVECTION_UNITY, b, T, k, f) for b in range(2, 6)]
    (r1, c1, v1, b1), _, _ = _system(p, g, adv)
    (r2, c2, v2, b2), _, _ = _system(p, g, uni)
    assert np.array_equal(r1, r2) and np.array_equal(c1, c2) and np.array_equal(v1, v2)
    # weight 2 rho: the unity operator on twice the mass flux
    scl, sb, flux, bnd_nid, _ = p.oracle_inputs(g)
    scl2 = dict(scl)
    w2 = max(scl) + 1
    scl2[w2] = (2.0 * scl[den][0], 2.0 * scl[den][1])
    ra, ca, va, _ = O.assemble(g, [O.Op(vfvm.OP_SCL_DOM_ADVECTION, 0, T, w2, k, f)], scl2, sb, flux, bnd_nid, flow_den=p.flow_den)
    rb, cb, vb, _ = O.assemble(g, [O.Op(vfvm.OP_SCL_DOM_ADVECTION_UNITY, 0, T, k, f)], scl, sb, {f: 2.0 * flux[f]}, bnd_nid)
    assert np.max(np.abs(va - vb)) <= 1e-14 * np.max(np.abs(vb))
    assert not np.allclose(va, v2[:va.size])


def test_time_unity_and_previous_reduce_to_time():
    mesh = vfvm.Mesh.lattice(5, 0, 0.3)
    p = heat_flow_c5like(mesh)
    g = O.dual(mesh.get("pos"), mesh.get("tets"), mesh.num_pts)
    T, rcp = p.unknown, p.rcp_id
    scl, sb, flux, bnd_nid, _ = p.oracle_inputs(g)
    one = max(scl) + 1
    scl1 = dict(scl)
    scl1[one] = (np.ones(p.N), np.zeros((3, p.N)))
    prev = {T: np.linspace(0.0, 1.0, p.N), rcp: scl[rcp][0], one: np.ones(p.N)}
    dt = 0.01
    a = O.assemble(g, [O.Op(vfvm.OP_SCL_DOM_TIME_UNITY, 0, T), O.Op(vfvm.OP_SCL_BND_TIME_UNITY, 3, T)], scl1, sb, flux, bnd_nid, dt, prev)
    b = O.assemble(g, [O.Op(vfvm.OP_SCL_DOM_TIME, 0, T, one), O.Op(vfvm.OP_SCL_BND_TIME, 3, T, one)], scl1, sb, flux, bnd_nid, dt, prev)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    a = O.assemble(g, [O.Op(vfvm.OP_SCL_DOM_TIME_PREVIOUS, 0, T, rcp)], scl, sb, flux, bnd_nid, dt, prev)
    b = O.assemble(g, [O.Op(vfvm.OP_SCL_DOM_TIME, 0, T, rcp)], scl, sb, flux, bnd_nid, dt, prev)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def _production_problem(mesh, c_values):
    from problems import Problem
    p = Problem(mesh)
    c = p.scldom(vfvm.FIELD_UNKNOWN, c_values)
    one = p.scldom(vfvm.FIELD_CONSTANT, 1.0)
    return p, c, one


def test_production_integrates_linear_fields_exactly():
    """OpSclDomProduction / OpSclBndProduction with a unit coefficient on a linear field c:
    the operator's action A c - b on vertex v is V_v c(x_v + cskew_v), the field at the
    control volume's centroid, so its sum over all vertices is the integral of c over the
    cube, c(1/2, 1/2, 1/2) (the first-moment identity of the dual; dropping the skew term or
    flipping its sign breaks it on the jittered lattice)."""
    mesh = vfvm.Mesh.lattice(6, 2, 0.3)
    pos = mesh.get("pos")
    c_lin = 0.3 + 1.1 * pos[:, 0] - 0.7 * pos[:, 1] + 0.4 * pos[:, 2]
    p, c, one = _production_problem(mesh, c_lin)
    p.ops = [(vfvm.OP_SCL_DOM_PRODUCTION, 0, c, one, -1)] + [(vfvm.OP_SCL_BND_PRODUCTION, b, c, one, -1) for b in range(6)]
    rp, col, val, b, g = p.oracle_system()
    action = np.add.reduceat(val * c_lin[col], rp[:-1]) - b
    exact = 0.3 + 1.1 * 0.5 - 0.7 * 0.5 + 0.4 * 0.5
    assert abs(action.sum() - exact) <= 1e-13
    # without the skew correction the sum is off on the jittered lattice
    assert abs((g["vol"] * c_lin).sum() - exact) > 1e-6


def test_time_and_production_give_backward_euler_closed_form():
    """m dc/dt = s c by OpSclDomTime (weight m) and OpSclDomProduction (coefficient s) on a
    uniform current state: each row is (-m V / dt + s V) c = -m c_prev V / dt, so the solve
    gives the implicit-Euler map c = c_prev / (1 - s dt / m) at every vertex."""
    from problems import Problem
    mesh = vfvm.Mesh.lattice(5, 1, 0.3)
    p = Problem(mesh)
    c = p.scldom(vfvm.FIELD_UNKNOWN, 0.25)
    m = p.scldom(vfvm.FIELD_CONSTANT, 2.0)
    s = p.scldom(vfvm.FIELD_CONSTANT, 0.5)
    dt = 0.1
    prev = np.random.default_rng(5).uniform(0.5, 1.5, p.N)
    p.ops = [(vfvm.OP_SCL_DOM_TIME, 0, c, m, -1), (vfvm.OP_SCL_DOM_PRODUCTION, 0, c, s, -1)]
    p.ops += [op for b in range(6) for op in ((vfvm.OP_SCL_BND_TIME, b, c, m, -1), (vfvm.OP_SCL_BND_PRODUCTION, b, c, s, -1))]
    rp, col, val, b, g = p.oracle_system(value_prev={c: prev, m: np.full(p.N, 2.0)}, dt=dt)
    assert np.array_equal(np.diff(rp), np.ones(p.N, np.int64)) and np.array_equal(col, np.arange(p.N))
    x = b / val
    assert np.max(np.abs(x - prev / (1.0 - 0.5 * dt / 2.0))) <= 1e-14
