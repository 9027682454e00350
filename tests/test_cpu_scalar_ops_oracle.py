"""CPU pins of the oracle's OpSclDomAdvection / OpSclBndAdvection, OpSclDomTimePrevious and
OpSclDomTimeUnity restatements against the operators they reduce to:
  advection with weight = density   == advection-unity (m_f / rho_f = 1 exactly)
  advection with weight = 2 density == advection-unity on the doubled mass flux
  time-unity                         == time with weight 1
  time-previous with prev = current  == time."""
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
