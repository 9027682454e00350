"""CPU tests of the multi-domain meshes and of the oracle's interface operators
(opscl_itf_*.rs), held to closed forms that do not restate them.

  join: the box lattice with the whole index range is the cube lattice; the two copies of an
      interface node coincide; the control volumes of the joined domains tile the cube; the two
      halves of an interface node's control volume close (their face vectors cancel).
  continuity + diffusion: slabs of different conductivity between T = 0 and T = 1 give the
      piecewise-linear profile with continuous flux to round-off; the same with side a's faces
      dropped from the interface diffusion (a planted mistake) does not.
  source / production: the source's right-hand sides over all rows integrate q over the cube
      (both halves of every interface node counted once); T = x + q x (1 - x) / (2 k) across
      the interface converges at second order.
  time (+ production): an implicit-Euler step of m dc/dt = s c on a uniform state integrates the
      ODE of every control volume, an interface node's combined volume with m and s weighted
      by its two halves.
  advection: the 1D convection-diffusion profile through an interface crossed by the flow
      converges at second order (advection-unity and weighted forms).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import vfvm
import vfvm_oracle as O
from problems import X0, X1, heat_c1, itf_advdiff, itf_heat, itf_transient, slab_domains


def _solve(p, **kw):
    rp, col, val, b, starts = p.oracle_system(**kw)
    A = sp.csr_matrix((val, col, rp), shape=(p.N, p.N)).tocsc()
    return spla.spsolve(A, b), starts


def _piecewise_linear(p, ks):
    """Exact T of 1D conduction through slabs of conductivity ks between T(0) = 0, T(1) = 1."""
    x = p.pos[:, 0]
    cuts = [p.parts[d].get("pos")[:, 0].min() for d in range(len(p.parts))] + [1.0]
    res = np.array([(cuts[d + 1] - cuts[d]) / ks[d] for d in range(len(ks))])
    flux = 1.0 / res.sum()
    T = np.empty(p.N)
    t0 = 0.0
    for d in range(len(ks)):
        s = p.dom_slice(d)
        T[s] = t0 + flux / ks[d] * (x[s] - cuts[d])
        t0 += flux * res[d]
    return T


def test_box_lattice_of_the_whole_range_is_the_cube_lattice():
    a = vfvm.Mesh.lattice(7, 3, 0.3)
    b = vfvm.Mesh.lattice_box(7, (0, 0, 0), (6, 6, 6), 3, 0.3)
    for k in ("pos", "tets", "bnd_ptr", "bnd_nid", "row_ptr", "ngh", "area", "dist", "vol", "cskew", "bsf_area"):
        assert np.array_equal(a.get(k), b.get(k)), k


def test_join_layout_and_closed_halves():
    parts, itfs = slab_domains(9, [3, 6])
    m = vfvm.Mesh.join(parts, itfs)
    assert m.num_dom == 3 and m.num_itf == 2 and m.num_pts == 0
    dp = m.get("dom_ptr")
    assert list(dp) == list(np.cumsum([0] + [q.num_tot for q in parts]))
    # every domain's boundaries but the interface sides, in domain order (6 + 6 + 6 - 4)
    assert m.num_bnd == 14
    pos = m.get("pos")
    va, vb = m.get("itf_va"), m.get("itf_vb")
    assert len(va) == 2 * 81
    assert np.array_equal(pos[va], pos[vb])
    assert list(m.get("itf_dom")) == [0, 1, 1, 2]
    # side a keeps its group order (domain 0's x1 face, domain 1's x1 face)
    assert np.array_equal(va[:81] - dp[0], parts[0].num_pts + parts[0].bnd_nodes(X1))
    assert np.array_equal(vb[81:] - dp[2], parts[2].num_pts + parts[2].bnd_nodes(X0))
    vol = m.get("vol")
    assert abs(vol.sum() - 1.0) < 1e-14
    # interior interface nodes: the face vectors of both halves cancel (one closed cell)
    rp, area, norm = m.get("row_ptr"), m.get("area"), m.get("norm")
    y, z = pos[va, 1], pos[va, 2]
    inner = (y > 0) & (y < 1) & (z > 0) & (z < 1)
    for q in np.nonzero(inner)[0]:
        tot = np.zeros(3)
        for v in (va[q], vb[q]):
            e = np.arange(rp[v], rp[v + 1])
            tot += (area[e] * norm[:, e]).sum(axis=1)
        assert np.abs(tot).max() < 1e-14, (q, tot)


@pytest.mark.parametrize("splits,ks", [([4], (1.0, 3.0)), ([3, 6], (2.0, 0.5, 5.0))])
def test_piecewise_linear_conduction_exact(splits, ks):
    p = itf_heat(9, splits=splits, k=ks)
    T, _ = _solve(p)
    assert np.max(np.abs(T - _piecewise_linear(p, ks))) < 1e-12


def test_planted_interface_mistake_is_caught(monkeypatch):
    """Side a's faces dropped from the interface diffusion: the flux balance of the combined
    control volume loses one half."""
    right = O.assemble_interface

    def wrong(doms, starts, itf, ops, fields, *a, **kw):
        r, c, v, br, bv = right(doms, starts, itf, ops, fields, *a, **kw)
        da = itf[0]
        lo, hi = starts[da], starts[da + 1]
        rows_b = set((starts[itf[1]] + np.asarray(itf[3])).tolist())
        keep = ~(np.isin(r, list(rows_b)) & (c >= lo) & (c < hi))
        return r[keep], c[keep], v[keep], br, bv

    monkeypatch.setattr(O, "assemble_interface", wrong)
    p = itf_heat(9, k=(1.0, 3.0))
    T, _ = _solve(p)
    assert np.max(np.abs(T - _piecewise_linear(p, (1.0, 3.0)))) > 1e-2


def test_source_integrates_over_the_cube():
    q = 2.5
    p = itf_heat(9, splits=[3, 6], k=(1.0, 1.0, 1.0), q=q)
    # only the source operators: every row's b is -q times its control volume(s)
    p.ops = [o for o in p.ops if o[0] in (vfvm.OP_SCL_DOM_SOURCE, vfvm.OP_SCL_BND_SOURCE, vfvm.OP_SCL_ITF_SOURCE,
                                          vfvm.OP_SCL_DOM_DIFFUSION, vfvm.OP_SCL_BND_DIFFUSION, vfvm.OP_SCL_ITF_DIFFUSION)]
    T, hs = p.unknown, p.fields["src"]
    p.ops += [(vfvm.OP_SCL_BND_SOURCE, p.jbnd[(0, X0)], T, hs, -1, -1), (vfvm.OP_SCL_BND_SOURCE, p.jbnd[(2, X1)], T, hs, -1, -1)]
    _, _, _, b, _ = p.oracle_system()
    assert abs(b.sum() + q) < 1e-13


def test_source_closed_form_converges_across_interfaces():
    """T = x + q x (1 - x) / (2 k) (heat_analytic.rs:9) through an interface: the error falls
    with h as on the uncut cube, and stays within 1.3x of the uncut cube's error."""
    errs, cube = [], []
    for n in (9, 17):
        p = itf_heat(n, splits=[(n - 1) // 2], k=(2.0, 2.0), q=3.0)
        T, _ = _solve(p)
        x = p.pos[:, 0]
        errs.append(np.max(np.abs(T - (x + 3.0 * x * (1 - x) / (2 * 2.0)))))
        c = heat_c1(vfvm.Mesh.lattice(n), k=2.0, q=3.0)
        rp, col, val, b, _ = c.oracle_system()
        Tc = spla.spsolve(sp.csr_matrix((val, col, rp), shape=(c.N, c.N)).tocsc(), b)
        xc = c.pos[:, 0]
        cube.append(np.max(np.abs(Tc - (xc + 3.0 * xc * (1 - xc) / (2 * 2.0)))))
    assert errs[0] / errs[1] > 3.0, errs
    assert errs[0] < 1.3 * cube[0] and errs[1] < 1.3 * cube[1], (errs, cube)


@pytest.mark.parametrize("previous", [False, True])
def test_implicit_euler_of_the_combined_control_volume(previous):
    p = itf_transient(7, m=(1.0, 2.5), s=(-0.5, -2.0), dt=0.1, previous=previous)
    prev = {p.unknown: np.ones(p.N)}
    if previous:
        prev[1] = np.concatenate([np.full(q.num_tot, m) for q, m in zip(p.parts, p.m)])
    T, _ = _solve(p, dt=p.dt, value_prev=prev)
    vol = p.mesh.get("vol")
    dom = np.concatenate([np.full(q.num_tot, d) for d, q in enumerate(p.parts)])
    m_of, s_of = np.array(p.m)[dom], np.array(p.s)[dom]
    # every vertex: c = 1 / (1 - dt s / m); an interface node: its two halves' volume-weighted m, s
    exact = 1.0 / (1.0 - p.dt * s_of / m_of)
    va, vb = p.mesh.get("itf_va"), p.mesh.get("itf_vb")
    mm = m_of[va] * vol[va] + m_of[vb] * vol[vb]
    ss = s_of[va] * vol[va] + s_of[vb] * vol[vb]
    exact[va] = exact[vb] = 1.0 / (1.0 - p.dt * ss / mm)
    assert np.max(np.abs(T - exact)) < 1e-13
    # and the interface nodes differ from either slab's own value (the test sees both halves)
    assert np.min(np.abs(exact[vb] - 1.0 / (1.0 - p.dt * p.s[1] / p.m[1]))) > 1e-3


@pytest.mark.parametrize("weighted", [False, True])
def test_advection_diffusion_through_an_interface_converges(weighted):
    errs = []
    for n in (9, 17):
        p = itf_advdiff(n, 5.0, weighted=weighted)
        c, _ = _solve(p)
        errs.append(np.max(np.abs(c - p.exact)))
    assert errs[0] < 4e-3 and errs[1] < 1.2e-3 and errs[0] / errs[1] > 3.0, errs


def test_advection_interface_side_matters():
    """The upwind coefficients of one side only (the other side's advection left out) break the
    profile: both sides' halves carry the flow."""
    p = itf_advdiff(9, 5.0)
    p.ops = [o for o in p.ops if not (o[0] == vfvm.OP_SCL_ITF_ADVECTION_UNITY and o[6] == 0.0)]
    c, _ = _solve(p)
    assert np.max(np.abs(c - p.exact)) > 1e-2
