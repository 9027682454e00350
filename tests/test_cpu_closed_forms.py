"""CPU tests: the oracle's advection, Neumann and flow operators held to closed forms that do
not restate them.

  advection (power-law upwind, opscl_dom_advunity.rs:46-115): the 1D convection-diffusion
      profile c = (exp(Pe x) - 1) / (exp(Pe) - 1) under uniform flow, converging at second
      order on the jittered lattice; the same mistake planted in the oracle (the upwind term
      taken on the downstream side) leaves an O(1) error that does not fall.
  Neumann (opscl_bnd_neu.rs:36-87): the linear profile T = -q x / k to round-off; the
      boundary right-hand sides of a linear flux summing to its surface integral; a regular
      tetrahedron whose fluxes fall back to the vertex's own value on neighbours of other
      boundaries (opscl_bnd_neu.rs:73-80), in closed form.
  flow (SteadyFlow operator set): plane Poiseuille flow (examples/flow/flow_analytic.rs:10)
      and the asymptotic suction profile (convective-diffusive balance of the momentum
      advection) converging under refinement; the planted upwind mistake breaks the latter.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import vfvm
import vfvm_oracle as O
from problems import (advdiff_1d, neumann_linear_profile, neumann_surface_integral, neumann_tetra, poiseuille,
                      suction_profile)


def _oracle_solve(p):
    rp, col, val, b, _ = p.oracle_system()
    A = sp.csr_matrix((val, col, rp), shape=(p.N, p.N)).tocsc()
    return spla.spsolve(A, b)


def _advdiff_errors(sizes, Pe):
    out = []
    for n in sizes:
        p = advdiff_1d(vfvm.Mesh.lattice(n, 0, 0.3), Pe)
        out.append(np.max(np.abs(_oracle_solve(p) - p.exact)))
    return out


def test_advection_diffusion_closed_form_converges():
    e8, e16, e24 = _advdiff_errors((8, 16, 24), 5.0)
    # second order: h halves from 1/7 to 1/15 (x 2.14), then 1/15 -> 1/23 (x 1.53)
    assert e8 < 2.5e-3 and e16 < 7e-4 and e24 < 3e-4, (e8, e16, e24)
    assert e8 / e16 > 3.0 and e16 / e24 > 1.8, (e8, e16, e24)


def test_advection_planted_upwind_mistake_is_caught(monkeypatch):
    right = O._adv_coeff
    monkeypatch.setattr(O, "_adv_coeff", lambda cond, F: right(cond, -F))  # upwind on the downstream side
    e8, e16 = _advdiff_errors((8, 16), 5.0)
    assert e8 > 0.5 and e16 > 0.5, (e8, e16)


def test_neumann_linear_profile_exact():
    p = neumann_linear_profile(7)
    T = _oracle_solve(p)
    assert np.max(np.abs(T - p.exact)) <= 1e-12


def test_neumann_surface_integral_of_linear_flux():
    coef = (0.3, 1.1, -0.7, 2.0)
    p = neumann_surface_integral(7, coef)
    _, _, _, b, _ = p.oracle_system()
    assert abs(b.sum() - p.exact_integral) <= 1e-12 * abs(p.exact_integral)
    # points carry no Neumann term
    assert np.all(b[:p.mesh.num_pts] == 0.0)


@pytest.mark.parametrize("groups", [((0, 1), (2,), (3,)), ((0, 1, 2, 3),), ((0,), (1,), (2,), (3,))])
def test_neumann_tetrahedron_fallback_closed_form(groups):
    p = neumann_tetra(groups=groups)
    _, _, _, b, _ = p.oracle_system()
    assert np.max(np.abs(b - p.exact_b)) <= 1e-14 * np.max(np.abs(p.exact_b))


def _flow_oracle(p):
    g = O.dual(p.mesh.get("pos"), p.mesh.get("tets"), p.mesh.num_pts)
    gc = O.grad_coeff(g)
    return O.picard_flow(g, gc, p.vel, p.pres, p.den, p.visc, p.fce, p.bnd_vel,
                         [p.mesh.bnd_nodes(k) for k in range(6)], p.pref, 40, 1e-10, 1.0)


def pressure_gradient(mesh, pres):
    """Least-squares linear fit of the pressure over the interior points."""
    NP = mesh.num_pts
    pos = mesh.get("pos")[:NP]
    X = np.c_[np.ones(NP), pos]
    return np.linalg.lstsq(X, pres[:NP], rcond=None)[0][1:]


def test_poiseuille_converges():
    errs, dpdx = [], []
    for n in (5, 7):
        p = poiseuille(vfvm.Mesh.lattice(n, 0, 0.3), U=1.0, mu=1.0)
        vel, pres, its, _ = _flow_oracle(p)
        errs.append(np.max(np.abs(vel - p.uex)))
        dpdx.append(pressure_gradient(p.mesh, pres))
    # velocity error falls; dp/dx approaches -8 mu U / H^2 = -8, dp/dy, dp/dz stay small
    assert errs[1] < 0.5 * errs[0] and errs[1] < 0.05, errs
    assert abs(dpdx[1][0] + 8.0) < 0.5 * abs(dpdx[0][0] + 8.0) and abs(dpdx[1][0] + 8.0) < 0.5, dpdx
    assert np.all(np.abs(dpdx[1][1:]) < 0.2), dpdx


def _suction_error(n, **kw):
    p = suction_profile(vfvm.Mesh.lattice(n, 0, 0.3), **kw)
    vel, _, _, _ = _flow_oracle(p)
    NP = p.mesh.num_pts
    return np.max(np.abs(vel[0][:NP] - p.uex[0][:NP]))


def test_suction_profile_converges_and_planted_mistake_is_caught(monkeypatch):
    e5, e7 = _suction_error(5), _suction_error(7)
    assert e7 < 0.75 * e5 and e7 < 0.04, (e5, e7)
    right = O._adv_coeff
    monkeypatch.setattr(O, "_adv_coeff", lambda cond, F: right(cond, -F))
    w7 = _suction_error(7)
    assert w7 > 3.0 * e7, (e7, w7)
