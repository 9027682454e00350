"""CPU tests: the oracle's coupled heat-flow driver (heatflow_solve: the SteadyHeatFlow
composition with a Boussinesq body force evaluated before every assembly) held to closed
forms and to the decoupled solvers.

Pins (each fails for a dropped coupling, a wrong shift of the temperature block, the flux of
another iterate or a wrong buoyancy sign):
  conduction:   no flow and no buoyancy give T = t_hot + (t_cold - t_hot) x exactly (the
                diffusion operator is exact for linear fields), velocity zero; likewise the
                concentration of a second scalar block (C = 1 - x).
  decoupled:    gamma = 0 with a moving lid: the velocity and pressure equal picard_flow's,
                and T equals a direct solve of the scalar system with the converged flow's
                flux (the temperature rows read the flux of the same iterate).
  buoyancy:     a heated x = 0 wall with gamma > 0 drives upward flow next to it and
                downward flow next to the cooled wall.
  residual:     scaled_residual_scalar equals its closed form on a two-entry case.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import vfvm
import vfvm_oracle as O
from problems import HeatFlowProblem, lid_cavity


@pytest.fixture(scope="module")
def mesh5():
    return vfvm.Mesh.lattice(5, 0, 0.3)


@pytest.fixture(scope="module")
def g5(mesh5):
    return O.dual(mesh5.get("pos"), mesh5.get("tets"), mesh5.num_pts)


@pytest.fixture(scope="module")
def gc5(g5):
    return O.grad_coeff(g5)


def _solve(hp, g, gc, max_iter=60, tol=1e-9, damp=1.0):
    hp.host_ids()
    fp = hp.fp
    bnd_nid = [hp.mesh.bnd_nodes(b) for b in range(6)]
    return O.heatflow_solve(g, gc, fp.vel, fp.pres, hp.S0(), fp.den, fp.visc, hp.fce_of, fp.bnd_vel, bnd_nid, fp.pref,
                            hp.oracle_heat_rows(g), hp.oracle_t_dirichlet(), max_iter, tol, damp)


def test_scaled_residual_scalar_closed_form():
    xo = np.array([1.0, -2.0])
    xn = np.array([1.5, -2.0])
    assert O.scaled_residual_scalar(xo, xn) == pytest.approx(np.sqrt(0.25 / 2) / 2.0, rel=1e-15)
    assert np.isnan(O.scaled_residual_scalar(xo, np.array([np.nan, 0.0])))


@pytest.mark.parametrize("species", [False, True])
def test_conduction_limit_linear_profile(mesh5, g5, gc5, species):
    fp = lid_cavity(mesh5, rho=1.0, mu=0.1, lid=0.0)
    hp = HeatFlowProblem(fp, rcp=1.0, k=0.2, q=0.0, gamma=0.0, species=species, D=0.3)
    vel, pres, S, its = _solve(hp, g5, gc5)
    x = mesh5.get("pos")[:, 0]
    N = mesh5.num_tot
    assert S.size == (2 if species else 1) * N
    assert np.max(np.abs(vel)) <= 1e-12
    assert np.max(np.abs(S[:N] - (1.0 - x))) <= 1e-11
    if species:
        assert np.max(np.abs(S[N:] - (1.0 - x))) <= 1e-11


def test_decoupled_matches_flow_and_scalar_solves(mesh5, g5, gc5):
    fp = lid_cavity(mesh5, rho=1.0, mu=0.05, lid=1.0)
    hp = HeatFlowProblem(fp, rcp=2.0, k=0.05, q=0.3, gamma=0.0)
    vel, pres, T, its = _solve(hp, g5, gc5, tol=1e-10, damp=0.8)
    bnd_nid = [mesh5.bnd_nodes(b) for b in range(6)]
    ovel, opres, _, _ = O.picard_flow(g5, gc5, fp.vel, fp.pres, fp.den, fp.visc, np.zeros((3, fp.N)), fp.bnd_vel, bnd_nid,
                                      fp.pref, 80, 1e-12, 0.8)
    assert np.max(np.abs(vel - ovel)) <= 1e-8 * np.max(np.abs(ovel))
    assert np.max(np.abs(pres - opres)) <= 1e-8 * np.max(np.abs(opres))
    # the temperature rows with the converged flow's flux, solved directly
    f = O.flow_fields(g5, gc5, vel, pres, fp.den, fp.visc, hp.fce_of(T), fce_const=False)
    r, c, v, b = hp.oracle_heat_rows(g5)(T, T, f)
    rp, col, val = O.coo_to_csr(r, c, v, fp.N)
    T_direct = spla.spsolve(sp.csr_matrix((val, col, rp), shape=(fp.N, fp.N)).tocsc(), b)
    assert np.max(np.abs(T - T_direct)) <= 1e-8 * np.max(np.abs(T_direct))


def test_buoyancy_drives_convection(mesh5, g5, gc5):
    fp = lid_cavity(mesh5, rho=1.0, mu=0.1, lid=0.0)
    hp = HeatFlowProblem(fp, rcp=1.0, k=0.1, q=0.0, gamma=20.0)
    vel, pres, T, its = _solve(hp, g5, gc5, tol=1e-9, damp=0.8)
    pos = mesh5.get("pos")
    assert T.size == mesh5.num_tot
    NP = mesh5.num_pts
    x = pos[:NP, 0]
    w = vel[2][:NP]
    assert np.mean(w[x < 0.4]) > 0.0 and np.mean(w[x > 0.6]) < 0.0
    # the wall temperatures hold and T stays within them (maximum principle, no source)
    assert np.min(T) >= -1e-9 and np.max(T) <= 1.0 + 1e-9
    assert its >= 2
