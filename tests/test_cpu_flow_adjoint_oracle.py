"""CPU tests: the oracle's Jacobian-transpose of the coupled flow system (O.flow_jt) against
central finite differences of its own residual (O.coupled_residual), which is the oracle's
bit-exact restatement of the assembly.

For the Picard fixed point R(x) = A(c(x)) x - b(c(x), x_prev), the identity
    d . (J(x)^T lam) = lam . (R(x + h d) - R(x - h d)) / (2 h) + O(h^2)
in random directions d, lam fixes every term of J^T = A^T + E^T: a dropped, mis-signed or
mis-indexed contribution of the Rhie-Chow flux, rc, the stabiliser, the advection
coefficients, the velocity / body-force gradients or the Boussinesq closure breaks it (the
planted mistakes below show the test notices them).  Bar: 1e-7 relative, at h = 1e-6 on a
state of order 1 (the central difference's truncation and rounding are ~1e-10).
"""
import numpy as np
import pytest

import vfvm
import vfvm_oracle as O
from problems import HeatFlowProblem, flow_general


def _problem(n, transient, species, seed=0):
    mesh = vfvm.Mesh.lattice(n, seed, 0.3)
    mesh.build_grad_coeff()
    pos = mesh.get("pos")
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    fp = flow_general(mesh, seed)
    hp = HeatFlowProblem(fp, rcp=2.0 + 0.5 * y, k=0.05 * (1.0 + 0.3 * z), q=0.3 * np.cos(3.0 * y), gamma=5.0,
                         T0=0.4 + 0.2 * z, transient=transient, species=species, D=0.02 * (1.0 + x * y), C0=0.5)
    hp.host_ids()
    g = O.dual(pos, mesh.get("tets"), mesh.num_pts)
    gc = O.grad_coeff(g)
    return mesh, hp, g, gc


def _state(hp, rng, vscale=1.0):
    """A state of order 1 with every term active (random velocities everywhere, so that no
    face flux sits exactly at a kink of the upwinding); vscale < 1: slow flow, so that every
    face Peclet number, also on the dual faces of negative area, stays below 10."""
    N, K = hp.N, hp.nscl
    vel = vscale * (hp.fp.vel + 0.3 * rng.standard_normal((3, N)))
    pres = hp.fp.pres + 0.2 * rng.standard_normal(N)
    S = 0.5 + 0.3 * rng.standard_normal(K * N)
    return np.concatenate([vel.ravel(), pres, S])


def _fd_check(res, jt, x, xp, rng, trials=3, h=1e-6):
    worst = 0.0
    for _ in range(trials):
        lam = rng.standard_normal(x.size)
        d = rng.standard_normal(x.size)
        y = jt(x, xp, lam)
        fd = lam @ (res(x + h * d, xp)[0] - res(x - h * d, xp)[0]) / (2.0 * h)
        ad = d @ y
        worst = max(worst, abs(ad - fd) / max(abs(fd), 1e-30))
    return worst


# (n, vscale): the 5^3 lattice (seed 0) has two dual faces of negative area next to the
# boundary (cond < 0 there); at vscale 0.01 their Peclet numbers are below 10
CASES = [(4, False, False, 1.0), (4, True, True, 1.0), (5, True, True, 0.01)]


@pytest.mark.parametrize("n,transient,species,vscale", CASES)
def test_flow_jt_matches_finite_differences(n, transient, species, vscale):
    mesh, hp, g, gc = _problem(n, transient, species)
    dt = 0.05 if transient else 0.0
    rng = np.random.default_rng(1)
    x = _state(hp, rng, vscale)
    xp = _state(hp, rng, vscale) if transient else None
    res = hp.oracle_residual(g, gc, dt)
    jt = hp.oracle_jt(g, gc, dt)
    err = _fd_check(res, jt, x, xp, rng)
    assert err < 1e-7, err


def test_flow_jt_residual_is_the_assembled_system():
    # R = A x - b with the A, b of heatflow_solve's own system at x
    mesh, hp, g, gc = _problem(4, False, False)
    rng = np.random.default_rng(2)
    x = _state(hp, rng)
    R, A, f = hp.oracle_residual(g, gc)(x)
    N = hp.N
    r, c, v, b = O.assemble_flow(g, gc, f, hp.fp.bnd_vel, [mesh.bnd_nodes(k) for k in range(6)], hp.fp.pref)
    rh, ch, vh, bh = hp.oracle_heat_rows(g)(x[4 * N:], x[4 * N:], f)
    full = np.zeros(5 * N)
    np.add.at(full, np.asarray(r), np.asarray(v) * x[np.asarray(c)])
    np.add.at(full, 4 * N + np.asarray(rh), np.asarray(vh) * x[4 * N + np.asarray(ch)])
    assert np.allclose(R, full - np.concatenate([b, bh]), rtol=0, atol=1e-12 * np.abs(full).max())


@pytest.mark.parametrize("mistake", ["drop_rc", "flip_advection_derivative", "drop_body", "sign_of_flux"])
def test_flow_jt_planted_mistakes_are_caught(monkeypatch, mistake):
    n, vscale = (5, 0.01) if mistake == "sign_of_flux" else (4, 1.0)
    mesh, hp, g, gc = _problem(n, False, False)
    rng = np.random.default_rng(3)
    x = _state(hp, rng, vscale)
    res = hp.oracle_residual(g, gc)
    if mistake == "sign_of_flux":
        # the power law's slope signed by F instead of Pe = F / cond: wrong only on the faces of
        # negative area
        def dphi(F, cond):
            with np.errstate(divide="ignore", invalid="ignore"):
                a4 = np.maximum(1.0 - 0.1 * np.abs(np.where(cond != 0.0, F / cond, 0.0)), 0.0) ** 4
            return np.where(F >= 0.0, -0.5 * a4, 0.5 * a4 - 1.0)
        monkeypatch.setattr(O, "_dphi_adv", dphi)
        jt = hp.oracle_jt(g, gc)
        # the faces of negative area carry a small share of a random direction: a full-size
        # check over the rows of their two vertices
        return _check_rows_catch(res, jt, x, g)
    if mistake == "drop_rc":
        # a_P's derivative left out: rc treated as fixed
        monkeypatch.setattr(O, "_dh_ap", lambda F, cond: np.zeros_like(F))
        jt = hp.oracle_jt(g, gc)
    elif mistake == "flip_advection_derivative":
        right = O._dphi_adv
        monkeypatch.setattr(O, "_dphi_adv", lambda F, cond: -right(F, cond))
        jt = hp.oracle_jt(g, gc)
    else:
        base = hp.oracle_jt(g, gc)
        hp.dfce = lambda: np.zeros((3, hp.N))
        jt = base
    assert _fd_check(res, jt, x, None, rng, trials=2) > 1e-3


def _check_rows_catch(res, jt, x, g):
    # the temperature rows of the vertices of the negative-area faces against the velocity
    # columns of the same vertices (the face flux's own dependence)
    N = g["N"]
    neg = np.nonzero(g["area"] < 0.0)[0]
    assert neg.size > 0
    verts = np.unique(np.concatenate([g["rows"][neg], g["ngh"][neg]])).astype(np.int64)
    worst = 0.0
    for v in verts:
        row = 4 * N + int(v)
        lam = np.zeros(x.size)
        lam[row] = 1.0
        y = jt(x, None, lam)
        for col in [c * N + int(u) for c in range(3) for u in verts]:
            h = 1e-6
            e = np.zeros(x.size)
            e[col] = h
            fd = (res(x + e)[0][row] - res(x - e)[0][row]) / (2 * h)
            worst = max(worst, abs(y[col] - fd) / max(abs(fd), 1e-3 * np.abs(y).max()))
    assert worst > 1e-3, worst
