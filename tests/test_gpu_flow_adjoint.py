"""The adjoint of the coupled flow system on the GPU (through the C ABI).

  residual / J^T: the device residual R(x) = A(c(x)) x - b(c(x), x_prev) and the device
      J(x)^T lam (transposed CSR + the vector-Jacobian kernels) against the oracle's
      coupled_residual / flow_jt (itself pinned by central finite differences on the CPU) on
      the most general coupled list: function-field density, viscosity, heat capacity,
      conductivity, source and diffusivity, the Boussinesq closure, heat and species,
      steady and transient.  Bars: residual 1e-12 and J^T 1e-11 relative to the largest entry
      (the summation orders of the products differ from scipy's).
  gradients: dJ/dtheta of the time-averaged hot-wall Nusselt number of a natural-convection
      cavity (heat + species + Boussinesq, implicit Euler with checkpoints and a replayed
      segment) with respect to the hot wall's Dirichlet values, and of a velocity functional
      of a steady lid-driven cavity with respect to the lid velocity, against central
      differences of the device forward pass in random directions.  Bar: 1e-5 relative.
"""
import numpy as np
import pytest

import vfvm
import vfvm_oracle as O
from problems import HeatFlowProblem, X0, Z1, flow_general, lid_cavity
from vfvm.objectives import WallFlux

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = vfvm.Context(0)
    yield c
    c.close()


def _general(mesh, transient, species):
    pos = mesh.get("pos")
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    fp = flow_general(mesh, 0)
    return HeatFlowProblem(fp, rcp=2.0 + 0.5 * y, k=0.05 * (1.0 + 0.3 * z), q=0.3 * np.cos(3.0 * y), gamma=5.0,
                           T0=0.4 + 0.2 * z, transient=transient, species=species, D=0.02 * (1.0 + x * y), C0=0.5)


def _state(hp, rng, vscale=1.0):
    N, K = hp.N, hp.nscl
    vel = vscale * (hp.fp.vel + 0.3 * rng.standard_normal((3, N)))
    pres = hp.fp.pres + 0.2 * rng.standard_normal(N)
    S = 0.5 + 0.3 * rng.standard_normal(K * N)
    return np.concatenate([vel.ravel(), pres, S])


# vscale 0.01: slow flow, every face Peclet number below 10, also on the two dual faces of
# negative area of this mesh (cond < 0 there)
@pytest.mark.parametrize("transient,species,vscale", [(False, False, 1.0), (True, True, 1.0), (True, True, 0.01)])
def test_residual_and_jt_vs_oracle(ctx, transient, species, vscale):
    print(f"[flow residual / J^T vs oracle transient={transient} species={species} vscale={vscale}]", flush=True)
    mesh = vfvm.Mesh.lattice(5, 0, 0.3)
    mesh.build_grad_coeff()
    hp = _general(mesh, transient, species)
    ops = hp.to_device(ctx)
    hp.affine_closure(ctx)
    ctx.pattern_build(ops)
    g = O.dual(mesh.get("pos"), mesh.get("tets"), mesh.num_pts)
    gc = O.grad_coeff(g)
    dt = 0.05 if transient else 0.0
    rng = np.random.default_rng(7)
    x = _state(hp, rng, vscale)
    xp = _state(hp, rng, vscale) if transient else None
    r = ctx.flow_residual(ops, x, xp, dt)
    R, _, _ = hp.oracle_residual(g, gc, dt)(x, xp)
    assert np.max(np.abs(r - R)) <= 1e-12 * np.max(np.abs(R))
    for _ in range(2):
        lam = rng.standard_normal(x.size)
        y = ctx.flow_jt_apply(ops, x, lam, xp, dt)
        yo = hp.oracle_jt(g, gc, dt)(x, xp, lam)
        assert np.max(np.abs(y - yo)) <= 1e-11 * np.max(np.abs(yo)), np.max(np.abs(y - yo)) / np.max(np.abs(yo))
    # and the device's own residual in a random direction
    lam, d = rng.standard_normal(x.size), rng.standard_normal(x.size)
    h = 1e-6
    fd = lam @ (ctx.flow_residual(ops, x + h * d, xp, dt) - ctx.flow_residual(ops, x - h * d, xp, dt)) / (2 * h)
    ad = d @ ctx.flow_jt_apply(ops, x, lam, xp, dt)
    assert abs(ad - fd) <= 1e-7 * abs(fd), (ad, fd)
    ctx.vecdom_func_affine(hp.fp.ids[4], hp.hid["T"], None, None)


LIN = dict(tol=1e-13, restart=200, max_outer=200, precond=vfvm.PRECOND_BJACOBI)
INNER = dict(tol=1e-3, restart=100, max_outer=10, precond=vfvm.PRECOND_BJACOBI)


def _cavity(mesh):
    fp = lid_cavity(mesh, rho=1.0, mu=0.1, lid=0.0)
    return HeatFlowProblem(fp, rcp=1.0, k=0.1, q=0.0, gamma=20.0, transient=True, species=True, D=0.07)


def test_nusselt_gradient_wrt_hot_wall_vs_fd(ctx):
    mesh = vfvm.Mesh.lattice(5, 0, 0.3)
    mesh.build_grad_coeff()
    hp = _cavity(mesh)
    ops = hp.to_device(ctx)
    hp.affine_closure(ctx)
    N = mesh.num_tot
    n = 6 * N
    wf = WallFlux(mesh, X0)
    seed = np.zeros(n)
    np.add.at(seed, 4 * N + wf.src, wf.coef)
    np.add.at(seed, 4 * N + wf.dst, -wf.coef)
    M, dt = 4, 0.05
    weights = np.full(M, 1.0 / M)
    t0 = hp.hid["t0"]
    cnt = len(mesh.bnd_nodes(X0))
    vel, pres = hp.fp.ids[0], hp.fp.ids[1]

    def run(theta):
        ctx.vecdom_set(vel, np.zeros((3, N)))
        ctx.scldom_set(pres, np.zeros(N))
        ctx.scldom_set(hp.hid["T"], hp.T0)
        ctx.scldom_set(hp.hid["C"], hp.C0)
        ctx.sclbnd_set(t0, theta)
        st = ctx.flow_adjoint_forward(ops, dt, M, max_iter=60, tol=1e-12, damp=1.0, checkpoint_every=2, lin=LIN, seed=seed,
                                      weights=weights)
        return st.objective

    theta = np.ones(cnt)
    print("[nusselt gradient: forward]", flush=True)
    J0 = run(theta)
    # the objective is the time-averaged hot-wall Nusselt number the host evaluates
    print("[nusselt gradient: backward]", flush=True)
    grads, st = ctx.flow_adjoint_backward(seed, [(vfvm.WRT_SCLBND, t0)], weights=weights, inner=INNER, tol=1e-12, max_iter=300,
                                          restart=30)
    grad = grads[0]
    print(f" J = {J0:.6f}, outer {st.outer_iterations}, inner steps {st.base.adjoint_iterations}, "
          f"replayed {st.base.recomputed_steps}, max outer residual {st.max_outer_residual:.2e}")
    assert st.base.recomputed_steps == 2 and st.base.replay_mismatches == 0
    assert st.max_outer_residual <= 1e-12
    assert np.isfinite(J0) and J0 > 0.0 and np.all(np.isfinite(grad))
    rng = np.random.default_rng(11)
    pos = mesh.get("pos")[mesh.num_pts + mesh.bnd_nodes(X0).astype(np.int64)]
    for trial in range(2):
        # a smooth perturbation profile and a rough one
        delta = np.cos(np.pi * pos[:, 1]) * np.sin(np.pi * pos[:, 2]) if trial == 0 else rng.standard_normal(cnt)
        eps = 1e-4
        print(f"[nusselt gradient: finite differences {trial}]", flush=True)
        fd = (run(theta + eps * delta) - run(theta - eps * delta)) / (2 * eps)
        ad = grad @ delta
        print(f" adjoint {ad:.10e} fd {fd:.10e}")
        assert abs(ad - fd) <= 1e-5 * abs(fd), (ad, fd)
    ctx.vecdom_func_affine(hp.fp.ids[4], hp.hid["T"], None, None)


def test_lid_velocity_gradient_steady_vs_fd(ctx):
    mesh = vfvm.Mesh.lattice(5, 1, 0.3)
    mesh.build_grad_coeff()
    fp = lid_cavity(mesh, rho=1.0, mu=0.05, lid=1.0)
    ops = fp.to_device(ctx)
    N = mesh.num_tot
    pos = mesh.get("pos")
    # J: a weighted sum of the velocity and pressure of the points
    rng = np.random.default_rng(5)
    seed = np.zeros(4 * N)
    NP = mesh.num_pts
    seed[:NP] = 1.0 + pos[:NP, 2]
    seed[2 * N:2 * N + NP] = rng.standard_normal(NP)
    seed[3 * N:3 * N + NP] = 0.1 * rng.standard_normal(NP)
    vid = fp.ids[6][Z1]
    lid = fp.bnd_vel[Z1][1].copy()

    def run(v):
        ctx.vecdom_set(fp.ids[0], np.zeros((3, N)))
        ctx.scldom_set(fp.ids[1], np.zeros(N))
        ctx.vecbnd_set(vid, v)
        return ctx.flow_adjoint_forward(ops, 0.0, 0, max_iter=80, tol=1e-12, damp=0.9, lin=LIN, seed=seed).objective

    print("[lid gradient: forward]", flush=True)
    J0 = run(lid)
    print("[lid gradient: backward]", flush=True)
    grads, st = ctx.flow_adjoint_backward(seed, [(vfvm.WRT_VECBND, vid)], inner=INNER, tol=1e-12, max_iter=300, restart=30)
    grad = grads[0]
    print(f" J = {J0:.6f}, outer {st.outer_iterations}, inner steps {st.base.adjoint_iterations}")
    assert grad.shape == lid.shape and np.all(np.isfinite(grad))
    for trial in range(2):
        delta = rng.standard_normal(lid.shape)
        eps = 1e-5
        print(f"[lid gradient: finite differences {trial}]", flush=True)
        fd = (run(lid + eps * delta) - run(lid - eps * delta)) / (2 * eps)
        ad = float(np.sum(grad * delta))
        print(f" adjoint {ad:.10e} fd {fd:.10e}")
        assert abs(ad - fd) <= 1e-5 * abs(fd), (ad, fd)
