"""The oracle's discrete adjoint pinned by central finite differences of its own forward
map (sparse-LU solves of the assembled systems), on small jittered lattices."""
import numpy as np
import pytest

import vfvm
import vfvm_oracle as O
from vfvm.fields import lognormal_conductivity
from problems import X0, X1, face_mean_weights, heat_c3, mass_c2


def _setup(p):
    g = O.dual(p.pos, p.mesh.get("tets"), p.mesh.num_pts)
    return g


def _objective(p, g, w, T0, dt, num_ts):
    scl, sb, flux, bnd_nid, ops = p.oracle_inputs(g)
    states = O.linear_map(g, ops, scl, sb, flux, bnd_nid, p.unknown, T0, dt, num_ts)
    return float(w @ states[-1]), states


def _adjoint(p, g, w, T0, dt, num_ts):
    scl, sb, flux, bnd_nid, ops = p.oracle_inputs(g)
    _, states = _objective(p, g, w, T0, dt, num_ts)
    acc = O.adjoint_accumulate(g, ops, scl, sb, flux, bnd_nid, p.unknown, states, dt, num_ts, w)
    return acc, (scl, sb, flux, bnd_nid, ops)


def _fd_field(p, g, w, T0, dt, num_ts, fid, verts, rel=1e-5):
    kind, base = p.scl[fid]
    out = []
    for v in verts:
        d = rel * max(abs(base[v]), 1e-3)
        vals = []
        for s in (+1, -1):
            x = base.copy()
            x[v] += s * d
            p.scl[fid] = (kind, x)
            vals.append(_objective(p, g, w, T0, dt, num_ts)[0])
        p.scl[fid] = (kind, base)
        out.append((vals[0] - vals[1]) / (2 * d))
    return np.array(out)


@pytest.fixture(scope="module")
def heat6():
    mesh = vfvm.Mesh.lattice(6, seed=3)
    k = lognormal_conductivity(mesh.get("pos"), ell=0.3, modes=64, seed=1)
    q = 0.5 + 0.25 * np.sin(3.0 * mesh.get("pos")[:, 1])
    wt = 1.0 + 0.2 * mesh.get("pos")[:, 2]
    p = heat_c3(mesh, k, w=wt, q=q)
    g = _setup(p)
    return p, g


def _check(adj, fd, tol=2e-6):
    scale = max(np.max(np.abs(adj)), 1e-300)
    assert np.max(np.abs(adj - fd)) <= tol * scale, (adj, fd)


@pytest.mark.parametrize("num_ts", [0, 4])
def test_conductivity_gradient_vs_fd(heat6, num_ts):
    p, g = heat6
    N = p.N
    w = face_mean_weights(p.mesh, X1)
    T0 = np.zeros(N)
    dt = 0.02
    if num_ts == 0:
        # steady: the time operators are not part of the steady map
        p_ops = p.ops
        p.ops = [o for o in p.ops if o[0] not in (vfvm.OP_SCL_DOM_TIME, vfvm.OP_SCL_BND_TIME)]
    try:
        acc, (scl, sb, flux, bnd_nid, ops) = _adjoint(p, g, w, T0, dt, num_ts)
        grad = O.adjoint_field_gradient(g, ops, scl, flux, bnd_nid, acc, dt, p.k_id, constant=False)
        NP = p.mesh.num_pts
        x1 = NP + p.mesh.bnd_nodes(X1).astype(int)
        x0 = NP + p.mesh.bnd_nodes(X0).astype(int)
        verts = [0, 7, NP // 2, NP - 1, int(x1[3]), int(x0[5]), N - 1]
        fd = _fd_field(p, g, w, T0, dt, num_ts, p.k_id, verts)
        _check(grad[verts], fd)
    finally:
        if num_ts == 0:
            p.ops = p_ops


def test_source_and_weight_gradients_vs_fd(heat6):
    p, g = heat6
    N = p.N
    w = face_mean_weights(p.mesh, X1)
    T0 = 0.3 * np.cos(2.0 * p.pos[:, 0])
    dt, num_ts = 0.02, 3
    acc, (scl, sb, flux, bnd_nid, ops) = _adjoint(p, g, w, T0, dt, num_ts)
    verts = [1, 30, p.mesh.num_pts - 2, N - 3]
    gq = O.adjoint_field_gradient(g, ops, scl, flux, bnd_nid, acc, dt, p.q_id, constant=False)
    _check(gq[verts], _fd_field(p, g, w, T0, dt, num_ts, p.q_id, verts))
    gw = O.adjoint_field_gradient(g, ops, scl, flux, bnd_nid, acc, dt, p.w_id, constant=False)
    _check(gw[verts], _fd_field(p, g, w, T0, dt, num_ts, p.w_id, verts))


def test_dirichlet_and_initial_gradients_vs_fd(heat6):
    p, g = heat6
    N = p.N
    w = face_mean_weights(p.mesh, X1)
    T0 = 0.3 * np.cos(2.0 * p.pos[:, 0])
    dt, num_ts = 0.02, 3
    acc, (scl, sb, flux, bnd_nid, ops) = _adjoint(p, g, w, T0, dt, num_ts)
    gd = O.adjoint_dirichlet_gradient(g, ops, bnd_nid, acc, p.tin_id)
    b, base = p.sbnd[p.tin_id]
    fd = []
    for i in (0, 4, len(base) - 1):
        vals = []
        for s in (+1, -1):
            x = base.copy()
            x[i] += s * 1e-5
            p.sbnd[p.tin_id] = (b, x)
            vals.append(_objective(p, g, w, T0, dt, num_ts)[0])
        p.sbnd[p.tin_id] = (b, base)
        fd.append((vals[0] - vals[1]) / 2e-5)
    _check(gd[[0, 4, len(base) - 1]], np.array(fd))
    gi = O.adjoint_initial_gradient(g, ops, bnd_nid, acc, num_ts)
    NP = p.mesh.num_pts
    verts = [2, NP - 5, NP + int(p.mesh.bnd_nodes(X0)[1]), NP + int(p.mesh.bnd_nodes(X1)[2])]
    fd = []
    for v in verts:
        vals = []
        for s in (+1, -1):
            x = T0.copy()
            x[v] += s * 1e-5
            vals.append(_objective(p, g, w, x, dt, num_ts)[0])
        fd.append((vals[0] - vals[1]) / 2e-5)
    fd = np.array(fd)
    assert gi[verts[2]] == 0.0 and abs(fd[2]) < 1e-12  # Dirichlet node: overwritten
    _check(gi[verts], fd)


def test_advection_diffusivity_gradient_vs_fd():
    mesh = vfvm.Mesh.lattice(6, seed=5)
    D = 1e-2 * np.exp(0.5 * np.sin(4.0 * mesh.get("pos")[:, 0] + mesh.get("pos")[:, 2]))
    p = mass_c2(mesh, D=D, c_out=0.0)  # c = 1 at x=0, c = 0 at x=1: J depends on D
    g = _setup(p)
    w = np.random.default_rng(2).uniform(size=p.N) / p.N
    T0 = np.zeros(p.N)
    acc, (scl, sb, flux, bnd_nid, ops) = _adjoint(p, g, w, T0, 0.0, 0)
    grad = O.adjoint_field_gradient(g, ops, scl, flux, bnd_nid, acc, 0.0, p.d_id, constant=False)
    NP = mesh.num_pts
    verts = [0, 11, NP // 3, NP - 1, NP + int(mesh.bnd_nodes(X1)[4]), NP + int(mesh.bnd_nodes(X0)[7]), p.N - 1]
    fd = _fd_field(p, g, w, T0, 0.0, 0, p.d_id, verts)
    _check(grad[verts], fd, tol=5e-6)


def test_transposed_solve_identity(heat6):
    """<seed, T_lin(dT0)> = <dJ/dT0, dT0>: the adjoint of the whole map is its transpose."""
    p, g = heat6
    N = p.N
    rng = np.random.default_rng(0)
    w = rng.normal(size=N)
    dt, num_ts = 0.02, 3
    T0 = np.zeros(N)
    acc, (scl, sb, flux, bnd_nid, ops) = _adjoint(p, g, w, T0, dt, num_ts)
    gi = O.adjoint_initial_gradient(g, ops, bnd_nid, acc, num_ts)
    dT0 = rng.normal(size=N)
    J1 = _objective(p, g, w, T0 + dT0, dt, num_ts)[0]
    J0 = _objective(p, g, w, T0, dt, num_ts)[0]
    # J is affine in T0: the difference is exact up to round-off
    assert abs((J1 - J0) - gi @ dT0) <= 1e-10 * max(abs(J1 - J0), 1.0)
