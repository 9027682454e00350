"""CPU tests: the oracle's flow restatement (grad_coeff, update_flow, assemble_flow) held to
closed forms and invariants, and the C++ grad_coeff builder to the oracle.

Pins (each one fails for a dropped term, a wrong sign or a transposed operand):
  grad_coeff:    sum_m c_m p_m = V grad(p) for linear p at every vertex (divergence theorem).
  update_flow:   F(v->n) = -F(n->v); linear p with body force grad(p) gives zero Rhie-Chow
                 correction, F = rho_f (u_f . n) A (opscl_dom_pstab.rs:31).
  assemble_flow: velocity-block row sums 0 (advection + diffusion); gradient block on linear
                 p = -V grad(p); divergence block on linear u = -rho V div(u) on points and
                 nodes; pressure-stabiliser block annihilates linear p; transpose and
                 dilatatory b vanish for linear u and constant viscosity; pattern = 2-ring
                 velocity columns and 3-ring pressure columns; identity reference row.
"""
import numpy as np
import pytest

import vfvm
import vfvm_oracle as O
from problems import FlowProblem, flow_general, lid_cavity


@pytest.fixture(scope="module")
def mesh6():
    m = vfvm.Mesh.lattice(6, 0, 0.3)
    m.build_grad_coeff()
    return m


@pytest.fixture(scope="module")
def g6(mesh6):
    return O.dual(mesh6.get("pos"), mesh6.get("tets"), mesh6.num_pts)


@pytest.fixture(scope="module")
def gc6(g6):
    return O.grad_coeff(g6)


def _lin(pos, a0, gvec):
    return a0 + pos @ np.asarray(gvec, np.float64)


def test_grad_coeff_builder_bitwise(mesh6, gc6):
    ptr, vid, c = gc6
    assert np.array_equal(mesh6.get("gc_ptr").astype(np.int64), ptr)
    assert np.array_equal(mesh6.get("gc_vid").astype(np.int64), vid)
    assert np.array_equal(mesh6.get("gc_c").reshape(3, -1), c)


def test_grad_coeff_linear_exact(g6, gc6):
    ptr, vid, c = gc6
    pos = vfvm.Mesh.lattice(6, 0, 0.3).get("pos")
    gvec = np.array([0.7, -1.3, 2.1])
    p = _lin(pos, 0.4, gvec)
    N = g6["N"]
    rows = np.repeat(np.arange(N), np.diff(ptr))
    for k in range(3):
        s = np.zeros(N)
        np.add.at(s, rows, c[k] * p[vid])
        scale = np.zeros(N)
        np.add.at(scale, rows, np.abs(c[k] * p[vid]))
        assert np.all(np.abs(s - g6["vol"] * gvec[k]) <= 1e-12 * (scale + 1.0))
    # every list holds v and its neighbours (2-ring closure of the stencils)
    for v in range(N):
        lst = set(vid[ptr[v]:ptr[v + 1]].tolist())
        assert v in lst and set(g6["ngh"][g6["row_ptr"][v]:g6["row_ptr"][v + 1]].tolist()) <= lst


def _fields(mesh, g, gc, vel, pres, den, visc, fce):
    return O.flow_fields(g, gc, vel, pres, den, visc, fce, vel_const=False, den_const=False, visc_const=False, fce_const=False)


def test_flux_antisymmetric_and_hydrostatic(mesh6, g6, gc6):
    pos = mesh6.get("pos")
    N = mesh6.num_tot
    p = flow_general(mesh6)
    f = _fields(mesh6, g6, gc6, p.vel, p.pres, p.den, p.visc, p.fce)
    rows, ngh = g6["rows"].astype(np.int64), g6["ngh"].astype(np.int64)
    # reverse edge of every directed edge
    key = {(int(a), int(b)): e for e, (a, b) in enumerate(zip(rows, ngh))}
    rev = np.array([key[(int(b), int(a))] for a, b in zip(rows, ngh)])
    F = f["flux"]
    assert np.max(np.abs(F + F[rev])) <= 1e-13 * np.max(np.abs(F))
    assert np.all(f["rc"] > 0.0)
    # hydrostatic balance: linear p, body force = grad p (constant) -> no Rhie-Chow correction
    gvec = np.array([0.5, -2.0, 1.5])
    pl = _lin(pos, 0.3, gvec)
    fce = np.repeat(gvec[:, None], N, axis=1)
    f2 = O.flow_fields(g6, gc6, p.vel, pl, p.den, p.visc, fce, den_const=False, visc_const=False, fce_const=True)
    sx, sy, sz = g6["skew"]
    nx, ny, nz = g6["norm"]
    den_f = O._face_mean(p.den, f2["deng"], rows, ngh, sx, sy, sz)
    uf = [0.5 * (p.vel[c][rows] + p.vel[c][ngh]) + 0.5 * (sx * (f2["velg"][3 * c][rows] + f2["velg"][3 * c][ngh]) + sy * (f2["velg"][3 * c + 1][rows] + f2["velg"][3 * c + 1][ngh]) + sz * (f2["velg"][3 * c + 2][rows] + f2["velg"][3 * c + 2][ngh])) for c in range(3)]
    F0 = den_f * (uf[0] * nx + uf[1] * ny + uf[2] * nz) * g6["area"]
    assert np.max(np.abs(f2["flux"] - F0)) <= 1e-11 * np.max(np.abs(F0))
    # any constant body force with a linear pressure balances (the correction is consistent);
    # a curved pressure makes it act
    f3 = O.flow_fields(g6, gc6, p.vel, pl, p.den, p.visc, np.zeros((3, N)), den_const=False, visc_const=False, fce_const=True)
    assert np.max(np.abs(f3["flux"] - F0)) <= 1e-11 * np.max(np.abs(F0))
    f4 = O.flow_fields(g6, gc6, p.vel, pl + 40.0 * pos[:, 0] ** 2, p.den, p.visc, fce, den_const=False, visc_const=False, fce_const=True)
    assert np.max(np.abs(f4["flux"] - F0)) > 1e-4 * np.max(np.abs(F0))


@pytest.fixture(scope="module")
def general6(mesh6, g6, gc6):
    p = flow_general(mesh6)
    f = _fields(mesh6, g6, gc6, p.vel, p.pres, p.den, p.visc, p.fce)
    bnd_nid = [mesh6.bnd_nodes(b) for b in range(6)]
    r, c, v, b = O.assemble_flow(g6, gc6, f, p.bnd_vel, bnd_nid, None)
    rp, col, val = O.coo_to_csr(r, c, v, 4 * mesh6.num_tot)
    return p, f, (rp, col, val, b)


def _dense_block(rp, col, val, rows, c0, c1):
    import scipy.sparse as sp
    A = sp.csr_matrix((val, col, rp))
    return A[rows][:, c0:c1]


def test_assemble_flow_invariants(mesh6, g6, gc6, general6):
    p, f, (rp, col, val, b) = general6
    N, NP = mesh6.num_tot, mesh6.num_pts
    pos = mesh6.get("pos")
    import scipy.sparse as sp
    A = sp.csr_matrix((val, col, rp), shape=(4 * N, 4 * N))
    pts = np.arange(NP)
    # momentum: the velocity block of each component row sums to zero
    for c in range(3):
        blk = A[c * N + pts][:, c * N:(c + 1) * N]
        rs = np.asarray(blk.sum(axis=1)).ravel()
        assert np.max(np.abs(rs)) <= 1e-12 * abs(blk).max()
        assert A[c * N + pts][:, [k * N + j for k in range(3) if k != c for j in range(N)]].nnz == 0
    # gradient block on linear p: -V grad p
    gvec = np.array([1.1, -0.6, 0.9])
    pl = _lin(pos, -0.2, gvec)
    for c in range(3):
        r = A[c * N + pts][:, 3 * N:] @ pl
        assert np.max(np.abs(r + g6["vol"][pts] * gvec[c])) <= 1e-11
    # divergence block on linear u: -rho V div u (constant density problem)
    q = FlowProblem(mesh6, p.vel, p.pres, 1.7, p.visc, p.fce, p.bnd_vel, None, const=("den",))
    fq = O.flow_fields(g6, gc6, q.vel, q.pres, q.den, q.visc, q.fce, den_const=True, visc_const=False, fce_const=False)
    r_, c_, v_, b_ = O.assemble_flow(g6, gc6, fq, q.bnd_vel, [mesh6.bnd_nodes(k) for k in range(6)], None)
    Aq = sp.csr_matrix(sp.coo_matrix((v_, (r_, c_)), shape=(4 * N, 4 * N)))
    Gu = np.array([[0.3, -1.2, 0.5], [0.8, 0.1, -0.7], [-0.4, 0.6, 0.9]])
    ulin = np.stack([_lin(pos, 0.1 * k, Gu[k]) for k in range(3)]).ravel()
    div = np.trace(Gu)
    r = Aq[3 * N:][:, :3 * N] @ ulin
    assert np.max(np.abs(r + 1.7 * g6["vol"] * div)) <= 1e-11
    # pressure-stabiliser block annihilates linear p (points and nodes)
    r = A[3 * N:][:, 3 * N:] @ pl
    scale = abs(A[3 * N:][:, 3 * N:]) @ np.abs(pl)
    assert np.all(np.abs(r) <= 1e-11 * (scale + 1e-30))
    # pattern: continuity rows hold 3 x grad_coeff(v) velocity columns and the 3-ring
    ptr, vid, _ = gc6
    for v in (0, NP // 2, NP, N - 1):
        cols = col[rp[3 * N + v]:rp[3 * N + v + 1]]
        gl = vid[ptr[v]:ptr[v + 1]]
        assert np.array_equal(cols[:3 * gl.size], np.concatenate([k * N + gl for k in range(3)]))
        ring = set()
        for u in [v] + g6["ngh"][g6["row_ptr"][v]:g6["row_ptr"][v + 1]].tolist():
            ring |= set(vid[ptr[u]:ptr[u + 1]].tolist())
        assert np.array_equal(cols[3 * gl.size:], 3 * N + np.array(sorted(ring)))


def test_transpose_dilatatory_vanish_for_linear_velocity(mesh6, g6, gc6):
    pos = mesh6.get("pos")
    N, NP = mesh6.num_tot, mesh6.num_pts
    Gu = np.array([[0.3, -1.2, 0.5], [0.8, 0.1, -0.7], [-0.4, 0.6, 0.9]])
    vel = np.stack([_lin(pos, 0.1 * k, Gu[k]) for k in range(3)])
    p = lid_cavity(mesh6)
    f = O.flow_fields(g6, gc6, vel, np.zeros(N), np.full(N, 1.0), np.full(N, 0.05), np.zeros((3, N)))
    _, _, _, b = O.assemble_flow(g6, gc6, f, p.bnd_vel, [mesh6.bnd_nodes(k) for k in range(6)], None)
    for c in range(3):
        assert np.max(np.abs(b[c * N:c * N + NP])) <= 1e-13
    # a body force enters as -f V
    fce = np.stack([np.full(N, 2.0), np.full(N, -1.0), np.full(N, 0.5)])
    f2 = O.flow_fields(g6, gc6, vel, np.zeros(N), np.full(N, 1.0), np.full(N, 0.05), fce)
    _, _, _, b2 = O.assemble_flow(g6, gc6, f2, p.bnd_vel, [mesh6.bnd_nodes(k) for k in range(6)], None)
    for c in range(3):
        assert np.max(np.abs(b2[c * N:c * N + NP] + fce[c, 0] * g6["vol"][:NP])) <= 1e-13


def test_reference_pressure_row(mesh6, g6, gc6):
    p = lid_cavity(mesh6, pref=(5, 2.5))
    N, NP = mesh6.num_tot, mesh6.num_pts
    f = O.flow_fields(g6, gc6, p.vel, p.pres, p.den, p.visc, p.fce)
    r, c, v, b = O.assemble_flow(g6, gc6, f, p.bnd_vel, [mesh6.bnd_nodes(k) for k in range(6)], p.pref)
    xid = 3 * N + NP + 5
    sel = r == xid
    assert sel.sum() == 1 and c[sel][0] == xid and v[sel][0] == 1.0 and b[xid] == 2.5


def test_scaled_residual_flow_closed_form():
    # N = 2: velocity (0,0,0) -> (3,4,0) at vertex 0; pressure (0,0) -> (1,-1); rho = 1
    xo = np.zeros(8)
    xn = np.array([3.0, 0.0, 4.0, 0.0, 0.0, 0.0, 1.0, -1.0])
    r = O.scaled_residual_flow(2, xo, xn, np.ones(2))
    # vector: rms = sqrt(25 / 2), scale = |(3,4,0)| = 5; pressure: rms ~ 1, scale max(2, 1 * 25)
    assert abs(r - np.sqrt(12.5) / 5.0) < 1e-15
    xn2 = np.array([0.3, 0.0, 0.4, 0.0, 0.0, 0.0, 10.0, -10.0])
    r2 = O.scaled_residual_flow(2, xo, xn2, np.ones(2))
    assert abs(r2 - max(np.sqrt(0.125) / 0.5, np.sqrt(((10.0 - 1e-10) ** 2 * 2) / 2) / 20.0)) < 1e-14
    assert np.isnan(O.scaled_residual_flow(2, xo, np.where(np.arange(8) == 5, np.inf, xn), np.ones(2)))


def test_transient_flow_terms(mesh6, g6, gc6):
    """TransientFlow adds -rho V / dt to the velocity diagonal of the points, -rho u_prev V / dt to
    their b, and (rho - rho_prev) V / dt to every continuity row's b."""
    import scipy.sparse as sp
    p = flow_general(mesh6)
    N, NP = mesh6.num_tot, mesh6.num_pts
    f = O.flow_fields(g6, gc6, p.vel, p.pres, p.den, p.visc, p.fce, den_const=False, visc_const=False, fce_const=False)
    bn = [mesh6.bnd_nodes(k) for k in range(6)]
    dt = 0.05
    vprev = 0.5 * p.vel + 0.1
    dprev = p.den * 0.9
    r0, c0, v0, b0 = O.assemble_flow(g6, gc6, f, p.bnd_vel, bn, None)
    r1, c1, v1, b1 = O.assemble_flow(g6, gc6, f, p.bnd_vel, bn, None, dt, vprev, dprev)
    A0 = sp.csr_matrix(sp.coo_matrix((v0, (r0, c0)), shape=(4 * N, 4 * N)))
    A1 = sp.csr_matrix(sp.coo_matrix((v1, (r1, c1)), shape=(4 * N, 4 * N)))
    D = (A1 - A0).tocoo()
    vol = g6["vol"]
    pts = np.arange(NP)
    expect = np.zeros(4 * N)
    for c in range(3):
        expect[c * N + pts] = -p.den[pts] * vol[pts] / dt
    dense_diag = np.zeros(4 * N)
    np.add.at(dense_diag, D.row[D.row == D.col], D.data[D.row == D.col])
    assert np.max(np.abs(D.data[D.row != D.col])) <= 1e-12 if np.any(D.row != D.col) else True
    assert np.max(np.abs(dense_diag - expect)) <= 1e-12 * np.max(np.abs(expect))
    db = b1 - b0
    for c in range(3):
        assert np.max(np.abs(db[c * N + pts] + p.den[pts] * vprev[c][pts] * vol[pts] / dt)) <= 1e-12
        assert np.all(db[c * N + NP:(c + 1) * N] == 0.0)
    assert np.max(np.abs(db[3 * N:] - (p.den - dprev) * vol / dt)) <= 1e-12
