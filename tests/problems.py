"""Problem set-ups shared by the parity tests: the same operator list is given to
the device path (through the C ABI) and to the oracle (numpy restatement)."""
import numpy as np

import vfvm
import vfvm_oracle as O

X0, X1, Y0, Y1, Z0, Z1 = range(6)


def velocity_cellular(pos, U0):
    """u = U0 (sin pi x cos pi y, -cos pi x sin pi y, 0), SoA (3, N)."""
    x, y = pos[:, 0], pos[:, 1]
    return np.stack([U0 * np.sin(np.pi * x) * np.cos(np.pi * y), -U0 * np.cos(np.pi * x) * np.sin(np.pi * y), 0.0 * x])


class Problem:
    """Fields + operator list for the device context and the oracle."""

    def __init__(self, mesh: vfvm.Mesh):
        self.mesh = mesh
        self.N = mesh.num_tot
        self.pos = mesh.get("pos")
        self.scl = []      # (kind, values)
        self.sbnd = []     # (bnd, values)
        self.vec = []      # (kind, values (3,N))
        self.flw = []      # (vel, pres, den, visc, fce)
        self.ops = []

    def scldom(self, kind, values):
        v = np.broadcast_to(np.asarray(values, np.float64), (self.N,)).copy()
        self.scl.append((kind, v))
        return len(self.scl) - 1

    def sclbnd(self, b, values):
        cnt = len(self.mesh.bnd_nodes(b))
        self.sbnd.append((b, np.broadcast_to(np.asarray(values, np.float64), (cnt,)).copy()))
        return len(self.sbnd) - 1

    def vecdom(self, kind, values):
        self.vec.append((kind, np.asarray(values, np.float64).reshape(3, self.N).copy()))
        return len(self.vec) - 1

    def sclflw(self, vel, pres, den, visc, fce):
        self.flw.append((vel, pres, den, visc, fce))
        return len(self.flw) - 1

    # ---- device
    def to_device(self, ctx: vfvm.Context):
        ctx.upload(self.mesh)
        for kind, v in self.scl:
            ctx.scldom_add(kind, 0.0, v)
        for b, v in self.sbnd:
            ctx.sclbnd_add(b, 0.0, v)
        for kind, v in self.vec:
            ctx.vecdom_add(kind, values=v)
        for f in self.flw:
            ctx.sclflw_add(*f)
        return ctx

    # ---- oracle
    def oracle_inputs(self, g, unknown_value=None):
        """Fields (value, gradient), boundary values, face fluxes, boundary node lists and
        operators in the oracle's form."""
        scl = {}
        for i, (kind, v) in enumerate(self.scl):
            if unknown_value is not None and kind == vfvm.FIELD_UNKNOWN:
                v = unknown_value
            grad = O.update_grad(g, v) if kind != vfvm.FIELD_CONSTANT else np.zeros((3, self.N))
            scl[i] = (v, grad)
        flux = {}
        for i, (vel, pres, den, visc, fce) in enumerate(self.flw):
            vk, vv = self.vec[vel]
            vg = O.update_grad_vec(g, vv) if vk != vfvm.FIELD_CONSTANT else np.zeros((9, self.N))
            flux[i] = O.flux_prescribed(g, vv, vg, scl[den][0], scl[den][1])
        bnd_nid = [self.mesh.bnd_nodes(b) for b in range(6)]
        sb = {i: v for i, (b, v) in enumerate(self.sbnd)}
        ops = [O.Op(*o) for o in self.ops]
        return scl, sb, flux, bnd_nid, ops

    def oracle_system(self, g=None, value_prev=None, dt=0.0, unknown_value=None):
        """Assembled CSR (row_ptr, col, val) and b of the oracle, and the dual used."""
        if g is None:
            g = O.dual(self.pos, self.mesh.get("tets"), self.mesh.num_pts)
        scl, sb, flux, bnd_nid, ops = self.oracle_inputs(g, unknown_value)
        r, c, v, b = O.assemble(g, ops, scl, sb, flux, bnd_nid, dt=dt, value_prev=value_prev)
        rp, col, val = O.coo_to_csr(r, c, v, self.N)
        return rp, col, val, b, g


def heat_c1(mesh, k=1.0, q=1.0):
    """Steady heat: k, source q, T=0 at x=0, T=1 at x=1, zero flux elsewhere
    (SteadyHeat operator order, steady_heat.rs:143-171)."""
    p = Problem(mesh)
    T = p.scldom(vfvm.FIELD_UNKNOWN, 0.0)
    kk = p.scldom(vfvm.FIELD_CONSTANT if np.isscalar(k) else vfvm.FIELD_FUNCTION, k)
    hp = p.scldom(vfvm.FIELD_CONSTANT, 0.0)
    hs = p.scldom(vfvm.FIELD_CONSTANT, q)
    t0 = p.sclbnd(X0, 0.0)
    t1 = p.sclbnd(X1, 1.0)
    fl = {b: p.sclbnd(b, 0.0) for b in (Y0, Y1, Z0, Z1)}
    p.ops = [(vfvm.OP_SCL_DOM_DIFFUSION, 0, T, kk, T), (vfvm.OP_SCL_DOM_PRODUCTION, 0, T, hp, -1),
             (vfvm.OP_SCL_DOM_SOURCE, 0, T, hs, -1),
             (vfvm.OP_SCL_BND_DIRICHLET, X0, T, t0, -1), (vfvm.OP_SCL_BND_DIRICHLET, X1, T, t1, -1)]
    for b in (Y0, Y1, Z0, Z1):
        p.ops += [(vfvm.OP_SCL_BND_DIFFUSION, b, T, kk, T), (vfvm.OP_SCL_BND_PRODUCTION, b, T, hp, -1),
                  (vfvm.OP_SCL_BND_SOURCE, b, T, hs, -1), (vfvm.OP_SCL_BND_NEUMANN, b, T, fl[b], -1)]
    p.unknown = T
    return p


def mass_c2(mesh, D=1e-2, U0=None, c_out=None, prod=0.0):
    """Steady convection-diffusion on the cellular flow (SteadyMassFlow mass part,
    steady_massflow.rs:455-545): c=1 at x=0 (and c=c_out at x=1 if given), zero
    diffusive flux elsewhere; U0 defaults to cell Peclet 5."""
    n = round(mesh.num_tot ** (1 / 3))
    h = 1.0 / (n - 1)
    if U0 is None:
        U0 = 5.0 * (D if np.isscalar(D) else float(np.mean(D))) / h
    p = Problem(mesh)
    c = p.scldom(vfvm.FIELD_UNKNOWN, 0.0)
    Dd = p.scldom(vfvm.FIELD_CONSTANT if np.isscalar(D) else vfvm.FIELD_FUNCTION, D)
    mp = p.scldom(vfvm.FIELD_CONSTANT, prod)
    ms = p.scldom(vfvm.FIELD_CONSTANT, 0.0)
    den = p.scldom(vfvm.FIELD_CONSTANT, 1.0)
    visc = p.scldom(vfvm.FIELD_CONSTANT, 1.0)
    pres = p.scldom(vfvm.FIELD_CONSTANT, 0.0)
    vel = p.vecdom(vfvm.FIELD_FUNCTION, velocity_cellular(p.pos, U0))
    fce = p.vecdom(vfvm.FIELD_CONSTANT, np.zeros((3, p.N)))
    f = p.sclflw(vel, pres, den, visc, fce)
    cin = p.sclbnd(X0, 1.0)
    dir_b = [(X0, cin)]
    flux_bnds = [Y0, Y1, Z0, Z1]
    if c_out is not None:
        dir_b.append((X1, p.sclbnd(X1, c_out)))
    else:
        flux_bnds = [X1] + flux_bnds
    q = {b: p.sclbnd(b, 0.0) for b in flux_bnds}
    p.ops = [(vfvm.OP_SCL_DOM_DIFFUSION, 0, c, Dd, c), (vfvm.OP_SCL_DOM_PRODUCTION, 0, c, mp, -1),
             (vfvm.OP_SCL_DOM_SOURCE, 0, c, ms, -1), (vfvm.OP_SCL_DOM_ADVECTION_UNITY, 0, c, Dd, f)]
    for b, v in dir_b:
        p.ops.append((vfvm.OP_SCL_BND_DIRICHLET, b, c, v, -1))
    for b in flux_bnds:
        p.ops += [(vfvm.OP_SCL_BND_DIFFUSION, b, c, Dd, c), (vfvm.OP_SCL_BND_PRODUCTION, b, c, mp, -1),
                  (vfvm.OP_SCL_BND_SOURCE, b, c, ms, -1), (vfvm.OP_SCL_BND_NEUMANN, b, c, q[b], -1)]
    for b in flux_bnds:
        p.ops.append((vfvm.OP_SCL_BND_ADVECTION_UNITY, b, c, Dd, f))
    p.unknown, p.d_id, p.cin_id = c, Dd, cin
    return p


def heat_c3(mesh, k, w=1.0, q=0.0, t_in=1.0):
    """Transient heat (TransientHeat operator order, transient_heat.rs:152-171): per-vertex
    conductivity k, weight rho c_p = w, source q, T = t_in at x=0, zero heat flux on the
    other five faces, T0 = 0.  k, q: scalars (constant fields) or per-vertex arrays."""
    p = Problem(mesh)
    T = p.scldom(vfvm.FIELD_UNKNOWN, 0.0)
    ww = p.scldom(vfvm.FIELD_CONSTANT if np.isscalar(w) else vfvm.FIELD_FUNCTION, w)
    kk = p.scldom(vfvm.FIELD_CONSTANT if np.isscalar(k) else vfvm.FIELD_FUNCTION, k)
    hp = p.scldom(vfvm.FIELD_CONSTANT, 0.0)
    hs = p.scldom(vfvm.FIELD_CONSTANT if np.isscalar(q) else vfvm.FIELD_FUNCTION, q)
    tin = p.sclbnd(X0, t_in)
    fl = {b: p.sclbnd(b, 0.0) for b in (X1, Y0, Y1, Z0, Z1)}
    p.ops = [(vfvm.OP_SCL_DOM_TIME, 0, T, ww, -1), (vfvm.OP_SCL_DOM_DIFFUSION, 0, T, kk, T),
             (vfvm.OP_SCL_DOM_PRODUCTION, 0, T, hp, -1), (vfvm.OP_SCL_DOM_SOURCE, 0, T, hs, -1),
             (vfvm.OP_SCL_BND_DIRICHLET, X0, T, tin, -1)]
    for b in (X1, Y0, Y1, Z0, Z1):
        p.ops += [(vfvm.OP_SCL_BND_TIME, b, T, ww, -1), (vfvm.OP_SCL_BND_DIFFUSION, b, T, kk, T),
                  (vfvm.OP_SCL_BND_PRODUCTION, b, T, hp, -1), (vfvm.OP_SCL_BND_SOURCE, b, T, hs, -1),
                  (vfvm.OP_SCL_BND_NEUMANN, b, T, fl[b], -1)]
    p.unknown, p.k_id, p.w_id, p.q_id, p.tin_id = T, kk, ww, hs, tin
    return p


def face_mean_weights(mesh, b=X1):
    """dJ/dT of J = mean of T over the nodes of boundary b."""
    w = np.zeros(mesh.num_tot)
    nodes = mesh.num_pts + mesh.bnd_nodes(b).astype(np.int64)
    w[nodes] = 1.0 / nodes.size
    return w
