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
        bnd_nid = [self.mesh.bnd_nodes(b) for b in range(self.mesh.num_bnd)]
        sb = {i: v for i, (b, v) in enumerate(self.sbnd)}
        ops = [O.Op(*o) for o in self.ops]
        self.flow_den = {i: scl[den] for i, (vel, pres, den, visc, fce) in enumerate(self.flw)}
        return scl, sb, flux, bnd_nid, ops

    def oracle_system(self, g=None, value_prev=None, dt=0.0, unknown_value=None):
        """Assembled CSR (row_ptr, col, val) and b of the oracle, and the dual used."""
        if g is None:
            g = O.dual(self.pos, self.mesh.get("tets"), self.mesh.num_pts)
        scl, sb, flux, bnd_nid, ops = self.oracle_inputs(g, unknown_value)
        r, c, v, b = O.assemble(g, ops, scl, sb, flux, bnd_nid, dt=dt, value_prev=value_prev, flow_den=self.flow_den)
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


class FlowProblem:
    """SteadyFlow composition (steady_flow.rs:133-240) of one domain with a velocity condition
    on every boundary and a reference pressure, for the device context and the oracle.
    Field arrays are host copies: vel (3,N), pres, den, visc (N), fce (3,N); `const` marks the
    fields uploaded as constants (zero gradient)."""

    def __init__(self, mesh: vfvm.Mesh, vel, pres, den, visc, fce, bnd_vel, pref=(0, 0.0), const=("den", "visc", "fce")):
        self.mesh = mesh
        self.N = mesh.num_tot
        self.vel = np.asarray(vel, np.float64).reshape(3, self.N).copy()
        self.pres = np.broadcast_to(np.asarray(pres, np.float64), (self.N,)).copy()
        self.den = np.broadcast_to(np.asarray(den, np.float64), (self.N,)).copy()
        self.visc = np.broadcast_to(np.asarray(visc, np.float64), (self.N,)).copy()
        self.fce = np.asarray(fce, np.float64).reshape(3, self.N).copy()
        self.bnd_vel = [(b, np.asarray(v, np.float64).reshape(3, -1)) for b, v in bnd_vel]
        self.pref = pref
        self.const = set(const)

    def ops(self, ids):
        vel, pres, den, visc, fce, flw, vb = ids
        o = [(vfvm.OP_VEC_DOM_TIME, 0, vel, den, -1)] if getattr(self, "transient", False) else []
        o += [(vfvm.OP_VEC_DOM_ADVECTION, 0, vel, den, visc, flw), (vfvm.OP_VEC_DOM_GRADIENT, 0, vel, pres, -1),
             (vfvm.OP_VEC_DOM_DIFFUSION, 0, vel, visc, -1), (vfvm.OP_VEC_DOM_TRANSPOSE, 0, vel, visc, -1),
             (vfvm.OP_VEC_DOM_DILATATORY, 0, vel, visc, -1), (vfvm.OP_VEC_DOM_SOURCE, 0, vel, fce, -1),
             (vfvm.OP_SCL_DOM_DIVERGENCE, 0, pres, vel, den), (vfvm.OP_SCL_DOM_PRESSURE_STABILIZER, 0, pres, den, fce, flw)]
        for (b, _), vid in zip(self.bnd_vel, vb):
            o += [(vfvm.OP_VEC_BND_DIRICHLET, b, vel, vid, -1), (vfvm.OP_SCL_BND_DIVERGENCE, b, pres, vel, den),
                  (vfvm.OP_SCL_BND_PRESSURE_STABILIZER, b, pres, den, fce, flw)]
        if self.pref is not None:
            o.append((vfvm.OP_SCL_DOM_PRESSURE_REFERENCE, 0, pres, int(self.pref[0]), -1, -1, float(self.pref[1])))
        return o

    def to_device(self, ctx: vfvm.Context):
        if self.mesh.num_grad_coeff == 0:
            self.mesh.build_grad_coeff()
        ctx.upload(self.mesh)
        kind = lambda name: vfvm.FIELD_CONSTANT if name in self.const else vfvm.FIELD_FUNCTION
        vel = ctx.vecdom_add(vfvm.FIELD_UNKNOWN, values=self.vel)
        pres = ctx.scldom_add(vfvm.FIELD_UNKNOWN, 0.0, self.pres)
        den = ctx.scldom_add(kind("den"), 0.0, self.den)
        visc = ctx.scldom_add(kind("visc"), 0.0, self.visc)
        fce = ctx.vecdom_add(kind("fce"), values=self.fce)
        flw = ctx.sclflw_add(vel, pres, den, visc, fce)
        vb = [ctx.vecbnd_add(b, values=v) for b, v in self.bnd_vel]
        self.ids = (vel, pres, den, visc, fce, flw, vb)
        return self.ops(self.ids)

    def oracle(self, g=None, gc=None, dt=0.0, vel_prev=None, den_prev=None):
        """Oracle dual, grad_coeff, fields and the assembled CSR system at the current values."""
        if g is None:
            g = O.dual(self.mesh.get("pos"), self.mesh.get("tets"), self.mesh.num_pts)
        if gc is None:
            gc = O.grad_coeff(g)
        f = O.flow_fields(g, gc, self.vel, self.pres, self.den, self.visc, self.fce, vel_const=False,
                          den_const="den" in self.const, visc_const="visc" in self.const, fce_const="fce" in self.const)
        bnd_nid = [self.mesh.bnd_nodes(b) for b in range(self.mesh.num_bnd)]
        r, c, v, b = O.assemble_flow(g, gc, f, self.bnd_vel, bnd_nid, self.pref, dt, vel_prev, den_prev)
        rp, col, val = O.coo_to_csr(r, c, v, 4 * self.N)
        return g, gc, f, (rp, col, val, b)


def lid_cavity(mesh, rho=1.0, mu=0.01, lid=1.0, pref=(0, 0.0)):
    """Lid-driven cavity (flow_lid.rs): no-slip walls, lid velocity (lid, 0, 0) on z = 1,
    zero body force, reference pressure at node pref[0]; zero initial fields."""
    N = mesh.num_tot
    bnd = []
    for b in range(6):
        cnt = len(mesh.bnd_nodes(b))
        v = np.zeros((3, cnt))
        if b == Z1:
            v[0] = lid
        bnd.append((b, v))
    return FlowProblem(mesh, np.zeros((3, N)), 0.0, rho, mu, np.zeros((3, N)), bnd, pref)


def flow_general(mesh, seed=0):
    """Flow data with every term active: smooth non-zero velocity and pressure, variable
    density, viscosity and body force (function fields), non-zero wall velocities."""
    rng = np.random.default_rng(seed)
    pos = mesh.get("pos")
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    N = mesh.num_tot
    vel = np.stack([np.sin(np.pi * x) * np.cos(np.pi * y) + 0.3 * z, -np.cos(np.pi * x) * np.sin(np.pi * y) + 0.2 * x * z,
                    0.5 * x * y - 0.1 * z * z])
    pres = np.cos(2.0 * x) + y * z - 0.5 * x * x
    den = 1.0 + 0.3 * x + 0.2 * y * z
    visc = 0.02 * (1.0 + 0.5 * np.sin(3.0 * y) + 0.2 * x)
    fce = np.stack([0.1 + 0.2 * y, -0.3 * x * z, -1.0 + 0.4 * x * y])
    bnd = []
    for b in range(6):
        nodes = mesh.num_pts + mesh.bnd_nodes(b).astype(np.int64)
        bnd.append((b, vel[:, nodes] + 0.05 * rng.standard_normal((3, nodes.size))))
    return FlowProblem(mesh, vel, pres, den, visc, fce, bnd, pref=(3, 0.25), const=())


def heat_flow_c5like(mesh, transient=False):
    """Temperature carried by a prescribed flow (the heat part of SteadyHeatFlow /
    TransientHeatFlow): OpSclDomAdvection with weight rho c_p, conductivity k and the flow's
    density (all function fields), source, T = 1 at x = 0 and T = 0 at x = 1, zero flux on
    the other faces; transient: OpSclDomTimePrevious (weight rho c_p) on the points and
    OpSclBndTimeUnity on the flux boundaries."""
    pos = mesh.get("pos")
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    p = Problem(mesh)
    T = p.scldom(vfvm.FIELD_UNKNOWN, 0.0)
    rcp = p.scldom(vfvm.FIELD_FUNCTION, 2.0 + 0.5 * np.sin(2.0 * x) * (1.0 + y))
    k = p.scldom(vfvm.FIELD_FUNCTION, 0.05 * (1.0 + 0.3 * z + 0.2 * x * y))
    q = p.scldom(vfvm.FIELD_FUNCTION, 0.3 * np.cos(3.0 * y))
    den = p.scldom(vfvm.FIELD_FUNCTION, 1.0 + 0.4 * x * z)
    visc = p.scldom(vfvm.FIELD_CONSTANT, 1.0)
    pres = p.scldom(vfvm.FIELD_CONSTANT, 0.0)
    vel = p.vecdom(vfvm.FIELD_FUNCTION, velocity_cellular(pos, 3.0))
    fce = p.vecdom(vfvm.FIELD_CONSTANT, np.zeros((3, p.N)))
    f = p.sclflw(vel, pres, den, visc, fce)
    t0 = p.sclbnd(X0, 1.0)
    t1 = p.sclbnd(X1, 0.0)
    fl = {b: p.sclbnd(b, 0.0) for b in (Y0, Y1, Z0, Z1)}
    ops = []
    if transient:
        ops.append((vfvm.OP_SCL_DOM_TIME_PREVIOUS, 0, T, rcp, -1))
    ops += [(vfvm.OP_SCL_DOM_DIFFUSION, 0, T, k, T), (vfvm.OP_SCL_DOM_SOURCE, 0, T, q, -1),
            (vfvm.OP_SCL_DOM_ADVECTION, 0, T, rcp, k, f),
            (vfvm.OP_SCL_BND_DIRICHLET, X0, T, t0, -1), (vfvm.OP_SCL_BND_DIRICHLET, X1, T, t1, -1)]
    for b in (Y0, Y1, Z0, Z1):
        if transient:
            ops.append((vfvm.OP_SCL_BND_TIME_UNITY, b, T, -1, -1))
        ops += [(vfvm.OP_SCL_BND_DIFFUSION, b, T, k, T), (vfvm.OP_SCL_BND_SOURCE, b, T, q, -1),
                (vfvm.OP_SCL_BND_NEUMANN, b, T, fl[b], -1), (vfvm.OP_SCL_BND_ADVECTION, b, T, rcp, k, f)]
    p.ops = ops
    p.unknown, p.rcp_id = T, rcp
    return p


class HeatFlowProblem:
    """SteadyHeatFlow / TransientHeatFlow / TransientHeatMassFlow composition
    (steady_heatflow.rs:154-536, transient_heatmassflow.rs:248-600) of one domain: the flow of
    a FlowProblem and a temperature unknown carried by the flow's own Rhie-Chow flux
    (OpSclDomAdvection with weight rho c_p, conductivity k and the flow's density), with a
    heat source q, T = t_hot on x = 0 and t_cold on x = 1, zero heat flux on the other faces
    (diffusion, source, Neumann and advection operators there), and a body force
    fce = (0, 0, gamma (T - t_ref)) evaluated on the host before every assembly
    (Boussinesq).  species: a concentration unknown C after T (diffusivity D, the flow's flux
    through OpSclDomAdvectionUnity, C = 1 on x = 0 and 0 on x = 1, zero flux elsewhere).
    transient: the velocity time operator first, OpSclDomTimePrevious / OpSclBndTimePrevious
    (weight rho c_p for T, the density for C) ahead of the scalar operators of each row class.
    Ids live in the dict `hid` once the fields are uploaded (or assigned for the oracle)."""

    def __init__(self, fp: FlowProblem, rcp, k, q, gamma=0.0, t_ref=0.5, t_hot=1.0, t_cold=0.0, T0=0.0, transient=False,
                 species=False, D=0.1, C0=0.0):
        self.fp = fp
        self.mesh = fp.mesh
        N = self.N = fp.N
        b = lambda v: np.broadcast_to(np.asarray(v, np.float64), (N,)).copy()
        self.rcp, self.k, self.q, self.T0 = b(rcp), b(k), b(q), b(T0)
        self.D, self.C0 = b(D), b(C0)
        self.gamma, self.t_ref, self.t_hot, self.t_cold = float(gamma), float(t_ref), float(t_hot), float(t_cold)
        self.transient = transient
        self.species = species
        self.nscl = 2 if species else 1
        fp.transient = transient
        fp.const = fp.const - {"fce"}  # the body force is a function field

    def fce_of(self, S):
        """Body force from the scalar unknowns S (T first)."""
        T = S[:self.N]
        return np.stack([np.zeros(self.N), np.zeros(self.N), self.gamma * (T - self.t_ref)])

    def host_ids(self):
        """Ids as to_device assigns them after FlowProblem.to_device's fields (oracle-only use)."""
        self.hid = dict(T=2, rcp=3, k=4, q=5, flw=0, den=1, t0=0, t1=1, fl={Y0: 2, Y1: 3, Z0: 4, Z1: 5}, C=6, D=7, c0=6,
                        c1=7, cfl={Y0: 8, Y1: 9, Z0: 10, Z1: 11})
        return self.hid

    def heat_ops(self, h):
        T, rcp, k, q, flw = h["T"], h["rcp"], h["k"], h["q"], h["flw"]
        o = [(vfvm.OP_SCL_DOM_TIME_PREVIOUS, 0, T, rcp, -1)] if self.transient else []
        o += [(vfvm.OP_SCL_DOM_DIFFUSION, 0, T, k, T), (vfvm.OP_SCL_DOM_SOURCE, 0, T, q, -1),
              (vfvm.OP_SCL_DOM_ADVECTION, 0, T, rcp, k, flw),
              (vfvm.OP_SCL_BND_DIRICHLET, X0, T, h["t0"], -1), (vfvm.OP_SCL_BND_DIRICHLET, X1, T, h["t1"], -1)]
        for bb in (Y0, Y1, Z0, Z1):
            if self.transient:
                o.append((vfvm.OP_SCL_BND_TIME_PREVIOUS, bb, T, rcp, -1))
            o += [(vfvm.OP_SCL_BND_DIFFUSION, bb, T, k, T), (vfvm.OP_SCL_BND_SOURCE, bb, T, q, -1),
                  (vfvm.OP_SCL_BND_NEUMANN, bb, T, h["fl"][bb], -1), (vfvm.OP_SCL_BND_ADVECTION, bb, T, rcp, k, flw)]
        return o

    def species_ops(self, h):
        C, D, flw, den = h["C"], h["D"], h["flw"], h["den"]
        o = [(vfvm.OP_SCL_DOM_TIME_PREVIOUS, 0, C, den, -1)] if self.transient else []
        o += [(vfvm.OP_SCL_DOM_DIFFUSION, 0, C, D, C), (vfvm.OP_SCL_DOM_ADVECTION_UNITY, 0, C, D, flw),
              (vfvm.OP_SCL_BND_DIRICHLET, X0, C, h["c0"], -1), (vfvm.OP_SCL_BND_DIRICHLET, X1, C, h["c1"], -1)]
        for bb in (Y0, Y1, Z0, Z1):
            if self.transient:
                o.append((vfvm.OP_SCL_BND_TIME_PREVIOUS, bb, C, den, -1))
            o += [(vfvm.OP_SCL_BND_DIFFUSION, bb, C, D, C), (vfvm.OP_SCL_BND_NEUMANN, bb, C, h["cfl"][bb], -1),
                  (vfvm.OP_SCL_BND_ADVECTION_UNITY, bb, C, D, flw)]
        return o

    def to_device(self, ctx: vfvm.Context):
        """Uploads every field; returns the coupled operator list in the reference's
        assemble_matrix order (flow, heat, species, then the reference pressure)."""
        fops = self.fp.to_device(ctx)
        self.fp.fce = self.fce_of(self.T0)
        ctx.vecdom_set(self.fp.ids[4], self.fp.fce)
        h = dict(flw=self.fp.ids[5], den=self.fp.ids[2])
        h["T"] = ctx.scldom_add(vfvm.FIELD_UNKNOWN, 0.0, self.T0)
        h["rcp"] = ctx.scldom_add(vfvm.FIELD_FUNCTION, 0.0, self.rcp)
        h["k"] = ctx.scldom_add(vfvm.FIELD_FUNCTION, 0.0, self.k)
        h["q"] = ctx.scldom_add(vfvm.FIELD_FUNCTION, 0.0, self.q)
        cnt = lambda bb: len(self.mesh.bnd_nodes(bb))
        h["t0"] = ctx.sclbnd_add(X0, values=np.full(cnt(X0), self.t_hot))
        h["t1"] = ctx.sclbnd_add(X1, values=np.full(cnt(X1), self.t_cold))
        h["fl"] = {bb: ctx.sclbnd_add(bb, values=np.zeros(cnt(bb))) for bb in (Y0, Y1, Z0, Z1)}
        sops = []
        if self.species:
            h["C"] = ctx.scldom_add(vfvm.FIELD_UNKNOWN, 0.0, self.C0)
            h["D"] = ctx.scldom_add(vfvm.FIELD_FUNCTION, 0.0, self.D)
            h["c0"] = ctx.sclbnd_add(X0, values=np.ones(cnt(X0)))
            h["c1"] = ctx.sclbnd_add(X1, values=np.zeros(cnt(X1)))
            h["cfl"] = {bb: ctx.sclbnd_add(bb, values=np.zeros(cnt(bb))) for bb in (Y0, Y1, Z0, Z1)}
            sops = self.species_ops(h)
        self.hid = h
        pref = [o for o in fops if o[0] == vfvm.OP_SCL_DOM_PRESSURE_REFERENCE]
        rest = [o for o in fops if o[0] != vfvm.OP_SCL_DOM_PRESSURE_REFERENCE]
        return rest + self.heat_ops(h) + sops + pref

    def hook(self, ctx):
        """The host closure: fce := fce_of(T) from the device's current temperature."""
        fce_id, T_id = self.fp.ids[4], self.hid["T"]
        return lambda t: ctx.vecdom_set(fce_id, self.fce_of(ctx.scldom_get(T_id)))

    def scalar_values(self, ctx):
        """The scalar unknowns' current values stacked [T, C]."""
        S = [ctx.scldom_get(self.hid["T"])]
        if self.species:
            S.append(ctx.scldom_get(self.hid["C"]))
        return np.concatenate(S)

    # ---- oracle
    def S0(self):
        return np.concatenate([self.T0, self.C0]) if self.species else self.T0.copy()

    def oracle_heat_rows(self, g, dt=0.0):
        """rows(S, S_prev, f) -> (rows, cols, vals, b) of the scalar unknowns' rows (T block, then
        C block at offset N) for O.heatflow_solve and the coupled assembly."""
        h = self.hid
        N = self.N
        bnd_nid = [self.mesh.bnd_nodes(bb) for bb in range(self.mesh.num_bnd)]
        gr = lambda v: (v, O.update_grad(g, v))
        fixed = {h["rcp"]: gr(self.rcp), h["k"]: gr(self.k), h["q"]: gr(self.q)}
        sb = {h["t0"]: np.full(len(bnd_nid[X0]), self.t_hot), h["t1"]: np.full(len(bnd_nid[X1]), self.t_cold)}
        sb.update({h["fl"][bb]: np.zeros(len(bnd_nid[bb])) for bb in (Y0, Y1, Z0, Z1)})
        hops = [O.Op(*o) for o in self.heat_ops(h)]
        if self.species:
            fixed[h["D"]] = gr(self.D)
            sb.update({h["c0"]: np.ones(len(bnd_nid[X0])), h["c1"]: np.zeros(len(bnd_nid[X1]))})
            sb.update({h["cfl"][bb]: np.zeros(len(bnd_nid[bb])) for bb in (Y0, Y1, Z0, Z1)})
            cops = [O.Op(*o) for o in self.species_ops(h)]

        def rows(S, S_prev, f):
            flw = h["flw"]
            scl = dict(fixed)
            scl[h["T"]] = gr(S[:N])
            scl[h["den"]] = (f["den"], f["deng"])
            prev = {h["T"]: S_prev[:N], h["rcp"]: self.rcp, h["den"]: f["den"]}
            r, c, v, b = O.assemble(g, hops, scl, sb, {flw: f["flux"]}, bnd_nid, dt=dt, value_prev=prev,
                                    flow_den={flw: (f["den"], f["deng"])})
            if not self.species:
                return r, c, v, b
            scl[h["C"]] = gr(S[N:])
            prev[h["C"]] = S_prev[N:]
            r2, c2, v2, b2 = O.assemble(g, cops, scl, sb, {flw: f["flux"]}, bnd_nid, dt=dt, value_prev=prev,
                                        flow_den={flw: (f["den"], f["deng"])})
            return (np.concatenate([np.asarray(r), N + np.asarray(r2)]), np.concatenate([np.asarray(c), N + np.asarray(c2)]),
                    np.concatenate([np.asarray(v), np.asarray(v2)]), np.concatenate([b, b2]))
        return rows

    def affine_closure(self, ctx):
        """The body force as a device closure fce = (0, 0, gamma (T - t_ref)) of the temperature
        (in place of the host hook)."""
        N = self.N
        f0 = np.zeros((3, N))
        f0[2] = -self.gamma * self.t_ref
        df = np.zeros((3, N))
        df[2] = self.gamma
        ctx.vecdom_func_affine(self.fp.ids[4], self.hid["T"], f0, df)

    def dfce(self):
        """d fce / d T of the Boussinesq closure (3, N)."""
        df = np.zeros((3, self.N))
        df[2] = self.gamma
        return df

    def oracle_residual(self, g, gc, dt=0.0):
        """R(x, x_prev) of the coupled system (O.coupled_residual) -> (R, A, f)."""
        fp = self.fp
        bnd_nid = [self.mesh.bnd_nodes(bb) for bb in range(self.mesh.num_bnd)]
        rows = self.oracle_heat_rows(g, dt)
        return lambda x, xp=None: O.coupled_residual(g, gc, x, xp, fp.den, fp.visc, self.fce_of, fp.bnd_vel, bnd_nid, fp.pref,
                                                      rows, dt, den_const="den" in fp.const, visc_const="visc" in fp.const)

    def oracle_jt(self, g, gc, dt=0.0):
        """(x, x_prev, lam) -> J(x)^T lam of the coupled system (O.flow_jt)."""
        res = self.oracle_residual(g, gc, dt)
        h = self.hid
        N = self.N
        bnd_nid = [self.mesh.bnd_nodes(bb) for bb in range(self.mesh.num_bnd)]
        gr = lambda v: (v, O.update_grad(g, v))
        scl = {h["rcp"]: gr(self.rcp), h["k"]: gr(self.k), h["q"]: gr(self.q)}
        if self.species:
            scl[h["D"]] = gr(self.D)
        pref_row = 3 * N + self.mesh.num_pts + int(self.fp.pref[0]) if self.fp.pref is not None else -1

        def jt(x, xp, lam):
            _, A, f = res(x, xp)
            fd = {h["flw"]: (f["den"], f["deng"])}
            blocks = [O.scalar_adv_terms(g, [O.Op(*o) for o in self.heat_ops(h)], scl, fd, bnd_nid)]
            if self.species:
                blocks.append(O.scalar_adv_terms(g, [O.Op(*o) for o in self.species_ops(h)], scl, fd, bnd_nid))
            return O.flow_jt(g, gc, f, x, lam, A, pref_row, blocks, 0, self.dfce())
        return jt

    def oracle_t_dirichlet(self):
        NP, N = self.mesh.num_pts, self.N
        n0 = NP + self.mesh.bnd_nodes(X0).astype(np.int64)
        n1 = NP + self.mesh.bnd_nodes(X1).astype(np.int64)

        def load(S):
            S = S.copy()
            S[n0] = self.t_hot
            S[n1] = self.t_cold
            if self.species:
                S[N + n0] = 1.0
                S[N + n1] = 0.0
            return S
        return load



# --------------------------------------------------------------------------- closed forms
def advdiff_1d(mesh, Pe):
    """Uniform flow u = (Pe, 0, 0) with unit diffusivity (global Peclet number Pe), c = 0 at
    x = 0 and c = 1 at x = 1, zero diffusive flux on the four other faces, in the SteadyMassFlow
    operator order (steady_massflow.rs:455-545).  The exact solution is the 1D profile
    c = (exp(Pe x) - 1) / (exp(Pe) - 1) of D c'' = U c'."""
    p = Problem(mesh)
    c = p.scldom(vfvm.FIELD_UNKNOWN, 0.0)
    D = p.scldom(vfvm.FIELD_CONSTANT, 1.0)
    den = p.scldom(vfvm.FIELD_CONSTANT, 1.0)
    visc = p.scldom(vfvm.FIELD_CONSTANT, 1.0)
    pres = p.scldom(vfvm.FIELD_CONSTANT, 0.0)
    vel = p.vecdom(vfvm.FIELD_CONSTANT, np.stack([np.full(p.N, float(Pe)), np.zeros(p.N), np.zeros(p.N)]))
    fce = p.vecdom(vfvm.FIELD_CONSTANT, np.zeros((3, p.N)))
    f = p.sclflw(vel, pres, den, visc, fce)
    c0, c1 = p.sclbnd(X0, 0.0), p.sclbnd(X1, 1.0)
    q = {b: p.sclbnd(b, 0.0) for b in (Y0, Y1, Z0, Z1)}
    p.ops = [(vfvm.OP_SCL_DOM_DIFFUSION, 0, c, D, c), (vfvm.OP_SCL_DOM_ADVECTION_UNITY, 0, c, D, f),
             (vfvm.OP_SCL_BND_DIRICHLET, X0, c, c0, -1), (vfvm.OP_SCL_BND_DIRICHLET, X1, c, c1, -1)]
    for b in (Y0, Y1, Z0, Z1):
        p.ops += [(vfvm.OP_SCL_BND_DIFFUSION, b, c, D, c), (vfvm.OP_SCL_BND_NEUMANN, b, c, q[b], -1),
                  (vfvm.OP_SCL_BND_ADVECTION_UNITY, b, c, D, f)]
    p.unknown = c
    p.exact = np.expm1(Pe * p.pos[:, 0]) / np.expm1(Pe)
    return p


def random_cube(n_int, n_face, seed=0, device=None):
    """A mesh of the unit cube on random points (no lattice): n_int uniform interior points,
    n_face uniform points inside each face, 7 points along each edge and the 8 corners, joined by
    the Delaunay tetrahedra of scipy (Qhull); boundary groups x0, x1, y0, y1, z0, z1 (an edge or
    corner node in the first face it lies on)."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    inner = rng.uniform(0.02, 0.98, size=(n_int, 3))
    nodes, grp = [], []
    for ax in range(3):
        for side in (0.0, 1.0):
            q = rng.uniform(0.02, 0.98, size=(n_face, 3))
            q[:, ax] = side
            nodes.append(q)
            grp += [2 * ax + int(side)] * n_face
    t = np.linspace(0.0, 1.0, 9)[1:-1]
    for a in range(3):
        b, c = [x for x in range(3) if x != a]
        for s1 in (0.0, 1.0):
            for s2 in (0.0, 1.0):
                q = np.zeros((len(t), 3))
                q[:, a] = t
                q[:, b] = s1
                q[:, c] = s2
                nodes.append(q)
                grp += [2 * b + int(s1)] * len(t)
    corners = np.array([[i, j, k] for i in (0.0, 1.0) for j in (0.0, 1.0) for k in (0.0, 1.0)])
    nodes.append(corners)
    grp += [int(q[0]) for q in corners]
    grp = np.array(grp)
    pos = np.concatenate([inner] + nodes)
    tets = Delaunay(pos).simplices.astype(np.int32)
    groups = [np.nonzero(grp == g)[0].astype(np.int32) for g in range(6)]
    return vfvm.Mesh.from_tets(pos, tets, n_int, groups, device=device)


def regroup_lattice(n, groups_of, seed=0, jitter=0.3, device=None):
    """The jittered lattice mesh with its nodes regrouped into boundaries: groups_of(pos of the
    nodes, lattice boundary of each node) returns a group index per node."""
    m = vfvm.Mesh.lattice(n, seed, jitter)
    pos = m.get("pos")
    nb = np.full(m.num_nds, -1)
    for b in range(m.num_bnd):
        nb[m.bnd_nodes(b)] = b
    gid = np.asarray(groups_of(pos[m.num_pts:], nb))
    groups = [np.nonzero(gid == k)[0].astype(np.int32) for k in range(int(gid.max()) + 1)]
    return vfvm.Mesh.from_tets(pos, m.get("tets"), m.num_pts, groups, device=device)


def neumann_linear_profile(n, k=2.0, flux=0.7):
    """T = 0 at x = 0, Neumann flux `flux` on the interior of the x = 1 face, zero flux on the
    four lateral faces, and the border of the x = 1 face held at the exact value: the solution
    of k lap T = 0 is T = -flux x / k (OpSclBndNeumann adds +A flux to b, so the flux is
    -k dT/dn), exact on the Voronoi dual because the scheme is exact for linear fields.
    Groups: 0 x0 (Dirichlet 0), 1 border of x = 1 (Dirichlet -flux/k), 2 interior of x = 1
    (Neumann flux), 3-6 y0, y1, z0, z1 (Neumann 0)."""
    def groups(pn, nb):
        g = np.where(nb == X0, 0, nb + 1)
        on_x1 = nb == X1
        border = on_x1 & ((pn[:, 1] == 0.0) | (pn[:, 1] == 1.0) | (pn[:, 2] == 0.0) | (pn[:, 2] == 1.0))
        g = np.where(on_x1, np.where(border, 1, 2), g)
        return g
    mesh = regroup_lattice(n, groups)
    p = Problem(mesh)
    T = p.scldom(vfvm.FIELD_UNKNOWN, 0.0)
    kk = p.scldom(vfvm.FIELD_CONSTANT, k)
    d0, d1 = p.sclbnd(0, 0.0), p.sclbnd(1, -flux / k)
    p.ops = [(vfvm.OP_SCL_DOM_DIFFUSION, 0, T, kk, T), (vfvm.OP_SCL_BND_DIRICHLET, 0, T, d0, -1),
             (vfvm.OP_SCL_BND_DIRICHLET, 1, T, d1, -1)]
    for b, q in ((2, flux), (3, 0.0), (4, 0.0), (5, 0.0), (6, 0.0)):
        p.ops += [(vfvm.OP_SCL_BND_DIFFUSION, b, T, kk, T), (vfvm.OP_SCL_BND_NEUMANN, b, T, p.sclbnd(b, q), -1)]
    p.unknown = T
    p.exact = -flux * p.pos[:, 0] / k
    return p


def neumann_surface_integral(n, coef):
    """Every node in one boundary group with a Neumann flux N = a + bx x + by y + bz z (linear):
    the boundary right-hand sides sum to the integral of N over the cube's surface,
    6 a + 3 (bx + by + bz), exactly (each sub-face weight is the barycentric coordinate of
    the sub-face centroid, so a sub-face adds the exact integral of a linear N over it, and
    the sub-faces of the three corners of a surface triangle tile it)."""
    def groups(pn, nb):
        return np.zeros(len(pn), np.int64)
    mesh = regroup_lattice(n, groups)
    p = Problem(mesh)
    T = p.scldom(vfvm.FIELD_UNKNOWN, 0.0)
    kk = p.scldom(vfvm.FIELD_CONSTANT, 1.0)
    a, bx, by, bz = coef
    pn = p.pos[mesh.num_pts + mesh.bnd_nodes(0)]
    q = p.sclbnd(0, a + bx * pn[:, 0] + by * pn[:, 1] + bz * pn[:, 2])
    p.ops = [(vfvm.OP_SCL_DOM_DIFFUSION, 0, T, kk, T), (vfvm.OP_SCL_BND_DIFFUSION, 0, T, kk, T),
             (vfvm.OP_SCL_BND_NEUMANN, 0, T, q, -1)]
    p.unknown = T
    p.exact_integral = 6.0 * a + 3.0 * (bx + by + bz)
    return p


REG_TET = np.array([[1.0, 1.0, 1.0], [1.0, -1.0, -1.0], [-1.0, 1.0, -1.0], [-1.0, -1.0, 1.0]])


def neumann_tetra(flux=(0.3, -1.1, 2.5, 0.7), groups=((0, 1), (2,), (3,))):
    """A single regular tetrahedron (edge 2 sqrt 2, face area 2 sqrt 3), all four vertices
    nodes, in boundary groups `groups`, with the Neumann fluxes `flux` per vertex.  Each face
    is equilateral, so its circumcentre is its centroid: every corner's sub-face has a third
    of the face area and the weights (11/18, 7/36, 7/36).  A neighbour on another boundary
    takes the vertex's own flux (opscl_bnd_neu.rs:73-80), so
    b_i = (A_f / 3) sum over the faces of i of (11/18 N_i + 7/36 (N~_j + N~_k))."""
    gid = np.zeros(4, np.int64)
    for k, gr in enumerate(groups):
        gid[list(gr)] = k
    mesh = vfvm.Mesh.from_tets(REG_TET, np.array([[0, 1, 2, 3]], np.int32), 0,
                               [np.array(gr, np.int32) for gr in groups])
    p = Problem(mesh)
    T = p.scldom(vfvm.FIELD_UNKNOWN, 0.0)
    kk = p.scldom(vfvm.FIELD_CONSTANT, 1.0)
    fl = np.asarray(flux, np.float64)
    p.ops = [(vfvm.OP_SCL_DOM_DIFFUSION, 0, T, kk, T)]
    for k, gr in enumerate(groups):
        nodes = mesh.bnd_nodes(k)
        p.ops += [(vfvm.OP_SCL_BND_DIFFUSION, k, T, kk, T), (vfvm.OP_SCL_BND_NEUMANN, k, T, p.sclbnd(k, fl[nodes]), -1)]
    p.unknown = T
    area_f = 2.0 * np.sqrt(3.0)
    exact = np.zeros(4)
    for i in range(4):
        for face in ([0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]):
            if i not in face:
                continue
            j, k2 = [x for x in face if x != i]
            nj = fl[j] if gid[j] == gid[i] else fl[i]
            nk = fl[k2] if gid[k2] == gid[i] else fl[i]
            exact[i] += area_f / 3.0 * (11.0 / 18.0 * fl[i] + 7.0 / 36.0 * (nj + nk))
    p.exact_b = exact
    return p


def poiseuille(mesh, U=1.0, mu=1.0, rho=1.0):
    """Plane Poiseuille flow between z = 0 and z = 1 through the whole Navier-Stokes operator
    set (the closed form of examples/flow/flow_analytic.rs:10): u = (4 U z (1 - z), 0, 0) held
    on every face, no body force; the exact fields are that velocity and p = -8 mu U x (zero at
    x = 0), the reference pressure pinned at node 0 to its exact value."""
    pos = mesh.get("pos")
    N = mesh.num_tot
    uex = np.stack([4.0 * U * pos[:, 2] * (1.0 - pos[:, 2]), np.zeros(N), np.zeros(N)])
    pex = -8.0 * mu * U * pos[:, 0]
    bnd = []
    for b in range(6):
        nodes = mesh.num_pts + mesh.bnd_nodes(b).astype(np.int64)
        bnd.append((b, uex[:, nodes]))
    nid0 = mesh.num_pts
    p = FlowProblem(mesh, np.zeros((3, N)), 0.0, rho, mu, np.zeros((3, N)), bnd, pref=(0, float(pex[nid0])))
    p.uex, p.pex = uex, pex
    return p


def suction_profile(mesh, U=1.0, V=1.0, mu=0.2, rho=1.0):
    """Asymptotic suction profile: a shear flow u(z) between z = 0 (u = 0) and z = 1 (u = U)
    with a uniform cross-flow w = -V through both walls; the momentum balance rho w u' = mu u''
    gives u = U (1 - exp(-V z / nu)) / (1 - exp(-V / nu)), nu = mu / rho, at constant pressure.
    The exact velocity is held on every face.  Pins the upwind side of the momentum advection
    (OpVecDomAdvection, opvec_dom_adv.rs:47-116): a flipped sign convects the layer the wrong way."""
    pos = mesh.get("pos")
    N = mesh.num_tot
    nu = mu / rho
    uex = np.stack([U * -np.expm1(-V * pos[:, 2] / nu) / -np.expm1(-V / nu), np.zeros(N), np.full(N, -V)])
    bnd = []
    for b in range(6):
        nodes = mesh.num_pts + mesh.bnd_nodes(b).astype(np.int64)
        bnd.append((b, uex[:, nodes]))
    p = FlowProblem(mesh, np.zeros((3, N)), 0.0, rho, mu, np.zeros((3, N)), bnd, pref=(0, 0.0))
    p.uex, p.pex = uex, np.zeros(N)
    return p


# ------------------------------------------------------------------ multi-domain systems
class MultiProblem:
    """Fields and operators of a multi-domain system (several domains joined by interfaces):
    the joined mesh and joined-mesh fields for the device path, the per-domain duals, fields and
    operator lists for the oracle (assemble_domains), which restates the reference's per-domain
    variables and its interface operators."""

    def __init__(self, parts, itfs):
        self.parts = list(parts)
        self.itfs = [tuple(f) for f in itfs]
        self.mesh = vfvm.Mesh.join(self.parts, self.itfs)
        self.N = self.mesh.num_tot
        self.pos = self.mesh.get("pos")
        self.dom_ptr = self.mesh.get("dom_ptr")
        side = {(f[0], f[1]) for f in self.itfs} | {(f[2], f[3]) for f in self.itfs}
        self.jbnd = {}  # (domain, local boundary) -> joined boundary id
        for d, m in enumerate(self.parts):
            for b in range(m.num_bnd):
                if (d, b) not in side:
                    self.jbnd[(d, b)] = len(self.jbnd)
        self.lbnd = {j: db for db, j in self.jbnd.items()}
        self.scl, self.sbnd, self.vec, self.flw, self.ops = [], [], [], [], []

    def dom_slice(self, d):
        return slice(int(self.dom_ptr[d]), int(self.dom_ptr[d + 1]))

    def scldom(self, kind, per_dom):
        """per_dom: one value (or array over the domain's vertices) per domain, or a callable
        of the joined positions."""
        if callable(per_dom):
            v = np.asarray(per_dom(self.pos), np.float64)
        else:
            v = np.concatenate([np.broadcast_to(np.asarray(x, np.float64), (m.num_tot,)) for x, m in zip(per_dom, self.parts)])
        self.scl.append((kind, v.copy()))
        return len(self.scl) - 1

    def sclbnd(self, d, b, values):
        j = self.jbnd[(d, b)]
        cnt = len(self.parts[d].bnd_nodes(b))
        self.sbnd.append((j, np.broadcast_to(np.asarray(values, np.float64), (cnt,)).copy()))
        return len(self.sbnd) - 1

    def vecdom(self, kind, values):
        self.vec.append((kind, np.asarray(values, np.float64).reshape(3, self.N).copy()))
        return len(self.vec) - 1

    def sclflw(self, vel, pres, den, visc, fce):
        self.flw.append((vel, pres, den, visc, fce))
        return len(self.flw) - 1

    def to_device(self, ctx):
        ctx.upload(self.mesh)
        for kind, v in self.scl:
            ctx.scldom_add(kind, 0.0, v)
        for j, v in self.sbnd:
            ctx.sclbnd_add(j, 0.0, v)
        for kind, v in self.vec:
            ctx.vecdom_add(kind, values=v)
        for f in self.flw:
            ctx.sclflw_add(*f)
        return ctx

    # ---- oracle
    def oracle_system(self, dt=0.0, value_prev=None, op_filter=None, unknown_value=None):
        """CSR (row_ptr, col, val), b and block starts of the oracle's per-domain assembly.
        value_prev: field id -> joined-mesh previous values; unknown_value: the unknown's
        values (joined mesh) the operators read (lagged terms of a Picard iteration)."""
        if getattr(self, "doms", None) is None:
            self.doms = [O.dual(m.get("pos"), m.get("tets"), m.num_pts) for m in self.parts]
        doms = self.doms
        sl = [self.dom_slice(d) for d in range(len(self.parts))]
        fields = {}
        for i, (kind, v) in enumerate(self.scl):
            if unknown_value is not None and kind == vfvm.FIELD_UNKNOWN:
                v = np.asarray(unknown_value, np.float64)
            fields[i] = {}
            for d, g in enumerate(doms):
                vd = v[sl[d]]
                gr = O.update_grad(g, vd) if kind != vfvm.FIELD_CONSTANT else np.zeros((3, g["N"]))
                fields[i][d] = (vd, gr)
        flux, den = {}, {}
        for i, (vel, pres, dn, visc, fce) in enumerate(self.flw):
            vk, vv = self.vec[vel]
            flux[i], den[i] = {}, {}
            for d, g in enumerate(doms):
                vd = vv[:, sl[d]]
                vg = O.update_grad_vec(g, vd) if vk != vfvm.FIELD_CONSTANT else np.zeros((9, g["N"]))
                flux[i][d] = O.flux_prescribed(g, vd, vg, fields[dn][d][0], fields[dn][d][1])
                den[i][d] = fields[dn][d]
        prev = None
        if value_prev is not None:
            prev = {fid: {d: np.asarray(v, np.float64)[sl[d]] for d in range(len(doms))} for fid, v in value_prev.items()}
        sclbnd = [{i: v for i, (j, v) in enumerate(self.sbnd) if self.lbnd[j][0] == d} for d in range(len(doms))]
        bnd_nids = [[m.bnd_nodes(b) for b in range(m.num_bnd)] for m in self.parts]
        dom_ops = [[] for _ in doms]
        itf_ops = [[] for _ in self.itfs]
        for o in self.ops:
            if op_filter is not None and not op_filter(o):
                continue
            kind, geom, unk, a, b, c = o[:6]
            value = o[6] if len(o) > 6 else 0.0
            if kind < 20:
                dom_ops[geom].append(O.Op(kind, 0, unk, a, b, c, value))
            elif kind < 40:
                d, lb = self.lbnd[geom]
                dom_ops[d].append(O.Op(kind, lb, unk, a, b, c, value))
            elif kind >= 70:
                itf_ops[geom].append(O.Op(kind, geom, unk, a, b, c, value))
            else:
                raise ValueError(f"operator kind {kind} on a joined mesh")
        va, vb = self.mesh.get("itf_va"), self.mesh.get("itf_vb")
        iptr = self.mesh.get("itf_ptr")
        itfs = []
        for k, (da, _, db, _) in enumerate(self.itfs):
            q = slice(int(iptr[k]), int(iptr[k + 1]))
            itfs.append((da, db, va[q].astype(np.int64) - self.dom_ptr[da], vb[q].astype(np.int64) - self.dom_ptr[db]))
        return O.assemble_domains(doms, bnd_nids, dom_ops, itfs, itf_ops, fields, sclbnd, dt=dt, value_prev=prev,
                                  flux=flux, flow_den=den)


def slab_domains(n, splits, seed=0, jitter=0.3, device=None):
    """The n^3 lattice cut at the x lattice planes `splits` into boxes joined by interfaces:
    domain d spans i in [s_d, s_(d+1)]; interface k joins domain k's x1 face (side a) to
    domain k+1's x0 face (side b)."""
    cuts = [0] + list(splits) + [n - 1]
    parts = [vfvm.Mesh.lattice_box(n, (cuts[d], 0, 0), (cuts[d + 1], n - 1, n - 1), seed, jitter, device=device)
             for d in range(len(cuts) - 1)]
    itfs = [(k, X1, k + 1, X0) for k in range(len(cuts) - 2)]
    return parts, itfs


def itf_heat(n, splits=None, k=(1.0, 3.0), q=0.0, prd=0.0, seed=0, device=None):
    """Steady heat through slabs of different conductivity (SteadyHeat with add_hcnt_itf,
    steady_heat.rs:100-171, as feat_itf.rs sets it up): T = 0 on the first slab's x = 0 face,
    T = 1 on the last slab's x = 1 face, zero flux on the other outer faces, temperature and
    flux continuity at every interface.  k, q, prd: one value per slab or callables of the
    joined positions."""
    if splits is None:
        splits = [(n - 1) // 2]
    parts, itfs = slab_domains(n, splits, seed, device=device)
    p = MultiProblem(parts, itfs)
    nd = len(parts)
    per = lambda x: x if callable(x) else (list(x) if np.ndim(x) else [x] * nd)
    kind = lambda x: vfvm.FIELD_FUNCTION if callable(x) else vfvm.FIELD_CONSTANT
    T = p.scldom(vfvm.FIELD_UNKNOWN, [0.0] * nd)
    kk = p.scldom(kind(k), per(k))
    hp = p.scldom(kind(prd), per(prd))
    hs = p.scldom(kind(q), per(q))
    ops = [(vfvm.OP_SCL_DOM_DIFFUSION, d, T, kk, T, -1) for d in range(nd)]
    for d in range(nd):
        ops += [(vfvm.OP_SCL_DOM_PRODUCTION, d, T, hp, -1, -1), (vfvm.OP_SCL_DOM_SOURCE, d, T, hs, -1, -1)]
    ops.append((vfvm.OP_SCL_BND_DIRICHLET, p.jbnd[(0, X0)], T, p.sclbnd(0, X0, 0.0), -1, -1))
    ops.append((vfvm.OP_SCL_BND_DIRICHLET, p.jbnd[(nd - 1, X1)], T, p.sclbnd(nd - 1, X1, 1.0), -1, -1))
    for d in range(nd):
        for b in (Y0, Y1, Z0, Z1):
            j = p.jbnd[(d, b)]
            ops += [(vfvm.OP_SCL_BND_DIFFUSION, j, T, kk, T, -1), (vfvm.OP_SCL_BND_PRODUCTION, j, T, hp, -1, -1),
                    (vfvm.OP_SCL_BND_SOURCE, j, T, hs, -1, -1), (vfvm.OP_SCL_BND_NEUMANN, j, T, p.sclbnd(d, b, 0.0), -1, -1)]
    for i in range(len(itfs)):
        ops += [(vfvm.OP_SCL_ITF_CONTINUITY, i, T, -1, -1, -1), (vfvm.OP_SCL_ITF_DIFFUSION, i, T, kk, T, -1),
                (vfvm.OP_SCL_ITF_PRODUCTION, i, T, hp, -1, -1), (vfvm.OP_SCL_ITF_SOURCE, i, T, hs, -1, -1)]
    p.ops = ops
    p.unknown = T
    p.fields = dict(k=kk, prd=hp, src=hs)
    return p


def itf_advdiff(n, Pe, splits=None, seed=0, device=None, weighted=False):
    """The 1D convection-diffusion profile c = (exp(Pe x) - 1) / (exp(Pe) - 1) through slabs
    joined by interfaces (SteadyMassFlow with add_mcnt_itf and a prescribed flow in every slab,
    steady_massflow.rs:300-345): uniform u = (Pe, 0, 0), unit diffusivity, the interface
    advection on both sides (the flow crosses the interface).  weighted: the temperature form
    (OpSclDomAdvection / OpSclItfAdvection with weight = density, SteadyHeatFlow)."""
    if splits is None:
        splits = [(n - 1) // 2]
    parts, itfs = slab_domains(n, splits, seed, device=device)
    p = MultiProblem(parts, itfs)
    nd = len(parts)
    c = p.scldom(vfvm.FIELD_UNKNOWN, [0.0] * nd)
    D = p.scldom(vfvm.FIELD_CONSTANT, [1.0] * nd)
    den = p.scldom(vfvm.FIELD_CONSTANT, [1.0] * nd)
    visc = p.scldom(vfvm.FIELD_CONSTANT, [1.0] * nd)
    pres = p.scldom(vfvm.FIELD_CONSTANT, [0.0] * nd)
    vel = p.vecdom(vfvm.FIELD_CONSTANT, np.stack([np.full(p.N, float(Pe)), np.zeros(p.N), np.zeros(p.N)]))
    fce = p.vecdom(vfvm.FIELD_CONSTANT, np.zeros((3, p.N)))
    f = p.sclflw(vel, pres, den, visc, fce)
    if weighted:
        adv = lambda kind, g: (kind, g, c, den, D, f)
        dom_adv, bnd_adv, itf_adv = vfvm.OP_SCL_DOM_ADVECTION, vfvm.OP_SCL_BND_ADVECTION, vfvm.OP_SCL_ITF_ADVECTION
    else:
        adv = lambda kind, g: (kind, g, c, D, f, -1)
        dom_adv, bnd_adv, itf_adv = vfvm.OP_SCL_DOM_ADVECTION_UNITY, vfvm.OP_SCL_BND_ADVECTION_UNITY, vfvm.OP_SCL_ITF_ADVECTION_UNITY
    ops = []
    for d in range(nd):
        ops += [(vfvm.OP_SCL_DOM_DIFFUSION, d, c, D, c, -1), adv(dom_adv, d)]
    ops.append((vfvm.OP_SCL_BND_DIRICHLET, p.jbnd[(0, X0)], c, p.sclbnd(0, X0, 0.0), -1, -1))
    ops.append((vfvm.OP_SCL_BND_DIRICHLET, p.jbnd[(nd - 1, X1)], c, p.sclbnd(nd - 1, X1, 1.0), -1, -1))
    for d in range(nd):
        for b in (Y0, Y1, Z0, Z1):
            j = p.jbnd[(d, b)]
            ops += [(vfvm.OP_SCL_BND_DIFFUSION, j, c, D, c, -1), (vfvm.OP_SCL_BND_NEUMANN, j, c, p.sclbnd(d, b, 0.0), -1, -1),
                    adv(bnd_adv, j)]
    for i in range(len(itfs)):
        ops += [(vfvm.OP_SCL_ITF_CONTINUITY, i, c, -1, -1, -1), (vfvm.OP_SCL_ITF_DIFFUSION, i, c, D, c, -1)]
        for side in (0, 1):
            ops.append(adv(itf_adv, i) + (float(side),))
    p.ops = ops
    p.unknown = c
    p.exact = np.expm1(Pe * p.pos[:, 0]) / np.expm1(Pe)
    return p


def itf_transient(n, m=(1.0, 2.5), s=(-0.5, -2.0), dt=0.1, splits=None, seed=0, device=None, previous=False):
    """One implicit-Euler step of m dc/dt = s c on slabs with different m and s per slab and no
    exchange between control volumes (zero conductivity): every control volume integrates the
    ODE by itself, so an
    interface node's combined volume gives c = c_prev / (1 - dt (s_a V_a + s_b V_b) /
    (m_a V_a + m_b V_b)) (TransientHeat with add_hcnt_itf: time, diffusion, production per
    slab and interface).  previous: the time-previous operators (the weight's previous values)."""
    if splits is None:
        splits = [(n - 1) // 2]
    parts, itfs = slab_domains(n, splits, seed, device=device)
    p = MultiProblem(parts, itfs)
    nd = len(parts)
    c = p.scldom(vfvm.FIELD_UNKNOWN, [1.0] * nd)
    w = p.scldom(vfvm.FIELD_CONSTANT, list(m))
    k = p.scldom(vfvm.FIELD_CONSTANT, [0.0] * nd)  # no exchange between control volumes
    sp_ = p.scldom(vfvm.FIELD_CONSTANT, list(s))
    tk = (vfvm.OP_SCL_DOM_TIME_PREVIOUS, vfvm.OP_SCL_BND_TIME_PREVIOUS, vfvm.OP_SCL_ITF_TIME_PREVIOUS) if previous else \
         (vfvm.OP_SCL_DOM_TIME, vfvm.OP_SCL_BND_TIME, vfvm.OP_SCL_ITF_TIME)
    ops = []
    for d in range(nd):
        ops += [(tk[0], d, c, w, -1, -1), (vfvm.OP_SCL_DOM_DIFFUSION, d, c, k, c, -1), (vfvm.OP_SCL_DOM_PRODUCTION, d, c, sp_, -1, -1)]
    for d in range(nd):
        for b in range(6):
            if (d, b) in p.jbnd:
                j = p.jbnd[(d, b)]
                ops += [(tk[1], j, c, w, -1, -1), (vfvm.OP_SCL_BND_DIFFUSION, j, c, k, c, -1),
                        (vfvm.OP_SCL_BND_PRODUCTION, j, c, sp_, -1, -1), (vfvm.OP_SCL_BND_NEUMANN, j, c, p.sclbnd(d, b, 0.0), -1, -1)]
    for i in range(len(itfs)):
        ops += [(tk[2], i, c, w, -1, -1), (vfvm.OP_SCL_ITF_CONTINUITY, i, c, -1, -1, -1),
                (vfvm.OP_SCL_ITF_DIFFUSION, i, c, k, c, -1), (vfvm.OP_SCL_ITF_PRODUCTION, i, c, sp_, -1, -1)]
    p.ops = ops
    p.unknown = c
    p.dt = dt
    p.m, p.s = tuple(m), tuple(s)
    return p
