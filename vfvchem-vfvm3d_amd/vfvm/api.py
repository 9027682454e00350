"""Host-side mirror of the reference's user API for the scalar transport path.

Names, argument order and error behaviour follow the reference:
  Variables            src/base/vars.rs:38-213
  SteadyHeat           src/steady/core/steady_heat.rs:9-183
  TransientHeat        src/transient/core/transient_heat.rs
  SteadyMass           src/steady/core/steady_mass.rs (one component)
  SteadyMassFlow       src/steady/core/steady_massflow.rs, mass part with a
                       prescribed velocity (add_flow_prescribed instead of an
                       unknown velocity)
  SteadyFlow           src/steady/core/steady_flow.rs:9-321 (velocity boundaries and
                       a reference pressure; pressure/traction boundaries and
                       interfaces are outside the built path)
  TransientFlow        src/transient/core/transient_flow.rs (the same, implicit Euler)
  SteadyHeatFlow       src/steady/core/steady_heatflow.rs:9-536 (one domain)
  TransientHeatMassFlow src/transient/core/transient_heatmassflow.rs (one domain, one
                       species component with self-diffusion)
  LinearGmres          src/linear/linear_gmres.rs:16-50
The operator lists these classes build are applied in the order of each
reference solver's assemble_matrix; the Picard loops run in the library
(vfvm_steady_solve / vfvm_transient_solve, mirroring steady_base.rs:16-135 and
transient_base.rs:16-160).
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .core import (OP_SCL_BND_ADVECTION, OP_SCL_BND_DIVERGENCE, OP_SCL_BND_PRESSURE_STABILIZER,
                   OP_SCL_BND_TIME_PREVIOUS, OP_SCL_DOM_ADVECTION, OP_SCL_DOM_DIVERGENCE, OP_SCL_DOM_PRESSURE_REFERENCE,
                   OP_SCL_DOM_PRESSURE_STABILIZER, OP_SCL_DOM_TIME_PREVIOUS, OP_VEC_BND_DIRICHLET, OP_VEC_DOM_ADVECTION,
                   OP_VEC_DOM_DIFFUSION, OP_VEC_DOM_DILATATORY, OP_VEC_DOM_GRADIENT, OP_VEC_DOM_SOURCE, OP_VEC_DOM_TIME,
                   OP_VEC_DOM_TRANSPOSE, PRECOND_BJACOBI)
from .core import (FIELD_CONSTANT, FIELD_FUNCTION, FIELD_UNKNOWN, FP32_REFINED, FP64, PRECOND_ILU0, PRECOND_JACOBI,
                   PRECOND_NONE, OP_SCL_BND_ADVECTION_UNITY, OP_SCL_BND_DIFFUSION, OP_SCL_BND_DIRICHLET,
                   OP_SCL_BND_NEUMANN, OP_SCL_BND_PRODUCTION, OP_SCL_BND_SOURCE, OP_SCL_BND_TIME,
                   OP_SCL_DOM_ADVECTION_UNITY, OP_SCL_DOM_DIFFUSION, OP_SCL_DOM_PRODUCTION, OP_SCL_DOM_SOURCE,
                   OP_SCL_DOM_TIME, Context, Mesh)
from ._lib import VFVChemError

BOUNDARY_NAMES = ("x0", "x1", "y0", "y1", "z0", "z1")


class Variables:
    """Mesh, fields and their device copies (Variables, src/base/vars.rs:7-34)."""

    def __init__(self, mesh: Mesh, device: int = 0):
        self.mesh = mesh
        self.ctx = Context(device)
        self.ctx.upload(mesh)
        self.pos = mesh.get("pos")
        self.dom = []
        self.bnd = []           # mesh boundary index per bnd_id
        self.scldom = []        # (kind, func, unkn_ids)
        self.sclbnd = []        # (bnd_id, func, unkn_ids)
        self.vecdom = []
        self.vecbnd = []
        self.sclflw = []
        self.outputs = []       # (kind, field id, file_path) of fields written by the drivers
        self.t_offset = 0.0     # time of the driver call's start (transient output in chunks)

    # convergence scaling (vars.rs:28-31, 203-213): reference scales of unknown fields (lower
    # bounds of their scales in the Picard residual and the adaptive error norm) and scale_off
    def set_scldom_scale(self, scldom_id: int, scale: float) -> None:
        self.ctx.set_scldom_scale(scldom_id, scale)

    def set_vecdom_scale(self, vecdom_id: int, scale: float) -> None:
        self.ctx.set_vecdom_scale(vecdom_id, scale)

    @property
    def scale_off(self) -> bool:
        return getattr(self, "_scale_off", False)

    @scale_off.setter
    def scale_off(self, off: bool) -> None:
        self._scale_off = bool(off)
        self.ctx.set_scale_off(self._scale_off)

    @staticmethod
    def new_lattice(n: int, seed: int = 0, jitter: float = 0.3, device: int = 0) -> "Variables":
        """Jittered unit-cube lattice (replaces Variables::new_bounds, vars.rs:45, for the lattice meshes);
        the Voronoi dual is built on the context's GPU (bitwise equal to the host build)."""
        return Variables(Mesh.lattice(n, seed, jitter, device=device), device)

    def add_dom(self, grp3d_id: str = "cube") -> int:
        if self.dom:
            raise VFVChemError(L.ERR_INVALID_ARGUMENT, "one domain per mesh on the built path")
        self.dom.append(grp3d_id)
        return 0

    def add_bnd(self, dom_id: int, grp2d_id: str) -> int:
        if grp2d_id not in BOUNDARY_NAMES or dom_id != 0:
            raise VFVChemError(L.ERR_INVALID_ARGUMENT, f"BoundaryGroupIDNotFound: {grp2d_id}")
        self.bnd.append(BOUNDARY_NAMES.index(grp2d_id))
        return len(self.bnd) - 1

    def _bnd_pos(self, bnd_id):
        nids = self.mesh.bnd_nodes(self.bnd[bnd_id])
        return self.pos[self.mesh.num_pts + nids]

    # scalar domain fields (vars.rs:80-102)
    def _output(self, kind: str, i: int, file_path: str) -> None:
        # file_path "" = no output; "<name>.csv" or "<name>.vtu" (scl_dom.rs:56-66, 132-138)
        if not file_path:
            return
        name, sep, typ = file_path.rpartition(".")
        if not sep:
            raise VFVChemError(L.ERR_INVALID_ARGUMENT, f"UnknownOutputType: {file_path}")
        if typ not in ("csv", "vtu"):
            raise VFVChemError(L.ERR_INVALID_ARGUMENT, f"UnsupportedOutputType: {typ}")
        self.outputs.append((kind, i, file_path))

    def add_scldom_cnst(self, dom_id: int, value: float, file_path: str = "") -> int:
        i = self.ctx.scldom_add(FIELD_CONSTANT, value)
        self.scldom.append((FIELD_CONSTANT, None, []))
        self._output("scldom", i, file_path)
        return i

    def add_scldom_func(self, dom_id: int, func, unkn_id=(), file_path: str = "") -> int:
        """func(t, pos[N,3], [values of each unknown]) -> values[N] (vectorised closure)."""
        vals = np.asarray(func(0.0, self.pos, [self.ctx.scldom_get(u) for u in unkn_id]), np.float64)
        i = self.ctx.scldom_add(FIELD_FUNCTION, 0.0, vals)
        self.scldom.append((FIELD_FUNCTION, func, list(unkn_id)))
        self._output("scldom", i, file_path)
        return i

    def add_scldom_unkn(self, dom_id: int, value_init: float, file_path: str = "") -> int:
        i = self.ctx.scldom_add(FIELD_UNKNOWN, value_init)
        self.scldom.append((FIELD_UNKNOWN, None, []))
        self._output("scldom", i, file_path)
        return i

    def add_sclbnd_cnst(self, bnd_id: int, value: float, file_path: str = "") -> int:
        i = self.ctx.sclbnd_add(self.bnd[bnd_id], value)
        self.sclbnd.append((bnd_id, None, [], float(value)))
        self._output("sclbnd", i, file_path)
        return i

    def _bnd_unkn(self, bnd_id, unkn):
        nodes = self.mesh.num_pts + self.mesh.bnd_nodes(self.bnd[bnd_id]).astype(np.int64)
        return [self.ctx.scldom_get(u)[nodes] for u in unkn]

    def add_sclbnd_func(self, bnd_id: int, func, unkn_id=(), file_path: str = "") -> int:
        """func(t, pos[count,3], [values of each unknown on the boundary's nodes]) -> values[count]."""
        vals = np.asarray(func(0.0, self._bnd_pos(bnd_id), self._bnd_unkn(bnd_id, unkn_id)), np.float64)
        i = self.ctx.sclbnd_add(self.bnd[bnd_id], 0.0, vals)
        self.sclbnd.append((bnd_id, func, list(unkn_id), None))
        self._output("sclbnd", i, file_path)
        return i

    # vector domain fields (vars.rs:145-167); values SoA (3, N)
    def add_vecdom_cnst(self, dom_id: int, value, file_path: str = "") -> int:
        i = self.ctx.vecdom_add(FIELD_CONSTANT, cnst3=np.asarray(value, np.float64))
        self.vecdom.append((FIELD_CONSTANT, None, []))
        self._output("vecdom", i, file_path)
        return i

    def add_vecdom_func(self, dom_id: int, func, unkn_id=(), file_path: str = "") -> int:
        """func(t, pos[N,3], [values of each scalar unknown]) -> values (3, N)."""
        vals = np.asarray(func(0.0, self.pos, [self.ctx.scldom_get(u) for u in unkn_id]), np.float64).reshape(3, -1)
        i = self.ctx.vecdom_add(FIELD_FUNCTION, values=vals)
        self.vecdom.append((FIELD_FUNCTION, func, list(unkn_id)))
        self._output("vecdom", i, file_path)
        return i

    def add_vecdom_unkn(self, dom_id: int, value_init, file_path: str = "") -> int:
        i = self.ctx.vecdom_add(FIELD_UNKNOWN, cnst3=np.asarray(value_init, np.float64))
        self.vecdom.append((FIELD_UNKNOWN, None, []))
        self._output("vecdom", i, file_path)
        return i

    # vector boundary fields (vars.rs:169-184); values SoA (3, count)
    def add_vecbnd_cnst(self, bnd_id: int, value, file_path: str = "") -> int:
        i = self.ctx.vecbnd_add(self.bnd[bnd_id], cnst3=np.asarray(value, np.float64))
        self.vecbnd.append((bnd_id, None, [], np.asarray(value, np.float64)))
        self._output("vecbnd", i, file_path)
        return i

    def add_vecbnd_func(self, bnd_id: int, func, unkn_id=(), file_path: str = "") -> int:
        vals = np.asarray(func(0.0, self._bnd_pos(bnd_id), self._bnd_unkn(bnd_id, unkn_id)), np.float64).reshape(3, -1)
        i = self.ctx.vecbnd_add(self.bnd[bnd_id], values=vals)
        self.vecbnd.append((bnd_id, func, list(unkn_id), None))
        self._output("vecbnd", i, file_path)
        return i

    def add_sclflw(self, dom_id: int, vel_id: int, pres_id: int, den_id: int, visc_id: int, fce_id: int) -> int:
        i = self.ctx.sclflw_add(vel_id, pres_id, den_id, visc_id, fce_id)
        self.sclflw.append((vel_id, pres_id, den_id, visc_id, fce_id))
        return i

    def update_func(self, t: float) -> None:
        """_update_func (steady_base.rs:242-261): re-evaluate function fields on the host, in
        the reference's order (scalar domain, scalar boundary, vector domain, vector boundary)."""
        for i, (kind, func, unkn) in enumerate(self.scldom):
            if kind == FIELD_FUNCTION:
                self.ctx.scldom_set(i, func(t, self.pos, [self.ctx.scldom_get(u) for u in unkn]))
        for i, (bnd_id, func, unkn, _) in enumerate(self.sclbnd):
            if func is not None:
                self.ctx.sclbnd_set(i, func(t, self._bnd_pos(bnd_id), self._bnd_unkn(bnd_id, unkn)))
        for i, (kind, func, unkn) in enumerate(self.vecdom):
            if kind == FIELD_FUNCTION:
                self.ctx.vecdom_set(i, np.asarray(func(t, self.pos, [self.ctx.scldom_get(u) for u in unkn]),
                                                  np.float64).reshape(3, -1))
        for i, (bnd_id, func, unkn, _) in enumerate(self.vecbnd):
            if func is not None:
                self.ctx.vecbnd_set(i, np.asarray(func(t, self._bnd_pos(bnd_id), self._bnd_unkn(bnd_id, unkn)),
                                                  np.float64).reshape(3, -1))

    def has_func(self) -> bool:
        return (any(k == FIELD_FUNCTION for k, _, _ in self.scldom) or any(e[1] is not None for e in self.sclbnd)
                or any(k == FIELD_FUNCTION for k, _, _ in self.vecdom) or any(e[1] is not None for e in self.vecbnd))

    def write(self, ts: int, t: float = 0.0) -> None:
        """The drivers' _write (steady_base.rs:213-233): every field with an output path, as
        CSV or VTU (write_csv.rs:5-100, write_vtu.rs) numbered ts; boundary fields on their boundary's nodes (function
        fields evaluated at time t)."""
        for kind, i, path in self.outputs:
            if kind == "scldom":
                self.mesh.write_csv(self.ctx.scldom_get(i), path, ts)
            elif kind == "vecdom":
                self.mesh.write_csv(self.ctx.vecdom_get(i), path, ts)
            else:
                bnd_id, func, unkn, cnst = (self.sclbnd if kind == "sclbnd" else self.vecbnd)[i]
                count = len(self.mesh.bnd_nodes(self.bnd[bnd_id]))
                if func is not None:
                    vals = np.asarray(func(t, self._bnd_pos(bnd_id), self._bnd_unkn(bnd_id, unkn)), np.float64)
                elif kind == "sclbnd":
                    vals = np.full(count, cnst)
                else:
                    vals = np.repeat(np.asarray(cnst, np.float64).reshape(3, 1), count, axis=1)
                self.mesh.write_csv(vals.reshape(-1, count) if kind == "vecbnd" else vals, path, ts, bnd=self.bnd[bnd_id])

    def value(self, scldom_id: int) -> np.ndarray:
        return self.ctx.scldom_get(scldom_id)

    def value_vec(self, vecdom_id: int) -> np.ndarray:
        """Vector field values, SoA (3, N)."""
        return self.ctx.vecdom_get(vecdom_id)


class LinearGmres:
    """Jacobi-preconditioned restarted GMRES (LinearGmres, linear_gmres.rs:16-50).

    Extra keyword options select the device variants: precision FP64 or
    FP32_REFINED, precond JACOBI / ILU0 / NONE."""

    def __init__(self, tol=1e-6, restart=50, max_outer=200, precision=FP64, precond=PRECOND_JACOBI, inner_tol=1e-5,
                 max_refine=0):
        self.opts = dict(tol=tol, restart=restart, max_outer=max_outer, precision=precision, precond=precond,
                         inner_tol=inner_tol, max_refine=max_refine)

    @staticmethod
    def new(num_thread: int) -> "LinearGmres":
        return LinearGmres.new_with_options(num_thread, 1e-6, 50, 200)

    @staticmethod
    def new_with_options(num_thread: int, tol: float, restart: int, max_outer: int, **kw) -> "LinearGmres":
        # input checks of linear_gmres.rs:26-37
        if num_thread == 0:
            raise VFVChemError(L.ERR_INVALID_ARGUMENT, "LinearGmresNonPositiveThreads")
        if not tol > 0.0:
            raise VFVChemError(L.ERR_LINEAR_NON_POSITIVE_TOLERANCE, "LinearGmresNonPositiveTolerance")
        if restart == 0:
            raise VFVChemError(L.ERR_LINEAR_NON_POSITIVE_RESTART, "LinearGmresNonPositiveRestart")
        if max_outer == 0:
            raise VFVChemError(L.ERR_LINEAR_NON_POSITIVE_MAX_OUTER, "LinearGmresNonPositiveMaxOuter")
        return LinearGmres(tol=tol, restart=restart, max_outer=max_outer, **kw)

    # LinearBase (linear_base.rs:5-10) for a caller's own triplet list: the device GMRES on a
    # private context of device 0.  Triplets are (rows, cols, vals) arrays.
    def _context(self) -> Context:
        if getattr(self, "_ctx", None) is None:
            self._ctx = Context(0)
        return self._ctx

    def update_sparse(self, a_triplet, mat_size: int) -> None:
        rows, cols, _ = a_triplet
        self._context().linear_update_sparse(mat_size, rows, cols)

    def build_matrix(self, a_triplet) -> "LinearGmres":
        self._context().linear_build_matrix(a_triplet[2])
        return self

    def solve_matrix(self, a_mat: "LinearGmres", b_vec, x_init, mat_size: int) -> np.ndarray:
        if a_mat is not self or len(b_vec) != mat_size:
            raise VFVChemError(L.ERR_INVALID_ARGUMENT, "solve_matrix: matrix of another solver or size mismatch")
        x, self.stats = self._context().linear_solve_matrix(b_vec, x_init, **self.opts)
        return x


class _UpdateHook:
    """Registers Variables.update_func as the library's update hook for one driver call: the
    library calls it with the time before every assembly (the reference's _update_func,
    steady_base.rs:50-54, 100; transient_base.rs:71-100)."""

    def __init__(self, vars: Variables):
        self.vars = vars

    def __enter__(self):
        if self.vars.has_func():
            self.vars.ctx.set_update_hook(lambda t: self.vars.update_func(t + self.vars.t_offset))
        return self

    def __exit__(self, *exc):
        self.vars.ctx.set_update_hook(None)
        return False


class SteadyBase:
    """SteadyBase::solve (steady_base.rs:16-135) over the operator list of assemble_matrix."""

    def operators(self, vars: Variables):
        raise NotImplementedError

    def solve(self, vars: Variables, lin: LinearGmres, max_iter: int, tol: float, damp: float):
        vars.update_func(0.0)
        ops = self.operators(vars)
        vars.t_offset = 0.0
        with _UpdateHook(vars):
            self.stats = vars.ctx.steady_solve(ops, max_iter, tol, damp, lin=lin.opts)
        vars.write(0)  # steady_base.rs:119-121
        return self.stats


class TransientBase:
    """TransientBase::solve (transient_base.rs:16-160), implicit Euler with a Picard loop per step."""

    def operators(self, vars: Variables):
        raise NotImplementedError

    def solve(self, vars: Variables, lin: LinearGmres, dt: float, num_ts: int, num_ts_write: int, max_iter: int, tol: float,
              damp: float):
        if num_ts_write == 0:
            raise VFVChemError(L.ERR_INVALID_ARGUMENT, "SolverNonPositiveWriteFrequency")
        vars.update_func(0.0)
        ops = self.operators(vars)
        vars.t_offset = 0.0
        vars.write(0)  # transient_base.rs:67
        if not vars.outputs:
            with _UpdateHook(vars):
                self.stats = vars.ctx.transient_solve(ops, dt, num_ts, max_iter, tol, damp, lin=lin.opts)
            return self.stats
        # with outputs: the steps in chunks that end at the output steps ((ts + 1) % num_ts_write
        # == 0, transient_base.rs:141-143); each call continues from the current state, and the
        # hook's time is offset by the steps already done
        done, picard = 0, 0
        try:
            while done < num_ts:
                chunk = min(num_ts_write - done % num_ts_write, num_ts - done)
                vars.t_offset = done * dt
                with _UpdateHook(vars):
                    self.stats = vars.ctx.transient_solve(ops, dt, chunk, max_iter, tol, damp, lin=lin.opts)
                picard += self.stats.picard_iterations
                done += chunk
                if done % num_ts_write == 0:
                    vars.write(done // num_ts_write, done * dt)
        finally:
            vars.t_offset = 0.0
        self.stats.picard_iterations = picard
        return self.stats


class SteadyHeat(SteadyBase):
    """SteadyHeat (steady_heat.rs:9-183), one domain."""

    def __init__(self):
        self.itr = []       # (dom, temp, cond, hprd, hsrc)
        self.temp_bnd = []  # (bnd, temp)
        self.hflx_bnd = []  # (bnd, hflx)

    @staticmethod
    def new():
        return SteadyHeat()

    def add_heat_dom(self, dom_id, temp_id, cond_id, hprd_id, hsrc_id):
        self.itr.append((dom_id, temp_id, cond_id, hprd_id, hsrc_id))

    def add_temp_bnd(self, bnd_id, temp_id):
        self.temp_bnd.append((bnd_id, temp_id))

    def add_hflx_bnd(self, bnd_id, hflx_id):
        self.hflx_bnd.append((bnd_id, hflx_id))

    def operators(self, vars: Variables):
        # assemble_matrix order, steady_heat.rs:143-171
        ops = []
        (dom, T, k, hp, hs) = self.itr[0]
        ops += [(OP_SCL_DOM_DIFFUSION, 0, T, k, T), (OP_SCL_DOM_PRODUCTION, 0, T, hp, -1), (OP_SCL_DOM_SOURCE, 0, T, hs, -1)]
        for (b, tv) in self.temp_bnd:
            ops.append((OP_SCL_BND_DIRICHLET, vars.bnd[b], T, tv, -1))
        for (b, q) in self.hflx_bnd:
            mb = vars.bnd[b]
            ops += [(OP_SCL_BND_DIFFUSION, mb, T, k, T), (OP_SCL_BND_PRODUCTION, mb, T, hp, -1),
                    (OP_SCL_BND_SOURCE, mb, T, hs, -1), (OP_SCL_BND_NEUMANN, mb, T, q, -1)]
        return ops


class TransientHeat(TransientBase):
    """TransientHeat (transient_heat.rs), one domain."""

    def __init__(self):
        self.itr = []
        self.temp_bnd = []
        self.hflx_bnd = []

    @staticmethod
    def new():
        return TransientHeat()

    def add_heat_dom(self, dom_id, temp_id, vlcp_id, cond_id, hprd_id, hsrc_id):
        self.itr.append((dom_id, temp_id, vlcp_id, cond_id, hprd_id, hsrc_id))

    def add_temp_bnd(self, bnd_id, temp_id):
        self.temp_bnd.append((bnd_id, temp_id))

    def add_hflx_bnd(self, bnd_id, hflx_id):
        self.hflx_bnd.append((bnd_id, hflx_id))

    def operators(self, vars: Variables):
        # assemble_matrix order, transient_heat.rs:152-171
        (dom, T, w, k, hp, hs) = self.itr[0]
        ops = [(OP_SCL_DOM_TIME, 0, T, w, -1), (OP_SCL_DOM_DIFFUSION, 0, T, k, T), (OP_SCL_DOM_PRODUCTION, 0, T, hp, -1),
               (OP_SCL_DOM_SOURCE, 0, T, hs, -1)]
        for (b, tv) in self.temp_bnd:
            ops.append((OP_SCL_BND_DIRICHLET, vars.bnd[b], T, tv, -1))
        for (b, q) in self.hflx_bnd:
            mb = vars.bnd[b]
            ops += [(OP_SCL_BND_TIME, mb, T, w, -1), (OP_SCL_BND_DIFFUSION, mb, T, k, T), (OP_SCL_BND_PRODUCTION, mb, T, hp, -1),
                    (OP_SCL_BND_SOURCE, mb, T, hs, -1), (OP_SCL_BND_NEUMANN, mb, T, q, -1)]
        return ops


class SteadyMassFlow(SteadyBase):
    """Mass part of SteadyMassFlow (steady_massflow.rs) for one component on a
    prescribed flow: diffusion, production, source, power-law advection with the
    unity weight (OpSclDomAdvectionUnity), Dirichlet, Neumann (mflx) and outflow
    (mout) boundaries, in the order of steady_massflow.rs:455-545."""

    def __init__(self):
        self.itr = None
        self.flw = None
        self.conc_bnd, self.mflx_bnd, self.mout_bnd = [], [], []

    @staticmethod
    def new(num_comp: int = 1):
        if num_comp != 1:
            raise VFVChemError(L.ERR_OPERATOR_UNSUPPORTED, "one component on the built path")
        return SteadyMassFlow()

    def add_mass_dom(self, comp_id, dom_id, conc_id, diff_ids: dict, mprd_id, msrc_id):
        if set(diff_ids) != {0}:
            raise VFVChemError(L.ERR_OPERATOR_UNSUPPORTED, "cross-diffusion is outside the built path")
        self.itr = (conc_id, diff_ids[0], mprd_id, msrc_id)

    def add_flow_prescribed(self, dom_id, flw_id):
        self.flw = flw_id

    def add_conc_bnd(self, comp_id, bnd_id, conc_id):
        self.conc_bnd.append((bnd_id, conc_id))

    def add_mflx_bnd(self, comp_id, bnd_id, mflx_id):
        self.mflx_bnd.append((bnd_id, mflx_id))

    def add_mout_bnd(self, comp_id, bnd_id):
        self.mout_bnd.append(bnd_id)

    def operators(self, vars: Variables):
        c, D, mp, ms = self.itr
        f = self.flw
        ops = [(OP_SCL_DOM_DIFFUSION, 0, c, D, c), (OP_SCL_DOM_PRODUCTION, 0, c, mp, -1), (OP_SCL_DOM_SOURCE, 0, c, ms, -1)]
        if f is not None:
            ops.append((OP_SCL_DOM_ADVECTION_UNITY, 0, c, D, f))
        for (b, cv) in self.conc_bnd:
            ops.append((OP_SCL_BND_DIRICHLET, vars.bnd[b], c, cv, -1))
        for (b, q) in self.mflx_bnd:
            mb = vars.bnd[b]
            ops += [(OP_SCL_BND_DIFFUSION, mb, c, D, c), (OP_SCL_BND_PRODUCTION, mb, c, mp, -1),
                    (OP_SCL_BND_SOURCE, mb, c, ms, -1), (OP_SCL_BND_NEUMANN, mb, c, q, -1)]
        if f is not None:
            for (b, q) in self.mflx_bnd:
                ops.append((OP_SCL_BND_ADVECTION_UNITY, vars.bnd[b], c, D, f))
        for b in self.mout_bnd:
            mb = vars.bnd[b]
            ops += [(OP_SCL_BND_DIFFUSION, mb, c, D, c), (OP_SCL_BND_ADVECTION_UNITY, mb, c, D, f),
                    (OP_SCL_BND_PRODUCTION, mb, c, mp, -1), (OP_SCL_BND_SOURCE, mb, c, ms, -1)]
        return ops


class SteadyMass(SteadyMassFlow):
    """SteadyMass (steady_mass.rs), one component: SteadyMassFlow without flow."""

    @staticmethod
    def new(num_comp: int = 1):
        if num_comp != 1:
            raise VFVChemError(L.ERR_OPERATOR_UNSUPPORTED, "one component on the built path")
        return SteadyMass()


class SteadyFlow(SteadyBase):
    """SteadyFlow (steady_flow.rs:9-321): steady incompressible Navier-Stokes with the
    Rhie-Chow pressure stabiliser; one domain, velocity boundaries, reference pressure.
    The unknown vector is [u, v, w, p] (initial_start, steady_flow.rs:94-104)."""

    def __init__(self):
        self.itr = []        # (dom, vel, pres, den, visc, fce)
        self.vel_bnd = []    # (bnd, vel)
        self.pref = []       # (dom, nid, pres)
        self.flw = None

    @staticmethod
    def new():
        return SteadyFlow()

    def add_flow_dom(self, dom_id, vel_id, pres_id, den_id, visc_id, fce_id):
        self.itr.append((dom_id, vel_id, pres_id, den_id, visc_id, fce_id))

    def add_vel_bnd(self, bnd_id, vel_id):
        self.vel_bnd.append((bnd_id, vel_id))

    def add_pres_bnd(self, bnd_id, pres_id):
        raise VFVChemError(L.ERR_OPERATOR_UNSUPPORTED, "pressure boundaries are outside the built path")

    def add_trac_bnd(self, bnd_id, pres_id):
        raise VFVChemError(L.ERR_OPERATOR_UNSUPPORTED, "traction boundaries are outside the built path")

    def set_pres_ref(self, dom_id, nid, pres):
        self.pref.append((dom_id, int(nid), float(pres)))

    def operators(self, vars: Variables):
        if len(self.itr) != 1:
            raise VFVChemError(L.ERR_INVALID_ARGUMENT, "one flow domain on the built path")
        dom, vel, pres, den, visc, fce = self.itr[0]
        if self.flw is None:
            # initial_start creates the flow data (steady_flow.rs:106-118)
            self.flw = vars.add_sclflw(dom, vel, pres, den, visc, fce)
        f = self.flw
        ops = [(OP_VEC_DOM_ADVECTION, 0, vel, den, visc, f), (OP_VEC_DOM_GRADIENT, 0, vel, pres, -1),
               (OP_VEC_DOM_DIFFUSION, 0, vel, visc, -1), (OP_VEC_DOM_TRANSPOSE, 0, vel, visc, -1),
               (OP_VEC_DOM_DILATATORY, 0, vel, visc, -1), (OP_VEC_DOM_SOURCE, 0, vel, fce, -1),
               (OP_SCL_DOM_DIVERGENCE, 0, pres, vel, den), (OP_SCL_DOM_PRESSURE_STABILIZER, 0, pres, den, fce, f)]
        for bnd_id, vb in self.vel_bnd:
            b = vars.bnd[bnd_id]
            ops += [(OP_VEC_BND_DIRICHLET, b, vel, vb, -1), (OP_SCL_BND_DIVERGENCE, b, pres, vel, den),
                    (OP_SCL_BND_PRESSURE_STABILIZER, b, pres, den, fce, f)]
        for _, nid, p in self.pref:
            ops.append((OP_SCL_DOM_PRESSURE_REFERENCE, 0, pres, nid, -1, -1, p))
        return ops


class TransientFlow(TransientBase):
    """TransientFlow (transient_flow.rs): SteadyFlow's operators with OpVecDomTime first and the
    divergence operators' time part, implicit Euler."""

    def __init__(self):
        self.flow = SteadyFlow()

    @staticmethod
    def new():
        return TransientFlow()

    def add_flow_dom(self, dom_id, vel_id, pres_id, den_id, visc_id, fce_id):
        self.flow.add_flow_dom(dom_id, vel_id, pres_id, den_id, visc_id, fce_id)

    def add_vel_bnd(self, bnd_id, vel_id):
        self.flow.add_vel_bnd(bnd_id, vel_id)

    def add_pres_bnd(self, bnd_id, pres_id):
        self.flow.add_pres_bnd(bnd_id, pres_id)

    def add_trac_bnd(self, bnd_id, pres_id):
        self.flow.add_trac_bnd(bnd_id, pres_id)

    def set_pres_ref(self, dom_id, nid, pres):
        self.flow.set_pres_ref(dom_id, nid, pres)

    def operators(self, vars: Variables):
        ops = self.flow.operators(vars)
        _, vel, _, den, _, _ = self.flow.itr[0]
        return [(OP_VEC_DOM_TIME, 0, vel, den, -1)] + ops


def _split_pref(ops):
    pref = [o for o in ops if o[0] == OP_SCL_DOM_PRESSURE_REFERENCE]
    return [o for o in ops if o[0] != OP_SCL_DOM_PRESSURE_REFERENCE], pref


class SteadyHeatFlow(SteadyBase):
    """SteadyHeatFlow (steady_heatflow.rs:9-536), one domain: the flow's operators, then the
    heat operators in assemble_matrix order (steady_heatflow.rs:466-512), then the reference
    pressure.  Unknowns [u, v, w, p, T] (initial_start, :154-186); the heat advection reads
    the flow's own Rhie-Chow flux."""

    transient = False

    def __init__(self):
        self.flow = SteadyFlow()
        self.itr_heat = []   # (dom, temp, vlcp, cond, hprd, hsrc)
        self.temp_bnd = []   # (bnd, temp)
        self.hflx_bnd = []   # (bnd, hflx)
        self.hout_bnd = []   # bnd

    @staticmethod
    def new():
        return SteadyHeatFlow()

    def add_heat_dom(self, dom_id, temp_id, vlcp_id, cond_id, hprd_id, hsrc_id):
        self.itr_heat.append((dom_id, temp_id, vlcp_id, cond_id, hprd_id, hsrc_id))

    def add_temp_bnd(self, bnd_id, temp_id):
        self.temp_bnd.append((bnd_id, temp_id))

    def add_hflx_bnd(self, bnd_id, hflx_id):
        self.hflx_bnd.append((bnd_id, hflx_id))

    def add_hout_bnd(self, bnd_id):
        self.hout_bnd.append(bnd_id)

    def add_hcnt_itf(self, itf_id):
        raise VFVChemError(L.ERR_OPERATOR_UNSUPPORTED, "interfaces are outside the built path")

    def add_flow_dom(self, dom_id, vel_id, pres_id, den_id, visc_id, fce_id):
        self.flow.add_flow_dom(dom_id, vel_id, pres_id, den_id, visc_id, fce_id)

    def add_vel_bnd(self, bnd_id, vel_id):
        self.flow.add_vel_bnd(bnd_id, vel_id)

    def add_pres_bnd(self, bnd_id, pres_id):
        self.flow.add_pres_bnd(bnd_id, pres_id)

    def add_trac_bnd(self, bnd_id, pres_id):
        self.flow.add_trac_bnd(bnd_id, pres_id)

    def add_fcnt_itf(self, itf_id):
        raise VFVChemError(L.ERR_OPERATOR_UNSUPPORTED, "interfaces are outside the built path")

    def set_pres_ref(self, dom_id, nid, pres):
        self.flow.set_pres_ref(dom_id, nid, pres)

    def heat_operators(self, vars: Variables):
        if len(self.itr_heat) != 1:
            raise VFVChemError(L.ERR_INVALID_ARGUMENT, "one heat domain on the built path")
        _, temp, vlcp, cond, hprd, hsrc = self.itr_heat[0]
        flw = self.flow.flw
        tp = self.transient
        ops = [(OP_SCL_DOM_TIME_PREVIOUS, 0, temp, vlcp, -1)] if tp else []
        ops += [(OP_SCL_DOM_DIFFUSION, 0, temp, cond, temp), (OP_SCL_DOM_PRODUCTION, 0, temp, hprd, -1),
                (OP_SCL_DOM_SOURCE, 0, temp, hsrc, -1), (OP_SCL_DOM_ADVECTION, 0, temp, vlcp, cond, flw)]
        for bnd_id, tv in self.temp_bnd:
            ops.append((OP_SCL_BND_DIRICHLET, vars.bnd[bnd_id], temp, tv, -1))
        for bnd_id, hf in self.hflx_bnd:
            b = vars.bnd[bnd_id]
            ops += [(OP_SCL_BND_TIME_PREVIOUS, b, temp, vlcp, -1)] if tp else []
            ops += [(OP_SCL_BND_DIFFUSION, b, temp, cond, temp), (OP_SCL_BND_PRODUCTION, b, temp, hprd, -1),
                    (OP_SCL_BND_SOURCE, b, temp, hsrc, -1), (OP_SCL_BND_NEUMANN, b, temp, hf, -1)]
        for bnd_id, _ in self.hflx_bnd:
            ops.append((OP_SCL_BND_ADVECTION, vars.bnd[bnd_id], temp, vlcp, cond, flw))
        for bnd_id in self.hout_bnd:
            b = vars.bnd[bnd_id]
            ops += [(OP_SCL_BND_TIME_PREVIOUS, b, temp, vlcp, -1)] if tp else []
            ops += [(OP_SCL_BND_DIFFUSION, b, temp, cond, temp), (OP_SCL_BND_ADVECTION, b, temp, vlcp, cond, flw),
                    (OP_SCL_BND_PRODUCTION, b, temp, hprd, -1), (OP_SCL_BND_SOURCE, b, temp, hsrc, -1)]
        return ops

    def mass_operators(self, vars: Variables):
        return []

    def operators(self, vars: Variables):
        flow_ops, pref = _split_pref(self.flow.operators(vars))
        if self.transient:
            _, vel, _, den, _, _ = self.flow.itr[0]
            flow_ops = [(OP_VEC_DOM_TIME, 0, vel, den, -1)] + flow_ops
        return flow_ops + self.heat_operators(vars) + self.mass_operators(vars) + pref


class TransientHeatFlow(SteadyHeatFlow):
    """TransientHeatFlow (transient_heatflow.rs): SteadyHeatFlow's composition with the
    velocity time operator and OpSclDomTimePrevious / OpSclBndTimePrevious (weight rho c_p)
    ahead of each heat row class, implicit Euler (TransientBase::solve)."""

    transient = True
    solve = TransientBase.solve

    @staticmethod
    def new():
        return TransientHeatFlow()


class TransientHeatMassFlow(TransientHeatFlow):
    """TransientHeatMassFlow (transient_heatmassflow.rs:9-900), one domain and one species
    component: TransientHeatFlow plus the species operators in assemble_matrix order
    (diffusion, production, source, OpSclDomTimePrevious weighted by the density,
    OpSclDomAdvectionUnity with the self-diffusivity; :785-840).  Unknowns
    [u, v, w, p, T, C] (initial_start, :248-268).  Cross-diffusion (another component
    driving the flux) is outside the built path."""

    def __init__(self, num_comp: int = 1):
        super().__init__()
        if num_comp != 1:
            raise VFVChemError(L.ERR_OPERATOR_UNSUPPORTED, "one species component on the built path (six unknowns per vertex)")
        self.itr_mass = []   # (comp, dom, conc, diff_ids, mprd, msrc)
        self.conc_bnd = []   # (comp, bnd, conc)
        self.mflx_bnd = []   # (comp, bnd, mflx)
        self.mout_bnd = []   # (comp, bnd)

    @staticmethod
    def new(num_comp: int = 1):
        return TransientHeatMassFlow(num_comp)

    def add_mass_dom(self, comp_id, dom_id, conc_id, diff_ids: dict, mprd_id, msrc_id):
        self.itr_mass.append((comp_id, dom_id, conc_id, dict(diff_ids), mprd_id, msrc_id))

    def add_conc_bnd(self, comp_id, bnd_id, conc_id):
        self.conc_bnd.append((comp_id, bnd_id, conc_id))

    def add_mflx_bnd(self, comp_id, bnd_id, mflx_id):
        self.mflx_bnd.append((comp_id, bnd_id, mflx_id))

    def add_mout_bnd(self, comp_id, bnd_id):
        self.mout_bnd.append((comp_id, bnd_id))

    def mass_operators(self, vars: Variables):
        if len(self.itr_mass) != 1:
            raise VFVChemError(L.ERR_INVALID_ARGUMENT, "one species domain on the built path")
        comp, _, conc, diff_ids, mprd, msrc = self.itr_mass[0]
        if set(diff_ids) - {comp}:
            raise VFVChemError(L.ERR_OPERATOR_UNSUPPORTED, "cross-diffusion into another unknown is outside the built path")
        diff = diff_ids.get(comp, -1)
        _, _, _, den, _, _ = self.flow.itr[0]
        flw = self.flow.flw
        ops = [(OP_SCL_DOM_DIFFUSION, 0, conc, diff, conc)] if diff >= 0 else []
        ops += [(OP_SCL_DOM_PRODUCTION, 0, conc, mprd, -1), (OP_SCL_DOM_SOURCE, 0, conc, msrc, -1),
                (OP_SCL_DOM_TIME_PREVIOUS, 0, conc, den, -1), (OP_SCL_DOM_ADVECTION_UNITY, 0, conc, diff, flw)]
        for _, bnd_id, cv in self.conc_bnd:
            ops.append((OP_SCL_BND_DIRICHLET, vars.bnd[bnd_id], conc, cv, -1))
        for _, bnd_id, mf in self.mflx_bnd:
            b = vars.bnd[bnd_id]
            ops += [(OP_SCL_BND_DIFFUSION, b, conc, diff, conc)] if diff >= 0 else []
            ops += [(OP_SCL_BND_PRODUCTION, b, conc, mprd, -1), (OP_SCL_BND_SOURCE, b, conc, msrc, -1),
                    (OP_SCL_BND_NEUMANN, b, conc, mf, -1)]
        for _, bnd_id, _ in self.mflx_bnd:
            ops.append((OP_SCL_BND_TIME_PREVIOUS, vars.bnd[bnd_id], conc, den, -1))
        for _, bnd_id, _ in self.mflx_bnd:
            ops.append((OP_SCL_BND_ADVECTION_UNITY, vars.bnd[bnd_id], conc, diff, flw))
        for _, bnd_id in self.mout_bnd:
            b = vars.bnd[bnd_id]
            ops += [(OP_SCL_BND_TIME_PREVIOUS, b, conc, den, -1)]
            ops += [(OP_SCL_BND_DIFFUSION, b, conc, diff, conc)] if diff >= 0 else []
            ops += [(OP_SCL_BND_ADVECTION_UNITY, b, conc, diff, flw), (OP_SCL_BND_PRODUCTION, b, conc, mprd, -1),
                    (OP_SCL_BND_SOURCE, b, conc, msrc, -1)]
        return ops
