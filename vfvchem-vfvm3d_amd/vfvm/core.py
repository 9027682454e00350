"""Thin object layer over the C ABI: meshes, device contexts, operators."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, lib, ptr

# operator kinds (include/vfvm.h, vfvm_op_kind)
OP_SCL_DOM_DIFFUSION = 1
OP_SCL_DOM_PRODUCTION = 2
OP_SCL_DOM_SOURCE = 3
OP_SCL_DOM_TIME = 4
OP_SCL_DOM_ADVECTION_UNITY = 5
OP_SCL_BND_DIRICHLET = 20
OP_SCL_BND_NEUMANN = 21
OP_SCL_BND_DIFFUSION = 22
OP_SCL_BND_PRODUCTION = 23
OP_SCL_BND_SOURCE = 24
OP_SCL_BND_TIME = 25
OP_SCL_BND_ADVECTION_UNITY = 26
OP_SCL_DOM_ADVECTION = 6
OP_SCL_DOM_TIME_PREVIOUS = 7
OP_SCL_DOM_TIME_UNITY = 8
OP_SCL_BND_ADVECTION = 27
OP_SCL_BND_TIME_PREVIOUS = 28
OP_SCL_BND_TIME_UNITY = 29
OP_VEC_DOM_ADVECTION = 40
OP_VEC_DOM_GRADIENT = 41
OP_VEC_DOM_DIFFUSION = 42
OP_VEC_DOM_TRANSPOSE = 43
OP_VEC_DOM_DILATATORY = 44
OP_VEC_DOM_SOURCE = 45
OP_SCL_DOM_DIVERGENCE = 46
OP_SCL_DOM_PRESSURE_STABILIZER = 47
OP_SCL_DOM_PRESSURE_REFERENCE = 48
OP_VEC_DOM_TIME = 49
OP_VEC_BND_DIRICHLET = 60
OP_SCL_BND_DIVERGENCE = 61
OP_SCL_BND_PRESSURE_STABILIZER = 62
OP_SCL_ITF_CONTINUITY = 70
OP_SCL_ITF_DIFFUSION = 71
OP_SCL_ITF_PRODUCTION = 72
OP_SCL_ITF_SOURCE = 73
OP_SCL_ITF_TIME = 74
OP_SCL_ITF_TIME_PREVIOUS = 75
OP_SCL_ITF_TIME_UNITY = 76
OP_SCL_ITF_ADVECTION = 77
OP_SCL_ITF_ADVECTION_UNITY = 78
FLW_FLUX, FLW_RC, FLW_MOM = 0, 1, 2

FIELD_CONSTANT, FIELD_FUNCTION, FIELD_UNKNOWN = 0, 1, 2
FP64, FP32_REFINED = 0, 1
PRECOND_NONE, PRECOND_JACOBI, PRECOND_ILU0, PRECOND_BJACOBI = 0, 1, 2, 3
K_SPMV, K_MDOT, K_ASSEMBLE, K_PRECOND, K_UPDATE = 0, 1, 2, 3, 4
ORTHO_AUTO, ORTHO_CGS, ORTHO_CGS2 = 0, 1, 2
WRT_SCLDOM, WRT_SCLBND, WRT_INITIAL, WRT_VECBND = 0, 1, 2, 3

_MESH_ARRAYS = {
    "pos": (0, np.float64), "tets": (1, np.int32), "bnd_ptr": (2, np.int64), "bnd_nid": (3, np.int32),
    "row_ptr": (4, np.int32), "ngh": (5, np.int32), "dist": (6, np.float64), "area": (7, np.float64),
    "norm": (8, np.float64), "skew": (9, np.float64), "gw": (10, np.float64), "vol": (11, np.float64),
    "cskew": (12, np.float64), "bsf_ptr": (13, np.int32), "bsf_vj": (14, np.int32), "bsf_vk": (15, np.int32),
    "bsf_area": (16, np.float64), "bsf_norm": (17, np.float64), "bsf_wgt": (18, np.float64),
    "gc_ptr": (19, np.int32), "gc_vid": (20, np.int32), "gc_c": (21, np.float64),
    "dom_ptr": (22, np.int64), "dom_npts": (23, np.int64), "itf_ptr": (24, np.int64), "itf_va": (25, np.int32),
    "itf_vb": (26, np.int32), "itf_dom": (27, np.int32),
}


class Mesh:
    """Host mesh with its Voronoi dual (Tetrahedron, src/base/geom_tet.rs:5-46)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        info = L.MeshInfo()
        check(lib.vfvm_mesh_info_get(self._h, C.byref(info)))
        self.num_pts = info.num_pts
        self.num_nds = info.num_nds
        self.num_tot = info.num_pts + info.num_nds
        self.num_tets = info.num_tets
        self.num_edges = info.num_edges
        self.num_subfaces = info.num_subfaces
        self.num_bnd = info.num_bnd
        self.max_degree = info.max_degree
        self.num_grad_coeff = info.num_grad_coeff
        self.num_dom = info.num_dom
        self.num_itf = info.num_itf
        self.num_itf_nodes = info.num_itf_nodes
        self._cache = {}
        self._parts = ()

    def build_grad_coeff(self, device=None):
        """Pressure-gradient coefficients of the flow operators (geom_tet.rs:295-347); built
        on GPU `device` when given (bitwise equal to the host build)."""
        if device is None:
            check(lib.vfvm_mesh_build_grad_coeff(self._h))
        else:
            check(lib.vfvm_mesh_build_grad_coeff_dev(self._h, int(device)))
        info = L.MeshInfo()
        check(lib.vfvm_mesh_info_get(self._h, C.byref(info)))
        self.num_grad_coeff = info.num_grad_coeff
        return self

    @classmethod
    def lattice(cls, n: int, seed: int = 0, jitter: float = 0.3, device=None) -> "Mesh":
        """Jittered lattice mesh; the dual is built on GPU `device` when given (bitwise equal)."""
        h = C.c_void_p()
        if device is None:
            check(lib.vfvm_mesh_lattice(int(n), int(seed), float(jitter), C.byref(h)))
        else:
            check(lib.vfvm_mesh_lattice_dev(int(n), int(seed), float(jitter), int(device), C.byref(h)))
        return cls(h)

    @classmethod
    def lattice_box(cls, n: int, lo, hi, seed: int = 0, jitter: float = 0.3, device=None) -> "Mesh":
        """The vertices lo <= (i, j, k) <= hi of the n^3 jittered lattice (a box of the unit cube;
        boxes sharing a face place its vertices identically); dual on GPU `device` when given."""
        lo = np.ascontiguousarray(lo, np.int32)
        hi = np.ascontiguousarray(hi, np.int32)
        h = C.c_void_p()
        check(lib.vfvm_mesh_lattice_box(int(n), int(seed), float(jitter), ptr(lo, C.c_int32), ptr(hi, C.c_int32),
                                        -1 if device is None else int(device), C.byref(h)))
        return cls(h)

    @classmethod
    def join(cls, doms, itfs) -> "Mesh":
        """Multi-domain mesh of the meshes `doms` joined by interfaces `itfs` = [(dom_a, bnd_a,
        dom_b, bnd_b), ...] (boundary bnd_a of dom_a and bnd_b of dom_b hold the same nodes)."""
        hs = (C.c_void_p * len(doms))(*[d._h for d in doms])
        q = np.ascontiguousarray(np.asarray(itfs, np.int32).reshape(-1), np.int32)
        h = C.c_void_p()
        check(lib.vfvm_mesh_join(len(doms), hs, len(itfs), ptr(q, C.c_int32), C.byref(h)))
        m = cls(h)
        m._parts = tuple(doms)
        return m

    @classmethod
    def from_tets(cls, pos, tets, num_pts: int, bnd_nids, device=None) -> "Mesh":
        pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
        tets = np.ascontiguousarray(tets, np.int32).reshape(-1, 4)
        bptr = np.zeros(len(bnd_nids) + 1, np.int64)
        for b, ids in enumerate(bnd_nids):
            bptr[b + 1] = bptr[b] + len(ids)
        bnid = np.ascontiguousarray(np.concatenate([np.asarray(x, np.int32) for x in bnd_nids]) if bnd_nids else np.zeros(0, np.int32), np.int32)
        h = C.c_void_p()
        if device is None:
            check(lib.vfvm_mesh_from_tets(int(num_pts), int(pos.shape[0] - num_pts), ptr(pos), int(tets.shape[0]), ptr(tets, C.c_int32),
                                          len(bnd_nids), ptr(bptr, C.c_int64), ptr(bnid, C.c_int32), C.byref(h)))
        else:
            check(lib.vfvm_mesh_from_tets_dev(int(num_pts), int(pos.shape[0] - num_pts), ptr(pos), int(tets.shape[0]),
                                              ptr(tets, C.c_int32), len(bnd_nids), ptr(bptr, C.c_int64), ptr(bnid, C.c_int32),
                                              int(device), C.byref(h)))
        return cls(h)

    def _size(self, name):
        N, E, S, T = self.num_tot, self.num_edges, self.num_subfaces, self.num_tets
        return {"pos": 3 * N, "tets": 4 * T, "bnd_ptr": self.num_bnd + 1, "bnd_nid": None, "row_ptr": N + 1, "ngh": E,
                "dist": E, "area": E, "norm": 3 * E, "skew": 3 * E, "gw": 3 * E, "vol": N, "cskew": 3 * N,
                "bsf_ptr": N + 1, "bsf_vj": S, "bsf_vk": S, "bsf_area": S, "bsf_norm": 3 * S, "bsf_wgt": 3 * S,
                "gc_ptr": N + 1, "gc_vid": self.num_grad_coeff, "gc_c": 3 * self.num_grad_coeff,
                "dom_ptr": self.num_dom + 1, "dom_npts": self.num_dom, "itf_ptr": self.num_itf + 1,
                "itf_va": self.num_itf_nodes, "itf_vb": self.num_itf_nodes, "itf_dom": 2 * self.num_itf}[name]

    def get(self, name: str) -> np.ndarray:
        if name in self._cache:
            return self._cache[name]
        aid, dt = _MESH_ARRAYS[name]
        n = self._size(name)
        if n is None:
            n = int(self.get("bnd_ptr")[-1])
        out = np.empty(max(n, 1), dt)
        check(lib.vfvm_mesh_get(self._h, aid, out.ctypes.data_as(C.c_void_p)))
        out = out[:n]
        if name == "pos":
            out = out.reshape(-1, 3)
        elif name == "tets":
            out = out.reshape(-1, 4)
        elif name in ("norm", "skew", "gw", "bsf_norm", "bsf_wgt", "cskew", "gc_c"):
            out = out.reshape(3, -1)
        self._cache[name] = out
        return out

    def write_csv(self, values, file_path: str, ts: int = 0, bnd: int = -1, digits: int = 9) -> str:
        """Output of a field as CSV or VTU by the type of `file_path` (write_csv.rs:5-100,
        write_vtu.rs): `values` of shape (n,) or (3, n) over the vertices (bnd < 0) or the nodes
        of boundary `bnd`; `file_path` is split at its last '.' into name and type as the
        reference does (scl_dom.rs:60); digits = 9 is the reference's format, 0 writes 17
        significant digits.  Returns the path written."""
        v = np.ascontiguousarray(values, np.float64)
        ncomp = 1 if v.ndim == 1 else v.shape[0]
        nb = len(self.get("bnd_ptr")) - 1
        n = self.num_tot if bnd < 0 else (len(self.bnd_nodes(bnd)) if bnd < nb else v.shape[-1])  # bad ids: the library reports
        if v.shape[-1] != n or ncomp not in (1, 3):
            raise ValueError(f"write_csv: values of shape {v.shape} for {n} vertices")
        name, sep, typ = file_path.rpartition(".")
        if not sep:
            raise ValueError(f"output path without a type: {file_path}")
        fn = lib.vfvm_write_vtu if typ == "vtu" else lib.vfvm_write_csv
        check(fn(self._h, int(bnd), int(ncomp), v.ctypes.data_as(C.POINTER(C.c_double)), name.encode(), typ.encode(), int(ts),
                 int(digits)))
        return f"{name}_{ts:04d}.{typ}"

    def bnd_nodes(self, b: int) -> np.ndarray:
        p = self.get("bnd_ptr")
        return self.get("bnd_nid")[p[b]:p[b + 1]]

    def __del__(self):
        try:
            if self._h:
                lib.vfvm_mesh_destroy(self._h)
                self._h = None
        except Exception:
            pass


def make_ops(ops):
    """Operator descriptors from tuples (kind, geom_id, unk_id, a, b[, c[, value]])."""
    arr = (L.OpStruct * len(ops))()
    for i, o in enumerate(ops):
        o = tuple(o)
        if len(o) < 5 or len(o) > 7:
            raise ValueError("operator tuple must be (kind, geom_id, unk_id, a, b[, c[, value]])")
        arr[i].kind, arr[i].geom_id, arr[i].unk_id, arr[i].a, arr[i].b = (int(v) for v in o[:5])
        arr[i].c = int(o[5]) if len(o) > 5 else -1
        arr[i].value = float(o[6]) if len(o) > 6 else 0.0
    return arr


def solve_opts(tol=1e-6, restart=50, max_outer=200, precision=FP64, precond=PRECOND_JACOBI, inner_tol=1e-5, max_refine=0,
               ortho=0):
    o = L.SolveOpts()
    lib.vfvm_solve_opts_default(C.byref(o))
    o.tol, o.restart, o.max_outer, o.precision, o.precond, o.inner_tol, o.max_refine, o.ortho = (
        float(tol), int(restart), int(max_outer), int(precision), int(precond), float(inner_tol), int(max_refine), int(ortho))
    return o


class Context:
    """One device context (one HIP stream on one GPU)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(lib.vfvm_create(int(device), C.byref(self._h)))
        self.mesh = None
        self._sclbnd_size = []
        self._vecbnd_size = []

    def close(self):
        if self._h:
            lib.vfvm_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, mesh: Mesh):
        check(lib.vfvm_mesh_upload(self._h, mesh._h))
        self.mesh = mesh
        self._sclbnd_size = []
        self._vecbnd_size = []

    # fields
    def scldom_add(self, kind, cnst=0.0, values=None) -> int:
        i = C.c_int32()
        v = None if values is None else np.ascontiguousarray(values, np.float64)
        check(lib.vfvm_scldom_add(self._h, int(kind), float(cnst), None if v is None else ptr(v), C.byref(i)))
        return i.value

    def scldom_set(self, fid, values):
        v = np.ascontiguousarray(values, np.float64)
        check(lib.vfvm_scldom_set(self._h, int(fid), ptr(v)))

    def scldom_get(self, fid) -> np.ndarray:
        out = np.empty(self.mesh.num_tot)
        check(lib.vfvm_scldom_get(self._h, int(fid), ptr(out)))
        return out

    def scldom_update_prev(self, fid):
        check(lib.vfvm_scldom_update_prev(self._h, int(fid)))

    def vecdom_add(self, kind, cnst3=None, values=None) -> int:
        i = C.c_int32()
        c3 = None if cnst3 is None else np.ascontiguousarray(cnst3, np.float64)
        v = None if values is None else np.ascontiguousarray(values, np.float64)
        check(lib.vfvm_vecdom_add(self._h, int(kind), None if c3 is None else ptr(c3), None if v is None else ptr(v), C.byref(i)))
        return i.value

    def vecdom_set(self, fid, values):
        v = np.ascontiguousarray(values, np.float64)
        check(lib.vfvm_vecdom_set(self._h, int(fid), ptr(v)))

    def sclbnd_add(self, bnd, cnst=0.0, values=None) -> int:
        i = C.c_int32()
        v = None if values is None else np.ascontiguousarray(values, np.float64)
        check(lib.vfvm_sclbnd_add(self._h, int(bnd), float(cnst), None if v is None else ptr(v), C.byref(i)))
        self._sclbnd_size.append(len(self.mesh.bnd_nodes(bnd)))
        return i.value

    def sclbnd_set(self, fid, values):
        v = np.ascontiguousarray(values, np.float64)
        check(lib.vfvm_sclbnd_set(self._h, int(fid), ptr(v)))

    def vecbnd_set(self, fid, values):
        """New boundary velocity values, SoA (3, count)."""
        v = np.ascontiguousarray(values, np.float64)
        check(lib.vfvm_vecbnd_set(self._h, int(fid), ptr(v)))

    def sclflw_add(self, vel, pres, den, visc, fce) -> int:
        i = C.c_int32()
        check(lib.vfvm_sclflw_add(self._h, int(vel), int(pres), int(den), int(visc), int(fce), C.byref(i)))
        return i.value

    def vecdom_get(self, fid) -> np.ndarray:
        out = np.empty(3 * self.mesh.num_tot)
        check(lib.vfvm_vecdom_get(self._h, int(fid), ptr(out)))
        return out.reshape(3, -1)

    def vecbnd_add(self, bnd, cnst3=None, values=None) -> int:
        """Boundary velocity (3 x count SoA, or one constant vector)."""
        i = C.c_int32()
        c3 = None if cnst3 is None else np.ascontiguousarray(cnst3, np.float64)
        v = None if values is None else np.ascontiguousarray(values, np.float64)
        check(lib.vfvm_vecbnd_add(self._h, int(bnd), None if c3 is None else ptr(c3), None if v is None else ptr(v), C.byref(i)))
        self._vecbnd_size.append(len(self.mesh.bnd_nodes(bnd)))
        return i.value

    def vecdom_func_affine(self, vec, scl, f0, df):
        """vec := f0 + df * scl on the device wherever the drivers evaluate function fields
        (Variables::add_vecdom_func for an affine closure); f0, df: (3, N).  None, None removes
        the closure."""
        if f0 is None or df is None:
            check(lib.vfvm_vecdom_func_affine(self._h, int(vec), int(scl), None, None))
            return
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(f0, np.float64).reshape(3, -1), (3, self.mesh.num_tot)))
        b = np.ascontiguousarray(np.broadcast_to(np.asarray(df, np.float64).reshape(3, -1), (3, self.mesh.num_tot)))
        check(lib.vfvm_vecdom_func_affine(self._h, int(vec), int(scl), ptr(a), ptr(b)))

    def update_flow(self, flw):
        check(lib.vfvm_update_flow(self._h, int(flw)))

    def sclflw_get(self, flw, what) -> np.ndarray:
        size = {FLW_FLUX: self.mesh.num_edges, FLW_RC: self.mesh.num_tot, FLW_MOM: 3 * self.mesh.num_tot}[what]
        out = np.empty(size)
        check(lib.vfvm_sclflw_get(self._h, int(flw), int(what), ptr(out)))
        return out

    # pattern / assembly
    def pattern_build(self, ops) -> int:
        arr = make_ops(ops)
        nnz = C.c_int64()
        check(lib.vfvm_pattern_build(self._h, arr, len(ops), C.byref(nnz)))
        # unknowns: one scalar, or the flow system [u, v, w, p, coupled scalars...]
        # (steady_flow.rs:94-104, steady_heatflow.rs:468-505)
        self.system_info()
        return nnz.value

    def system_info(self):
        """(n, nnz) of the built system."""
        n, nnz = C.c_int64(), C.c_int64()
        check(lib.vfvm_system_info(self._h, C.byref(n), C.byref(nnz)))
        self.n, self.nnz = n.value, nnz.value
        return n.value, nnz.value

    def _size(self):
        n = getattr(self, "n", None)
        return n if n is not None else self.mesh.num_tot

    def pattern_get(self):
        rp = np.empty(self._size() + 1, np.int32)
        col = np.empty(max(self.nnz, 1), np.int32)
        check(lib.vfvm_pattern_get(self._h, ptr(rp, C.c_int32), ptr(col, C.c_int32)))
        return rp, col[:self.nnz]

    def assemble(self, ops, dt=0.0):
        arr = make_ops(ops)
        check(lib.vfvm_assemble(self._h, arr, len(ops), float(dt)))

    def matrix_get(self):
        vals = np.empty(max(self.nnz, 1))
        b = np.empty(self._size())
        check(lib.vfvm_matrix_get(self._h, ptr(vals), ptr(b)))
        return vals[:self.nnz], b

    def solve(self, x0=None, **kw):
        x = np.zeros(self._size()) if x0 is None else np.array(x0, np.float64)
        st = L.SolveStats()
        o = solve_opts(**kw)
        code = lib.vfvm_solve(self._h, C.byref(o), ptr(x), C.byref(st))
        check(code)
        return x, st

    def solve_device(self, x_dev_ptr: int, **kw):
        st = L.SolveStats()
        o = solve_opts(**kw)
        check(lib.vfvm_solve_device(self._h, C.byref(o), C.c_void_p(x_dev_ptr), C.byref(st)))
        return st

    # LinearBase for caller-assembled triplets (linear_base.rs:5-10)
    def linear_update_sparse(self, mat_size: int, rows, cols) -> int:
        """update_sparse: the pattern of the triplets (rows[k], cols[k]); returns nnz."""
        rows = np.ascontiguousarray(rows, np.int64)
        cols = np.ascontiguousarray(cols, np.int64)
        if rows.shape != cols.shape:
            raise ValueError("rows and cols differ in length")
        nnz = C.c_int64()
        check(lib.vfvm_linear_update_sparse(self._h, int(mat_size), int(rows.size), ptr(rows, C.c_int64), ptr(cols, C.c_int64),
                                            C.byref(nnz)))
        self.system_info()
        return self.nnz

    def linear_build_matrix(self, vals) -> None:
        """build_matrix: values of the triplets in the update_sparse order, duplicates summed."""
        vals = np.ascontiguousarray(vals, np.float64)
        check(lib.vfvm_linear_build_matrix(self._h, int(vals.size), ptr(vals)))

    def linear_solve_matrix(self, b, x_init=None, **kw):
        """solve_matrix: x of A x = b from x_init (zero if None) with the GMRES options kw."""
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        xi = None if x_init is None else np.ascontiguousarray(x_init, np.float64)
        st = L.SolveStats()
        o = solve_opts(**kw)
        check(lib.vfvm_linear_solve_matrix(self._h, C.byref(o), ptr(b), None if xi is None else ptr(xi), ptr(x), C.byref(st)))
        return x, st

    # convergence scaling (vars.rs:28-31,203-213)
    def set_scldom_scale(self, fid: int, scale: float) -> None:
        check(lib.vfvm_set_scldom_scale(self._h, int(fid), float(scale)))

    def set_vecdom_scale(self, fid: int, scale: float) -> None:
        check(lib.vfvm_set_vecdom_scale(self._h, int(fid), float(scale)))

    def set_scale_off(self, off: bool) -> None:
        check(lib.vfvm_set_scale_off(self._h, 1 if off else 0))

    def solve_device_noalloc(self, x_dev_ptr: int, **kw):
        """vfvm_solve_device_noalloc: the same solve with no device allocation (FP64, Jacobi or
        no preconditioner; workspace in the library's fixed per-GPU scratch array)."""
        st = L.SolveStats()
        o = solve_opts(**kw)
        check(lib.vfvm_solve_device_noalloc(self._h, C.byref(o), C.c_void_p(x_dev_ptr), C.byref(st)))
        return st

    def residual(self, x):
        x = np.ascontiguousarray(x, np.float64)
        r = C.c_double()
        check(lib.vfvm_residual(self._h, ptr(x), C.byref(r)))
        return r.value

    def precond_apply(self, r, **kw):
        """z = M^-1 r with the preconditioner of the solve options kw."""
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty_like(r)
        o = solve_opts(**kw)
        check(lib.vfvm_precond_apply(self._h, C.byref(o), ptr(r), ptr(z)))
        return z

    def ilu_order(self):
        """(perm, color_ptr): the ILU(0) row order of the built pattern."""
        n = self._size()
        perm = np.empty(n, np.int32)
        cptr = np.empty(n + 1, np.int32)
        nc = C.c_int32()
        check(lib.vfvm_ilu_order(self._h, ptr(perm, C.c_int32), ptr(cptr, C.c_int32), C.byref(nc)))
        return perm, cptr[:nc.value + 1].copy()

    def ilu_wait_limit(self, polls):
        """Sets the polls of a one-launch ILU(0) sweep's wait for an adjacent tile (< 0: every such
        wait gives up at once); returns the number of waits that gave up so far."""
        lapsed = C.c_uint64(0)
        check(lib.vfvm_ilu_wait_limit(self._h, int(polls), C.byref(lapsed)))
        return int(lapsed.value)

    def ilu_tiles(self):
        """(tile_ptr, level): the tiles of the tiled ILU(0) order over the permuted rows and the
        level of each position inside its tile (tile_ptr empty for the multicolour order)."""
        n = self._size()
        tp = np.empty(n + 1, np.int32)
        lev = np.empty(n, np.int32)
        nt = C.c_int32()
        check(lib.vfvm_ilu_tiles(self._h, ptr(tp, C.c_int32), C.byref(nt), ptr(lev, C.c_int32)))
        return (tp[:nt.value + 1].copy() if nt.value > 0 else np.zeros(0, np.int32)), lev

    def steady_solve(self, ops, max_iter, tol, damp, lin=None):
        """SteadyBase::solve; lin: dict of linear-solver options (solve_opts keywords)."""
        arr = make_ops(ops)
        st = L.DriverStats()
        o = solve_opts(**(lin or {}))
        check(lib.vfvm_steady_solve(self._h, arr, len(ops), C.byref(o), int(max_iter), float(tol), float(damp), C.byref(st)))
        self.system_info()
        return st

    def transient_solve(self, ops, dt, num_ts, max_iter, tol, damp, lin=None):
        """TransientBase::solve; lin: dict of linear-solver options (solve_opts keywords)."""
        arr = make_ops(ops)
        st = L.DriverStats()
        o = solve_opts(**(lin or {}))
        check(lib.vfvm_transient_solve(self._h, arr, len(ops), C.byref(o), float(dt), int(num_ts), int(max_iter), float(tol),
                                       float(damp), C.byref(st)))
        self.system_info()
        return st

    def transient_adaptive(self, ops, dt_init, t_end, lin=None, hook=None, max_steps=100000, **opts):
        """TransientAdaptive::solve_adaptive_with (transient_adaptive.rs:81-350).  opts: fields
        of vfvm_adaptive_opts (out_times as a sequence, scl_exclude as a sequence of ids).
        hook(t, dt) is called after each accepted step.  Returns (stats, steps (k, 3) array of
        (t, dt, err))."""
        arr = make_ops(ops)
        o = L.AdaptiveOpts()
        lib.vfvm_adaptive_opts_default(float(dt_init), float(t_end), C.byref(o))
        keep = []
        for k, v in opts.items():
            if k == "out_times":
                a = np.ascontiguousarray(v, np.float64)
                keep.append(a)
                o.out_times, o.num_out = ptr(a), a.size
            elif k == "scl_exclude":
                a = np.ascontiguousarray(v, np.int32)
                keep.append(a)
                o.scl_exclude, o.num_exclude = ptr(a, C.c_int32), a.size
            else:
                setattr(o, k, v)
        lo = solve_opts(**(lin or {}))
        st = L.AdaptiveStats()
        steps = np.zeros((max_steps, 3))
        cb = L.STEP_HOOK((lambda user, t, dt: hook(t, dt)) if hook else (lambda user, t, dt: None))
        check(lib.vfvm_transient_adaptive(self._h, arr, len(ops), C.byref(lo), C.byref(o), cb, None, ptr(steps), int(max_steps),
                                          C.byref(st)))
        self.system_info()
        return st, steps[:min(st.accepted, max_steps)].copy()

    def set_update_hook(self, hook):
        """Variables::_update_func (steady_base.rs:242-261): hook(t) runs on the host before
        every assembly of the drivers, with the unknowns holding the current iterate; it may
        read fields and set function fields.  None removes it."""
        if hook is None:
            self._update_cb = None
            check(lib.vfvm_set_update_hook(self._h, C.cast(None, L.UPDATE_HOOK), None))
            return
        self._update_cb = L.UPDATE_HOOK(lambda user, t: hook(t))
        check(lib.vfvm_set_update_hook(self._h, self._update_cb, None))

    # discrete adjoint
    def adjoint_forward(self, ops, dt=0.0, num_ts=0, checkpoint_every=10, lin=None):
        """Forward pass of the linear steady (num_ts=0) or implicit-Euler map, keeping
        every checkpoint_every-th state on the device; leaves T^num_ts in the unknown."""
        arr = make_ops(ops)
        st = L.AdjointStats()
        o = solve_opts(**(lin or {}))
        check(lib.vfvm_adjoint_forward(self._h, arr, len(ops), C.byref(o), float(dt), int(num_ts), int(checkpoint_every),
                                       C.byref(st)))
        return st

    def adjoint_backward(self, seed, wrt, lin=None):
        """Backward pass for dJ/dT^num_ts = seed.  wrt: list of (kind, id) with kind one
        of WRT_SCLDOM, WRT_SCLBND, WRT_INITIAL.  Returns (list of gradients, stats)."""
        n = self.mesh.num_tot
        seed = np.ascontiguousarray(seed, np.float64)
        if seed.shape != (n,):
            raise ValueError(f"seed must have {n} entries")
        outs = []
        arr = (L.AdjointWrt * max(len(wrt), 1))()
        for i, (kind, fid) in enumerate(wrt):
            size = self._sclbnd_size[fid] if kind == WRT_SCLBND else n
            g = np.zeros(size)
            outs.append(g)
            arr[i].kind, arr[i].id, arr[i].grad = int(kind), int(fid), ptr(g)
        st = L.AdjointStats()
        o = solve_opts(**(lin or {}))
        check(lib.vfvm_adjoint_backward(self._h, C.byref(o), ptr(seed), len(wrt), arr, C.byref(st)))
        return outs, st

    # adjoint of the coupled flow system
    def flow_adjoint_forward(self, ops, dt=0.0, num_ts=0, max_iter=20, tol=1e-8, damp=1.0, checkpoint_every=10, lin=None,
                             seed=None, weights=None):
        """Driver of the flow list (steady for num_ts = 0, else num_ts implicit-Euler steps) with
        checkpoints every checkpoint_every states; J = sum_n w_n seed . x^n is returned in
        stats.objective."""
        arr = make_ops(ops)
        st = L.FlowAdjointStats()
        o = solve_opts(**(lin or {}))
        sd = None if seed is None else np.ascontiguousarray(seed, np.float64)
        wt = None if weights is None else np.ascontiguousarray(weights, np.float64)
        if wt is not None and wt.shape != (int(num_ts),):
            raise ValueError(f"weights must have num_ts = {num_ts} entries")
        check(lib.vfvm_flow_adjoint_forward(self._h, arr, len(ops), C.byref(o), float(dt), int(num_ts), int(max_iter), float(tol),
                                            float(damp), int(checkpoint_every), None if sd is None else ptr(sd),
                                            None if wt is None else ptr(wt), C.byref(st)))
        self.system_info()
        return st

    def flow_adjoint_backward(self, seed, wrt, weights=None, inner=None, tol=1e-10, max_iter=100, restart=20, on_state=None):
        """dJ/dtheta for the last flow_adjoint_forward: wrt is a list of (WRT_SCLBND, id) or
        (WRT_VECBND, id); returns (list of gradients, stats).  inner: options of the transposed
        solves that precondition the outer GCR on J^T.  on_state(step, outer, inner_steps,
        rel_residual) is called after each adjoint state."""
        n = self._size()
        seed = np.ascontiguousarray(seed, np.float64)
        if seed.shape != (n,):
            raise ValueError(f"seed must have {n} entries")
        wt = None if weights is None else np.ascontiguousarray(weights, np.float64)
        outs = []
        arr = (L.AdjointWrt * max(len(wrt), 1))()
        for i, (kind, fid) in enumerate(wrt):
            size = self._sclbnd_size[fid] if kind == WRT_SCLBND else 3 * self._vecbnd_size[fid]
            g = np.zeros(size)
            outs.append(g)
            arr[i].kind, arr[i].id, arr[i].grad = int(kind), int(fid), ptr(g)
        o = L.FlowAdjointOpts()
        o.inner = solve_opts(**(inner or {}))
        o.tol, o.max_iter, o.restart = float(tol), int(max_iter), int(restart)
        cb = L.ADJOINT_HOOK(lambda _u, s_, a_, b_, r_: on_state(s_, a_, b_, r_)) if on_state else L.ADJOINT_HOOK()
        o.on_state = cb
        st = L.FlowAdjointStats()
        check(lib.vfvm_flow_adjoint_backward(self._h, C.byref(o), ptr(seed), None if wt is None else ptr(wt), len(wrt), arr,
                                             C.byref(st)))
        return [g.reshape(3, -1) if kind == WRT_VECBND else g for g, (kind, _) in zip(outs, wrt)], st

    def flow_residual(self, ops, x, x_prev=None, dt=0.0):
        """r = A(c(x)) x - b(c(x), x_prev) of the built flow pattern at the state x."""
        arr = make_ops(ops)
        x = np.ascontiguousarray(x, np.float64)
        xp = None if x_prev is None else np.ascontiguousarray(x_prev, np.float64)
        r = np.empty(self._size())
        check(lib.vfvm_flow_residual_eval(self._h, arr, len(ops), float(dt), ptr(x), None if xp is None else ptr(xp), ptr(r)))
        return r

    def flow_jt_apply(self, ops, x, lam, x_prev=None, dt=0.0):
        """y = J(x)^T lam, J the full Jacobian of flow_residual at x."""
        arr = make_ops(ops)
        x = np.ascontiguousarray(x, np.float64)
        lam = np.ascontiguousarray(lam, np.float64)
        xp = None if x_prev is None else np.ascontiguousarray(x_prev, np.float64)
        y = np.empty(self._size())
        check(lib.vfvm_flow_jt_apply(self._h, arr, len(ops), float(dt), ptr(x), None if xp is None else ptr(xp), ptr(lam), ptr(y)))
        return y

    # timing
    def timing_enable(self, mask=-1):
        """Bit k of mask times kernel class k (K_SPMV, K_MDOT, ...); -1: all, 0: off."""
        check(lib.vfvm_timing_enable(self._h, int(mask)))

    def timing_reset(self):
        check(lib.vfvm_timing_reset(self._h))

    def timing_get(self, kclass):
        n = C.c_int64()
        ms = C.c_double()
        by = C.c_double()
        check(lib.vfvm_timing_get(self._h, int(kclass), C.byref(n), C.byref(ms), C.byref(by)))
        return n.value, ms.value, by.value
