"""CPU oracle for the vfvm3d hot path — TEST INFRASTRUCTURE ONLY.

This module restates the reference's algorithms (msl-multiphysics/vfvchem-vfvm3d,
pure Rust) in numpy so that the HIP path can be checked against it.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg may import it, and only as
the checker or the timed CPU baseline; the product path never calls into it.

Every function cites the reference file:line it follows.  Floating-point
expressions keep the reference's operand order (elementwise numpy operations are
IEEE-rounded one at a time, like the reference's f64 arithmetic), and every
accumulation that the reference performs sequentially is performed sequentially
here (np.add.at in the reference's order), so results are reproducible to the bit.

Pinning (how each function is checked independently of itself):
  dual()            invariants: closed interior cells (sum A n = 0), sum of CV
                    volumes = mesh volume, LSQ exactness for linear fields,
                    brute-force single-tetrahedron closed form; Delaunay input
                    checked against scipy.spatial.Delaunay (Qhull).
  update_grad()     exact for linear fields (closed form).
  flux_prescribed() antisymmetry F(v->n) = -F(n->v) and F = rho (u.n) A for a
                    uniform velocity (closed form).
  assemble()        row sums, symmetry of diffusion, exactness on linear fields,
                    1D closed form T = x + q x(1-x)/(2k) of heat_analytic.rs:9,
                    power-law limits; pattern {v} u ngh(v) / {v}; weighted advection,
                    time-previous and time-unity reduce bitwise to advection-unity and
                    time; production integrates linear fields exactly; time + production
                    give the implicit-Euler map of m dc/dt = s c.
  solve (C, oracle/gmres_ref.c) and scipy SuperLU: residual and cross-check.
  ilu0()            LU = A on the pattern; exact where ILU(0) drops no fill.
  linear_map / adjoint_*   central finite differences of forward re-solves; the
                    affine identity of the map.
  grad_coeff()      sum_m c_m p_m = V grad(p) for linear p at every vertex.
  update_flow()     flux antisymmetry; hydrostatic balance (linear p with matching body
                    force gives no Rhie-Chow correction); rc > 0.
  assemble_flow()   velocity-block row sums 0; gradient block on linear p = -V grad(p);
                    divergence block on linear u = -rho V div(u); stabiliser annihilates
                    linear p; transpose / dilatatory b vanish for linear u; time terms.
  scaled_residual_flow / scaled_residual_scalar   hand-computed cases.
  picard_flow / transient_flow   exercised by the GPU drivers' parity (same iteration
                    counts) and by heatflow_solve's decoupled limit.
  heatflow_solve()  conduction limit T = C = 1 - x exactly; decoupled limit equals
                    picard_flow and a direct scalar solve; buoyancy sign.
  transient_adaptive()   fixed mode, output-time landing, error falls with rtol.
  coupled_residual / flow_jt   J^T of the Picard fixed point against central finite
                    differences of coupled_residual in random directions.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "dual", "update_grad", "update_grad_vec", "flux_prescribed", "assemble",
    "coo_to_csr", "Op", "assemble_interface", "assemble_domains",
]


# --------------------------------------------------------------------------- helpers
def _dot(a, b):
    # dot3 of src/base/delaunay.rs:190-192: a.0*b.0 + a.1*b.1 + a.2*b.2 (left to right)
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def _cross(a, b):
    # cross3 of src/base/delaunay.rs:195-197
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def _scale(s, a):
    return np.asarray(s)[..., None] * a


def _len(a):
    return np.sqrt(_dot(a, a))


def _circumcircle(a, b, c):
    # src/base/delaunay.rs:204-211
    u, v = b - a, c - a
    w = _cross(u, v)
    num = _scale(_dot(u, u), _cross(v, w)) + _scale(_dot(v, v), _cross(w, u))
    return a + _scale(0.5 / _dot(w, w), num)


def _circumsphere(a, b, c, d):
    # src/base/delaunay.rs:213-218
    u, v, w = b - a, c - a, d - a
    num = (_scale(_dot(u, u), _cross(v, w)) + _scale(_dot(v, v), _cross(w, u))) + _scale(_dot(w, w), _cross(u, v))
    return a + _scale(0.5 / _dot(u, _cross(v, w)), num)


# --------------------------------------------------------------------------- dual
def dual(pos: np.ndarray, tets: np.ndarray, num_pts: int) -> dict:
    """Voronoi dual of a tetrahedralisation; src/base/geom_tet.rs:58-293.

    pos: (N,3) float64; tets: (T,4) int in the reference orientation
    (geom_tet.rs:13).  Returns the arrays in CSR layout (neighbours ascending,
    geom_tet.rs:174-196), SoA vectors shaped (3, E) / (3, N) / (3, S).
    Raises ValueError('points on boundary') as TetraPointsOnBoundary
    (geom_tet.rs:119-120) and ValueError('singular gradient') as
    TetraSingularGradient (geom_tet.rs:284).
    """
    pos = np.asarray(pos, np.float64)
    tets = np.asarray(tets, np.int64)
    N = pos.shape[0]
    T = tets.shape[0]

    # step 1: boundary faces (geom_tet.rs:99-120)
    fkeys = np.stack([np.sort(tets[:, [(k + 1) % 4, (k + 2) % 4, (k + 3) % 4]], axis=1) for k in range(4)], axis=1)
    flat = fkeys.reshape(-1, 3)
    _, inv, counts = np.unique(flat, axis=0, return_inverse=True, return_counts=True)
    once = (counts[inv.ravel()] == 1).reshape(T, 4)
    surf_order = [[1, 3, 2], [0, 2, 3], [0, 3, 1], [0, 1, 2]]
    surf = []
    tt, kk = np.nonzero(once)  # row-major: tet order, then k order (geom_tet.rs:110-116)
    for t, k in zip(tt, kk):
        surf.append(tets[t, surf_order[k]])
    surf = np.array(surf, np.int64).reshape(-1, 3)
    if np.any(surf < num_pts):
        raise ValueError("points on boundary")

    # step 2: face pieces (geom_tet.rs:122-171), vectorised over tets, accumulated in order
    c_t = _circumsphere(pos[tets[:, 0]], pos[tets[:, 1]], pos[tets[:, 2]], pos[tets[:, 3]])
    pieces_key = []
    pieces_area = []
    pieces_sk = []
    cent_idx = []
    cent_val = []
    edges = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    rec = []  # per (edge, side) in loop order: key_min, key_max, area, area*(cent-m), (vid,val) pairs
    for (li, lj) in edges:
        other = [l for l in range(4) if l != li and l != lj]
        vi, vj = tets[:, li], tets[:, lj]
        vk, vl = tets[:, other[0]], tets[:, other[1]]
        pi, pj = pos[vi], pos[vj]
        dist = _len(pj - pi)
        e = _scale(1.0 / dist, pj - pi)
        m = _scale(0.5, pi + pj)
        sides = []
        for (va, vb) in [(vk, vl), (vl, vk)]:
            pa, pb = pos[va], pos[vb]
            c_f = _circumcircle(pi, pj, pa)
            u = (pa - m) - _scale(_dot(pa - m, e), e)
            u = _scale(1.0 / _len(u), u)
            n = _cross(e, u)
            flip = _dot(pb - m, n) < 0.0
            n = np.where(flip[:, None], -1.0 * n, n)
            s1 = _dot(c_f - m, u)
            s2 = _dot(c_t - c_f, n)
            area = 0.5 * s1 * s2
            cent = _scale(1.0 / 3.0, m + (c_f + c_t))
            skc = _scale(area, cent - m)
            vol = area * dist / 6.0
            ci = _scale(vol * 0.75, cent - pi)
            cj = _scale(vol * 0.75, cent - pj)
            sides.append((area, skc, ci, cj))
        rec.append((np.minimum(vi, vj), np.maximum(vi, vj), vi, vj, sides))
    # sequence index so that the sequential accumulation follows tet, edge, side order
    K = len(edges) * 2
    kmin = np.empty((T, K), np.int64)
    kmax = np.empty((T, K), np.int64)
    ar = np.empty((T, K))
    sk = np.empty((T, K, 3))
    cv = np.empty((T, 2 * K), np.int64)
    cc = np.empty((T, 2 * K, 3))
    for q, (a, b, vi, vj, sides) in enumerate(rec):
        for s, (area, skc, ci, cj) in enumerate(sides):
            col = 2 * q + s
            kmin[:, col] = a
            kmax[:, col] = b
            ar[:, col] = area
            sk[:, col] = skc
            cv[:, 2 * col] = vi
            cc[:, 2 * col] = ci
            cv[:, 2 * col + 1] = vj
            cc[:, 2 * col + 1] = cj
    key = kmin.ravel() * N + kmax.ravel()
    ukey, kinv = np.unique(key, return_inverse=True)
    e_area = np.zeros(ukey.size)
    e_sk = np.zeros((ukey.size, 3))
    np.add.at(e_area, kinv, ar.ravel())
    for d in range(3):
        col = np.zeros(ukey.size)
        np.add.at(col, kinv, sk.reshape(-1, 3)[:, d])
        e_sk[:, d] = col
    cent_sum = np.zeros((N, 3))
    for d in range(3):
        col = np.zeros(N)
        np.add.at(col, cv.ravel(), cc.reshape(-1, 3)[:, d])
        cent_sum[:, d] = col

    # step 3: sorted edges -> neighbour lists, face data, CV volume (geom_tet.rs:173-201)
    ea = ukey // N
    eb = ukey % N
    pa, pb = pos[ea], pos[eb]
    dist = _len(pb - pa)
    norm = _scale(1.0 / dist, pb - pa)
    big = np.abs(e_area) > 1e-12 * dist * dist
    with np.errstate(divide="ignore", invalid="ignore"):
        skew = np.where(big[:, None], _scale(1.0 / e_area, e_sk), 0.0)
    # directed edges: rows ascending, neighbours ascending
    src = np.concatenate([ea, eb])
    dst = np.concatenate([eb, ea])
    und = np.concatenate([np.arange(ukey.size), np.arange(ukey.size)])
    sgn = np.concatenate([np.ones(ukey.size), -np.ones(ukey.size)])
    order = np.lexsort((dst, src))
    src, dst, und, sgn = src[order], dst[order], und[order], sgn[order]
    E = src.size
    row_ptr = np.zeros(N + 1, np.int64)
    np.add.at(row_ptr, src + 1, 1)
    row_ptr = np.cumsum(row_ptr)
    # CV volume accumulated over the sorted edge list: for each vertex, edges in key order
    vol = np.zeros(N)
    contrib = e_area * dist / 6.0
    seq_v = np.stack([ea, eb], axis=1).ravel()
    seq_c = np.stack([contrib, contrib], axis=1).ravel()
    np.add.at(vol, seq_v, seq_c)
    cskew = _scale(1.0 / vol, cent_sum)

    out = dict(
        num_pts=int(num_pts), N=N,
        row_ptr=row_ptr.astype(np.int32), ngh=dst.astype(np.int32),
        dist=dist[und], area=e_area[und],
        norm=(norm[und] * sgn[:, None]).T.copy(),
        skew=skew[und].T.copy(), vol=vol, cskew=cskew.T.copy(),
    )
    if np.any(sgn < 0):
        # scale3(-1.0, norm) is an exact sign flip, as multiplying by -1 here
        pass

    # step 4: boundary sub-faces (geom_tet.rs:208-259), per surface triangle in order
    bvid, bvj, bvk, barea, bnorm, bwgt = [], [], [], [], [], []
    bnd_sum = np.zeros(N)
    if surf.size:
        p0, p1, p2 = pos[surf[:, 0]], pos[surf[:, 1]], pos[surf[:, 2]]
        nrm = _cross(p1 - p0, p2 - p0)
        nrm = _scale(1.0 / _len(nrm), nrm)
        c = _circumcircle(p0, p1, p2)
        sq = [_dot(p2 - p1, p2 - p1), _dot(p0 - p2, p0 - p2), _dot(p1 - p0, p1 - p0)]
        lam_raw = [sq[0] * (sq[1] + sq[2] - sq[0]), sq[1] * (sq[2] + sq[0] - sq[1]), sq[2] * (sq[0] + sq[1] - sq[2])]
        lam_sum = lam_raw[0] + lam_raw[1] + lam_raw[2]
        lam = [lr / lam_sum for lr in lam_raw]
        per_loc = []
        for loc in range(3):
            vi, vj, vk = surf[:, loc], surf[:, (loc + 1) % 3], surf[:, (loc + 2) % 3]
            pi, pj, pk = pos[vi], pos[vj], pos[vk]
            area = []
            for (pa_, pb_) in [(pj, pk), (pk, pj)]:
                m = _scale(0.5, pi + pa_)
                e = pa_ - pi
                ln = _len(e)
                u = (pb_ - m) - _scale(_dot(pb_ - m, e) / (ln * ln), e)
                area.append(0.25 * ln * _dot(c - m, _scale(1.0 / _len(u), u)))
            area_sum = area[0] + area[1]
            l0, l1, l2 = lam[loc], lam[(loc + 1) % 3], lam[(loc + 2) % 3]
            cent_j = np.stack([(1.5 + l0) / 3.0, (0.5 + l1) / 3.0, l2 / 3.0], axis=-1)
            cent_k = np.stack([(1.5 + l0) / 3.0, l1 / 3.0, (0.5 + l2) / 3.0], axis=-1)
            ok = np.abs(area_sum) > 1e-12 * _len(pj - pi) * _len(pk - pi)
            with np.errstate(divide="ignore", invalid="ignore"):
                w = _scale(1.0 / area_sum, _scale(area[0], cent_j) + _scale(area[1], cent_k))
            w = np.where(ok[:, None], w, np.array([1.0, 0.0, 0.0]))
            per_loc.append((vi, vj, vk, area_sum, w))
        # push order: triangle by triangle, loc 0..2 (geom_tet.rs:213-258)
        nf = surf.shape[0]
        vi_all = np.stack([p[0] for p in per_loc], axis=1).ravel()
        vj_all = np.stack([p[1] for p in per_loc], axis=1).ravel()
        vk_all = np.stack([p[2] for p in per_loc], axis=1).ravel()
        ar_all = np.stack([p[3] for p in per_loc], axis=1).ravel()
        w_all = np.stack([p[4] for p in per_loc], axis=1).reshape(-1, 3)
        n_all = np.repeat(nrm, 3, axis=0)
        np.add.at(bnd_sum, vi_all, ar_all)
        o = np.argsort(vi_all, kind="stable")
        bsf_ptr = np.zeros(N + 1, np.int64)
        np.add.at(bsf_ptr, vi_all + 1, 1)
        bsf_ptr = np.cumsum(bsf_ptr)
        out.update(bsf_ptr=bsf_ptr.astype(np.int32), bsf_vj=vj_all[o].astype(np.int32),
                   bsf_vk=vk_all[o].astype(np.int32), bsf_area=ar_all[o],
                   bsf_norm=n_all[o].T.copy(), bsf_wgt=w_all[o].T.copy())
    else:
        out.update(bsf_ptr=np.zeros(N + 1, np.int32), bsf_vj=np.zeros(0, np.int32), bsf_vk=np.zeros(0, np.int32),
                   bsf_area=np.zeros(0), bsf_norm=np.zeros((3, 0)), bsf_wgt=np.zeros((3, 0)))
    out["bnd_sum"] = bnd_sum
    out["surf"] = surf

    # step 5: least-squares gradient weights (geom_tet.rs:261-293)
    rows = np.repeat(np.arange(N), np.diff(row_ptr))
    d = pos[dst] - pos[rows]
    omega = 1.0 / (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    M = np.zeros((N, 3, 3))
    for r in range(3):
        for s in range(3):
            col = np.zeros(N)
            np.add.at(col, rows, omega * d[:, r] * d[:, s])
            M[:, r, s] = col
    m_ = M
    adj = np.empty((N, 3, 3))
    adj[:, 0, 0] = m_[:, 1, 1] * m_[:, 2, 2] - m_[:, 1, 2] * m_[:, 2, 1]
    adj[:, 0, 1] = m_[:, 0, 2] * m_[:, 2, 1] - m_[:, 0, 1] * m_[:, 2, 2]
    adj[:, 0, 2] = m_[:, 0, 1] * m_[:, 1, 2] - m_[:, 0, 2] * m_[:, 1, 1]
    adj[:, 1, 0] = m_[:, 1, 2] * m_[:, 2, 0] - m_[:, 1, 0] * m_[:, 2, 2]
    adj[:, 1, 1] = m_[:, 0, 0] * m_[:, 2, 2] - m_[:, 0, 2] * m_[:, 2, 0]
    adj[:, 1, 2] = m_[:, 0, 2] * m_[:, 1, 0] - m_[:, 0, 0] * m_[:, 1, 2]
    adj[:, 2, 0] = m_[:, 1, 0] * m_[:, 2, 1] - m_[:, 1, 1] * m_[:, 2, 0]
    adj[:, 2, 1] = m_[:, 0, 1] * m_[:, 2, 0] - m_[:, 0, 0] * m_[:, 2, 1]
    adj[:, 2, 2] = m_[:, 0, 0] * m_[:, 1, 1] - m_[:, 0, 1] * m_[:, 1, 0]
    det = m_[:, 0, 0] * adj[:, 0, 0] + m_[:, 0, 1] * adj[:, 1, 0] + m_[:, 0, 2] * adj[:, 2, 0]
    tr = m_[:, 0, 0] + m_[:, 1, 1] + m_[:, 2, 2]
    if not np.all(det > 1e-12 * (tr * tr * tr)):
        raise ValueError("singular gradient")
    om2 = 1.0 / _dot(d, d)
    gw = np.empty((3, E))
    for r in range(3):
        gw[r] = om2 * (adj[rows, r, 0] * d[:, 0] + adj[rows, r, 1] * d[:, 1] + adj[rows, r, 2] * d[:, 2]) / det[rows]
    out["gw"] = gw
    out["rows"] = rows.astype(np.int32)
    return out


# --------------------------------------------------------------------------- gradients
def update_grad(g: dict, value: np.ndarray) -> np.ndarray:
    """LSQ gradient, ScalarDomain::update_grad (src/base/scl_dom.rs:146-164).

    grad_c = sum_loc w_{c,loc} (u_n - u_c), accumulated in loc order from 0.0.
    Returns (3, N).
    """
    rows, ngh = g["rows"], g["ngh"]
    delta = value[ngh] - value[rows]
    out = np.zeros((3, g["N"]))
    for r in range(3):
        col = np.zeros(g["N"])
        np.add.at(col, rows, g["gw"][r] * delta)
        out[r] = col
    return out


def update_grad_vec(g: dict, value3: np.ndarray) -> np.ndarray:
    """9-component gradient, VectorDomain::update_grad (src/base/vec_dom.rs:146-178).

    value3: (3, N) SoA.  Returns (9, N) ordered (u_x,u_y,u_z,v_x,...,w_z).
    """
    out = np.empty((9, g["N"]))
    for c in range(3):
        out[3 * c:3 * c + 3] = update_grad(g, value3[c])
    return out


# --------------------------------------------------------------------------- flow
def flux_prescribed(g: dict, vel: np.ndarray, vel_grad: np.ndarray, den: np.ndarray, den_grad: np.ndarray) -> np.ndarray:
    """Face mass flux of flow_helper::update_flux (src/operator/core/flow_helper.rs:118-171)
    for a prescribed velocity with uniform pressure and zero body force.

    With p constant the pressure-difference term (p_n - p_v)/dn is exactly zero and
    fce = 0, so vel_rc (flow_helper.rs:167) reduces to the momentum interpolation term
    mom_e, which is rc * grad(p) with grad(p) = p * sum_m c_m / V = 0 for a closed
    control volume; it is taken as 0 here (a round-off-level difference from the
    reference, which evaluates sum_m c_m in floating point).
    F = den_f * (u_f nx + v_f ny + w_f nz) * A, den_f and u_f skew-corrected
    arithmetic means (flow_helper.rs:158-161).  Returns (E,).
    """
    rows, ngh = g["rows"], g["ngh"]
    sx, sy, sz = g["skew"]
    nx, ny, nz = g["norm"]
    A = g["area"]
    dv, dn = den[rows], den[ngh]
    den_f = 0.5 * (dv + dn) + 0.5 * (sx * (den_grad[0][rows] + den_grad[0][ngh]) + sy * (den_grad[1][rows] + den_grad[1][ngh]) + sz * (den_grad[2][rows] + den_grad[2][ngh]))

    def face(c):
        val = vel[c]
        gx, gy, gz = vel_grad[3 * c], vel_grad[3 * c + 1], vel_grad[3 * c + 2]
        return 0.5 * (val[rows] + val[ngh]) + 0.5 * (sx * (gx[rows] + gx[ngh]) + sy * (gy[rows] + gy[ngh]) + sz * (gz[rows] + gz[ngh]))

    u_f, v_f, w_f = face(0), face(1), face(2)
    vel_rc = 0.0
    return den_f * (u_f * nx + v_f * ny + w_f * nz + vel_rc) * A


# --------------------------------------------------------------------------- assembly
class Op:
    """Operator descriptor mirroring the reference operator structs (same kinds and
    id meanings as include/vfvm.h's vfvm_op)."""

    SCL_DOM_DIFFUSION = 1
    SCL_DOM_PRODUCTION = 2
    SCL_DOM_SOURCE = 3
    SCL_DOM_TIME = 4
    SCL_DOM_ADVECTION_UNITY = 5
    SCL_BND_DIRICHLET = 20
    SCL_BND_NEUMANN = 21
    SCL_BND_DIFFUSION = 22
    SCL_BND_PRODUCTION = 23
    SCL_BND_SOURCE = 24
    SCL_BND_TIME = 25
    SCL_BND_ADVECTION_UNITY = 26
    SCL_DOM_ADVECTION = 6
    SCL_DOM_TIME_PREVIOUS = 7
    SCL_DOM_TIME_UNITY = 8
    SCL_BND_ADVECTION = 27
    SCL_BND_TIME_PREVIOUS = 28
    SCL_BND_TIME_UNITY = 29

    def __init__(self, kind, geom_id, unk_id, a=-1, b=-1, c=-1, value=0.0):
        self.kind, self.geom_id, self.unk_id, self.a, self.b = kind, geom_id, unk_id, a, b
        self.c, self.value = c, value


def _diff_face(G, Gg, v, n, skew):
    # harmonic mean + skew correction + guard; opscl_dom_diff.rs:78-80
    Gv, Gn = G[v], G[n]
    ssum = Gv + Gn
    with np.errstate(divide="ignore", invalid="ignore"):
        diff_m = np.where(ssum != 0.0, 2.0 * Gv * Gn / ssum, 0.0)
    sx, sy, sz = skew
    diff_f = diff_m + 0.5 * (sx * (Gg[0][v] + Gg[0][n]) + sy * (Gg[1][v] + Gg[1][n]) + sz * (Gg[2][v] + Gg[2][n]))
    return np.where(diff_f > 0.0, diff_f, diff_m)


def _pow5(x):
    # powi(5) as binary exponentiation: x^4 * x
    x2 = x * x
    x4 = x2 * x2
    return x4 * x


def _adv_coeff(cond, F):
    """Power-law upwind coefficient of a face (opscl_dom_advunity.rs:96-108, opscl_dom_adv.rs:
    106-118): Pe = F / cond (0 when cond = 0), A = max(0, 1 - 0.1 |Pe|)^5 and
    coeff = cond (A - 1) + max(-F, 0), added on (v, n) and subtracted on (v, v)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        pe = np.where(cond != 0.0, F / cond, 0.0)
    pl = _pow5(np.maximum(1.0 - 0.1 * np.abs(pe), 0.0))
    return cond * (pl - 1.0) + np.maximum(-F, 0.0)


def assemble(g: dict, ops, scl: dict, sclbnd: dict, flux: dict, bnd_nid: list, dt: float = 0.0,
             value_prev: dict | None = None, flow_den: dict | None = None):
    """Triplets and b of assemble_matrix for the scalar operators, in push order.

    g: dual(); ops: list of Op in the caller's assemble_matrix order; scl: id ->
    (value[N], grad[3,N]); sclbnd: id -> values per boundary node; flux: flw id ->
    F[E]; bnd_nid: list of node-id arrays per boundary; value_prev: field id -> previous
    values; flow_den: flw id -> (rho[N], grad rho[3,N]) of the flow (OpSclDomAdvection).
    The single unknown's block starts at 0 (SteadyHeat::initial_start, steady_heat.rs:68-78).
    Returns rows, cols, vals (push order), b.
    """
    N, NP = g["N"], g["num_pts"]
    rp, ngh = g["row_ptr"], g["ngh"]
    rowsE = g["rows"]
    R, C, V = [], [], []
    b = np.zeros(N)
    pts = np.arange(NP)
    edge_of_pts = rowsE < NP

    def edges_of(vset_mask_rows):
        # directed edges whose row is in the vertex set, in (row, loc) order
        return np.nonzero(vset_mask_rows)[0]

    def push_pairs(e_idx, coeff):
        # per face: (v, n) +coeff then (v, v) -coeff (opscl_dom_diff.rs:84-85)
        v = rowsE[e_idx]
        n = ngh[e_idx]
        rr = np.stack([v, v], axis=1).ravel()
        cc = np.stack([n, v], axis=1).ravel()
        vv = np.stack([coeff, -coeff], axis=1).ravel()
        R.append(rr); C.append(cc); V.append(vv)

    def b_add(vids, vals):
        np.add.at(b, vids, vals)

    for op in ops:
        k = op.kind
        if k < 20:
            vids = pts
            emask = edge_of_pts
        else:
            vids = NP + np.asarray(bnd_nid[op.geom_id], np.int64)
            m = np.zeros(N, bool)
            m[vids] = True
            emask = m[rowsE]
        if k in (Op.SCL_DOM_DIFFUSION, Op.SCL_BND_DIFFUSION):
            # opscl_dom_diff.rs:40-91, opscl_bnd_diff.rs:40-96
            G, Gg = scl[op.a]
            e = edges_of(emask)
            df = _diff_face(G, Gg, rowsE[e], ngh[e], g["skew"][:, e])
            push_pairs(e, df * g["area"][e] / g["dist"][e])
        elif k in (Op.SCL_DOM_ADVECTION_UNITY, Op.SCL_BND_ADVECTION_UNITY):
            # opscl_dom_advunity.rs:46-115, opscl_bnd_advunity.rs:52-..
            e = edges_of(emask)
            if op.a >= 0:
                G, Gg = scl[op.a]
                df = _diff_face(G, Gg, rowsE[e], ngh[e], g["skew"][:, e])
            else:
                df = np.zeros(e.size)
            cond = df * g["area"][e] / g["dist"][e]
            push_pairs(e, _adv_coeff(cond, flux[op.b][e]))
        elif k in (Op.SCL_DOM_ADVECTION, Op.SCL_BND_ADVECTION):
            # opscl_dom_adv.rs:49-125, opscl_bnd_adv.rs: flux = m_f / rho_f * F
            e = edges_of(emask)
            v_, n_ = rowsE[e], ngh[e]
            sx, sy, sz = g["skew"][:, e]
            G, Gg = scl[op.b]
            df = _diff_face(G, Gg, v_, n_, g["skew"][:, e])
            W, Wg = scl[op.a]
            Rr, Rg = flow_den[op.c]
            wgt_f = 0.5 * (W[v_] + W[n_]) + 0.5 * (sx * (Wg[0][v_] + Wg[0][n_]) + sy * (Wg[1][v_] + Wg[1][n_]) + sz * (Wg[2][v_] + Wg[2][n_]))
            den_f = 0.5 * (Rr[v_] + Rr[n_]) + 0.5 * (sx * (Rg[0][v_] + Rg[0][n_]) + sy * (Rg[1][v_] + Rg[1][n_]) + sz * (Rg[2][v_] + Rg[2][n_]))
            cond = df * g["area"][e] / g["dist"][e]
            with np.errstate(divide="ignore", invalid="ignore"):
                F = np.where(den_f != 0.0, wgt_f / den_f * flux[op.c][e], 0.0)
            push_pairs(e, _adv_coeff(cond, F))
        elif k in (Op.SCL_DOM_TIME_PREVIOUS, Op.SCL_BND_TIME_PREVIOUS, Op.SCL_DOM_TIME_UNITY, Op.SCL_BND_TIME_UNITY):
            # opscl_dom_timeprev.rs:36-52, opscl_dom_timeunity.rs:35-50 (apply_time)
            if not dt > 0.0:
                raise ValueError("OperatorUseApplyTime")
            vol = g["vol"][vids]
            prev = value_prev[op.unk_id]
            if k in (Op.SCL_DOM_TIME_PREVIOUS, Op.SCL_BND_TIME_PREVIOUS):
                Wp = value_prev[op.a]
                R.append(vids); C.append(vids); V.append(-Wp[vids] * vol / dt)
                b_add(vids, -Wp[vids] * prev[vids] * vol / dt)
            else:
                R.append(vids); C.append(vids); V.append(-vol / dt)
                b_add(vids, -prev[vids] * vol / dt)
        elif k in (Op.SCL_DOM_PRODUCTION, Op.SCL_BND_PRODUCTION):
            # opscl_dom_prod.rs:39-76
            P, Pg = scl[op.a]
            _, Ug = scl[op.unk_id]
            sx, sy, sz = g["cskew"][:, vids]
            vol = g["vol"][vids]
            prd_c = P[vids] + Pg[0][vids] * sx + Pg[1][vids] * sy + Pg[2][vids] * sz
            unk_skew = Ug[0][vids] * sx + Ug[1][vids] * sy + Ug[2][vids] * sz
            R.append(vids); C.append(vids); V.append(prd_c * vol)
            b_add(vids, -P[vids] * unk_skew * vol)
        elif k in (Op.SCL_DOM_SOURCE, Op.SCL_BND_SOURCE):
            # opscl_dom_src.rs:37-68
            S, Sg = scl[op.a]
            sx, sy, sz = g["cskew"][:, vids]
            src_c = S[vids] + Sg[0][vids] * sx + Sg[1][vids] * sy + Sg[2][vids] * sz
            b_add(vids, -src_c * g["vol"][vids])
        elif k in (Op.SCL_DOM_TIME, Op.SCL_BND_TIME):
            # opscl_dom_time.rs:35-62 (apply_time)
            if not dt > 0.0:
                raise ValueError("OperatorUseApplyTime")
            W, _ = scl[op.a]
            prev = value_prev[op.unk_id]
            vol = g["vol"][vids]
            R.append(vids); C.append(vids); V.append(-W[vids] * vol / dt)
            b_add(vids, -W[vids] * prev[vids] * vol / dt)
        elif k == Op.SCL_BND_DIRICHLET:
            # opscl_bnd_dir.rs:56-78
            val = np.asarray(sclbnd[op.a], np.float64)
            R.append(vids); C.append(vids); V.append(np.ones(vids.size))
            b_add(vids, val)
        elif k == Op.SCL_BND_NEUMANN:
            # opscl_bnd_neu.rs:36-87: per node, per sub-face, b += A (w_v N_v + w_j N_j + w_k N_k)
            val = np.asarray(sclbnd[op.a], np.float64)
            idx_of = -np.ones(N, np.int64)
            idx_of[vids] = np.arange(vids.size)
            bp = g["bsf_ptr"]
            for ii, v in enumerate(vids):
                fv = val[ii]
                for s in range(bp[v], bp[v + 1]):
                    j, kk = g["bsf_vj"][s], g["bsf_vk"][s]
                    fj = val[idx_of[j]] if idx_of[j] >= 0 else fv
                    fk = val[idx_of[kk]] if idx_of[kk] >= 0 else fv
                    wv, wj, wk = g["bsf_wgt"][:, s]
                    b[v] += g["bsf_area"][s] * (wv * fv + wj * fj + wk * fk)
        else:
            raise ValueError(f"unsupported operator kind {k}")
    rows = np.concatenate(R) if R else np.zeros(0, np.int64)
    cols = np.concatenate(C) if C else np.zeros(0, np.int64)
    vals = np.concatenate(V) if V else np.zeros(0)
    return rows.astype(np.int64), cols.astype(np.int64), vals, b


def coo_to_csr(rows, cols, vals, n):
    """Pattern + duplicate summation of SparseColMat::try_new_from_indices /
    new_from_argsort (linear_gmres.rs:59,71), stored row-major.

    Duplicates are summed sequentially in push order (a stable sort keeps it).
    Returns row_ptr[n+1], col[nnz], val[nnz].
    """
    order = np.lexsort((cols, rows))  # stable: push order among equal (row, col)
    r, c = rows[order], cols[order]
    key = r * n + c
    uk, first = np.unique(key, return_index=True)
    seg = np.searchsorted(uk, key)
    out = np.zeros(uk.size)
    np.add.at(out, seg, vals[order])
    ur = uk // n
    uc = uk % n
    row_ptr = np.zeros(n + 1, np.int64)
    np.add.at(row_ptr, ur + 1, 1)
    return np.cumsum(row_ptr), uc.astype(np.int64), out


# --------------------------------------------------------------------------- interfaces
# Multi-domain systems (Variables with several domains, src/base/vars.rs:7-34): each domain
# has its own dual (its own Tetrahedron) and its own unknown scalar, whose block starts where
# the previous domain's ends (SteadyBase::initial_start); an interface node has one vertex in
# each of its two domains.  The interface operators (src/operator/core/opscl_itf_*.rs) load
# the continuity row into side a's rows and the combined control volume into side b's rows.

ITF_CONTINUITY, ITF_DIFFUSION, ITF_PRODUCTION, ITF_SOURCE = 70, 71, 72, 73
ITF_TIME, ITF_TIME_PREVIOUS, ITF_TIME_UNITY, ITF_ADVECTION, ITF_ADVECTION_UNITY = 74, 75, 76, 77, 78


def assemble_interface(doms, starts, itf, ops, fields, dt=0.0, value_prev=None, flux=None, flow_den=None):
    """Triplets (push order) and b contributions of the scalar interface operators.

    doms: per-domain dual() dicts; starts: first row of each domain's unknown block;
    itf: (dom_a, dom_b, va, vb) with va[q], vb[q] the local vertex of node q in domain a / b
    (the interface's node order, geom_itf.rs:37-44); ops: Op list (kinds 70-78, value = flowing
    side 0/1 of the advection operators); fields[id][d] = (value, grad[3]) on domain d;
    value_prev[id][d]; flux[flw][d] = face flux of domain d's directed edges; flow_den[flw][d] =
    (rho, grad rho).  Returns rows, cols, vals, b_rows, b_vals in global numbering."""
    da, db, va, vb = itf
    va = np.asarray(va, np.int64)
    vb = np.asarray(vb, np.int64)
    sides = ((da, va), (db, vb))
    row_a = starts[da] + va
    row_b = starts[db] + vb
    R, C, V, BR, BV = [], [], [], [], []

    def face_edges(g, vids):
        # directed edges of the vertices vids in (node, loc) order, with the node index of each
        rp = g["row_ptr"]
        cnt = rp[vids + 1] - rp[vids]
        q = np.repeat(np.arange(vids.size), cnt)
        e = np.concatenate([np.arange(rp[v], rp[v + 1]) for v in vids]) if vids.size else np.zeros(0, np.int64)
        return q, e.astype(np.int64)

    for op in ops:
        k = op.kind
        if k == ITF_CONTINUITY:
            # opscl_itf_cont.rs:44-56: (a_v, a_v) +1, (a_v, b_v) -1 per node
            R.append(np.stack([row_a, row_a], 1).ravel())
            C.append(np.stack([row_a, starts[db] + vb], 1).ravel())
            V.append(np.tile([1.0, -1.0], va.size))
            continue
        for si, (d, vids) in enumerate(sides):
            g = doms[d]
            s0 = starts[d]
            if k in (ITF_DIFFUSION, ITF_ADVECTION, ITF_ADVECTION_UNITY):
                if k != ITF_DIFFUSION and si != int(op.value):
                    continue
                q, e = face_edges(g, vids)
                v_, n_ = g["rows"][e], g["ngh"][e]
                skew = g["skew"][:, e]
                cond_of = lambda fid: (_diff_face(*fields[fid][d], v_, n_, skew) * g["area"][e] / g["dist"][e])
                if k == ITF_DIFFUSION:
                    # opscl_itf_diff.rs:73-108: both sides' faces, rows of side b
                    coeff = cond_of(op.a)
                elif k == ITF_ADVECTION_UNITY:
                    # opscl_itf_advunity.rs:85-135: the flowing side's faces
                    cond = cond_of(op.a) if op.a >= 0 else np.zeros(e.size)
                    coeff = _adv_coeff(cond, flux[op.b][d][e])
                else:
                    # opscl_itf_adv.rs:86-148: flux = m_f / rho_f F
                    sx, sy, sz = skew
                    W, Wg = fields[op.a][d]
                    Rr, Rg = flow_den[op.c][d]
                    wgt_f = 0.5 * (W[v_] + W[n_]) + 0.5 * (sx * (Wg[0][v_] + Wg[0][n_]) + sy * (Wg[1][v_] + Wg[1][n_]) + sz * (Wg[2][v_] + Wg[2][n_]))
                    den_f = 0.5 * (Rr[v_] + Rr[n_]) + 0.5 * (sx * (Rg[0][v_] + Rg[0][n_]) + sy * (Rg[1][v_] + Rg[1][n_]) + sz * (Rg[2][v_] + Rg[2][n_]))
                    with np.errstate(divide="ignore", invalid="ignore"):
                        F = np.where(den_f != 0.0, wgt_f / den_f * flux[op.c][d][e], 0.0)
                    coeff = _adv_coeff(cond_of(op.b), F)
                rb = row_b[q]
                R.append(np.stack([rb, rb], 1).ravel())
                C.append(np.stack([s0 + n_, s0 + v_], 1).ravel())
                V.append(np.stack([coeff, -coeff], 1).ravel())
            elif k == ITF_PRODUCTION:
                # opscl_itf_prod.rs:60-92: both halves on (b_v, b_v)
                P, Pg = fields[op.a][d]
                _, Ug = fields[op.unk_id][d]
                sx, sy, sz = g["cskew"][:, vids]
                vol = g["vol"][vids]
                prd_c = P[vids] + Pg[0][vids] * sx + Pg[1][vids] * sy + Pg[2][vids] * sz
                unk_skew = Ug[0][vids] * sx + Ug[1][vids] * sy + Ug[2][vids] * sz
                R.append(row_b); C.append(row_b); V.append(prd_c * vol)
                BR.append(row_b); BV.append(-P[vids] * unk_skew * vol)
            elif k == ITF_SOURCE:
                # opscl_itf_src.rs:57-83
                S, Sg = fields[op.a][d]
                sx, sy, sz = g["cskew"][:, vids]
                src_c = S[vids] + Sg[0][vids] * sx + Sg[1][vids] * sy + Sg[2][vids] * sz
                BR.append(row_b); BV.append(-src_c * g["vol"][vids])
            elif k in (ITF_TIME, ITF_TIME_PREVIOUS, ITF_TIME_UNITY):
                # opscl_itf_time.rs:50-76, opscl_itf_timeprev.rs:50-78, opscl_itf_timeunity.rs:47-70:
                # each side's term on that side's copy of the unknown
                if not dt > 0.0:
                    raise ValueError("OperatorUseApplyTime")
                vol = g["vol"][vids]
                prev = value_prev[op.unk_id][d][vids]
                if k == ITF_TIME_UNITY:
                    w = np.ones(vids.size)
                elif k == ITF_TIME:
                    w = fields[op.a][d][0][vids]
                else:
                    w = value_prev[op.a][d][vids]
                R.append(row_b); C.append(s0 + vids); V.append(-w * vol / dt)
                BR.append(row_b); BV.append(-w * prev * vol / dt)
            else:
                raise ValueError(f"unsupported interface operator kind {k}")
    cat = lambda xs, dt_: np.concatenate(xs).astype(dt_) if xs else np.zeros(0, dt_)
    return cat(R, np.int64), cat(C, np.int64), cat(V, np.float64), cat(BR, np.int64), cat(BV, np.float64)


def assemble_domains(doms, bnd_nids, dom_ops, itfs, itf_ops, fields, sclbnd, dt=0.0, value_prev=None, flux=None,
                     flow_den=None):
    """The whole multi-domain system: per domain d, assemble() of its domain and boundary
    operators dom_ops[d] on its own dual with its fields (fields[id][d]), shifted to its block;
    then the interface operators itf_ops[k] of every interface itfs[k].  Returns the CSR
    (row_ptr, col, val), b and the block starts."""
    starts = np.concatenate([[0], np.cumsum([g["N"] for g in doms])]).astype(np.int64)
    Ntot = int(starts[-1])
    R, C, V = [], [], []
    b = np.zeros(Ntot)
    for d, g in enumerate(doms):
        scl_d = {fid: per[d] for fid, per in fields.items() if d in per}
        prev_d = {fid: per[d] for fid, per in (value_prev or {}).items() if d in per}
        flux_d = {f: per[d] for f, per in (flux or {}).items() if d in per}
        den_d = {f: per[d] for f, per in (flow_den or {}).items() if d in per}
        r, c, v, bd = assemble(g, dom_ops[d], scl_d, sclbnd[d], flux_d, bnd_nids[d], dt=dt, value_prev=prev_d,
                               flow_den=den_d)
        R.append(r + starts[d]); C.append(c + starts[d]); V.append(v)
        b[starts[d]:starts[d + 1]] += bd
    for itf, ops in zip(itfs, itf_ops):
        r, c, v, br, bv = assemble_interface(doms, starts, itf, ops, fields, dt, value_prev, flux, flow_den)
        R.append(r); C.append(c); V.append(v)
        np.add.at(b, br, bv)
    rp, col, val = coo_to_csr(np.concatenate(R), np.concatenate(C), np.concatenate(V), Ntot)
    return rp, col, val, b, starts


# --------------------------------------------------------------------------- CPU GMRES (C)
def _liboracle():
    import ctypes as C
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "liboracle.so")
    lib = C.CDLL(path)
    P64 = C.POINTER(C.c_int64)
    PD = C.POINTER(C.c_double)
    lib.orc_gmres.restype = C.c_int
    lib.orc_gmres.argtypes = [C.c_int64, P64, P64, PD, PD, PD, C.c_int, C.c_int, C.c_double, C.c_int64, P64, PD]
    lib.orc_gmres_threads.restype = C.c_int
    lib.orc_gmres_threads.argtypes = [C.c_int64, P64, P64, PD, PD, PD, C.c_int, C.c_int, C.c_double, C.c_int64, C.c_int, P64, PD]
    return lib


def gmres(rp, col, val, b, x0=None, restart=50, max_outer=200, tol=1e-6, max_steps=0, threads=1):
    """Jacobi-preconditioned restarted GMRES in FP64 (oracle/gmres_ref.c;
    LinearGmres::solve_matrix, linear_gmres.rs:76-84), on one thread unless `threads` > 1
    (OpenMP vector loops; used to time the CPU baseline on all host cores).
    Returns (x, converged, steps, rel_residual)."""
    import ctypes as C
    lib = _liboracle()
    n = len(rp) - 1
    rp = np.ascontiguousarray(rp, np.int64)
    col = np.ascontiguousarray(col, np.int64)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    it = C.c_int64()
    rel = C.c_double()
    P64 = C.POINTER(C.c_int64)
    PD = C.POINTER(C.c_double)
    if threads > 1:
        st = lib.orc_gmres_threads(n, rp.ctypes.data_as(P64), col.ctypes.data_as(P64), val.ctypes.data_as(PD),
                                   b.ctypes.data_as(PD), x.ctypes.data_as(PD), int(restart), int(max_outer), float(tol),
                                   int(max_steps), int(threads), C.byref(it), C.byref(rel))
    else:
        st = lib.orc_gmres(n, rp.ctypes.data_as(P64), col.ctypes.data_as(P64), val.ctypes.data_as(PD), b.ctypes.data_as(PD),
                           x.ctypes.data_as(PD), int(restart), int(max_outer), float(tol), int(max_steps), C.byref(it), C.byref(rel))
    return x, st == 0, it.value, rel.value


# --------------------------------------------------------------------------- discrete adjoint
# The reference has no adjoint; these functions restate the textbook discrete adjoint of
# the implicit-Euler map built from assemble() (one linear solve per step, operator
# coefficients independent of the unknown).  They are pinned by central finite differences
# of linear_map() itself (tests/test_cpu_adjoint_oracle.py), not by a second copy of the
# formulas.  Sparse LU (scipy SuperLU) solves both the forward and the transposed systems.

def _op_rows(g, op, bnd_nid):
    """Vertices of the operator's row class: points for domain operators, the boundary's
    nodes for boundary operators (operator_base.rs, xid = start + vid)."""
    if op.kind < 20:
        return np.arange(g["num_pts"])
    return g["num_pts"] + np.asarray(bnd_nid[op.geom_id], np.int64)


def _csr_system(g, ops, scl, sclbnd, flux, bnd_nid, unk, prev, dt):
    import scipy.sparse as sp
    N = g["N"]
    r, c, v, b = assemble(g, ops, scl, sclbnd, flux, bnd_nid, dt=dt, value_prev={unk: prev})
    rp, col, val = coo_to_csr(r, c, v, N)
    return sp.csr_matrix((val, col, rp), shape=(N, N)), b


def apply_initial_dirichlet(g, ops, sclbnd, bnd_nid, T0):
    """OpSclBndDirichlet::apply_initial (opscl_bnd_dir.rs:34-53) on a copy of T0."""
    T = np.array(T0, np.float64)
    for op in ops:
        if op.kind == Op.SCL_BND_DIRICHLET:
            T[g["num_pts"] + np.asarray(bnd_nid[op.geom_id], np.int64)] = sclbnd[op.a]
    return T


def linear_map(g, ops, scl, sclbnd, flux, bnd_nid, unk, T0, dt=0.0, num_ts=0):
    """States [T^0, ..., T^num_ts] of the linear map (steady: [T^0, T]) solved by sparse LU;
    T^0 is T0 with the Dirichlet values applied (transient_base.rs:16-160, one solve per step)."""
    from scipy.sparse.linalg import spsolve
    T = apply_initial_dirichlet(g, ops, sclbnd, bnd_nid, T0)
    states = [T]
    for _ in range(max(num_ts, 1)):
        A, b = _csr_system(g, ops, scl, sclbnd, flux, bnd_nid, unk, T, dt if num_ts > 0 else 0.0)
        T = spsolve(A.tocsc(), b)
        states.append(T)
    return states


def time_diag(g, ops, scl, bnd_nid, dt):
    """B = d b / d T^(n-1): diag of -w V / dt over the TIME operators (opscl_dom_time.rs:46-58)."""
    bd = np.zeros(g["N"])
    for op in ops:
        if op.kind in (Op.SCL_DOM_TIME, Op.SCL_BND_TIME):
            rows = _op_rows(g, op, bnd_nid)
            W = scl[op.a][0]
            np.add.at(bd, rows, -W[rows] * g["vol"][rows] / dt)
    return bd


def adjoint_accumulate(g, ops, scl, sclbnd, flux, bnd_nid, unk, states, dt, num_ts, seed):
    """Transposed solves A^T l^N = seed, A^T l^n = B l^(n+1) and the accumulators
    a_e = -sum_n l_v (T_n - T_v) (e = (v, n)), ls = sum_n l, tv = sum_n l (T^n - T^(n-1)).
    Returns dict(lams={n: l^n}, aE, ls, tv, bd)."""
    from scipy.sparse.linalg import splu
    A, _ = _csr_system(g, ops, scl, sclbnd, flux, bnd_nid, unk, states[-1], dt if num_ts > 0 else 0.0)
    lu = splu(A.T.tocsc())
    bd = time_diag(g, ops, scl, bnd_nid, dt) if num_ts > 0 else np.zeros(g["N"])
    rows, ngh = g["rows"].astype(np.int64), g["ngh"].astype(np.int64)
    aE = np.zeros(rows.size)
    ls = np.zeros(g["N"])
    tv = np.zeros(g["N"])
    last = len(states) - 1
    lams = {}
    lam = lu.solve(np.asarray(seed, np.float64))
    for k in range(last, 0, -1):
        if k != last:
            lam = lu.solve(bd * lams[k + 1])
        lams[k] = lam
        Tn = states[k]
        aE -= lam[rows] * (Tn[ngh] - Tn[rows])
        ls += lam
        if num_ts > 0:
            tv += lam * (Tn - states[k - 1])
    return dict(lams=lams, aE=aE, ls=ls, tv=tv, bd=bd)


def grad_matrices(g):
    """update_grad (scl_dom.rs:146-164) as sparse matrices: (grad G)_c = W_c G."""
    import scipy.sparse as sp
    N = g["N"]
    rows, ngh = g["rows"].astype(np.int64), g["ngh"].astype(np.int64)
    out = []
    for c in range(3):
        w = g["gw"][c]
        W = sp.csr_matrix((w, (rows, ngh)), shape=(N, N))
        dsum = np.zeros(N)
        np.add.at(dsum, rows, w)
        out.append((W - sp.diags(dsum)).tocsr())
    return out


def adjoint_field_gradient(g, ops, scl, flux, bnd_nid, acc, dt, fid, constant):
    """dJ/dG for the per-vertex values of scalar field fid, summed over its uses as a
    diffusion coefficient (DIFFUSION, ADVECTION_UNITY), a SOURCE and a TIME weight.
    Scatter form: every directed edge adds its partials to both of its vertices."""
    N = g["N"]
    rows, ngh = g["rows"].astype(np.int64), g["ngh"].astype(np.int64)
    G, Gg = scl[fid]
    grad = np.zeros(N)
    adj_gG = np.zeros((3, N))
    ad_all = g["area"] / g["dist"]
    # d(matrix entry)/d(face value), summed over the operators of each edge's row
    dface = np.zeros(rows.size)
    for op in ops:
        if op.a != fid:
            continue
        if op.kind in (Op.SCL_DOM_DIFFUSION, Op.SCL_BND_DIFFUSION, Op.SCL_DOM_ADVECTION_UNITY, Op.SCL_BND_ADVECTION_UNITY):
            m = np.zeros(N, bool)
            m[_op_rows(g, op, bnd_nid)] = True
            e = np.nonzero(m[rows])[0]
            if op.kind in (Op.SCL_DOM_DIFFUSION, Op.SCL_BND_DIFFUSION):
                dface[e] += ad_all[e]
            else:
                df = _diff_face(G, Gg, rows[e], ngh[e], g["skew"][:, e])
                cond = df * ad_all[e]
                F = flux[op.b][e]
                with np.errstate(divide="ignore", invalid="ignore"):
                    pe = np.where(cond != 0.0, F / cond, 0.0)
                x = np.maximum(1.0 - 0.1 * np.abs(pe), 0.0)
                # d/dcond [cond (x^5 - 1)] with d|Pe|/dcond = -|Pe|/cond
                dface[e] += ad_all[e] * (x ** 5 - 1.0 + 0.5 * x ** 4 * np.abs(pe))
        elif op.kind in (Op.SCL_DOM_SOURCE, Op.SCL_BND_SOURCE):
            r = _op_rows(g, op, bnd_nid)
            w = -acc["ls"][r] * g["vol"][r]
            np.add.at(grad, r, w)
            np.add.at(adj_gG.T, r, (w * g["cskew"][:, r]).T)
        elif op.kind in (Op.SCL_DOM_TIME, Op.SCL_BND_TIME):
            r = _op_rows(g, op, bnd_nid)
            np.add.at(grad, r, acc["tv"][r] * g["vol"][r] / dt)
    s = acc["aE"] * dface  # adjoint of each edge's face value
    Gv, Gn = G[rows], G[ngh]
    ssum = Gv + Gn
    with np.errstate(divide="ignore", invalid="ignore"):
        dm_dv = np.where(ssum != 0.0, 2.0 * Gn * Gn / (ssum * ssum), 0.0)
        dm_dn = np.where(ssum != 0.0, 2.0 * Gv * Gv / (ssum * ssum), 0.0)
        diff_m = np.where(ssum != 0.0, 2.0 * Gv * Gn / ssum, 0.0)
    sk = g["skew"]
    diff_s = diff_m + 0.5 * (sk[0] * (Gg[0][rows] + Gg[0][ngh]) + sk[1] * (Gg[1][rows] + Gg[1][ngh]) + sk[2] * (Gg[2][rows] + Gg[2][ngh]))
    chi = (diff_s > 0.0).astype(np.float64)
    np.add.at(grad, rows, s * dm_dv)
    np.add.at(grad, ngh, s * dm_dn)
    for c in range(3):
        np.add.at(adj_gG[c], rows, 0.5 * s * chi * sk[c])
        np.add.at(adj_gG[c], ngh, 0.5 * s * chi * sk[c])
    if not constant:
        for c, W in enumerate(grad_matrices(g)):
            grad += W.T @ adj_gG[c]
    return grad


def adjoint_dirichlet_gradient(g, ops, bnd_nid, acc, sid):
    """dJ/d(values of boundary field sid) through its Dirichlet rows (b_v = value)."""
    out = None
    for op in ops:
        if op.kind == Op.SCL_BND_DIRICHLET and op.a == sid:
            r = g["num_pts"] + np.asarray(bnd_nid[op.geom_id], np.int64)
            out = acc["ls"][r] if out is None else out + acc["ls"][r]
    return out


def adjoint_initial_gradient(g, ops, bnd_nid, acc, num_ts):
    """dJ/dT^0 = B l^1, zero where the Dirichlet values overwrite T^0."""
    if num_ts == 0:
        return np.zeros(g["N"])
    out = acc["bd"] * acc["lams"][1]
    for op in ops:
        if op.kind == Op.SCL_BND_DIRICHLET:
            out[g["num_pts"] + np.asarray(bnd_nid[op.geom_id], np.int64)] = 0.0
    return out


# --------------------------------------------------------------------------- ILU(0)
def ilu0(A, perm):
    """ILU(0) of P A P^T (row t of P A P^T = row perm[t] of A) by the IKJ variant
    (Saad, Iterative Methods for Sparse Linear Systems, Alg. 10.4): returns (L, U) with a
    unit lower L and LU = P A P^T on the pattern of P A P^T.  Dense row updates; small n."""
    import scipy.sparse as sp
    n = A.shape[0]
    P = sp.csr_matrix((np.ones(n), (np.arange(n), perm)), shape=(n, n))
    Ap = (P @ A @ P.T).tocsr()
    Ap.sort_indices()
    W = Ap.toarray()
    S = Ap.copy()
    S.data[:] = 1.0
    pat = S.toarray() != 0.0  # the stored pattern, explicit zeros included
    for i in range(Ap.shape[0]):
        for k in Ap.indices[Ap.indptr[i]:Ap.indptr[i + 1]]:
            if k >= i:
                break
            W[i, k] /= W[k, k]
            upd = pat[i, k + 1:]
            W[i, k + 1:][upd] -= W[i, k] * W[k, k + 1:][upd]
    L = sp.csr_matrix(np.tril(W, -1) * pat + np.eye(n))
    U = sp.csr_matrix(np.triu(W) * pat)
    return L, U


def ilu0_rows(A, perm):
    """ilu0() with sparse rows, for n where a dense copy does not fit: the same IKJ updates in
    the same order (pivot columns ascending, each pivot row's final U entries), so the factors
    equal ilu0()'s bit for bit (pinned by that in the tests)."""
    import scipy.sparse as sp
    n = A.shape[0]
    P = sp.csr_matrix((np.ones(n), (np.arange(n), perm)), shape=(n, n))
    Ap = (P @ A @ P.T).tocsr()
    Ap.sort_indices()
    ip, ix, dv = Ap.indptr, Ap.indices.tolist(), Ap.data.tolist()
    urow = [None] * n  # row k's U entries right of the diagonal: list of (j, u_kj)
    diag = [0.0] * n
    lr, lc, lv, ur, uc, uv = [], [], [], [], [], []
    for i in range(n):
        cols = ix[ip[i]:ip[i + 1]]
        w = dict(zip(cols, dv[ip[i]:ip[i + 1]]))
        for k in cols:
            if k >= i:
                break
            w[k] /= diag[k]
            wk = w[k]
            for j, ukj in urow[k]:
                if j in w:
                    w[j] -= wk * ukj
        diag[i] = w[i]
        urow[i] = [(j, w[j]) for j in cols if j > i]
        for j in cols:
            if j < i:
                lr.append(i); lc.append(j); lv.append(w[j])
            else:
                ur.append(i); uc.append(j); uv.append(w[j])
    L = sp.csr_matrix((lv + [1.0] * n, (lr + list(range(n)), lc + list(range(n)))), shape=(n, n))
    U = sp.csr_matrix((uv, (ur, uc)), shape=(n, n))
    return L, U


# --------------------------------------------------------------------------- flow (SteadyFlow)
# Pinning of this section (tests/test_cpu_flow_oracle.py), independent of the restatement:
#   grad_coeff()     sum_m c_m p_m = V grad(p) for linear p at every vertex (divergence
#                    theorem, closed form); C++ builder bitwise equal.
#   update_flow()    flux antisymmetry F(v->n) = -F(n->v); hydrostatic balance (linear p,
#                    body force = grad p) gives F = rho_f (u_f . n) A exactly
#                    (opscl_dom_pstab.rs:31); rc > 0 and rc = V / a_P with a_P >= sum cond.
#   assemble_flow()  momentum rows: velocity row sums 0, gradient block on linear p = -V grad p;
#                    continuity rows: divergence block on linear u = -rho V div(u) (points and
#                    nodes), pressure-stabiliser block annihilates linear p; transpose and
#                    dilatatory b vanish for linear u with constant viscosity; pattern
#                    2-ring / 3-ring; the reference-pressure row is the identity.

def grad_coeff(g: dict):
    """Pressure-gradient coefficients, Tetrahedron step 6 (src/base/geom_tet.rs:295-347).

    For vertex v the divergence-theorem gradient with skew-corrected face values,
    pushed per face (arithmetic mean, then the (p_m - p_s) skew terms of s = v, n) and per
    boundary sub-face (linear interpolation), stably sorted by vertex and merged by
    sequential summation.  Returns gc_ptr[N+1], gc_vid[G], gc_c (3, G).
    """
    N = g["N"]
    rp, ngh = g["row_ptr"], g["ngh"]
    area, norm, skew, gw = g["area"], g["norm"], g["skew"], g["gw"]
    bp = g["bsf_ptr"]
    ptr = [0]
    vid_out, c_out = [], []
    for v in range(N):
        t = []
        for e in range(rp[v], rp[v + 1]):
            n = int(ngh[e])
            A = float(area[e])
            nx, ny, nz = float(norm[0, e]), float(norm[1, e]), float(norm[2, e])
            sx, sy, sz = float(skew[0, e]), float(skew[1, e]), float(skew[2, e])
            t.append((v, 0.5 * A * nx, 0.5 * A * ny, 0.5 * A * nz))
            t.append((n, 0.5 * A * nx, 0.5 * A * ny, 0.5 * A * nz))
            for s in (v, n):
                for q in range(rp[s], rp[s + 1]):
                    wx, wy, wz = float(gw[0, q]), float(gw[1, q]), float(gw[2, q])
                    ws = 0.5 * (wx * sx + wy * sy + wz * sz) * A
                    t.append((int(ngh[q]), ws * nx, ws * ny, ws * nz))
                    t.append((s, -ws * nx, -ws * ny, -ws * nz))
        for s in range(bp[v], bp[v + 1]):
            A = float(g["bsf_area"][s])
            nx, ny, nz = (float(x) for x in g["bsf_norm"][:, s])
            wv, wj, wk = (float(x) for x in g["bsf_wgt"][:, s])
            t.append((v, wv * A * nx, wv * A * ny, wv * A * nz))
            t.append((int(g["bsf_vj"][s]), wj * A * nx, wj * A * ny, wj * A * nz))
            t.append((int(g["bsf_vk"][s]), wk * A * nx, wk * A * ny, wk * A * nz))
        t.sort(key=lambda x: x[0])  # stable, as sort_by_key
        merged = []
        for m, cx, cy, cz in t:
            if merged and merged[-1][0] == m:
                pm = merged[-1]
                merged[-1] = (m, pm[1] + cx, pm[2] + cy, pm[3] + cz)
            else:
                merged.append((m, cx, cy, cz))
        for m, cx, cy, cz in merged:
            vid_out.append(m)
            c_out.append((cx, cy, cz))
        ptr.append(len(vid_out))
    return (np.asarray(ptr, np.int64), np.asarray(vid_out, np.int64),
            np.asarray(c_out, np.float64).reshape(-1, 3).T.copy())


def _face_mean(f, fg, v, n, sx, sy, sz):
    # skew-corrected arithmetic mean (opscl_dom_div.rs:111-112, flow_helper.rs:47)
    m = 0.5 * (f[v] + f[n])
    return m + 0.5 * (sx * (fg[0][v] + fg[0][n]) + sy * (fg[1][v] + fg[1][n]) + sz * (fg[2][v] + fg[2][n]))


def _visc_face(f, fg, v, n, sx, sy, sz):
    # harmonic mean + skew correction + guard (opvec_dom_adv.rs:93-95, flow_helper.rs:48-50)
    fv, fn = f[v], f[n]
    s = fv + fn
    with np.errstate(divide="ignore", invalid="ignore"):
        m = np.where(s != 0.0, 2.0 * fv * fn / s, 0.0)
    ff = m + 0.5 * (sx * (fg[0][v] + fg[0][n]) + sy * (fg[1][v] + fg[1][n]) + sz * (fg[2][v] + fg[2][n]))
    return np.where(ff > 0.0, ff, m)


def _fce_centroid(g, fce, fceg, vids):
    # body force at the control-volume centroid (opvec_dom_src.rs:58-61)
    cx, cy, cz = g["cskew"][:, vids]
    return [fce[c][vids] + fceg[3 * c][vids] * cx + fceg[3 * c + 1][vids] * cy + fceg[3 * c + 2][vids] * cz for c in range(3)]


def update_flow(g: dict, gc, vel, velg, pres, den, deng, visc, viscg, fce, fceg):
    """flow_helper::update_flow + update_flux (src/operator/core/flow_helper.rs:3-179).

    rc_v = V_v / a_P, a_P = sum_faces cond A(|Pe|) + max(F_c, 0) accumulated in face order;
    mom_v = rc_v (grad p_v - fce_c,v) with grad p_v = sum_gc c p / V (sequential); face
    mass flux F = rho_f (u_f . n + vel_rc) A with vel_rc the Rhie-Chow correction.
    vel (3,N), velg (9,N), den/visc (N), grads (3,N), fce (3,N), fceg (9,N).
    Returns rc (N), mom (3,N), flux (E).
    """
    N = g["N"]
    rows, ngh = g["rows"].astype(np.int64), g["ngh"].astype(np.int64)
    sx, sy, sz = g["skew"]
    nx, ny, nz = g["norm"]
    A, d = g["area"], g["dist"]
    den_f = 0.5 * (den[rows] + den[ngh]) + 0.5 * (sx * (deng[0][rows] + deng[0][ngh]) + sy * (deng[1][rows] + deng[1][ngh]) + sz * (deng[2][rows] + deng[2][ngh]))
    visc_f = _visc_face(visc, viscg, rows, ngh, sx, sy, sz)
    cond = visc_f * A / d
    Fc = den_f * 0.5 * ((vel[0][rows] + vel[0][ngh]) * nx + (vel[1][rows] + vel[1][ngh]) * ny + (vel[2][rows] + vel[2][ngh]) * nz) * A
    with np.errstate(divide="ignore", invalid="ignore"):
        pe = np.where(cond != 0.0, Fc / cond, 0.0)
    term = cond * _pow5(np.maximum(1.0 - 0.1 * np.abs(pe), 0.0)) + np.maximum(Fc, 0.0)
    coeff_p = np.zeros(N)
    np.add.at(coeff_p, rows, term)
    with np.errstate(divide="ignore", invalid="ignore"):
        rc = np.where(coeff_p > 0.0, g["vol"] / coeff_p, 0.0)
    gc_ptr, gc_vid, gc_c = gc
    grow = np.repeat(np.arange(N), np.diff(gc_ptr))
    vol = g["vol"]
    pg = []
    for c in range(3):
        acc = np.zeros(N)
        np.add.at(acc, grow, gc_c[c] * pres[gc_vid] / vol[grow])
        pg.append(acc)
    allv = np.arange(N)
    fcc = _fce_centroid(g, fce, fceg, allv)
    mom = np.stack([rc * (pg[c] - fcc[c]) for c in range(3)])

    def uface(c):
        gx, gy, gz = velg[3 * c], velg[3 * c + 1], velg[3 * c + 2]
        return 0.5 * (vel[c][rows] + vel[c][ngh]) + 0.5 * (sx * (gx[rows] + gx[ngh]) + sy * (gy[rows] + gy[ngh]) + sz * (gz[rows] + gz[ngh]))

    u_f, v_f, w_f = uface(0), uface(1), uface(2)
    fce_e = 0.5 * (nx * (fce[0][rows] + fce[0][ngh]) + ny * (fce[1][rows] + fce[1][ngh]) + nz * (fce[2][rows] + fce[2][ngh]))
    mom_e = 0.5 * (nx * (mom[0][rows] + mom[0][ngh]) + ny * (mom[1][rows] + mom[1][ngh]) + nz * (mom[2][rows] + mom[2][ngh]))
    vel_rc = -0.5 * (rc[rows] + rc[ngh]) * ((pres[ngh] - pres[rows]) / d - fce_e) + mom_e
    flux = den_f * (u_f * nx + v_f * ny + w_f * nz + vel_rc) * A
    return rc, mom, flux


def assemble_flow(g: dict, gc, fields: dict, bnd_vel: list, bnd_nid: list, pref=None, dt=0.0, vel_prev=None, den_prev=None):
    """Triplets and b of SteadyFlow::assemble_matrix (src/steady/core/steady_flow.rs:244-303)
    in push order, unknowns [u, v, w, p] (steady_flow.rs:94-104).

    fields: vel (3,N), velg (9,N), den, deng, visc, viscg, fce (3,N), fceg (9,N), rc, flux
    (from update_flow).  bnd_vel: list of (bnd_id, values (3, count)) in add_vel_bnd order,
    each boundary getting OpVecBndDirichlet, OpSclBndDivergence, OpSclBndPressureStabilizer.
    pref: (nid, p_ref) or None (OpSclDomPressureReference, applied last).
    dt > 0: TransientFlow (transient_flow.rs:248-303): OpVecDomTime first (opvec_dom_time.rs:37-68,
    previous velocity vel_prev) and the divergence operators' apply_time (den_prev).
    Operators: opvec_dom_adv.rs:47-116, opvec_dom_grad.rs:39-70, opvec_dom_diff.rs:39-94,
    opvec_dom_dtrn.rs:39-113, opvec_dom_ddil.rs:41-112, opvec_dom_src.rs:37-73,
    opscl_dom_div.rs:68-153, opscl_dom_pstab.rs:52-151, opvec_bnd_dir.rs:56-80,
    opscl_bnd_div.rs:73-187, opscl_bnd_pstab.rs:49-153, opscl_dom_pref.rs:48-63.
    Returns rows, cols, vals (push order), b (4N).
    """
    N, NP = g["N"], g["num_pts"]
    rp, ngh = g["row_ptr"], g["ngh"]
    rowsE = g["rows"].astype(np.int64)
    area, dist, norm, skew, gw = g["area"], g["dist"], g["norm"], g["skew"], g["gw"]
    vel, velg = fields["vel"], fields["velg"]
    den, deng = fields["den"], fields["deng"]
    visc, viscg = fields["visc"], fields["viscg"]
    fce, fceg = fields["fce"], fields["fceg"]
    rc, flux = fields["rc"], fields["flux"]
    gc_ptr, gc_vid, gc_c = gc
    b = np.zeros(4 * N)
    R, C, V = [], [], []
    P = 3 * N  # pressure block start
    pts = np.arange(NP)
    e_pts = np.nonzero(rowsE < NP)[0]
    ev, en = rowsE[e_pts], ngh[e_pts].astype(np.int64)
    sx, sy, sz = skew[:, e_pts]
    nx, ny, nz = norm[:, e_pts]
    A, d = area[e_pts], dist[e_pts]
    visc_f = _visc_face(visc, viscg, ev, en, sx, sy, sz)
    transient = dt > 0.0
    if transient:
        # OpVecDomTime::apply_time: diag -rho V / dt, b -rho u_prev V / dt per component
        vol_p = g["vol"][pts]
        for c in range(3):
            R.append(c * N + pts); C.append(c * N + pts); V.append(-den[pts] * vol_p / dt)
        for c in range(3):
            b[c * N + pts] += -den[pts] * vel_prev[c][pts] * vol_p / dt

    def push_vec_pairs(coeff):
        # per face, per component: (c v, c n) +coeff, (c v, c v) -coeff
        rr = np.stack([c * N + ev for c in range(3) for _ in range(2)], axis=1).ravel()
        cc = np.stack([x for c in range(3) for x in (c * N + en, c * N + ev)], axis=1).ravel()
        vv = np.stack([x for c in range(3) for x in (coeff, -coeff)], axis=1).ravel()
        R.append(rr); C.append(cc); V.append(vv)

    # OpVecDomAdvection
    # (opvec_dom_adv.rs:47-116: the scalar power-law coefficient with the viscosity as the
    # diffusivity, on each component block)
    cond = visc_f * A / d
    push_vec_pairs(_adv_coeff(cond, flux[e_pts]))
    # OpVecDomGradient
    grow = np.repeat(np.arange(N), np.diff(gc_ptr))
    gsel = np.nonzero(grow < NP)[0]
    gv, gm = grow[gsel], gc_vid[gsel]
    R.append(np.stack([c * N + gv for c in range(3)], axis=1).ravel())
    C.append(np.stack([P + gm] * 3, axis=1).ravel())
    V.append(np.stack([-gc_c[c][gsel] for c in range(3)], axis=1).ravel())
    # OpVecDomDiffusion
    push_vec_pairs(visc_f * A / d)
    # OpVecDomTranspose (b): per vertex sums in face order
    fg = [[0.5 * (velg[3 * j + c][ev] + velg[3 * j + c][en]) for c in range(3)] for j in range(3)]  # fg[j][c] = d u_j / d x_c
    for c in range(3):
        t = visc_f * (fg[0][c] * nx + fg[1][c] * ny + fg[2][c] * nz) * A
        s = np.zeros(N)
        np.add.at(s, ev, t)
        b[c * N + pts] += -s[pts]
    # OpVecDomDilatatory (b)
    divv = velg[0] + velg[4] + velg[8]
    lam_f = (-2.0 / 3.0) * visc_f
    div_f = 0.5 * (divv[ev] + divv[en])
    co = lam_f * div_f * A
    for c, n_c in enumerate((nx, ny, nz)):
        s = np.zeros(N)
        np.add.at(s, ev, co * n_c)
        b[c * N + pts] += -s[pts]
    # OpVecDomSource (b)
    fcc = _fce_centroid(g, fce, fceg, pts)
    for c in range(3):
        b[c * N + pts] += -fcc[c] * g["vol"][pts]

    def continuity(v, R_, C_, V_, bsub):
        """Divergence + pressure-stabiliser pushes of vertex v's row (pure Python, push
        order).  bsub: also push the boundary sub-face terms of the divergence operator."""
        row = P + v
        den_v = den[v]
        rc_v, vol_v = rc[v], g["vol"][v]
        out_r, out_c, out_v = [], [], []
        pst_r, pst_c, pst_v = [], [], []
        bacc = []
        fcv = [float(x[0]) for x in _fce_centroid(g, fce, fceg, np.array([v]))]
        for e in range(rp[v], rp[v + 1]):
            n = int(ngh[e])
            face, dd = float(area[e]), float(dist[e])
            enx, eny, enz = (float(x) for x in norm[:, e])
            esx, esy, esz = (float(x) for x in skew[:, e])
            den_f = float(_face_mean(den, deng, v, n, esx, esy, esz))
            coeff = 0.5 * den_f * face
            for c, n_c in enumerate((enx, eny, enz)):
                out_r += [row, row]; out_c += [c * N + n, c * N + v]; out_v += [-coeff * n_c, -coeff * n_c]
            for q in range(rp[v], rp[v + 1]):
                m = int(ngh[q])
                ws = float(gw[0, q]) * esx + float(gw[1, q]) * esy + float(gw[2, q]) * esz
                for c, n_c in enumerate((enx, eny, enz)):
                    val = -coeff * ws * n_c
                    out_r += [row, row]; out_c += [c * N + m, c * N + v]; out_v += [val, -val]
            for q in range(rp[n], rp[n + 1]):
                m = int(ngh[q])
                ws = float(gw[0, q]) * esx + float(gw[1, q]) * esy + float(gw[2, q]) * esz
                for c, n_c in enumerate((enx, eny, enz)):
                    val = -coeff * ws * n_c
                    out_r += [row, row]; out_c += [c * N + m, c * N + n]; out_v += [val, -val]
        if bsub:
            for s in range(g["bsf_ptr"][v], g["bsf_ptr"][v + 1]):
                a_ = float(g["bsf_area"][s])
                snx, sny, snz = (float(x) for x in g["bsf_norm"][:, s])
                wv, wj, wk = (float(x) for x in g["bsf_wgt"][:, s])
                j, k = int(g["bsf_vj"][s]), int(g["bsf_vk"][s])
                for m, cf in ((v, wv * den_v * a_), (j, wj * den[j] * a_), (k, wk * den[k] * a_)):
                    for c, n_c in enumerate((snx, sny, snz)):
                        out_r.append(row); out_c.append(c * N + m); out_v.append(-cf * n_c)
        for e in range(rp[v], rp[v + 1]):
            n = int(ngh[e])
            face, dd = float(area[e]), float(dist[e])
            enx, eny, enz = (float(x) for x in norm[:, e])
            esx, esy, esz = (float(x) for x in skew[:, e])
            den_f = float(_face_mean(den, deng, v, n, esx, esy, esz))
            rc_n, vol_n = rc[n], g["vol"][n]
            coeff = den_f * 0.5 * (rc_v + rc_n) * face
            pst_r += [row, row]; pst_c += [P + n, P + v]; pst_v += [coeff / dd, -coeff / dd]
            coeff_v = den_f * 0.5 * rc_v / vol_v * face
            for q in range(gc_ptr[v], gc_ptr[v + 1]):
                pst_r.append(row); pst_c.append(P + int(gc_vid[q]))
                pst_v.append(-coeff_v * (gc_c[0, q] * enx + gc_c[1, q] * eny + gc_c[2, q] * enz))
            coeff_n = den_f * 0.5 * rc_n / vol_n * face
            for q in range(gc_ptr[n], gc_ptr[n + 1]):
                pst_r.append(row); pst_c.append(P + int(gc_vid[q]))
                pst_v.append(-coeff_n * (gc_c[0, q] * enx + gc_c[1, q] * eny + gc_c[2, q] * enz))
            fcn = [float(x[0]) for x in _fce_centroid(g, fce, fceg, np.array([n]))]
            frc = [0.5 * (rc_v + rc_n) * 0.5 * (fce[c][v] + fce[c][n]) - 0.5 * (rc_v * fcv[c] + rc_n * fcn[c]) for c in range(3)]
            bacc.append(den_f * (frc[0] * enx + frc[1] * eny + frc[2] * enz) * face)
        return (out_r, out_c, out_v), (pst_r, pst_c, pst_v), bacc

    # OpSclDomDivergence, then OpSclDomPressureStabilizer, over the points
    div_t, pst_t, bq = [], [], []
    for v in range(NP):
        dt_, pt_, ba = continuity(v, R, C, V, False)
        div_t.append(dt_); pst_t.append(pt_); bq.append((v, ba))
    for r_, c_, v_ in div_t:
        R.append(np.asarray(r_, np.int64)); C.append(np.asarray(c_, np.int64)); V.append(np.asarray(v_, np.float64))
    if transient:
        # OpSclDomDivergence::apply_time (opscl_dom_div.rs:40-66)
        b[P + pts] += (den[pts] - den_prev[pts]) * g["vol"][pts] / dt
    for (r_, c_, v_), (v, ba) in zip(pst_t, bq):
        R.append(np.asarray(r_, np.int64)); C.append(np.asarray(c_, np.int64)); V.append(np.asarray(v_, np.float64))
        for t in ba:
            b[P + v] += t
    # velocity boundaries: OpVecBndDirichlet, OpSclBndDivergence, OpSclBndPressureStabilizer
    for bid, val in bnd_vel:
        nodes = NP + np.asarray(bnd_nid[bid], np.int64)
        val = np.asarray(val, np.float64).reshape(3, -1)
        for i, v in enumerate(nodes):
            for c in range(3):
                R.append(np.array([c * N + v])); C.append(np.array([c * N + v])); V.append(np.array([1.0]))
            for c in range(3):
                b[c * N + v] += val[c, i]
        div_t, pst_t, bq = [], [], []
        for v in nodes:
            dt_, pt_, ba = continuity(int(v), R, C, V, True)
            div_t.append(dt_); pst_t.append(pt_); bq.append((int(v), ba))
        for r_, c_, v_ in div_t:
            R.append(np.asarray(r_, np.int64)); C.append(np.asarray(c_, np.int64)); V.append(np.asarray(v_, np.float64))
        if transient:
            # OpSclBndDivergence::apply_time (opscl_bnd_div.rs:40-70)
            b[P + nodes] += (den[nodes] - den_prev[nodes]) * g["vol"][nodes] / dt
        for (r_, c_, v_), (v, ba) in zip(pst_t, bq):
            R.append(np.asarray(r_, np.int64)); C.append(np.asarray(c_, np.int64)); V.append(np.asarray(v_, np.float64))
            for t in ba:
                b[P + v] += t
    rows = np.concatenate(R)
    cols = np.concatenate(C)
    vals = np.concatenate(V)
    if pref is not None:
        xid = P + NP + int(pref[0])
        keep = rows != xid
        rows, cols, vals = np.append(rows[keep], xid), np.append(cols[keep], xid), np.append(vals[keep], 1.0)
        b[xid] = float(pref[1])
    return rows, cols, vals, b


def flow_fields(g, gc, vel, pres, den, visc, fce, vel_const=False, den_const=True, visc_const=True, fce_const=True):
    """Gradients (update_grad of every field, zero for constants) and the flow data of
    update_flow, as assemble_flow reads them."""
    N = g["N"]
    velg = np.zeros((9, N)) if vel_const else update_grad_vec(g, vel)
    deng = np.zeros((3, N)) if den_const else update_grad(g, den)
    viscg = np.zeros((3, N)) if visc_const else update_grad(g, visc)
    fceg = np.zeros((9, N)) if fce_const else update_grad_vec(g, fce)
    rc, mom, flux = update_flow(g, gc, vel, velg, pres, den, deng, visc, viscg, fce, fceg)
    return dict(vel=vel, velg=velg, pres=pres, den=den, deng=deng, visc=visc, viscg=viscg, fce=fce, fceg=fceg, rc=rc,
                mom=mom, flux=flux)


def unscaled_residual(x_old, x_new):
    """The Picard residual with Variables::scale_off (src/base/scaling.rs:116-122):
    |x_new - x_old|_2 / (|x_old|_2 + 1e-10) over the whole vector, NaN if x_new is not finite."""
    if not np.all(np.isfinite(x_new)):
        return float("nan")
    return float(np.sqrt(((x_new - x_old) ** 2).sum()) / (np.sqrt((x_old * x_old).sum()) + 1e-10))


def scaled_residual_flow(N, x_old, x_new, den, scale_vel=0.0, scale_pres=0.0, scale_off=False):
    """Picard residual of the flow unknowns [vel (vector field), pres (pressure field)]
    (src/base/scaling.rs:88-135): max over the two fields of rms(x_new - x_old) / scale, the
    rms over N entries (per component for the vector), pressure changes below the round-off
    dead band 1e-11 max|p| not counted, scales max(max|vel|, scale_vel) and max(range p,
    rho_max U_max^2, scale_pres) (user reference scales, vars.rs:203-213, are lower bounds);
    with scale_off the unscaled residual of the whole vector."""
    if scale_off:
        return unscaled_residual(x_old, x_new)
    if not np.all(np.isfinite(x_new)):
        return float("nan")
    vo, vn = x_old[:3 * N].reshape(3, N), x_new[:3 * N].reshape(3, N)
    po, pn = x_old[3 * N:], x_new[3 * N:]
    vmax = max(np.sqrt((vo * vo).sum(axis=0)).max(), np.sqrt((vn * vn).sum(axis=0)).max())
    floor = 1e-30
    r_v = np.sqrt(((np.abs(vn - vo)) ** 2).sum() / max(N, 1)) / max(vmax, scale_vel, floor)
    dead = 1e-11 * max(np.abs(po).max(), np.abs(pn).max())
    r_p = np.sqrt((np.maximum(np.abs(pn - po) - dead, 0.0) ** 2).sum() / max(N, 1))
    rng = max(po.max(), pn.max()) - min(po.min(), pn.min())
    rho = np.abs(den).max()
    r_p /= max(rng, rho * vmax * vmax, scale_pres, floor)
    return float(max(r_v, r_p))


def transient_flow(g, gc, vel0, pres0, den, visc, fce, bnd_vel, bnd_nid, pref, dt, num_ts, max_iter, tol, damp):
    """TransientBase::solve (src/transient/transient_base.rs:16-160) for TransientFlow with
    constant density, viscosity and body force: implicit Euler, damped Picard iterations per
    step with SuperLU solves, previous state updated after each step.  Returns vel, pres and
    the Picard iterations of every step."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    N, NP = g["N"], g["num_pts"]
    vel = np.array(vel0, np.float64).reshape(3, N).copy()
    pres = np.array(pres0, np.float64).copy()
    for bid, val in bnd_vel:
        nodes = NP + np.asarray(bnd_nid[bid], np.int64)
        vel[:, nodes] = np.asarray(val).reshape(3, -1)
    vel_prev = vel.copy()

    def system(vel, pres):
        f = flow_fields(g, gc, vel, pres, den, visc, fce)
        r, c, v, b = assemble_flow(g, gc, f, bnd_vel, bnd_nid, pref, dt, vel_prev, den)
        rp, col, val = coo_to_csr(r, c, v, 4 * N)
        return sp.csr_matrix((val, col, rp), shape=(4 * N, 4 * N)).tocsc(), b

    iters = []
    x_iter = np.concatenate([vel.ravel(), pres])
    for ts in range(num_ts):
        A, b = system(vel, pres)
        for it in range(max_iter):
            x_u = spla.spsolve(A, b)
            x_d = (1.0 - damp) * x_iter + damp * x_u
            vel, pres = x_d[:3 * N].reshape(3, N).copy(), x_d[3 * N:].copy()
            A, b = system(vel, pres)
            res = scaled_residual_flow(N, x_iter, x_d, den)
            x_iter = x_d
            if res < tol:
                iters.append(it + 1)
                break
        else:
            raise RuntimeError("TransientDidNotConverge")
        vel_prev = vel.copy()
    return vel, pres, iters


def picard_flow(g, gc, vel0, pres0, den, visc, fce, bnd_vel, bnd_nid, pref, max_iter, tol, damp, scales=None):
    """SteadyBase::solve (src/steady/steady_base.rs:16-135) for SteadyFlow with constant
    density, viscosity and body force: Dirichlet velocities into the initial state
    (opvec_bnd_dir.rs:34-53), then Picard iterations solving each linear system with
    SuperLU, damping, re-assembly and the scaled residual (scales: keyword arguments of
    scaled_residual_flow).  Returns vel, pres, iterations, residuals."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    N, NP = g["N"], g["num_pts"]
    vel = np.array(vel0, np.float64).reshape(3, N).copy()
    pres = np.array(pres0, np.float64).copy()
    for bid, val in bnd_vel:
        nodes = NP + np.asarray(bnd_nid[bid], np.int64)
        vel[:, nodes] = np.asarray(val).reshape(3, -1)

    def system(vel, pres):
        f = flow_fields(g, gc, vel, pres, den, visc, fce)
        r, c, v, b = assemble_flow(g, gc, f, bnd_vel, bnd_nid, pref)
        rp, col, val = coo_to_csr(r, c, v, 4 * N)
        return sp.csr_matrix((val, col, rp), shape=(4 * N, 4 * N)).tocsc(), b

    x_iter = np.concatenate([vel.ravel(), pres])
    A, b = system(vel, pres)
    res_hist = []
    for it in range(max_iter):
        x_u = spla.spsolve(A, b)
        x_d = (1.0 - damp) * x_iter + damp * x_u
        vel, pres = x_d[:3 * N].reshape(3, N).copy(), x_d[3 * N:].copy()
        A, b = system(vel, pres)
        res = scaled_residual_flow(N, x_iter, x_d, den, **(scales or {}))
        res_hist.append(res)
        if not np.isfinite(res):
            raise FloatingPointError("SteadyNonFiniteResidual")
        x_iter = x_d
        if res < tol:
            return vel, pres, it + 1, res_hist
    raise RuntimeError("SteadyDidNotConverge")


def scaled_residual_scalar(x_old, x_new, scale=0.0, scale_off=False):
    """Picard residual of one scalar field (src/base/scaling.rs:88-135): rms(x_new - x_old)
    over its entries divided by max(|x| over both states, the user scale, 1e-30); with
    scale_off the unscaled residual."""
    if scale_off:
        return unscaled_residual(x_old, x_new)
    if not np.all(np.isfinite(x_new)):
        return float("nan")
    n = max(x_new.size, 1)
    rms = np.sqrt(((np.abs(x_new - x_old)) ** 2).sum() / n)
    return float(rms / max(np.abs(x_old).max(), np.abs(x_new).max(), scale, 1e-30))


def heatflow_solve(g, gc, vel0, pres0, T0, den, visc, fce_of, bnd_vel, bnd_nid, pref, heat_rows, t_dirichlet,
                   max_iter, tol, damp, dt=0.0, num_ts=0):
    """SteadyBase::solve / TransientBase::solve (src/steady/steady_base.rs:16-135,
    src/transient/transient_base.rs:16-160) for SteadyHeatFlow / TransientHeatFlow /
    TransientHeatMassFlow (src/steady/core/steady_heatflow.rs:154-536,
    src/transient/core/transient_heatmassflow.rs:248-600) with constant density and
    viscosity and a body force that depends on the scalar unknowns through a host closure
    (Boussinesq; the reference evaluates such function fields in _update_func before every
    assembly).

    The solution vector is [vel (3N), pres (N), S] with S the K scalar unknowns (temperature,
    then concentrations; K = len(T0) / N) (initial_start, steady_heatflow.rs:154-186).
    Before every assembly: fce = fce_of(S), gradients, update_flow (Rhie-Chow flux), then the
    flow rows (assemble_flow) and the scalar rows heat_rows(S, S_prev, f) -> (rows, cols,
    vals, b) over K N rows with the flow data f of the same iterate (its flux and density),
    shifted by 4N.  t_dirichlet(S) returns S with the Dirichlet values loaded
    (apply_initial).  SuperLU solves, damping, the scaled residual as the max over the
    fields.  dt = 0: steady; else num_ts implicit-Euler steps with the previous state
    updated after each step.  Returns vel, pres, S and the Picard iterations (total)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    N, NP = g["N"], g["num_pts"]
    vel = np.array(vel0, np.float64).reshape(3, N).copy()
    pres = np.array(pres0, np.float64).copy()
    for bid, val in bnd_vel:
        nodes = NP + np.asarray(bnd_nid[bid], np.int64)
        vel[:, nodes] = np.asarray(val).reshape(3, -1)
    T = t_dirichlet(np.array(T0, np.float64).copy())
    K = T.size // N
    vel_prev, T_prev = vel.copy(), T.copy()
    transient = dt > 0.0

    def system(vel, pres, T):
        fce = fce_of(T)
        f = flow_fields(g, gc, vel, pres, den, visc, fce, fce_const=False)
        r, c, v, b = assemble_flow(g, gc, f, bnd_vel, bnd_nid, pref, dt, vel_prev if transient else None,
                                   den if transient else None)
        rh, ch, vh, bh = heat_rows(T, T_prev, f)
        rows = np.concatenate([np.asarray(r, np.int64), 4 * N + np.asarray(rh, np.int64)])
        cols = np.concatenate([np.asarray(c, np.int64), 4 * N + np.asarray(ch, np.int64)])
        vals = np.concatenate([np.asarray(v, np.float64), np.asarray(vh, np.float64)])
        n = (4 + K) * N
        rp, col, val = coo_to_csr(rows, cols, vals, n)
        return sp.csr_matrix((val, col, rp), shape=(n, n)).tocsc(), np.concatenate([b, bh])

    def residual(xo, xn):
        r = [scaled_residual_flow(N, xo[:4 * N], xn[:4 * N], den)]
        r += [scaled_residual_scalar(xo[(4 + k) * N:(5 + k) * N], xn[(4 + k) * N:(5 + k) * N]) for k in range(K)]
        if any(np.isnan(x) for x in r):
            return float("nan")
        return max(r)

    total = 0
    x_iter = np.concatenate([vel.ravel(), pres, T])
    for ts in range(max(num_ts, 1) if transient else 1):
        A, b = system(vel, pres, T)
        for it in range(max_iter):
            x_u = spla.spsolve(A, b)
            x_d = (1.0 - damp) * x_iter + damp * x_u
            vel, pres, T = x_d[:3 * N].reshape(3, N).copy(), x_d[3 * N:4 * N].copy(), x_d[4 * N:].copy()
            A, b = system(vel, pres, T)
            res = residual(x_iter, x_d)
            if not np.isfinite(res):
                raise FloatingPointError("NonFiniteResidual")
            x_iter = x_d
            total += 1
            if res < tol:
                break
        else:
            raise RuntimeError("DidNotConverge")
        vel_prev, T_prev = vel.copy(), T.copy()
    return vel, pres, T, total


# --------------------------------------------------------------------------- flow adjoint
# The Jacobian of the Picard fixed point R(x) = A(c(x)) x - b(c(x), x_prev) of the coupled
# flow system, J = A + E with E the derivative through the state-dependent coefficients c
# (the Rhie-Chow flux, rc, the velocity and body-force gradients, the body force).  There is
# no reference counterpart; flow_jt() is pinned by central finite differences of
# coupled_residual() in random directions (tests/test_cpu_flow_adjoint_oracle.py): any
# dropped, mis-signed or mis-indexed term of E fails that identity.

def coupled_residual(g, gc, x, x_prev, den, visc, fce_of, bnd_vel, bnd_nid, pref, heat_rows, dt=0.0, den_const=True,
                     visc_const=True):
    """R(x) = A(x) x - b(x) of the system heatflow_solve iterates (its system() at the state
    x = [vel (3N), pres (N), S], previous state x_prev for dt > 0).  Returns R, A (CSR), the
    flow fields f."""
    import scipy.sparse as sp
    N = g["N"]
    x = np.asarray(x, np.float64)
    vel, pres, S = x[:3 * N].reshape(3, N), x[3 * N:4 * N], x[4 * N:]
    f = flow_fields(g, gc, vel, pres, den, visc, fce_of(S), den_const=den_const, visc_const=visc_const, fce_const=False)
    transient = dt > 0.0
    vp = np.asarray(x_prev, np.float64)[:3 * N].reshape(3, N) if transient else None
    r, c, v, b = assemble_flow(g, gc, f, bnd_vel, bnd_nid, pref, dt, vp, den if transient else None)
    Sp = np.asarray(x_prev, np.float64)[4 * N:] if transient else S
    rh, ch, vh, bh = heat_rows(S, Sp, f)
    K = S.size // N
    rows = np.concatenate([np.asarray(r, np.int64), 4 * N + np.asarray(rh, np.int64)])
    cols = np.concatenate([np.asarray(c, np.int64), 4 * N + np.asarray(ch, np.int64)])
    vals = np.concatenate([np.asarray(v, np.float64), np.asarray(vh, np.float64)])
    n = (4 + K) * N
    rp, col, val = coo_to_csr(rows, cols, vals, n)
    A = sp.csr_matrix((val, col, rp), shape=(n, n))
    return A @ x - np.concatenate([b, bh]), A, f


def scalar_adv_terms(g, ops, scl, flow_den, bnd_nid):
    """Per advection operator of a scalar list: (directed edges of its rows, kappa, cond) with
    the operator's coefficient phi(kappa F) = cond (P(kappa F / cond) - 1) + max(-kappa F, 0)
    (opscl_dom_advunity.rs:70-112: kappa = 1; opscl_dom_adv.rs:86-118: kappa = m_f / rho_f)."""
    N = g["N"]
    rowsE, ngh = g["rows"].astype(np.int64), g["ngh"].astype(np.int64)
    out = []
    for op in ops:
        if op.kind not in (Op.SCL_DOM_ADVECTION_UNITY, Op.SCL_BND_ADVECTION_UNITY, Op.SCL_DOM_ADVECTION, Op.SCL_BND_ADVECTION):
            continue
        m = np.zeros(N, bool)
        m[_op_rows(g, op, bnd_nid)] = True
        e = np.nonzero(m[rowsE])[0]
        v_, n_ = rowsE[e], ngh[e]
        sk = g["skew"][:, e]
        if op.kind in (Op.SCL_DOM_ADVECTION_UNITY, Op.SCL_BND_ADVECTION_UNITY):
            df = _diff_face(*scl[op.a], v_, n_, sk) if op.a >= 0 else np.zeros(e.size)
            kap = np.ones(e.size)
        else:
            df = _diff_face(*scl[op.b], v_, n_, sk)
            W, Wg = scl[op.a]
            Rr, Rg = flow_den[op.c]
            wgt_f = _face_mean(W, Wg, v_, n_, *sk)
            den_f = _face_mean(Rr, Rg, v_, n_, *sk)
            with np.errstate(divide="ignore", invalid="ignore"):
                kap = np.where(den_f != 0.0, wgt_f / den_f, 0.0)
        out.append((e, kap, df * g["area"][e] / g["dist"][e]))
    return out


def _dphi_adv(F, cond):
    """d/dF of cond (P(F / cond) - 1) + max(-F, 0), P(x) = max(0, 1 - 0.1|x|)^5:
    P'(Pe) - H(-F) with P'(Pe) = -0.5 a^4 sign(Pe), Pe = F / cond (cond < 0 on dual faces of
    negative area); -0.5 at F = 0 (the derivative for cond > 0, the mean of the one-sided
    ones for cond < 0); max(-F, 0) alone when cond = 0."""
    with np.errstate(divide="ignore", invalid="ignore"):
        pe = np.where(cond != 0.0, F / cond, 0.0)
    a4 = np.maximum(1.0 - 0.1 * np.abs(pe), 0.0) ** 4
    d = -0.5 * a4 * np.sign(pe) - (F < 0.0)
    return np.where(cond != 0.0, np.where(F == 0.0, -0.5, d), np.where(F < 0.0, -1.0, 0.0))


def _dh_ap(F, cond):
    """d/dF of cond P(F / cond) + max(F, 0), the face term of a_P (flow_helper.rs:3-70):
    P'(Pe) + H(F), 0.5 at F = 0."""
    with np.errstate(divide="ignore", invalid="ignore"):
        pe = np.where(cond != 0.0, F / cond, 0.0)
    a4 = np.maximum(1.0 - 0.1 * np.abs(pe), 0.0) ** 4
    d = -0.5 * a4 * np.sign(pe) + (F > 0.0)
    return np.where(cond != 0.0, np.where(F == 0.0, 0.5, d), np.where(F > 0.0, 1.0, 0.0))


def flow_jt(g, gc, f, x, lam, A, pref_row, adv_blocks, body_blk=-1, dfce=None):
    """y = J(x)^T lam = A^T lam + E^T lam for the coupled system of coupled_residual (A and f
    from it).  adv_blocks[k]: scalar_adv_terms of scalar block k; body_blk, dfce (3, N): the
    block the body force depends on pointwise, fce = f0 + dfce s.  E^T is written as scatters
    over the directed edges e = (v, n), following the chain
        F_e = rho_f (u_f . n + vel_rc) A  (flow_helper.rs:114-171),
        vel_rc = -(rc_v + rc_n)/2 ((p_n - p_v)/d - fce_e) + n . (mom_v + mom_n)/2,
        mom = rc (grad p - fce_c), rc = V / a_P, a_P = sum_e h(Fc_e)  (flow_helper.rs:3-112),
    through the rows that read them: momentum advection (opvec_dom_adv.rs:72-108) and the
    transpose / dilatatory / source right-hand sides (opvec_dom_dtrn.rs, opvec_dom_ddil.rs,
    opvec_dom_src.rs), the scalar advection rows, and the continuity rows, which equal
    -sum_e F_e (their velocity and pressure columns at fixed rc are in A)."""
    N, NP = g["N"], g["num_pts"]
    x = np.asarray(x, np.float64)
    lam = np.asarray(lam, np.float64)
    rows, ngh = g["rows"].astype(np.int64), g["ngh"].astype(np.int64)
    Ae, d, nv, sk, gw = g["area"], g["dist"], g["norm"], g["skew"], g["gw"]
    vol, cs = g["vol"], g["cskew"]
    vel, pres, rc, F = f["vel"], f["pres"], f["rc"], f["flux"]
    fce, fceg = f["fce"], f["fceg"]
    lu = lam[:3 * N].reshape(3, N)
    lp = lam[3 * N:4 * N]
    K = (x.size - 4 * N) // N
    S = [x[(4 + k) * N:(5 + k) * N] for k in range(K)]
    Ls = [lam[(4 + k) * N:(5 + k) * N] for k in range(K)]
    den_f = _face_mean(f["den"], f["deng"], rows, ngh, *sk)
    visc_f = _visc_face(f["visc"], f["viscg"], rows, ngh, *sk)
    # adjoint of the face mass flux
    gF = np.zeros(F.size)
    pm = rows < NP
    dph = _dphi_adv(F, visc_f * Ae / d)
    for c in range(3):
        gF += np.where(pm, lu[c][rows] * dph * (vel[c][ngh] - vel[c][rows]), 0.0)
    for k, terms in enumerate(adv_blocks):
        for e, kap, cond in terms:
            gF[e] += Ls[k][rows[e]] * kap * _dphi_adv(kap * F[e], cond) * (S[k][ngh[e]] - S[k][rows[e]])
    wF = gF * den_f * Ae
    wC = -lp[rows] * den_f * Ae
    wC[3 * N + rows == pref_row] = 0.0
    wT = wF + wC
    gu, gvelg = np.zeros((3, N)), np.zeros((9, N))
    gp, ggp, grc = np.zeros(N), np.zeros((3, N)), np.zeros(N)
    gfce, gfceg = np.zeros((3, N)), np.zeros((9, N))
    both = lambda arr, val_v, val_n: (np.add.at(arr, rows, val_v), np.add.at(arr, ngh, val_n))
    # u_f . n
    for c in range(3):
        both(gu[c], wF * 0.5 * nv[c], wF * 0.5 * nv[c])
        for q in range(3):
            both(gvelg[3 * c + q], wF * 0.5 * nv[c] * sk[q], wF * 0.5 * nv[c] * sk[q])
    # vel_rc: pressure (flux path only), grad p through mom, rc, body force
    rcs = rc[rows] + rc[ngh]
    both(gp, wF * 0.5 * rcs / d, -wF * 0.5 * rcs / d)
    for c in range(3):
        both(ggp[c], wF * 0.5 * rc[rows] * nv[c], wF * 0.5 * rc[ngh] * nv[c])
    gm = np.repeat(np.arange(N), np.diff(gc[0]))
    pg = []
    for c in range(3):
        acc = np.zeros(N)
        np.add.at(acc, gm, gc[2][c] * pres[gc[1]] / vol[gm])
        pg.append(acc)
    fcc = _fce_centroid(g, fce, fceg, np.arange(N))
    pg = np.stack([pg[c] - fcc[c] for c in range(3)])
    fce_e = 0.5 * sum(nv[c] * (fce[c][rows] + fce[c][ngh]) for c in range(3))
    dp = (pres[ngh] - pres[rows]) / d
    both(grc, wT * (-0.5 * (dp - fce_e) + 0.5 * sum(nv[c] * pg[c][rows] for c in range(3))),
         wT * (-0.5 * (dp - fce_e) + 0.5 * sum(nv[c] * pg[c][ngh] for c in range(3))))
    for c in range(3):
        both(gfce[c], wT * (0.25 * rcs - 0.5 * rc[rows]) * nv[c], wT * (0.25 * rcs - 0.5 * rc[ngh]) * nv[c])
        for q in range(3):
            both(gfceg[3 * c + q], wT * (-0.5 * rc[rows] * nv[c] * cs[q][rows]), wT * (-0.5 * rc[ngh] * nv[c] * cs[q][ngh]))
    # momentum right-hand sides of the points: +sum st + sum sd + fce_c V
    gdiv = np.zeros(N)
    for c in range(3):
        lc = np.where(pm, lu[c][rows], 0.0)
        for q in range(3):
            t = lc * visc_f * Ae * nv[q] * 0.5
            both(gvelg[3 * q + c], t, t)
        t = lc * (-2.0 / 3.0) * visc_f * 0.5 * Ae * nv[c]
        both(gdiv, t, t)
    for q in (0, 4, 8):
        gvelg[q] += gdiv
    for c in range(3):
        gfce[c][:NP] += lu[c][:NP] * vol[:NP]
        for q in range(3):
            gfceg[3 * c + q][:NP] += lu[c][:NP] * vol[:NP] * cs[q][:NP]
    # rc = V / a_P
    gaP = np.where(rc > 0.0, -grc * rc * rc / vol, 0.0)
    Fc = den_f * 0.5 * sum((vel[c][rows] + vel[c][ngh]) * nv[c] for c in range(3)) * Ae
    w = gaP[rows] * _dh_ap(Fc, visc_f * Ae / d) * den_f * 0.5 * Ae
    for c in range(3):
        both(gu[c], w * nv[c], w * nv[c])
    # least-squares gradients grad_q x_c,v = sum_e gw_q (x_c,n - x_c,v)
    for c in range(3):
        for q in range(3):
            t = gvelg[3 * c + q][rows] * gw[q]
            both(gu[c], -t, t)
            t = gfceg[3 * c + q][rows] * gw[q]
            both(gfce[c], -t, t)
    # grad p_v = sum_q gc_q p_q / V_v
    for c in range(3):
        np.add.at(gp, gc[1], ggp[c][gm] * gc[2][c] / vol[gm])
    y = A.T @ lam
    y[:3 * N] += gu.ravel()
    y[3 * N:4 * N] += gp
    if body_blk >= 0:
        y[(4 + body_blk) * N:(5 + body_blk) * N] += sum(dfce[c] * gfce[c] for c in range(3))
    return y


def bjacobi_apply(rp, col, val, r, N):
    """z = B^-1 r for the block diagonal B of the unknowns of each vertex (v, N+v, ...)."""
    n = rp.size - 1
    nb = n // N
    z = np.empty(n)
    for v in range(N):
        idx = np.arange(nb) * N + v
        B = np.zeros((nb, nb))
        for i, row in enumerate(idx):
            cs = col[rp[row]:rp[row + 1]]
            vs = val[rp[row]:rp[row + 1]]
            for j, cc in enumerate(idx):
                k = np.nonzero(cs == cc)[0]
                if k.size:
                    B[i, j] = vs[k[0]]
        z[idx] = np.linalg.solve(B, r[idx])
    return z


def transient_adaptive(system, x0, t_end, dt_init, dt_min=None, dt_max=None, rtol=1e-3, atol=1e-12, safety=0.9,
                       fac_min=0.2, fac_max=2.0, fac_fail=0.25, max_reject=30, out_times=None, max_iter=30, tol=1e-6,
                       damp=1.0, fixed=False, hook=None):
    """TransientAdaptive::solve_adaptive_with (src/transient/transient_adaptive.rs:81-350) for one
    scalar unknown.  system(x, x_prev, dt) -> (A, b) assembles at state x with previous state
    x_prev; x0 already carries the Dirichlet values.  Scalar field norm (scaling.rs:88-135):
    Picard residual rms(dx) / max|x|, error norm max |LTE| / (atol + rtol max|x|).
    Returns (x, stats dict, accepted steps [(t, dt, err)])."""
    from scipy.sparse.linalg import spsolve
    dt_min = 1e-12 * t_end if dt_min is None else dt_min
    dt_max = t_end if dt_max is None else dt_max
    t_eps = 1e-12 * t_end
    outs = [t_end * i / 10.0 for i in range(11)] if out_times is None else list(out_times)
    outs = sorted(t for t in outs if t > t_eps and t <= t_end)
    ded = []
    for t in outs:
        if not ded or abs(t - ded[-1]) > t_eps:
            ded.append(t)
    outs = ded
    if outs and outs[-1] >= t_end - t_eps:
        outs[-1] = t_end
    else:
        outs.append(t_end)

    def resid(xo, xn):
        if not np.all(np.isfinite(xn)):
            return float("nan")
        rms = np.sqrt(((np.abs(xn - xo)) ** 2).sum() / max(xn.size, 1))
        return rms / max(np.abs(xo).max(), np.abs(xn).max(), 1e-30)

    st = dict(accepted=0, rejected_error=0, rejected_picard=0, picard=0, outputs=0)
    steps = []
    x_n = np.array(x0, np.float64)
    x_nm1, dt_prev = None, None
    t, out_idx = 0.0, 0
    dt = min(max(dt_init, dt_min), dt_max)
    err_prev, n_reject, no_grow = 1.0, 0, False
    while t < t_end - t_eps:
        t_out = outs[out_idx]
        dt_step, hits = dt, False
        if t + dt_step >= t_out - t_eps:
            dt_step, hits = t_out - t, True
        elif t + 1.5 * dt_step > t_out:
            dt_step = 0.5 * (t_out - t)
        t_new = t_out if hits else t + dt_step
        x_iter = x_n.copy()
        A, b = system(x_iter, x_n, dt_step)
        conv, res = False, float("nan")
        for _ in range(max_iter):
            x_u = spsolve(A.tocsc(), b)
            x_d = (1.0 - damp) * x_iter + damp * x_u
            A, b = system(x_d, x_n, dt_step)
            res = resid(x_iter, x_d)
            st["picard"] += 1
            x_iter = x_d
            if not np.isfinite(res):
                break
            if res < tol:
                conv = True
                break
        if not conv:
            st["rejected_picard"] += 1
            n_reject += 1
            if fixed or n_reject > max_reject or dt_step <= dt_min * (1.0 + 1e-12):
                raise RuntimeError("TransientDidNotConverge")
            dt = max(dt_step * fac_fail, dt_min)
            no_grow = True
            continue
        err = 0.0
        if not fixed:
            if x_nm1 is not None:
                yp = x_n + dt_step / dt_prev * (x_n - x_nm1)
                fac = dt_step / (2.0 * dt_step + dt_prev)
            else:
                yp, fac = x_n, 0.5
            emax = (fac * np.abs(x_iter - yp)).max()
            scale = max(np.abs(x_n).max(), np.abs(x_iter).max(), 1e-30)
            err = emax / (atol + rtol * scale)
        if err > 1.0 and dt_step > dt_min * (1.0 + 1e-12):
            st["rejected_error"] += 1
            n_reject += 1
            if n_reject > max_reject:
                raise RuntimeError("TransientDidNotConverge")
            dt = max(dt_step * max(safety * err ** -0.5, fac_min), dt_min)
            no_grow = True
            continue
        if hook is not None:
            hook(t_new, dt_step, x_iter)
        st["accepted"] += 1
        steps.append((t_new, dt_step, err))
        x_nm1, dt_prev = x_n, dt_step
        x_n = x_iter
        t = t_new
        n_reject = 0
        if hits:
            st["outputs"] += 1
            out_idx += 1
            if out_idx >= len(outs):
                break
        if fixed:
            dt = dt_init
            continue
        e = max(err, 1e-10)
        fac = safety * e ** -0.35 * max(err_prev, 1e-10) ** 0.2
        fac = min(max(fac, fac_min), fac_max)
        if no_grow:
            fac, no_grow = min(fac, 1.0), False
        base = max(dt, dt_step) if hits else dt_step
        dt = min(max(base * fac, dt_min), dt_max)
        err_prev = e
    return x_n, st, steps
