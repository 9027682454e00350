"""CPU tests: mesh, exact Delaunay, Voronoi dual, and the oracle's pins.

Each oracle function is held to something other than itself: the scipy Qhull
tetrahedralisation, closed forms (regular tetrahedron, linear fields, the 1D heat
solution of heat_analytic.rs:9, uniform convected concentration), invariants
(closed control volumes, zero row sums, symmetry) and SuperLU.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import vfvm
import vfvm_oracle as O
from problems import heat_c1, mass_c2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _sorted_tets(t):
    return set(map(tuple, np.sort(np.asarray(t), axis=1)))


# ------------------------------------------------------------------ Delaunay
def test_delaunay_matches_qhull_golden():
    g = np.load(os.path.join(GOLD, "lattice6_qhull.npz"))
    m = vfvm.Mesh.lattice(6, 0, 0.3)
    assert np.array_equal(m.get("pos"), g["pos"])
    assert _sorted_tets(m.get("tets")) == _sorted_tets(g["tets"])


@pytest.mark.parametrize("n,seed", [(5, 1), (9, 0), (12, 7)])
def test_delaunay_matches_qhull(n, seed):
    from scipy.spatial import Delaunay
    m = vfvm.Mesh.lattice(n, seed, 0.3)
    pos = m.get("pos")
    assert _sorted_tets(m.get("tets")) == _sorted_tets(Delaunay(pos).simplices)


def test_tet_orientation_and_volume():
    m = vfvm.Mesh.lattice(8)
    P = m.get("pos")[m.get("tets")]
    det = np.einsum("ij,ij->i", np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), P[:, 3] - P[:, 0])
    # reference orientation: p3 below the counterclockwise face p0 p1 p2 (geom_tet.rs:13)
    assert np.all(det < 0)
    assert abs(-det.sum() / 6 - 1.0) < 1e-13


def test_lattice_layout():
    n = 7
    m = vfvm.Mesh.lattice(n)
    assert m.num_pts == (n - 2) ** 3 and m.num_tot == n ** 3
    pos = m.get("pos")
    # nodes lie on the cube faces, points strictly inside (geom_tet.rs:118-120)
    on = np.any((pos == 0.0) | (pos == 1.0), axis=1)
    assert not on[:m.num_pts].any() and on[m.num_pts:].all()
    # every node on exactly one boundary; x faces own their edges and corners
    nid = np.concatenate([m.bnd_nodes(b) for b in range(6)])
    assert np.array_equal(np.sort(nid), np.arange(m.num_nds))
    x0 = pos[m.num_pts + m.bnd_nodes(0)]
    assert np.all(x0[:, 0] == 0.0) and len(x0) == n * n
    # coordinates on the 2^-22 grid
    assert np.all(pos * 2 ** 22 == np.round(pos * 2 ** 22))


# ------------------------------------------------------------------ dual: product vs oracle, oracle pins
@pytest.mark.parametrize("n", [5, 9])
def test_dual_product_equals_oracle_bitwise(n):
    m = vfvm.Mesh.lattice(n)
    g = O.dual(m.get("pos"), m.get("tets"), m.num_pts)
    for k in ["row_ptr", "ngh", "dist", "area", "norm", "skew", "gw", "vol", "cskew", "bsf_ptr", "bsf_vj", "bsf_vk",
              "bsf_area", "bsf_norm", "bsf_wgt"]:
        assert np.array_equal(m.get(k), g[k]), k


def test_dual_from_tets_equals_lattice():
    m = vfvm.Mesh.lattice(6)
    m2 = vfvm.Mesh.from_tets(m.get("pos"), m.get("tets")[:, [0, 2, 1, 3]], m.num_pts, [m.bnd_nodes(b) for b in range(6)])
    for k in ["ngh", "area", "gw", "vol", "bsf_area"]:
        assert np.array_equal(m.get(k), m2.get(k)), k


def test_dual_invariants():
    m = vfvm.Mesh.lattice(10)
    g = O.dual(m.get("pos"), m.get("tets"), m.num_pts)
    rows = g["rows"]
    # closed interior control volumes: sum_faces A n = 0 (geom_tet.rs:186-196)
    s = np.zeros((3, g["N"]))
    for d in range(3):
        np.add.at(s[d], rows, g["area"] * g["norm"][d])
    assert np.abs(s[:, :m.num_pts]).max() < 1e-15
    # boundary cells close with their sub-faces
    bs = np.zeros((3, g["N"]))
    vid = np.repeat(np.arange(g["N"]), np.diff(g["bsf_ptr"]))
    for d in range(3):
        np.add.at(bs[d], vid, g["bsf_area"] * g["bsf_norm"][d])
    assert np.abs((s + bs)[:, m.num_pts:]).max() < 1e-14
    # control volumes tile the cube (geom_tet.rs:198-200)
    assert abs(g["vol"].sum() - 1.0) < 1e-13
    # boundary areas tile the six faces
    assert abs(g["bnd_sum"].sum() - 6.0) < 1e-12
    # LSQ weights reproduce linear fields exactly (geom_tet.rs:261-293)
    a = np.array([0.3, -1.7, 2.2])
    u = m.get("pos") @ a + 0.5
    grad = O.update_grad(g, u)
    assert np.abs(grad - a[:, None]).max() < 1e-11


def test_dual_regular_tetrahedron_closed_form():
    """One regular tetrahedron of edge 1: each Voronoi face piece is the right triangle with
    legs r_face = 1/(2 sqrt 3) and r_tet = 1/(2 sqrt 6), so every edge carries
    A = 2 * (1/2) r_face r_tet = 1/(12 sqrt 2); each control volume is a quarter of
    V = 1/(6 sqrt 2)."""
    pos = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float) / (2 * np.sqrt(2))
    tets = np.array([[0, 2, 1, 3]])
    g = O.dual(pos, tets, 0)
    assert np.allclose(g["area"], 1 / (12 * np.sqrt(2)), rtol=1e-14, atol=0)
    assert np.allclose(g["vol"], 1 / (24 * np.sqrt(2)), rtol=1e-14, atol=0)
    assert np.allclose(g["dist"], 1.0, rtol=1e-15)
    # face of edge (i, j) = triangles (m, c_f1, c_T), (m, c_f2, c_T) of equal area, c_T = 0 and
    # c_f1 + c_f2 = 2m/3, so its centre is m + (c_f1 + c_f2 + 2 c_T - 4m)/6 = m - 5m/9
    P = pos
    mid = 0.5 * (P[g["rows"]] + P[g["ngh"]])
    assert np.abs(g["skew"].T - (-5.0 / 9.0) * mid).max() < 1e-15
    # each control volume's centroid lies on the ray from the vertex toward the centre
    cs = g["cskew"].T
    assert np.abs(np.cross(cs, P)).max() < 1e-16 and np.all(np.einsum("ij,ij->i", cs, P) < 0)
    m = vfvm.Mesh.from_tets(pos, tets, 0, [np.arange(4)])
    assert np.array_equal(m.get("area"), g["area"]) and np.array_equal(m.get("vol"), g["vol"])


def test_points_on_boundary_rejected():
    # a point (vid < num_pts) on the hull is TetraPointsOnBoundary (geom_tet.rs:119-120)
    pos = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)
    with pytest.raises(vfvm.VFVChemError) as e:
        vfvm.Mesh.from_tets(pos, np.array([[0, 1, 2, 3]]), 1, [np.arange(3)])
    assert e.value.code == 11
    with pytest.raises(ValueError):
        O.dual(pos, np.array([[0, 2, 1, 3]]), 1)


def test_degenerate_and_invalid_input_rejected():
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float)  # flat
    with pytest.raises(vfvm.VFVChemError) as e:
        vfvm.Mesh.from_tets(pos, np.array([[0, 1, 2, 3]]), 0, [np.arange(4)])
    assert e.value.code == 10
    with pytest.raises(vfvm.VFVChemError) as e:
        vfvm.Mesh.from_tets(pos, np.array([[0, 1, 2, 7]]), 0, [np.arange(4)])
    assert e.value.code == 1
    with pytest.raises(vfvm.VFVChemError) as e:
        vfvm.Mesh.lattice(2)
    assert e.value.code == 1


# ------------------------------------------------------------------ operators (oracle)
def test_oracle_heat_golden_and_closed_form():
    gold = np.load(os.path.join(GOLD, "heat_c1_n6.npz"))
    p = heat_c1(vfvm.Mesh.lattice(6))
    rp, col, val, b, _ = p.oracle_system()
    assert np.array_equal(rp, gold["row_ptr"]) and np.array_equal(col, gold["col"])
    assert np.array_equal(val, gold["val"]) and np.array_equal(b, gold["b"])


def test_oracle_heat_converges_to_1d_solution():
    """T = x + q x (1 - x) / (2k) (heat_analytic.rs:9 with T(0)=0, T(1)=1): the error of
    the discrete solution falls at second order under refinement."""
    errs = []
    for n in (8, 16):
        p = heat_c1(vfvm.Mesh.lattice(n))
        rp, col, val, b, _ = p.oracle_system()
        x = spla.spsolve(sp.csr_matrix((val, col, rp), shape=(p.N, p.N)).tocsc(), b)
        xx = p.pos[:, 0]
        errs.append(np.abs(x - (xx + xx * (1 - xx) / 2)).max())
    assert errs[1] < errs[0] / 3.0 and errs[1] < 1e-3


def test_oracle_diffusion_row_sums_and_symmetry():
    p = heat_c1(vfvm.Mesh.lattice(9))
    rp, col, val, b, _ = p.oracle_system()
    A = sp.csr_matrix((val, col, rp), shape=(p.N, p.N))
    NP = p.mesh.num_pts
    rs = np.asarray(A.sum(axis=1)).ravel()
    # uniform T is in the null space of diffusion rows (opscl_dom_diff.rs:84-85)
    assert np.abs(rs[:NP]).max() < 1e-15
    # constant conductivity: the point-point block is symmetric
    App = A[:NP, :NP].toarray()
    assert np.abs(App - App.T).max() < 1e-16
    # interior rows are exact for linear T: sum A/d (T_n - T_v) = grad T . sum A n = 0
    lin = p.pos @ np.array([1.0, 2.0, -0.5])
    assert np.abs((A @ lin)[:NP]).max() < 1e-14
    # the source enters b as -q V (opscl_dom_src.rs:61)
    g = O.dual(p.pos, p.mesh.get("tets"), NP)
    assert np.allclose(b[:NP], -g["vol"][:NP], rtol=1e-15)
    # Dirichlet rows are the identity (opscl_bnd_dir.rs:73-74)
    x0 = NP + p.mesh.bnd_nodes(0)
    assert np.all(np.diff(rp)[x0] == 1) and np.all(A[x0, x0] == 1.0)


def test_oracle_power_law_limits():
    """A(|Pe|) = max(0, 1 - 0.1|Pe|)^5: |Pe| >= 10 gives pure upwinding, F = 0 no advection
    (opscl_dom_advunity.rs:104-107)."""
    cond = np.array([1.0, 1.0, 1.0, 2.0])
    F = np.array([20.0, -20.0, 0.0, 10.0])
    pe = F / cond
    pl = O._pow5(np.maximum(1.0 - 0.1 * np.abs(pe), 0.0))
    coeff = cond * (pl - 1.0) + np.maximum(-F, 0.0)
    assert np.array_equal(coeff, np.array([-1.0, 19.0, 0.0, -2.0 + 0.0 + 2.0 * (0.5 ** 5)]))
    assert O._pow5(0.5) == 0.5 ** 5


def test_oracle_flux_closed_form_and_antisymmetry():
    m = vfvm.Mesh.lattice(7)
    g = O.dual(m.get("pos"), m.get("tets"), m.num_pts)
    N = g["N"]
    u = np.array([0.3, -0.2, 0.7])
    vel = np.repeat(u[:, None], N, axis=1)
    F = O.flux_prescribed(g, vel, np.zeros((9, N)), np.ones(N), np.zeros((3, N)))
    assert np.allclose(F, (u @ g["norm"]) * g["area"], rtol=1e-14, atol=1e-18)
    # antisymmetry for a varying field: F(v->n) = -F(n->v)
    vel2 = np.stack([np.sin(m.get("pos")[:, 1]), m.get("pos")[:, 0] ** 2, np.cos(m.get("pos")[:, 2])])
    F2 = O.flux_prescribed(g, vel2, O.update_grad_vec(g, vel2), np.ones(N), np.zeros((3, N)))
    key = {(r, c): k for k, (r, c) in enumerate(zip(g["rows"], g["ngh"]))}
    rev = np.array([key[(c, r)] for r, c in zip(g["rows"], g["ngh"])])
    assert np.abs(F2 + F2[rev]).max() < 1e-15


def test_oracle_convection_uniform_solution():
    """u.n = 0 on every wall and c = 1 at the inlet: c = 1 solves the assembled system exactly
    (row sums vanish: opscl_dom_advunity.rs:59-60)."""
    p = mass_c2(vfvm.Mesh.lattice(9))
    rp, col, val, b, _ = p.oracle_system()
    A = sp.csr_matrix((val, col, rp), shape=(p.N, p.N))
    assert np.abs(A @ np.ones(p.N) - b).max() < 1e-15


def test_oracle_gmres_vs_superlu():
    p = mass_c2(vfvm.Mesh.lattice(10), c_out=0.0)
    rp, col, val, b, _ = p.oracle_system()
    A = sp.csr_matrix((val, col, rp), shape=(p.N, p.N))
    xl = spla.spsolve(A.tocsc(), b)
    x, ok, it, rel = O.gmres(rp, col, val, b, tol=1e-13, max_outer=400)
    assert ok and rel <= 1e-13
    assert abs(np.linalg.norm(b - A @ x) / np.linalg.norm(b) - rel) < 1e-15 + 1e-3 * rel
    assert np.linalg.norm(x - xl) / np.linalg.norm(xl) < 1e-9


def test_planted_mistake_is_caught():
    """A sign error in the oracle's skew correction (diff_f = diff_m - ... instead of +) must
    break the linear-field exactness pin for a varying conductivity."""
    m = vfvm.Mesh.lattice(8)
    g = O.dual(m.get("pos"), m.get("tets"), m.num_pts)
    k = 1.0 + m.get("pos")[:, 0]
    kg = O.update_grad(g, k)
    v, n = g["rows"], g["ngh"]
    good = O._diff_face(k, kg, v, n, g["skew"])
    bad = O._diff_face(k, -kg, v, n, g["skew"])
    # on a linear k the skew-corrected face value equals k at the face centre
    fc = 0.5 * (m.get("pos")[v] + m.get("pos")[n]) + g["skew"].T
    exact = 1.0 + fc[:, 0]
    hm = 2 * k[v] * k[n] / (k[v] + k[n])
    assert np.abs(good - (hm + (exact - 0.5 * (k[v] + k[n])))).max() < 1e-12
    assert np.abs(bad - (hm + (exact - 0.5 * (k[v] + k[n])))).max() > 1e-4
