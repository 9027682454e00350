"""Voronoi dual built on the GPU (vfvm_mesh_lattice_dev / vfvm_mesh_from_tets_dev) against
the host build (dual.cpp, restating Tetrahedron::update_tet, src/base/geom_tet.rs:58-293, and
pinned against the oracle in test_cpu_mesh_oracle.py): every array bitwise equal."""
import time

import numpy as np
import pytest

import vfvm

ARRAYS = ["row_ptr", "ngh", "dist", "area", "norm", "skew", "gw", "vol", "cskew", "bsf_ptr", "bsf_vj", "bsf_vk",
          "bsf_area", "bsf_norm", "bsf_wgt"]


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64) if a.dtype == np.float64 else a


def assert_same_dual(m_host, m_dev):
    assert m_host.num_edges == m_dev.num_edges and m_host.num_subfaces == m_dev.num_subfaces
    for name in ARRAYS:
        a, b = m_host.get(name), m_dev.get(name)
        assert a.shape == b.shape, name
        assert np.array_equal(bits(a), bits(b)), f"{name} differs"


@pytest.mark.gpu
@pytest.mark.parametrize("n", [3, 6, 11, 20])
def test_device_dual_equals_host_lattice(n):
    print(f"device dual: lattice n={n}", flush=True)
    assert_same_dual(vfvm.Mesh.lattice(n, 0, 0.3), vfvm.Mesh.lattice(n, 0, 0.3, device=0))


@pytest.mark.gpu
def test_device_dual_equals_host_from_tets_and_seed():
    print("device dual: from_tets, seed 7", flush=True)
    m = vfvm.Mesh.lattice(9, 7, 0.3)
    bnd = [m.bnd_nodes(b) for b in range(6)]
    tets = m.get("tets")[:, [0, 2, 1, 3]]  # any orientation is accepted
    h = vfvm.Mesh.from_tets(m.get("pos"), tets, m.num_pts, bnd)
    d = vfvm.Mesh.from_tets(m.get("pos"), tets, m.num_pts, bnd, device=0)
    assert_same_dual(h, d)
    assert_same_dual(m, d)


@pytest.mark.gpu
def test_device_dual_single_tetrahedron_and_errors():
    print("device dual: single tetrahedron, error codes", flush=True)
    pos = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], float)
    tets = np.array([[0, 1, 2, 3]])
    assert_same_dual(vfvm.Mesh.from_tets(pos, tets, 0, [np.arange(4)]),
                     vfvm.Mesh.from_tets(pos, tets, 0, [np.arange(4)], device=0))
    # a point on the hull: TetraPointsOnBoundary (geom_tet.rs:119-120), as on the host
    with pytest.raises(vfvm.VFVChemError) as e:
        vfvm.Mesh.from_tets(pos, tets, 1, [np.arange(3)], device=0)
    assert e.value.code == 11


@pytest.mark.gpu
def test_device_dual_large_lattice_and_speed():
    n = 50
    print(f"device dual: lattice n={n}", flush=True)
    t0 = time.perf_counter()
    h = vfvm.Mesh.lattice(n, 0, 0.3)
    t1 = time.perf_counter()
    d = vfvm.Mesh.lattice(n, 0, 0.3, device=0)
    t2 = time.perf_counter()
    assert_same_dual(h, d)
    print(f"lattice {n}^3 with host dual {t1 - t0:.2f} s, with device dual {t2 - t1:.2f} s", flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("n,seed", [(4, 0), (7, 3), (12, 0)])
def test_device_grad_coeff_equals_host(n, seed):
    print(f"device grad_coeff: lattice n={n} seed={seed}", flush=True)
    h = vfvm.Mesh.lattice(n, seed, 0.3).build_grad_coeff()
    d = vfvm.Mesh.lattice(n, seed, 0.3).build_grad_coeff(device=0)
    assert h.num_grad_coeff == d.num_grad_coeff
    for name in ("gc_ptr", "gc_vid", "gc_c"):
        assert np.array_equal(bits(h.get(name)), bits(d.get(name))), f"{name} differs"


@pytest.mark.gpu
def test_device_grad_coeff_large_and_speed():
    n = 40
    print(f"device grad_coeff: lattice n={n}", flush=True)
    m = vfvm.Mesh.lattice(n, 0, 0.3, device=0)
    pos, tets = m.get("pos"), m.get("tets")
    bnd = [m.bnd_nodes(b) for b in range(6)]
    h = vfvm.Mesh.from_tets(pos, tets, m.num_pts, bnd)
    t0 = time.perf_counter()
    h.build_grad_coeff()
    t1 = time.perf_counter()
    m.build_grad_coeff(device=0)
    t2 = time.perf_counter()
    for name in ("gc_ptr", "gc_vid", "gc_c"):
        assert np.array_equal(bits(h.get(name)), bits(m.get(name))), f"{name} differs"
    print(f"grad_coeff {n}^3: host {t1 - t0:.2f} s, device {t2 - t1:.2f} s", flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("n,seed", [(5, 0), (8, 3)])
def test_device_dual_equals_oracle(n, seed):
    """The device-built dual against the oracle's restatement of geom_tet.rs:58-293 directly
    (not through the host build): every array bit for bit."""
    import vfvm_oracle as O
    print(f"device dual vs oracle: lattice n={n} seed={seed}", flush=True)
    m = vfvm.Mesh.lattice(n, seed, 0.3, device=0)
    g = O.dual(m.get("pos"), m.get("tets"), m.num_pts)
    for name in ARRAYS:
        assert np.array_equal(bits(m.get(name)), bits(np.asarray(g[name], m.get(name).dtype))), name
