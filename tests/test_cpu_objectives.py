"""The wall Nusselt number of the natural-convection workload (vfvm.objectives) on the CPU:
1 up to round-off for the conduction profile T = 1 - x on jittered lattices (the control
volumes of the wall nodes close with their boundary faces), 0 for a uniform field, linear in
T, and the flux through the x = 1 wall of the same profile equal and opposite."""
import numpy as np
import pytest

import vfvm
from vfvm.objectives import WallFlux, hot_wall_nusselt


@pytest.mark.parametrize("n,seed", [(6, 0), (9, 4), (14, 1)])
def test_conduction_profile_gives_one(n, seed):
    m = vfvm.Mesh.lattice(n, seed, 0.3)
    x = m.get("pos")[:, 0]
    assert abs(hot_wall_nusselt(m, 1.0 - x, bnd=0) - 1.0) < 1e-12
    # the x = 1 wall carries the same flux out of the domain: -1 with T = 1 - x
    assert abs(hot_wall_nusselt(m, 1.0 - x, bnd=1) + 1.0) < 1e-12


def test_uniform_and_linear():
    m = vfvm.Mesh.lattice(7, 2, 0.3)
    f = WallFlux(m, 0, scale=2.0)
    assert f(np.full(m.num_tot, 3.5)) == 0.0
    rng = np.random.default_rng(0)
    a, b = rng.standard_normal(m.num_tot), rng.standard_normal(m.num_tot)
    assert abs(f(2.0 * a - b) - (2.0 * f(a) - f(b))) < 1e-12 * (1 + abs(f(a)) + abs(f(b)))


def test_wall_profile_map_and_transpose():
    from vfvm.objectives import WallProfile, wall_flux_seed
    m = vfvm.Mesh.lattice(7, 2, 0.3)
    wp = WallProfile(m, 0, shape=(5, 4))
    rng = np.random.default_rng(0)
    # partition of unity: a constant profile gives constant node values
    assert np.allclose(wp.values(np.full((5, 4), 0.7), base=1.0), 1.7, atol=1e-14)
    # a profile linear in the grid coordinates is reproduced exactly at the nodes
    pos = m.get("pos")[m.num_pts + m.bnd_nodes(0).astype(np.int64)]
    I, J = np.meshgrid(np.arange(5) / 4.0, np.arange(4) / 3.0, indexing="ij")
    assert np.allclose(wp.values(0.3 * I - 0.2 * J), 0.3 * pos[:, 1] - 0.2 * pos[:, 2], atol=1e-12)
    # grad is the transpose of values (dot-product test)
    th, g = rng.standard_normal((5, 4)), rng.standard_normal(wp.count)
    assert abs(g @ wp.values(th) - np.sum(wp.grad(g) * th)) < 1e-12 * np.abs(g).sum()
    # the flux seed reproduces WallFlux on any field
    T = rng.standard_normal(m.num_tot)
    seed = wall_flux_seed(m, 0, 0, m.num_tot)
    assert abs(seed @ T - hot_wall_nusselt(m, T, bnd=0)) < 1e-12 * np.abs(seed).sum()
