"""Objectives evaluated on solved fields (host side, over the mesh's dual).

hot_wall_nusselt: the wall Nusselt number of a differentially heated cavity (the objective
of the natural-convection workload): the conductive heat flux out of the control volumes
of a wall's nodes through their faces, sum over the wall nodes v and their neighbours n of
(A_vn / d_vn)(T_v - T_n), divided by k dT / L times the wall area (the caller's scale).
Faces between two wall nodes cancel in the sum.  For the conduction profile T = 1 - x and
the x = 0 wall of the unit cube it is 1 up to round-off: the faces of a wall node's control
volume close with its boundary face, so sum_n (A / d)(x_n - x_v) = sum_n A n_x = its
boundary area, and the boundary areas of the wall sum to 1.
"""
from __future__ import annotations

import numpy as np


class WallFlux:
    """Precomputed face list of one boundary's nodes; __call__(T) gives the flux sum."""

    def __init__(self, mesh, bnd: int, scale: float = 1.0):
        rp, ngh = mesh.get("row_ptr"), mesh.get("ngh")
        wall = mesh.num_pts + mesh.bnd_nodes(bnd).astype(np.int64)
        deg = rp[wall + 1] - rp[wall]
        eidx = np.concatenate([np.arange(rp[v], rp[v + 1]) for v in wall]) if len(wall) else np.zeros(0, np.int64)
        self.src = np.repeat(wall, deg)
        self.dst = ngh[eidx]
        self.coef = mesh.get("area")[eidx] / mesh.get("dist")[eidx]
        self.scale = scale

    def __call__(self, T) -> float:
        T = np.asarray(T, np.float64)
        return float(np.sum(self.coef * (T[self.src] - T[self.dst]))) / self.scale


def hot_wall_nusselt(mesh, T, bnd: int = 0, scale: float = 1.0) -> float:
    """Nusselt number of boundary `bnd` for the temperature T (vertex order)."""
    return WallFlux(mesh, bnd, scale)(T)


class WallProfile:
    """A perturbation profile of a wall's Dirichlet values on a tensor grid: value at wall
    node v = base + sum_ij theta_ij phi_ij(s_v, t_v), phi_ij the bilinear hat functions of a
    shape[0] x shape[1] grid spanning the wall's two in-plane coordinates (axes, default the
    two other than the wall normal's; grid lines at i / (shape[0] - 1)).  values() maps the
    profile to the nodes' values, grad() maps dJ/d(node values) to dJ/dtheta (the transpose)."""

    def __init__(self, mesh, bnd: int, shape=(32, 32), axes=None, lo=(0.0, 0.0), hi=(1.0, 1.0)):
        nodes = mesh.bnd_nodes(bnd).astype(np.int64)
        pos = mesh.get("pos")[mesh.num_pts + nodes]
        if axes is None:
            normal = int(np.argmin(np.ptp(pos, axis=0)))
            axes = tuple(a for a in range(3) if a != normal)
        self.shape = tuple(int(s) for s in shape)
        ns, nt = self.shape
        s = (pos[:, axes[0]] - lo[0]) / (hi[0] - lo[0]) * (ns - 1)
        t = (pos[:, axes[1]] - lo[1]) / (hi[1] - lo[1]) * (nt - 1)
        i0 = np.clip(np.floor(s).astype(np.int64), 0, ns - 2)
        j0 = np.clip(np.floor(t).astype(np.int64), 0, nt - 2)
        ts, tt = np.clip(s - i0, 0.0, 1.0), np.clip(t - j0, 0.0, 1.0)
        self.cols = np.stack([i0 * nt + j0, (i0 + 1) * nt + j0, i0 * nt + j0 + 1, (i0 + 1) * nt + j0 + 1], axis=1)
        self.w = np.stack([(1 - ts) * (1 - tt), ts * (1 - tt), (1 - ts) * tt, ts * tt], axis=1)
        self.count = nodes.size

    def values(self, theta, base=0.0):
        th = np.asarray(theta, np.float64).reshape(-1)
        return base + np.sum(self.w * th[self.cols], axis=1)

    def grad(self, g_nodes):
        out = np.zeros(self.shape[0] * self.shape[1])
        np.add.at(out, self.cols.ravel(), (self.w * np.asarray(g_nodes, np.float64)[:, None]).ravel())
        return out.reshape(self.shape)


def wall_flux_seed(mesh, bnd: int, offset: int, n: int, scale: float = 1.0):
    """d(WallFlux)/dT as a vector of n entries with the temperature block at `offset`
    (the seed of a linear objective in the flow adjoint)."""
    wf = WallFlux(mesh, bnd, scale)
    seed = np.zeros(n)
    np.add.at(seed, offset + wf.src, wf.coef / scale)
    np.add.at(seed, offset + wf.dst, -wf.coef / scale)
    return seed
