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
