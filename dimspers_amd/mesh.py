"""Annulus mesh of quadratic (P2) triangles, two per (r, theta) cell.

Host-side setup, run once per mesh.  Restates the connectivity of
``Grid_split1`` (src/Grid/mesher.jl:152-233): radial-fastest vertex numbering
(:168-172), each quad [v(i,j) v(i,j+1) v(i+1,j+1) v(i+1,j)] split into local
triangles [1,2,4] and [2,3,4] (:175-200), then edge midpoints appended in
first-seen order by ``add_midpoints``/``edge_connectivity`` (:349-417).  The
reference constructor itself cannot run (it calls ``Grid`` with 14 of its 17
fields), so only its connectivity rules are followed here.

Also restates ``fixangles!`` (src/Utilities.jl:45-61), ``point_ids``
(src/Grid/mesher.jl:419-441), ``free_slip``/``velocity_bcs``
(src/Setup/boundary_conditions.jl:15-37) and the node angles used by
``rotation_matrix`` (src/Stokes/StokesSolver.jl:203-245).

Index arrays returned here are 0-based int64; ``e2n_1based`` gives the
reference's 1-based column-major convention.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

R_IN = 1.22
R_OUT = 2.22


def linrange(a: float, b: float, n: int) -> np.ndarray:
    """Julia ``LinRange(a, b, n)`` element values: (1-t)*a + t*b, t = j/(n-1)."""
    t = np.arange(n, dtype=np.float64) / float(n - 1)
    return (1.0 - t) * a + t * b


def fixangles(th: np.ndarray) -> np.ndarray:
    """Unwrap element vertex angles across theta = 2*pi.

    src/Utilities.jl:45-61: if |t1-t2|+|t2-t3|+|t3-t1| > pi, add 2*pi to every
    vertex angle below pi.  ``th`` is (3, nel); a fixed copy is returned.
    """
    th = np.array(th, dtype=np.float64, copy=True)
    s = np.abs(th[0] - th[1]) + np.abs(th[1] - th[2]) + np.abs(th[2] - th[0])
    wrap = s > np.pi
    low = (th < np.pi) & wrap[None, :]
    th[low] += 2.0 * np.pi
    return th


@dataclass
class AnnulusMesh:
    """P2 triangle mesh of the annulus r in [R_IN, R_OUT] with nr x ntheta cells."""

    ncell_r: int
    ncell_t: int
    theta: np.ndarray          # (nnod,) node angle, midpoints may exceed 2*pi - dtheta
    r: np.ndarray              # (nnod,) node radius
    e2n: np.ndarray            # (6, nel) int64, 0-based, local order v1 v2 v3 m12 m23 m31
    nvert: int
    el_theta: np.ndarray = field(repr=False, default=None)  # (3, nel) vertex angles after fixangles
    el_r: np.ndarray = field(repr=False, default=None)      # (3, nel)
    split: int = 2             # triangles per cell: 2 (Grid_split1) or 4 (Grid)

    @property
    def nel(self) -> int:
        return self.e2n.shape[1]

    @property
    def nnod(self) -> int:
        return self.theta.shape[0]

    @property
    def nU(self) -> int:
        return 2 * self.nnod

    @property
    def nP(self) -> int:
        return self.nvert

    @property
    def e2nP(self) -> np.ndarray:
        return self.e2n[:3]

    def e2n_1based(self) -> np.ndarray:
        return np.asfortranarray(self.e2n + 1)

    def boundary_nodes(self):
        """``point_ids`` (mesher.jl:419-441): exact equality with the radius extrema."""
        rmin, rmax = self.r.min(), self.r.max()
        top = np.flatnonzero(self.r == rmax)
        bot = np.flatnonzero(self.r == rmin)
        return top, bot

    def free_slip(self):
        """``free_slip`` + ``velocity_bcs`` (boundary_conditions.jl:15-37), 0-based.

        Fixed dofs are the radial dof (second of the pair) of every boundary
        node, outer ring first then inner ring; ``ifree`` is the sorted
        complement.  Returns (fixed_dofs, ifree).
        """
        top, bot = self.boundary_nodes()
        nodes = np.concatenate([top, bot])
        return self._with_free(2 * nodes + 1)

    def free_slip_pinned(self, nr: int):
        """``free_slip(nr, ids)`` (boundary_conditions.jl:40-53): free slip plus the tangential
        dof of node ``nr`` (1-based, so dof 2nr - 1), appended last."""
        fixed, _ = self.free_slip()
        return self._with_free(np.concatenate([fixed, [2 * nr - 2]]))

    def no_slip(self):
        """``no_slip`` (boundary_conditions.jl:55-66): both dofs of every boundary node, the
        radial dofs of all of them first, then the tangential ones (nodes in id order)."""
        top, bot = self.boundary_nodes()
        nodes = np.sort(np.concatenate([top, bot]))
        return self._with_free(np.concatenate([2 * nodes + 1, 2 * nodes]))

    def _with_free(self, fixed):
        fixed = np.asarray(fixed, dtype=np.int64)
        mask = np.ones(self.nU, dtype=bool)
        mask[fixed] = False
        return fixed, np.flatnonzero(mask).astype(np.int64)


def build_annulus(ncell_r: int, ncell_t: int) -> AnnulusMesh:
    """2-triangle-per-cell P2 annulus mesh (Grid_split1 connectivity)."""
    nr = ncell_r + 1
    nt = ncell_t
    nvert = nr * nt
    dth = 2.0 * np.pi / nt
    # vertex v(i, j) = i + nr*j, radial fastest (mesher.jl:168-172)
    rv = np.tile(linrange(R_IN, R_OUT, nr), nt)
    tv = np.repeat(dth * np.arange(nt, dtype=np.float64), nr)

    i = np.tile(np.arange(ncell_r), nt)
    j = np.repeat(np.arange(nt), ncell_r)
    jp = (j + 1) % nt
    q1 = i + nr * j          # v(i, j)
    q2 = i + nr * jp         # v(i, j+1)
    q3 = q2 + 1              # v(i+1, j+1)
    q4 = q1 + 1              # v(i+1, j)
    nquad = ncell_r * nt
    tri = np.empty((3, 2 * nquad), dtype=np.int64)
    tri[:, 0::2] = np.stack([q1, q2, q4])   # local [1,2,4]
    tri[:, 1::2] = np.stack([q2, q3, q4])   # local [2,3,4]
    nel = tri.shape[1]

    return _with_midpoints(ncell_r, ncell_t, tri, tv, rv, nvert)


def build_grid(ncell_r: int, ncell_t: int) -> AnnulusMesh:
    """4-triangle-per-cell P2 annulus mesh: the connectivity of ``Grid`` (src/Grid/mesher.jl:59-150),
    the mesh ``init_grid(N; split=2)`` builds and MAIN.jl:33 runs on.

    Corner vertices as in ``build_annulus``; each quad [v(i,j) v(i,j+1) v(i+1,j+1) v(i+1,j)] gets a
    centre vertex nr*nt + q (q = i + ncell_r*j, radial fastest; angle + dtheta/2, radius + dr/2 with
    dr the first radial spacing, :101-118) and the four triangles [corner k, corner k+1, centre],
    k = 1..4 (:108-115); edge midpoints follow in first-seen order."""
    nr = ncell_r + 1
    nt = ncell_t
    ncorner = nr * nt
    dth = 2.0 * np.pi / nt
    rc = linrange(R_IN, R_OUT, nr)
    rv_c = np.tile(rc, nt)
    tv_c = np.repeat(dth * np.arange(nt, dtype=np.float64), nr)
    i = np.tile(np.arange(ncell_r), nt)
    j = np.repeat(np.arange(nt), ncell_r)
    jp = (j + 1) % nt
    q1 = i + nr * j
    q2 = i + nr * jp
    corners = np.stack([q1, q2, q2 + 1, q1 + 1])   # (4, ncell)
    ncell = ncell_r * nt
    centre = ncorner + np.arange(ncell)
    tri = np.empty((3, 4 * ncell), dtype=np.int64)
    for k in range(4):
        tri[0, k::4] = corners[k]
        tri[1, k::4] = corners[(k + 1) % 4]
        tri[2, k::4] = centre
    dr = rc[1] - rc[0]
    tv = np.concatenate([tv_c, tv_c[q1] + dth / 2.0])
    rv = np.concatenate([rv_c, rv_c[q1] + dr / 2.0])
    mesh = _with_midpoints(ncell_r, ncell_t, tri, tv, rv, ncorner + ncell)
    mesh.split = 4
    return mesh


def _with_midpoints(ncell_r, ncell_t, tri, tv, rv, nvert) -> AnnulusMesh:
    """Append the edge midpoints of the P1 triangles ``tri`` (vertex angles tv, radii rv)."""
    nel = tri.shape[1]
    # edges (1,2), (2,3), (3,1) of each element, numbered in first-seen order
    a = np.stack([tri[0], tri[1], tri[2]])        # (3, nel) edge start
    b = np.stack([tri[1], tri[2], tri[0]])        # (3, nel) edge end
    lo = np.minimum(a, b).T.reshape(-1)           # element-major, edge-minor scan order
    hi = np.maximum(a, b).T.reshape(-1)
    key = lo * nvert + hi
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    order = np.argsort(first, kind="stable")      # unique keys by first occurrence
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    edge_id = rank[inv].reshape(nel, 3).T         # (3, nel)
    nedge = uniq.size
    e_lo = (uniq // nvert)[order]
    e_hi = (uniq % nvert)[order]

    # midpoint coordinates (mesher.jl:394-417), with the 2*pi wrap rule
    ta, tb = tv[e_lo].copy(), tv[e_hi].copy()
    wrap = np.abs(ta - tb) > np.pi
    ta = np.where(wrap & (ta < np.pi), ta + 2.0 * np.pi, ta)
    tb = np.where(wrap & (tb < np.pi), tb + 2.0 * np.pi, tb)
    tmid = 0.5 * (ta + tb)
    rmid = 0.5 * (rv[e_lo] + rv[e_hi])

    e2n = np.concatenate([tri, edge_id + nvert], axis=0)
    theta = np.concatenate([tv, tmid])
    r = np.concatenate([rv, rmid])

    el_theta = fixangles(tv[tri])
    el_r = rv[tri]
    return AnnulusMesh(ncell_r, ncell_t, theta, r, e2n, nvert, el_theta, el_r)


def structured_ids(mesh: AnnulusMesh):
    """(i_r, i_theta, triangle type) of every element of ``build_annulus`` / ``build_grid``."""
    e = np.arange(mesh.nel)
    q = e // mesh.split
    return q % mesh.ncell_r, q // mesh.ncell_r, e % mesh.split
