"""Reference-triangle tables (host side).

Restates ``ip_triangle6``/``ip_triangle7`` (src/Algebra/Quadrature.jl:80-138;
the symmetric 6-point degree-4 and 7-point degree-5 rules, weights including
the factor 1/2) and the P1/P2 shape functions ``sf_N_tri3``/``sf_N_tri6``/
``sf_dN_tri6`` (Quadrature.jl:166-307).  Local P2 node order: vertices 1,2,3,
then midpoints of edges 1-2, 2-3, 3-1.
"""
from __future__ import annotations

import numpy as np


def ip_triangle6():
    g1 = (8.0 - np.sqrt(10.0) + np.sqrt(38.0 - 44.0 * np.sqrt(2.0 / 5.0))) / 18.0
    g2 = (8.0 - np.sqrt(10.0) - np.sqrt(38.0 - 44.0 * np.sqrt(2.0 / 5.0))) / 18.0
    x = np.array([
        [1.0 - 2.0 * g1, g1],
        [g1, 1.0 - 2.0 * g1],
        [g1, g1],
        [1.0 - 2.0 * g2, g2],
        [g2, 1.0 - 2.0 * g2],
        [g2, g2],
    ])
    w1 = (620.0 + np.sqrt(213125.0 - 53320.0 * np.sqrt(10.0))) / 3720.0
    w2 = (620.0 - np.sqrt(213125.0 - 53320.0 * np.sqrt(10.0))) / 3720.0
    w = 0.5 * np.array([w1, w1, w1, w2, w2, w2])
    return x, w


def ip_triangle7():
    g1 = (6.0 - np.sqrt(15.0)) / 21.0
    g2 = (6.0 + np.sqrt(15.0)) / 21.0
    x = np.array([
        [1.0 / 3.0, 1.0 / 3.0],
        [1.0 - 2.0 * g1, g1],
        [g1, 1.0 - 2.0 * g1],
        [g1, g1],
        [1.0 - 2.0 * g2, g2],
        [g2, 1.0 - 2.0 * g2],
        [g2, g2],
    ])
    w1 = (155.0 - np.sqrt(15.0)) / 1200.0
    w2 = (155.0 + np.sqrt(15.0)) / 1200.0
    w = 0.5 * np.array([0.225, w1, w1, w1, w2, w2, w2])
    return x, w


def sf_N_tri3(r, s):
    r = np.asarray(r, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    t = 1.0 - r - s
    return np.stack([t, r, s], axis=-1)


def sf_N_tri6(r, s):
    r = np.asarray(r, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    t = 1.0 - r - s
    return np.stack([t * (2 * t - 1), r * (2 * r - 1), s * (2 * s - 1), 4 * r * t, 4 * r * s, 4 * s * t], axis=-1)


def sf_dN_tri6(r, s):
    """(..., 2, 6): derivatives w.r.t. the two local coordinates."""
    r = np.asarray(r, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    t = 1.0 - r - s
    z = np.zeros_like(r)
    dr = np.stack([-(4 * t - 1), 4 * r - 1, z, 4 * (t - r), 4 * s, -4 * s], axis=-1)
    ds = np.stack([-(4 * t - 1), z, 4 * s - 1, -4 * r, 4 * r, 4 * (t - s)], axis=-1)
    return np.stack([dr, ds], axis=-2)
