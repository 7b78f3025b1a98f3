"""FP64 CPU oracle of the fabric path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module; the product path never calls it.

Restates what the reference tree holds of the chain deformation gradient -> finite strain
ellipse -> viscous tensor (MAIN.jl:184-185, 266-269):

* ``getFSE`` for the three healing types that run (src/Stress/FSE.jl): ``NoHealing``
  (:131-161, ``_FSE``), ``FabricDestructionOnly`` (:191-231, ``healing_FSE!``; MAIN.jl:114-120
  builds this one, eps = 1e3) and ``AnnealingFabricDestruction`` (:297-355, after the global
  ``normalize_F!`` of :163-178).  ``AnnealingOnly`` (:233-241) reads undefined names and cannot
  run, so it has no restatement.  ``FiniteStrainEllipsoid`` (:9-16) is the 6-tuple
  (x1, x2, y1, y2, a1, a2) per point.
* the symmetric 2x2 eigen solve those call, ``eigen`` of StaticArrays 1.4.3
  (Manifest.toml:624-628; not in the reference tree), restated from its published Hermitian
  2x2 method in ``sym_eigen2``, and ``eigval_order`` (FSE.jl:189).
* ``getDEM`` (src/Rheology/anisotropy.jl:60-93) without the HDF5 read: ``get_stride`` (:95-101),
  ``sort_axes`` (:103-112); the cut-off w = ``weakening(dη, φ)`` is an input.
* ``anisotropic_tensor`` (:140-222) with ``get_tensor_and_rotate!`` (:244-285): nearest DEM entry
  by ``argminsorted`` (:114-122), table tensor with its shear terms clamped to w, ``rotmatrix``
  (:287-292), ``directRotation2D`` (:503-536), eta55 clamped to w, 7th column = mean of the first
  six.  ``fabric_shape(a1, 1, parameterization)`` of Fabrics.jl 0.1.0 (Manifest.toml:160-165, an
  external package) is an input: per point, or sampled at increasing a1 nodes and interpolated
  linearly between them (the form the C ABI accepts).

Pinned (tests/test_fabric_oracle.py): ellipses of F = R(alpha) diag(s1, s2) R(beta)^T against
their closed form and against LAPACK's ``eigh``; the healing rules on constructed cases; the
table lookup against a brute-force nearest search on the raw table; the rotation against the
rotation of the fourth-order tensor; the DEM tables of the reference (tests/golden) through their
block structure.  The eigenvector direction of an exactly isotropic F (equal eigenvalues: every
direction is one) follows StaticArrays' diagonal branch, which the mathematics does not fix and no
Julia run checked: parity unpinned for that case.
"""
from __future__ import annotations

import numpy as np

NO_HEALING, FABRIC_DESTRUCTION, ANNEALING_DESTRUCTION = 0, 1, 2
ISOTROPIC_FSE = (1.0, 0.0, 0.0, 1.0, 1.0, 1.0)


def sym_eigen2(a11, a12, a22):
    """Eigen-decomposition of [[a11, a12], [a12, a22]] (arrays) as StaticArrays computes it for a
    Hermitian 2x2 stored in its upper triangle.

    a12 != 0: t = (a11 + a22) / 2, d = a11 a22 - a12^2, s = sqrt(t^2 - d) (0 if negative),
    values (t - s, t + s), eigenvector k = (value_k - a22, a12) / its norm.
    a12 == 0: a11 < a22 gives values (a11, a22) with the unit vectors in order, else (a22, a11)
    with them swapped.  Returns vals (2, n) and vecs (2, 2, n), column k = eigenvector k.
    """
    a11, a12, a22 = (np.asarray(v, dtype=np.float64) for v in (a11, a12, a22))
    t = (a11 + a22) / 2
    d = a11 * a22 - a12 * a12
    s2 = t * t - d
    s = np.sqrt(np.where(s2 < 0, 0.0, s2))
    v1, v2 = t - s, t + s
    with np.errstate(invalid="ignore", divide="ignore"):
        u1 = v1 - a22
        n1 = np.sqrt(u1 * u1 + a12 * a12)
        u2 = v2 - a22
        n2 = np.sqrt(u2 * u2 + a12 * a12)
        vecs = np.stack([np.stack([u1 / n1, u2 / n2]), np.stack([a12 / n1, a12 / n2])])
    vals = np.stack([v1, v2])
    diag = a12 == 0
    if np.any(diag):
        lt = a11 < a22
        vals[0] = np.where(diag, np.where(lt, a11, a22), vals[0])
        vals[1] = np.where(diag, np.where(lt, a22, a11), vals[1])
        one_lt = np.where(lt, 1.0, 0.0)
        for i, j, v in ((0, 0, one_lt), (1, 0, 1 - one_lt), (0, 1, 1 - one_lt), (1, 1, one_lt)):
            vecs[i, j] = np.where(diag, v, vecs[i, j])
    return vals, vecs


def normalize_F(F):
    """``normalize_F`` (FSE.jl:180-187): F / |max_ij F_ij| where that exceeds 1e30."""
    fmax = np.abs(F.max(axis=(-2, -1)))
    big = fmax > 1e30
    return np.where(big[:, None, None], F / np.where(big, fmax, 1.0)[:, None, None], F)


def ellipse(Fi):
    """Semi-axes and axes of F F^T (``_eigen`` + ``eigval_order``, FSE.jl:143-161, 189, 357-362)."""
    L11 = Fi[:, 0, 0] * Fi[:, 0, 0] + Fi[:, 0, 1] * Fi[:, 0, 1]
    L12 = Fi[:, 0, 0] * Fi[:, 1, 0] + Fi[:, 0, 1] * Fi[:, 1, 1]
    L22 = Fi[:, 1, 0] * Fi[:, 1, 0] + Fi[:, 1, 1] * Fi[:, 1, 1]
    vals, vecs = sym_eigen2(L11, L12, L22)
    second = vals[1] > vals[0]                       # imax = 2
    imax = np.where(second, 1, 0)
    imin = 1 - imax
    n = np.arange(Fi.shape[0])
    a1 = np.sqrt(np.abs(vals[imax, n]))
    a2 = np.sqrt(np.abs(vals[imin, n]))
    x1, y1 = vecs[0, imax, n], vecs[1, imax, n]
    x2, y2 = vecs[0, imin, n], vecs[1, imin, n]
    return np.stack([x1, x2, y1, y2, a1, a2], axis=1)


def get_fse(F, mode, eps=1e3):
    """``getFSE`` on points F (npt, 2, 2).  Returns (fse (npt, 6), F after the call)."""
    F = np.array(F, dtype=np.float64, copy=True)
    if mode == ANNEALING_DESTRUCTION:
        # normalize_F! (:163-178): one factor for all points, from |F11|
        normalizer = -np.inf
        f11 = np.abs(F[:, 0, 0])
        f11 = f11[~np.isnan(f11)]
        if f11.size:
            normalizer = f11.max()
        if normalizer > 1e30:
            F *= 1.0 / normalizer
    Fi = F if mode == NO_HEALING else normalize_F(F)
    fse = ellipse(Fi)
    if mode == NO_HEALING:
        return fse, F
    a1, a2 = fse[:, 4].copy(), fse[:, 5].copy()
    if mode == ANNEALING_DESTRUCTION:
        # unstretch_axes is overwritten by force_isotropy(a1, a2) (:323-326, 289-295)
        iso = (a1 <= 1) | (a2 >= 1)
        a1 = np.where(iso, 1.0, a1)
        a2 = np.where(iso, 0.0, a2)
        fse[:, 4], fse[:, 5] = a1, a2
    with np.errstate(invalid="ignore", divide="ignore"):
        keep = a1 / a2 < eps
    fse[~keep] = ISOTROPIC_FSE
    F[~keep] = np.eye(2)
    return fse, F


def get_stride(v):
    """``get_stride`` (anisotropy.jl:95-101): length of the leading run of equal values."""
    i = 0
    while v[i] == v[i + 1]:
        i += 1
    return i + 1


def get_dem(tensor, axes):
    """``getDEM`` (anisotropy.jl:60-93) on the arrays ``DEMloader`` reads: tensor n x 9, axes n x 3.
    Returns a dict with a1a2_blk, a2a3_blk, perm (0-based), sblk and the table."""
    tensor = np.asarray(tensor, dtype=np.float64)
    axes = np.asarray(axes, dtype=np.float64)
    a1a2 = np.log10(axes[:, 0] / 1.0)
    a2a3 = np.log10(1.0 / axes[:, 2])
    sblk = get_stride(a1a2)
    perm = np.argsort(a2a3[:sblk], kind="stable")     # sortperm (stable)
    return {"a1a2_blk": a1a2[::sblk].copy(), "a2a3_blk": a2a3[:sblk][perm], "perm": perm, "sblk": sblk,
            "tensor": tensor}


def argminsorted(v):
    """``argminsorted`` (anisotropy.jl:114-122), 0-based: first i with v[i] < v[i+1], else the last."""
    for i in range(len(v) - 1):
        if v[i] < v[i + 1]:
            return i
    return len(v) - 1


def _argminsorted_rows(v):
    """``argminsorted`` of every row of v (npt, m)."""
    up = v[:, :-1] < v[:, 1:]
    return np.where(up.any(axis=1), up.argmax(axis=1), v.shape[1] - 1)


def dem_index(dem, r1, r2):
    """Rows of the DEM table get_tensor_and_rotate! picks for fabric shapes (r1, r2) (:250-264)."""
    r1 = np.atleast_1d(np.asarray(r1, dtype=np.float64))
    r2 = np.atleast_1d(np.asarray(r2, dtype=np.float64))
    i1 = _argminsorted_rows(np.abs(r1[:, None] - dem["a1a2_blk"][None, :]))
    i2 = dem["perm"][_argminsorted_rows(np.abs(r2[:, None] - dem["a2a3_blk"][None, :]))]
    return dem["sblk"] * i1 + i2


def rotmatrix(x, y):
    """``rotmatrix`` (anisotropy.jl:287-292): (cos a, sin a, -sin a) of the direction (x, y)."""
    h = 1.0 / np.sqrt(y * y + x * x)
    return x * h, y * h, -y * h


def direct_rotation_2d(R, a, c, e, h):
    """``directRotation2D`` (anisotropy.jl:503-536) for the fabric-frame C11 = a, C33 = c, C55 = e,
    C13 = h; returns (eta11, eta33, eta55, eta13, eta15, eta35)."""
    x1, x2, y1 = R
    y2 = x1
    x1x1, x2x2, y1y1, y2y2 = x1 * x1, x2 * x2, y1 * y1, y2 * y2
    s = x1 * y2 + x2 * y1
    n11 = (c * x2x2 + h * x1x1 + 4 * e * x1x1) * x2x2 + (a * x1x1 + h * x2x2) * x1x1
    n33 = (c * y2y2 + h * y1y1 + 4 * e * y1y1) * y2y2 + (a * y1y1 + h * y2y2) * y1y1
    n55 = (c * x2 * y2 + h * x1 * y1) * x2 * y2 + s * s * e + (a * x1 * y1 + h * x2 * y2) * x1 * y1
    n13 = ((c * x2x2 + h * x1x1) * y2 + 4 * e * x1 * x2 * y1) * y2 + (a * x1x1 + h * x2x2) * y1y1
    n15 = ((c * x2x2 + h * x1x1) * y2 + 2 * s * e * x1) * x2 + (a * x1x1 + h * x2x2) * x1 * y1
    n35 = ((c * y2y2 + h * y1y1) * x2 + 2 * s * e * y1) * y2 + (a * y1y1 + h * y2y2) * x1 * y1
    return n11, n33, n55, n13, n15, n35


def shape_from_table(a1, nodes, r1, r2):
    """fabric_shape sampled at increasing ``nodes``: linear between nodes, the end values beyond.
    Written as the device evaluates it: k = last node <= a1, t = (a1 - x_k) / (x_k+1 - x_k),
    r = r_k + t (r_k+1 - r_k)."""
    nodes, r1, r2 = (np.asarray(v, dtype=np.float64) for v in (nodes, r1, r2))
    a1 = np.asarray(a1, dtype=np.float64)
    k = np.clip(np.searchsorted(nodes, a1, side="right") - 1, 0, nodes.size - 2)
    t = (a1 - nodes[k]) / (nodes[k + 1] - nodes[k])
    o1 = r1[k] + t * (r1[k + 1] - r1[k])
    o2 = r2[k] + t * (r2[k + 1] - r2[k])
    lo, hi = a1 <= nodes[0], a1 >= nodes[-1]
    o1 = np.where(lo, r1[0], np.where(hi, r1[-1], o1))
    o2 = np.where(lo, r2[0], np.where(hi, r2[-1], o2))
    return o1, o2


def anisotropic_tensor(fse, nel, dem, w, r1, r2):
    """``anisotropic_tensor`` (anisotropy.jl:140-222, 244-285).

    fse: (nel * nip, 6) points column-major (nip = 6 or 7); r1, r2: fabric shape at every point.
    Returns the six nel x 7 arrays (eta11, eta33, eta55, eta13, eta15, eta35).
    """
    fse = np.asarray(fse, dtype=np.float64)
    npt = fse.shape[0]
    T = dem["tensor"]
    im = dem_index(dem, r1, r2)
    a, c, h = T[im, 0], T[im, 2], T[im, 7]
    e = np.maximum(T[im, 4], w)
    n = list(direct_rotation_2d(rotmatrix(fse[:, 0], fse[:, 2]), a, c, e, h))
    n[2] = np.maximum(n[2], w)
    out = np.zeros((6, nel * 7))
    out[:, :npt] = np.stack(n)
    out = out.reshape(6, 7, nel)
    out[:, 6] = (out[:, 0] + out[:, 1] + out[:, 2] + out[:, 3] + out[:, 4] + out[:, 5]) * (1 / 6)
    return [np.asfortranarray(out[k].T) for k in range(6)]
