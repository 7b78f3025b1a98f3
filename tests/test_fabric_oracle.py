"""Pins of the fabric oracle (oracle/fabric_oracle.py): closed forms, LAPACK, brute force and the
reference's own DEM tables (tests/golden/dem_tables.npz, written by make_dem_tables.py)."""
from pathlib import Path

import numpy as np
import pytest

from oracle import fabric_oracle as fo

GOLDEN = Path(__file__).parent / "golden" / "dem_tables.npz"


def _rot(a):
    c, s = np.cos(a), np.sin(a)
    return np.stack([np.stack([c, -s], -1), np.stack([s, c], -1)], -2)


def _stretch(alpha, beta, s1, s2):
    """F = R(alpha) diag(s1, s2) R(beta)^T: left stretch axes R(alpha), semi-axes s1, s2."""
    D = np.zeros(alpha.shape + (2, 2))
    D[..., 0, 0], D[..., 1, 1] = s1, s2
    return _rot(alpha) @ D @ np.swapaxes(_rot(beta), -1, -2)


def _tables():
    d = np.load(GOLDEN)
    return {k[:-1]: (d[k], d[k[:-1] + "ax"], float(d[k[:-1] + "vol"])) for k in d.files if k.endswith("_C")}


def test_ellipse_of_a_rotated_stretch_is_its_closed_form():
    """F F^T = R(alpha) diag(s1^2, s2^2) R(alpha)^T: a1 = s1, a2 = s2, major axis (cos, sin) alpha,
    minor axis perpendicular; the right-stretch angle beta must not show (F^T F would).  The
    closed-form eigenvectors lose digits to cancellation where an axis is near vertical
    (value - a22 small), hence 1e-10 on the directions."""
    rng = np.random.default_rng(1)
    n = 4000
    alpha, beta = rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)
    s2 = np.exp(rng.uniform(-2, 2, n))
    s1 = s2 * np.exp(rng.uniform(np.log(1.5), np.log(30), n))
    fse, F = fo.get_fse(_stretch(alpha, beta, s1, s2), fo.NO_HEALING)
    x1, x2, y1, y2, a1, a2 = fse.T
    assert np.allclose(a1, s1, rtol=1e-13, atol=0)
    assert np.allclose(a2, s2, rtol=4e-16 * 30 ** 2 * 8, atol=0)
    assert np.abs(x1 * np.sin(alpha) - y1 * np.cos(alpha)).max() < 1e-10      # parallel
    assert np.abs(x1 * x1 + y1 * y1 - 1).max() < 1e-14
    assert np.abs(x2 * np.cos(alpha) + y2 * np.sin(alpha)).max() < 1e-10      # perpendicular
    assert np.abs(x1 * x2 + y1 * y2).max() < 1e-10


def test_ellipse_matches_lapack_eigh():
    rng = np.random.default_rng(2)
    F = rng.normal(size=(3000, 2, 2))
    fse, _ = fo.get_fse(F, fo.NO_HEALING)
    w, V = np.linalg.eigh(F @ np.swapaxes(F, 1, 2))               # ascending
    sep = (w[:, 1] - w[:, 0]) / w[:, 1]
    ok = sep > 1e-3
    assert np.allclose(fse[:, 4], np.sqrt(w[:, 1]), rtol=1e-13)
    assert np.allclose(fse[ok, 5], np.sqrt(w[ok, 0]), rtol=1e-13 / sep[ok].min() ** 2)
    cross = fse[:, 0] * V[:, 1, 1] - fse[:, 2] * V[:, 0, 1]       # major axis parallel to LAPACK's
    assert np.abs(cross[ok]).max() < 1e-10


def test_diagonal_branch():
    F = np.array([np.diag([2.0, 1.0]), np.diag([1.0, 3.0]), np.eye(2)])
    fse, _ = fo.get_fse(F, fo.NO_HEALING)
    assert np.array_equal(fse[0], [1, 0, 0, 1, 2, 1])       # major axis x
    assert np.array_equal(fse[1], [0, 1, 1, 0, 3, 1])       # major axis z
    # equal eigenvalues: StaticArrays' diagonal branch takes the swapped unit vectors
    assert np.array_equal(fse[2], [0, 1, 1, 0, 1, 1])


def test_fabric_destruction_rule():
    """healing_FSE! (FSE.jl:198-231): a1/a2 >= eps resets F and the ellipse; below it F stays."""
    F = _stretch(np.array([0.3, 0.3]), np.array([1.1, 1.1]), np.array([2e3, 5e2]) ** 0.5,
                 np.array([2e3, 5e2]) ** -0.5)
    fse, F1 = fo.get_fse(F, fo.FABRIC_DESTRUCTION, eps=1e3)
    assert np.array_equal(fse[0], fo.ISOTROPIC_FSE) and np.array_equal(F1[0], np.eye(2))
    assert np.isclose(fse[1, 4] / fse[1, 5], 5e2) and np.array_equal(F1[1], F[1])
    fse0, F0 = fo.get_fse(F, fo.NO_HEALING)
    assert np.isclose(fse0[0, 4] / fse0[0, 5], 2e3) and np.array_equal(F0, F)


def test_fabric_destruction_at_exactly_the_ratio():
    """The aggregate stays anisotropic only while a1/a2 < eps (FSE.jl:207, strict): F = diag(1000, 1)
    has a1/a2 = 1000 exactly and is reset at eps = 1e3, and kept at eps = 1000.5."""
    F = np.array([[[1000.0, 0.0], [0.0, 1.0]]])
    fse, F1 = fo.get_fse(F, fo.FABRIC_DESTRUCTION, eps=1e3)
    assert np.array_equal(fse[0], fo.ISOTROPIC_FSE) and np.array_equal(F1[0], np.eye(2))
    fse, F1 = fo.get_fse(F, fo.FABRIC_DESTRUCTION, eps=1000.5)
    assert fse[0, 4] == 1000.0 and fse[0, 5] == 1.0 and np.array_equal(F1, F)


def test_annealing_destruction_rule():
    """_FSE(F, destruction, annealing) (FSE.jl:309-355): force_isotropy turns a1 <= 1 or a2 >= 1
    into (1, 0), which the ratio test then destroys; other points keep their axes."""
    F = np.array([np.diag([0.9, 0.5]), np.diag([3.0, 1.2]), np.diag([2.0, 0.5]), np.eye(2)])
    fse, F1 = fo.get_fse(F, fo.ANNEALING_DESTRUCTION, eps=1e3)
    for i in (0, 1, 3):
        assert np.array_equal(fse[i], fo.ISOTROPIC_FSE) and np.array_equal(F1[i], np.eye(2))
    assert np.array_equal(fse[2], [1, 0, 0, 1, 2.0, 0.5]) and np.array_equal(F1[2], F[2])


def test_normalisation_of_large_F():
    """normalize_F (per point, by |max F_ij|) and normalize_F! (all points, by max |F11|)."""
    base = _stretch(np.array([0.7, 0.2]), np.array([0.1, 0.5]), np.array([2.0, 3.0]), np.array([1.0, 0.5]))
    F = base * np.array([1e40, 1.0])[:, None, None]
    fse, F1 = fo.get_fse(F, fo.FABRIC_DESTRUCTION)
    m = np.abs(F[0].max())
    assert np.isclose(fse[0, 4], 2e40 / m, rtol=1e-14) and np.array_equal(F1, F)
    assert np.isclose(fse[1, 4], 3.0, rtol=1e-14)
    fse2, F2 = fo.get_fse(F, fo.ANNEALING_DESTRUCTION)
    scale = 1.0 / np.abs(F[:, 0, 0]).max()
    assert np.allclose(F2[0], F[0] * scale, rtol=1e-15, atol=0)
    assert np.isclose(fse2[0, 4], 2e40 * scale, rtol=1e-14)
    # the second point, scaled by ~1e-40, has a1 <= 1: destroyed
    assert np.array_equal(fse2[1], fo.ISOTROPIC_FSE) and np.array_equal(F2[1], np.eye(2))


@pytest.mark.parametrize("name", ["DEM_1em3_vol10_", "DEM_1em3_vol20_", "DEM_1em3_vol30_"])
def test_dem_preparation_on_the_reference_tables(name):
    T, ax, vol = _tables()[name]
    assert T.shape == (100, 9) and ax.shape == (100, 3)
    dem = fo.get_dem(T, ax)
    assert dem["sblk"] == 10
    assert np.array_equal(dem["perm"], np.arange(9, -1, -1))     # ax3 rises, a2a3 falls in a block
    assert np.allclose(dem["a1a2_blk"], np.log10(ax[::10, 0]), rtol=0, atol=0)
    assert np.all(np.diff(dem["a1a2_blk"]) > 0) and np.all(np.diff(dem["a2a3_blk"]) > 0)
    assert np.allclose(dem["a2a3_blk"], np.log10(1.0 / ax[:10, 2])[::-1])


def _brute_nearest(dem, ax, r1, r2):
    """Row index of the nearest raw table entry: nearest a1a2 block value, then nearest a2a3 in
    that block; a tie goes to the larger table value."""
    a1a2 = np.log10(ax[:, 0])
    a2a3 = np.log10(1.0 / ax[:, 2])
    rows = []
    for p, q in zip(r1, r2):
        d1 = np.abs(p - a1a2)
        cand = np.flatnonzero(d1 == d1.min())
        cand = cand[a1a2[cand] == a1a2[cand].max()]
        d2 = np.abs(q - a2a3[cand])
        best = cand[d2 == d2.min()]
        rows.append(best[np.argmax(a2a3[best])])
    return np.array(rows)


def test_lookup_matches_brute_force_nearest_entry():
    T, ax, _ = _tables()["DEM_1em3_vol30_"]
    dem = fo.get_dem(T, ax)
    rng = np.random.default_rng(3)
    r1, r2 = rng.uniform(-0.3, 2.0, 2000), rng.uniform(-0.3, 2.3, 2000)
    assert np.array_equal(fo.dem_index(dem, r1, r2), _brute_nearest(dem, ax, r1, r2))


def test_lookup_tie_goes_to_the_later_entry():
    """argminsorted (anisotropy.jl:114-122) moves past equal distances."""
    ax = np.array([[a1, 1.0, a3] for a1 in (1.0, 10.0, 100.0) for a3 in (1.0, 0.1, 0.01)])
    T = np.arange(27 * 9, dtype=float).reshape(27, 9)
    dem = fo.get_dem(T, ax)
    assert np.array_equal(dem["a1a2_blk"], [0.0, 1.0, 2.0]) and np.array_equal(dem["a2a3_blk"], [0.0, 1.0, 2.0])
    # r1 = 0.5 ties blocks 0 and 1 -> block 1; r2 = 1.5 ties a2a3 = 1 and 2 -> a2a3 = 2 (raw entry 2)
    assert fo.dem_index(dem, 0.5, 1.5)[0] == 3 * 1 + 2
    assert np.array_equal(fo.dem_index(dem, 0.5, 1.5), _brute_nearest(dem, ax, [0.5], [1.5]))


def _fourth_order(a, c, e, h, cs, sn):
    """The (x, z) block of the orthotropic tensor rotated by -phi, (cos, sin) phi = (cs, sn)."""
    Cf = np.zeros((2, 2, 2, 2))
    Cf[0, 0, 0, 0], Cf[1, 1, 1, 1] = a, c
    Cf[0, 0, 1, 1] = Cf[1, 1, 0, 0] = h
    for idx in [(0, 1, 0, 1), (0, 1, 1, 0), (1, 0, 0, 1), (1, 0, 1, 0)]:
        Cf[idx] = e
    R = np.array([[cs, sn], [-sn, cs]])
    Cr = np.einsum("ia,jb,kc,ld,abcd->ijkl", R, R, R, R, Cf)
    return [Cr[0, 0, 0, 0], Cr[1, 1, 1, 1], Cr[0, 1, 0, 1], Cr[0, 0, 1, 1], Cr[0, 0, 0, 1], Cr[1, 1, 0, 1]]


def test_tensor_is_the_rotated_table_entry():
    T, ax, _ = _tables()["DEM_1em3_vol30_"]
    dem = fo.get_dem(T, ax)
    rng = np.random.default_rng(4)
    nel = 40
    F = rng.normal(size=(6 * nel, 2, 2))
    fse, _ = fo.get_fse(F, fo.NO_HEALING)
    r1, r2 = rng.uniform(0, 1.7, 6 * nel), rng.uniform(0, 2, 6 * nel)
    w = 0.03
    out = fo.anisotropic_tensor(fse, nel, dem, w, r1, r2)
    rows = _brute_nearest(dem, ax, r1, r2)
    clamped = 0
    for p in range(6 * nel):
        e, q = p % nel, p // nel
        t = T[rows[p]]
        nx = np.hypot(fse[p, 0], fse[p, 2])
        want = _fourth_order(t[0], t[2], max(t[4], w), t[7], fse[p, 0] / nx, fse[p, 2] / nx)
        clamped += t[4] < w
        want[2] = max(want[2], w)
        got = [o[e, q] for o in out]
        assert np.allclose(got, want, rtol=0, atol=1e-14), p
    assert clamped > 0
    for o in out:
        assert np.allclose(o[:, 6], o[:, :6].mean(axis=1), rtol=1e-15, atol=0)
    assert np.all(out[2] >= w)


def test_rotated_shear_term_is_clamped_to_the_cut_off():
    """get_tensor_and_rotate! clamps eta55 after the rotation too (:283): at 45 degrees the
    rotated eta55 of a = c = h is zero, so it comes back as w."""
    ax = np.array([[a1, 1.0, a3] for a1 in (1.0, 10.0) for a3 in (1.0, 0.1)])
    T = np.zeros((4, 9))
    T[:, [0, 2, 7]] = 0.1
    T[:, 4] = 1.0
    dem = fo.get_dem(T, ax)
    fse = np.tile([np.sqrt(0.5), 0.0, np.sqrt(0.5), 0.0, 2.0, 1.0], (6, 1))
    out = fo.anisotropic_tensor(fse, 1, dem, 0.5, np.zeros(6), np.zeros(6))
    assert np.allclose(out[2][0, :6], 0.5, rtol=0, atol=0)
    want = _fourth_order(0.1, 0.1, 1.0, 0.1, np.sqrt(0.5), np.sqrt(0.5))
    assert abs(want[2]) < 1e-15
    assert np.allclose([o[0, 0] for o in (out[0], out[1], out[3], out[4], out[5])],
                       [want[0], want[1], want[3], want[4], want[5]], atol=1e-15)


def test_shape_table_interpolation():
    nodes = np.array([1.0, 2.0, 5.0, 50.0])
    r1, r2 = np.array([0.0, 0.4, 0.9, 1.7]), np.array([2.0, 1.5, 0.5, 0.0])
    a1 = np.concatenate([nodes, np.random.default_rng(5).uniform(0.5, 60, 500)])
    g1, g2 = fo.shape_from_table(a1, nodes, r1, r2)
    assert np.allclose(g1, np.interp(a1, nodes, r1), rtol=0, atol=1e-15)
    assert np.allclose(g2, np.interp(a1, nodes, r2), rtol=0, atol=1e-15)
    assert np.array_equal(g1[:4], r1) and np.array_equal(g2[:4], r2)
