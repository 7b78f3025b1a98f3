"""Pins of the statistics oracle (oracle/stats_oracle.py) by closed forms."""
import numpy as np
import pytest

from dimspers_amd.mesh import build_annulus
from oracle import stats_oracle as so


@pytest.fixture(scope="module")
def mesh():
    return build_annulus(8, 64)


def test_area_of_the_inscribed_polygonal_annulus(mesh):
    """Straight triangles between the vertex rings: area = (1/2) N_theta sin(2 pi / N_theta)
    (r_out^2 - r_in^2), whatever the radial spacing."""
    got = so.volume_integral_cartesian(mesh, np.ones(mesh.theta.size))
    nt = mesh.ncell_t
    want = 0.5 * nt * np.sin(2 * np.pi / nt) * (mesh.r.max() ** 2 - mesh.r.min() ** 2)
    assert abs(got - want) < 1e-13 * want


def test_rms_velocity_of_a_uniform_field(mesh):
    """U = (ux, uz) constant in Cartesian node values: |U|^2 is constant, so Urms = |U| times the
    square root of the ratio of the inscribed polygonal area to the exact annulus area; both
    components count."""
    nt = mesh.ncell_t
    r_in, r_out = mesh.r.min(), mesh.r.max()
    ratio = 0.5 * nt * np.sin(2 * np.pi / nt) * (r_out ** 2 - r_in ** 2) / (np.pi * (r_out ** 2 - r_in ** 2))
    T = np.zeros(mesh.theta.size)
    for ux, uz in ((0.0, 2.0), (3.0, 0.0), (1.5, -0.5)):
        U = np.tile([ux, uz], mesh.theta.size)
        urms = so.write_stats(mesh, U, T)[0]
        want = np.hypot(ux, uz) * np.sqrt(ratio)
        assert abs(urms - want) < 1e-13 * want


def test_p2_indicator_integrals(mesh):
    """int N_vertex = 0 and int N_midpoint = area / 3 on a straight triangle (the 6-point
    rule is exact for quadratics)."""
    x = mesh.r * np.sin(mesh.theta)
    z = mesh.r * np.cos(mesh.theta)
    e2n = mesh.e2n
    tri_area = 0.5 * ((x[e2n[1]] - x[e2n[0]]) * (z[e2n[2]] - z[e2n[0]])
                      - (z[e2n[1]] - z[e2n[0]]) * (x[e2n[2]] - x[e2n[0]]))
    V = np.zeros(mesh.theta.size)
    v = int(e2n[0, 5])
    V[v] = 1.0
    assert abs(so.volume_integral_cartesian(mesh, V)) < 1e-15
    V[:] = 0.0
    m = int(e2n[4, 7])
    V[m] = 1.0
    owners = np.flatnonzero(np.any(e2n[3:] == m, axis=0))
    assert owners.size == 2
    want = tri_area[owners].sum() / 3.0
    assert abs(so.volume_integral_cartesian(mesh, V) - want) < 1e-14


def test_nusselt_of_a_linear_profile(mesh):
    """T = (r_out - r) / (r_out - r_in): each node one level inside the walls contributes
    dr / (r_out - r_in) / dr = 1 / (r_out - r_in), so Nu = -A dth n / (r_out - r_in) r_wall
    (top) and Nu = -A dth n / (r_out - r_in) r_wall (bottom: T - 1 = -(r - r_in)/(...))."""
    r_in, r_out = mesh.r.min(), mesh.r.max()
    T = (r_out - mesh.r) / (r_out - r_in)
    U = np.zeros(2 * mesh.theta.size)
    urms, tav, nu_top, nu_bot = so.write_stats(mesh, U, T)
    A, dr, dth, top, bot = so.scratch_nusselt(mesh)
    # P2 rows: 2 ncell_r + 1 radius levels, 2 ncell_t nodes on each midpoint row
    assert top.size == 2 * mesh.ncell_t and bot.size == 2 * mesh.ncell_t
    assert abs(dr - (r_out - r_in) / (2 * mesh.ncell_r)) < 1e-14
    g = 1.0 / (r_out - r_in)
    assert abs(nu_top - A * dth * (-top.size * g) * r_out) < 1e-12 * abs(nu_top)
    assert abs(nu_bot - A * dth * (-bot.size * g) * r_in) < 1e-12 * abs(nu_bot)
    assert urms == 0.0
    # the average of T over the annulus with the linear profile, by the exact polar integral,
    # differs from the oracle only by the polygon and P2 errors
    exact = 2 * np.pi * (r_out * (r_out ** 2 - r_in ** 2) / 2 - (r_out ** 3 - r_in ** 3) / 3) / (r_out - r_in)
    exact /= np.pi * (r_out ** 2 - r_in ** 2)
    assert abs(tav - exact) < 5e-3
