"""Pins of the thermal oracle (oracle/thermal_oracle.py) by closed forms."""
import numpy as np
import pytest

from dimspers_amd.mesh import build_annulus
from oracle import thermal_oracle as to


@pytest.fixture(scope="module")
def mesh():
    return build_annulus(8, 64)


def test_mass_sums_to_the_area(mesh):
    """rho = Cp = 1: 1^T M_T 1 = int r dr dtheta over the annulus, exact for the 7-point rule."""
    nn = mesh.theta.size
    _, MT, _ = to.thermal_matrices(mesh, np.ones(nn), 3.3, 1.0)
    area = np.pi * (mesh.r.max() ** 2 - mesh.r.min() ** 2)
    assert abs(MT.sum() - area) < 1e-12 * area
    # rho Cp scales it
    _, MT2, _ = to.thermal_matrices(mesh, 2.0 * np.ones(nn), 3.3, 1.5)
    assert abs(MT2.sum() - 3.0 * area) < 1e-12 * area


def test_a_step_without_fixed_nodes_conserves_heat_plus_the_source(mesh):
    """(M + dt K) T = M T0 + f with insulated walls: K 1 = 0, so 1^T M T = 1^T M T0 + sum f, and
    sum f = dQdT dt area (the P2 functions sum to one; the 7-point rule is exact for r)."""
    nn = mesh.theta.size
    rng = np.random.default_rng(4)
    T0 = 1.0 + 0.1 * rng.standard_normal(nn)
    dt, dqdt = 0.01, 0.3
    T = to.thermal_step(mesh, T0, np.ones(nn), 2.0, 1.0, dt, np.array([], dtype=np.int64), np.array([]), dQdT=dqdt)
    _, MT, _ = to.thermal_matrices(mesh, np.ones(nn), 2.0, 1.0)
    area = np.pi * (mesh.r.max() ** 2 - mesh.r.min() ** 2)
    gained = MT.sum(axis=0) @ (T - T0)
    assert abs(gained - dqdt * dt * area) < 1e-11 * dqdt * dt * area
    assert np.abs(T - T0).max() > 1e-4          # the step did diffuse and heat


def test_diffusion_annihilates_constants_and_measures_grad_r(mesh):
    """K_T 1 = 0; for u = r (held exactly by P2, |grad r| = 1 in Cartesian coordinates)
    u^T K_T u = K * area."""
    nn = mesh.theta.size
    K = 3.3
    KT, _, _ = to.thermal_matrices(mesh, np.ones(nn), K, 1.0)
    assert np.abs(KT @ np.ones(nn)).max() < 1e-11 * abs(KT).max()
    u = mesh.r.copy()
    area = np.pi * (mesh.r.max() ** 2 - mesh.r.min() ** 2)
    assert abs(u @ (KT @ u) - K * area) < 1e-11 * K * area
    # and a field with a different gradient does not give that (the pin can fail)
    assert abs((2 * u) @ (KT @ (2 * u)) - K * area) > 1.0


def test_source_term(mesh):
    """dQdT: f = dQdT dt int N, so sum f = dQdT dt area."""
    nn = mesh.theta.size
    _, _, F = to.thermal_matrices(mesh, np.ones(nn), 1.0, 1.0, dQdT=0.7, dt=2.0)
    area = np.pi * (mesh.r.max() ** 2 - mesh.r.min() ** 2)
    assert abs(F.sum() - 0.7 * 2.0 * area) < 1e-12 * area


def test_long_step_gives_steady_conduction():
    """dt -> infinity: the step solves K T = 0 with T = 1 on the inner and 0 on the outer
    wall, T = ln(r / r_out) / ln(r_in / r_out)."""
    m = build_annulus(16, 128)
    nn = m.theta.size
    r_in, r_out = m.r.min(), m.r.max()
    top = np.flatnonzero(m.r == r_out)
    bot = np.flatnonzero(m.r == r_in)
    fixed = np.concatenate([top, bot])
    vfix = np.concatenate([np.zeros(top.size), np.ones(bot.size)])
    T = to.thermal_step(m, np.zeros(nn), np.ones(nn), 1.0, 1.0, 1e12, fixed, vfix)
    exact = np.log(m.r / r_out) / np.log(r_in / r_out)
    err = np.abs(T - exact).max()
    assert err < 1e-5, err
