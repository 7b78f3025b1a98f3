"""Pins of the stress oracle (oracle/stress_oracle.py) by closed forms and identities."""
import numpy as np
import pytest

from dimspers_amd.mesh import build_annulus
from oracle import stress_oracle as so


@pytest.fixture(scope="module")
def mesh():
    return build_annulus(8, 64)


def _linear_field(mesh, a, b, c, d, e, f):
    """(x, z) per node: U_x = a theta + b r + c, U_z = d theta + e r + f."""
    th, r = mesh.theta, mesh.r
    U = np.empty(2 * th.size)
    U[0::2] = a * th + b * r + c
    U[1::2] = d * th + e * r + f
    return U


def _interior(mesh):
    """Elements whose six nodes lie away from the theta = 0 seam (the field is not periodic)."""
    th = mesh.theta[mesh.e2n]                                    # (6, nel)
    return np.all((th > 0.5) & (th < 2 * np.pi - 0.5), axis=0)


def test_linear_field_gradient_is_exact(mesh):
    """P2 holds a field linear in (theta, r) exactly, so with x = r sin(theta), z = r cos(theta)
    its Cartesian gradient at every point is d/dx = sin d/dr + cos/r d/dtheta and
    d/dz = cos d/dr - sin/r d/dtheta of the closed form (StressCalculation.jl:190-205)."""
    a, b, c, d, e, f = 0.7, -1.3, 0.2, 2.1, 0.4, -0.5
    U = _linear_field(mesh, a, b, c, d, e, f)
    uxx, uzz, uxz, uzx = so.velocity_gradient(mesh, U)
    _, _, thq, rq = so.cartesian_gradients(mesh)
    sn, cs = np.sin(thq), np.cos(thq)
    ok = _interior(mesh)
    assert ok.sum() > 0.7 * ok.size
    for got, want in ((uxx, b * sn + a * cs / rq), (uxz, b * cs - a * sn / rq),
                      (uzx, e * sn + d * cs / rq), (uzz, e * cs - d * sn / rq)):
        assert np.abs(got[ok] - want[ok]).max() < 1e-12 * (1 + np.abs(want).max())
    # the 6- and 7-point rules both reproduce it
    uxx6, *_ = so.velocity_gradient(mesh, U, nip=6)
    _, _, th6, r6 = so.cartesian_gradients(mesh, nip=6)
    want6 = b * np.sin(th6) + a * np.cos(th6) / r6
    assert np.abs(uxx6[ok] - want6[ok]).max() < 1e-12


def test_strain_rate_is_the_symmetric_part_of_the_gradient(mesh):
    """eps = (grad u + grad u^T) / 2 and eII = sqrt((exx^2 + ezz^2) / 2 + exz^2), from the closed-form
    Cartesian gradient of a field linear in (theta, r); for the rigid rotation u = (z, -x) in
    node values the symmetric part of the exact gradient vanishes."""
    a, b, c, d, e, f = 0.7, -1.3, 0.2, 2.1, 0.4, -0.5
    U = _linear_field(mesh, a, b, c, d, e, f)
    out = so.stress(mesh, U, np.ones((mesh.nel, 7)))
    _, _, thq, rq = so.cartesian_gradients(mesh)
    sn, cs = np.sin(thq), np.cos(thq)
    uxx, uxz = b * sn + a * cs / rq, b * cs - a * sn / rq
    uzx, uzz = e * sn + d * cs / rq, e * cs - d * sn / rq
    ok = _interior(mesh)
    exz = 0.5 * (uxz + uzx)
    eII = np.sqrt(0.5 * (uxx ** 2 + uzz ** 2) + exz ** 2)
    for k, want in (("exx", uxx), ("ezz", uzz), ("exz", exz), ("eII", eII)):
        assert np.abs(out[k][ok] - want[ok]).max() < 1e-12 * (1 + np.abs(want).max()), k


def test_rigid_translation_has_no_strain_rate(mesh):
    U = np.tile([0.3, -1.7], mesh.theta.size)
    out = so.stress(mesh, U, np.ones((mesh.nel, 7)))
    for k in ("exx", "ezz", "exz", "txx", "tzz", "txz"):
        assert np.abs(out[k]).max() < 1e-11


def test_rk4_is_the_fourth_order_exponential(mesh):
    """For F' = L F with L constant over the step, classic RK4 gives exactly
    (I + L + L^2/2 + L^3/6 + L^4/24) F0; L = dt (grad U)^T from the closed form above."""
    a, b, c, d, e, f = 0.7, -1.3, 0.2, 2.1, 0.4, -0.5
    U = _linear_field(mesh, a, b, c, d, e, f)
    rng = np.random.default_rng(7)
    F0 = rng.normal(size=(mesh.nel, 7, 2, 2)) + np.eye(2)
    dt = 0.05
    F1 = so.deformation_gradient(mesh, U, F0, dt)
    _, _, thq, rq = so.cartesian_gradients(mesh)
    sn, cs = np.sin(thq), np.cos(thq)
    L = dt * np.stack([np.stack([b * sn + a * cs / rq, b * cs - a * sn / rq], -1),
                       np.stack([e * sn + d * cs / rq, e * cs - d * sn / rq], -1)], -2)
    I = np.eye(2)
    L2 = L @ L
    P = I + L + L2 / 2 + L2 @ L / 6 + L2 @ L2 / 24
    want = P @ F0
    ok = _interior(mesh)
    assert np.abs(F1[ok] - want[ok]).max() < 1e-12
    # and it is not the first-order (Euler) update
    assert np.abs(F1[ok] - ((I + L) @ F0)[ok]).max() > 1e-6


def test_isotropic_law_is_twice_eta_deviator(mesh):
    """tau = 2 eta (eps - tr(eps)/3 I) for the three-dimensional deviator of the plane
    strain rate (StressCalculation.jl:117-120)."""
    rng = np.random.default_rng(3)
    U = rng.normal(size=2 * mesh.theta.size)
    eta = np.exp(rng.normal(size=(mesh.nel, 7)))
    out = so.stress(mesh, U, eta)
    tr = (out["exx"] + out["ezz"]) / 3.0
    scale = eta * (np.abs(out["exx"]) + np.abs(out["ezz"]) + np.abs(out["exz"]))   # size of the terms
    assert np.all(np.abs(out["txx"] - 2 * eta * (out["exx"] - tr)) <= 1e-14 * scale)
    assert np.all(np.abs(out["tzz"] - 2 * eta * (out["ezz"] - tr)) <= 1e-14 * scale)
    assert np.all(np.abs(out["txz"] - 2 * eta * out["exz"]) <= 1e-14 * scale)
    # for this law tau = 2 eta eps holds for the shear and tau_xx + tau_zz = (2/3) eta (exx + ezz)
    assert np.all(np.abs(out["txx"] + out["tzz"] - 2 / 3 * eta * (out["exx"] + out["ezz"])) <= 1e-14 * scale)


def test_anisotropic_law_matches_matrix_transcription(mesh):
    """:223-235 written as tau = A(C) [exx, ezz, exz] with the naming of :215-221
    (h11 = C.eta11, h22 = C.eta33, h33 = C.eta55, h12 = C.eta13, h13 = C.eta15,
    h23 = C.eta35, all times eta)."""
    rng = np.random.default_rng(5)
    U = rng.normal(size=2 * mesh.theta.size)
    eta = np.exp(rng.normal(size=(mesh.nel, 7)))
    C = [rng.normal(size=(mesh.nel, 7)) for _ in range(6)]
    out = so.stress(mesh, U, eta, C)
    eps = np.stack([out["exx"], out["ezz"], out["exz"]], -1)
    h11, h22, h33, h12, h13, h23 = (c * eta for c in C)
    A = np.stack([
        np.stack([2 * h11 / 3 - h12 / 3, 2 * h12 / 3 - h11 / 3, h13], -1),
        np.stack([2 * h12 / 3 - h22 / 3, 2 * h22 / 3 - h12 / 3, 2 * h23], -1),
        np.stack([2 * h13 / 3 - h23 / 3, 2 * h23 / 3 - h13 / 3, h33], -1),
    ], -2)
    tau = np.einsum("eqij,eqj->eqi", A, eps)
    scale = np.abs(tau).max()
    for i, k in enumerate(("txx", "tzz", "txz")):
        assert np.abs(out[k] - tau[..., i]).max() < 1e-13 * scale


def test_anisotropic_law_with_the_initial_isotropic_tensor(mesh):
    """With 𝓒init's tensor (anisotropy.jl:13-22: eta11 = eta33 = 4/3, eta55 = 2, eta13 = -2/3,
    eta15 = eta35 = 0) the anisotropic law (:223-235) differs from the isotropic one (:117-120)
    by exactly -(2/9) eta tr(eps) on the normal stresses and not at all on the shear, for any
    velocity: the two agree for incompressible flow.  Pins the eta11, eta33, eta13 and eta55 terms
    of the anisotropic law independently of its transcription."""
    rng = np.random.default_rng(6)
    U = rng.normal(size=2 * mesh.theta.size)
    eta = np.exp(rng.normal(size=(mesh.nel, 7)))
    init = [np.full((mesh.nel, 7), v) for v in (4 / 3, 4 / 3, 2.0, -2 / 3, 0.0, 0.0)]
    an = so.stress(mesh, U, eta, init)
    iso = so.stress(mesh, U, eta)
    shift = -2 / 9 * eta * (iso["exx"] + iso["ezz"])
    scale = np.abs(iso["txx"]).max()
    assert np.abs(an["txx"] - iso["txx"] - shift).max() < 1e-13 * scale
    assert np.abs(an["tzz"] - iso["tzz"] - shift).max() < 1e-13 * scale
    assert np.abs(an["txz"] - iso["txz"]).max() < 1e-13 * scale
    assert np.abs(shift).max() > 1e-3 * scale           # the velocity is compressible: the shift shows
