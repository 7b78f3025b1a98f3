"""Pins of the FP64 oracle (CPU only).

The reference ships no tests or golden vectors and Julia is absent, so each
oracle function is pinned here by what the mathematics fixes: exactness of the
quadrature, closed-form areas, null spaces of element matrices, an
independent fourth-order tensor rotation, the isotropic special case, and a
dense LAPACK solve of the whole saddle-point system.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from dimspers_amd.inputs import (StokesInputs, direct_rotation_2d, make_inputs, rotmatrix, scaled_workload,
                                 splitmix64_uniform)
from dimspers_amd.mesh import R_IN, R_OUT, build_annulus, build_grid, fixangles
from dimspers_amd.quadrature import ip_triangle6, ip_triangle7, sf_dN_tri6, sf_N_tri3, sf_N_tri6
from oracle import stokes_oracle as so


def _monomial_integral(a, b):
    # integral of xi^a eta^b over the reference triangle = a! b! / (a+b+2)!
    return math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)


@pytest.mark.parametrize("rule,degree", [(ip_triangle6, 4), (ip_triangle7, 5)])
def test_quadrature_exact_to_degree(rule, degree):
    x, w = rule()
    assert w.sum() == pytest.approx(0.5, abs=1e-15)
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            q = np.sum(w * x[:, 0] ** a * x[:, 1] ** b)
            assert q == pytest.approx(_monomial_integral(a, b), rel=1e-13, abs=1e-16)
    # and not exact one degree higher (the rule is the one of its stated degree)
    errs = [abs(np.sum(w * x[:, 0] ** a * x[:, 1] ** (degree + 1 - a)) - _monomial_integral(a, degree + 1 - a))
            for a in range(degree + 2)]
    assert max(errs) > 1e-8


def test_quadrature_decimals_of_reference_comments():
    # Quadrature.jl:98-103 lists the weights 0.223381589678011 and 0.109951743655322 (before the 1/2)
    _, w = ip_triangle6()
    assert 2 * w[0] == pytest.approx(0.223381589678011, abs=1e-14)
    assert 2 * w[3] == pytest.approx(0.109951743655322, abs=1e-14)


def test_shape_functions():
    rng = np.random.default_rng(1)
    r, s = rng.random(50) * 0.5, rng.random(50) * 0.5
    N = sf_N_tri6(r, s)
    assert np.allclose(N.sum(-1), 1.0, atol=1e-14)
    dN = sf_dN_tri6(r, s)
    assert np.allclose(dN.sum(-1), 0.0, atol=1e-13)
    h = 1e-6
    fd_r = (sf_N_tri6(r + h, s) - sf_N_tri6(r - h, s)) / (2 * h)
    fd_s = (sf_N_tri6(r, s + h) - sf_N_tri6(r, s - h)) / (2 * h)
    assert np.allclose(dN[:, 0], fd_r, atol=1e-8) and np.allclose(dN[:, 1], fd_s, atol=1e-8)
    nodes = np.array([[0, 0], [1, 0], [0, 1], [0.5, 0], [0.5, 0.5], [0, 0.5]])
    assert np.allclose(sf_N_tri6(nodes[:, 0], nodes[:, 1]), np.eye(6), atol=1e-15)
    assert np.allclose(sf_N_tri3(0.2, 0.3), [0.5, 0.2, 0.3])


@pytest.mark.parametrize("nr,nt", [(4, 16), (8, 32), (32, 256)])
def test_mesh_counts_and_patterns(nr, nt):
    m = build_annulus(nr, nt)
    assert m.nel == 2 * nr * nt
    assert m.nvert == (nr + 1) * nt
    assert m.nnod == (4 * nr + 2) * nt
    # node-pair patterns: nnz formulas measured on the reference connectivity
    dof = so.el2dof(m.e2n)
    I = np.repeat(dof, 12, axis=0).ravel()
    J = np.tile(dof, (12, 1)).ravel()
    K = sp.coo_matrix((np.ones(I.size), (I, J)), shape=(m.nU, m.nU)).tocsr()
    K.sum_duplicates()
    assert K.nnz == 184 * nr * nt + 32 * nt
    G = sp.coo_matrix((np.ones(6 * 3 * m.nel), (np.repeat(m.e2n, 3, axis=0).ravel(),
                                                 np.tile(m.e2n[:3], (6, 1)).ravel())), shape=(m.nnod, m.nP)).tocsr()
    assert 2 * G.nnz == 38 * nr * nt + 10 * nt
    M = sp.coo_matrix((np.ones(9 * m.nel), (np.repeat(m.e2n[:3], 3, axis=0).ravel(),
                                            np.tile(m.e2n[:3], (3, 1)).ravel())), shape=(m.nP, m.nP)).tocsr()
    assert M.nnz == 7 * nr * nt + 3 * nt
    fixed, ifree = m.free_slip()
    assert fixed.size == 4 * nt and ifree.size == m.nU - 4 * nt
    assert np.all(fixed % 2 == 1)


def test_mesh_connectivity_follows_grid_split1():
    nr, nt = 4, 16
    m = build_annulus(nr, nt)
    v = lambda i, j: i + (nr + 1) * (j % nt)
    # cell (i, j) = (2, 15): elements 2q, 2q+1 with q = i + Nr*j (mesher.jl:175-200)
    q = 2 + nr * 15
    assert list(m.e2n[:3, 2 * q]) == [v(2, 15), v(2, 16), v(3, 15)]
    assert list(m.e2n[:3, 2 * q + 1]) == [v(2, 16), v(3, 16), v(3, 15)]
    # midpoints are numbered in first-seen order: element 0 introduces the first three
    assert list(m.e2n[3:, 0]) == [m.nvert, m.nvert + 1, m.nvert + 2]
    # every midpoint sits at the mean (theta, r) of its edge (with the 2 pi wrap)
    th, r = m.el_theta, m.el_r
    for k, (a, b) in enumerate([(0, 1), (1, 2), (2, 0)]):
        mid = m.e2n[3 + k]
        assert np.allclose(m.r[mid], 0.5 * (r[a] + r[b]), atol=1e-15)
        assert np.allclose(np.cos(m.theta[mid]), np.cos(0.5 * (th[a] + th[b])), atol=1e-14)
    # fixangles: every element spans less than pi after unwrapping
    span = th.max(0) - th.min(0)
    assert np.all(span < 1.0) and np.all(span > 0)


def test_grid_connectivity_follows_the_four_triangle_grid():
    """``Grid`` (mesher.jl:59-150): corners radial fastest, centre nr*nt + q at angle + dtheta/2 and
    radius + dr/2, triangles [corner k, corner k+1, centre] of the quad [v(i,j) v(i,j+1) v(i+1,j+1)
    v(i+1,j)], midpoints in first-seen order; counts; every element positively oriented in (theta, r);
    the annulus area from the quadrature; free slip fixes only wall corners and wall midpoints."""
    nr, nt = 4, 16
    m = build_grid(nr, nt)
    v = lambda i, j: i + (nr + 1) * (j % nt)
    ncorner = (nr + 1) * nt
    assert m.split == 4 and m.nel == 4 * nr * nt and m.nvert == ncorner + nr * nt
    assert m.nnod == m.nvert + nr * nt * 4 + nr * nt + (nr + 1) * nt   # vertices + internal, radial, angular edges
    q = 2 + nr * 15                                 # cell (i, j) = (2, 15), the last column
    quad = [v(2, 15), v(2, 16), v(3, 16), v(3, 15)]
    for k in range(4):
        assert list(m.e2n[:3, 4 * q + k]) == [quad[k], quad[(k + 1) % 4], ncorner + q]
    dth, dr = 2 * np.pi / nt, (R_OUT - R_IN) / nr
    assert m.theta[ncorner + q] == pytest.approx(15 * dth + dth / 2, abs=1e-15)
    assert m.r[ncorner + q] == pytest.approx(R_IN + 2 * dr + dr / 2, abs=1e-15)
    assert list(m.e2n[3:, 0]) == [m.nvert, m.nvert + 1, m.nvert + 2]
    th, r = m.el_theta, m.el_r
    det = (th[1] - th[0]) * (r[2] - r[0]) - (th[2] - th[0]) * (r[1] - r[0])
    assert np.all(det > 0)
    for k, (a, b) in enumerate([(0, 1), (1, 2), (2, 0)]):
        mid = m.e2n[3 + k]
        assert np.allclose(m.r[mid], 0.5 * (r[a] + r[b]), atol=1e-15)
        assert np.allclose(np.cos(m.theta[mid]), np.cos(0.5 * (th[a] + th[b])), atol=1e-14)
    total = 0.0
    for qq in range(6):
        _, _, om, _, _ = so.ip_derivatives(m.el_theta, m.el_r, qq)
        total += om.sum()
    assert total == pytest.approx(np.pi * (R_OUT ** 2 - R_IN ** 2), rel=1e-13)
    fixed, _ = m.free_slip()
    assert fixed.size == 4 * nt and np.all(fixed % 2 == 1)
    assert not np.any((fixed // 2 >= ncorner) & (fixed // 2 < m.nvert))   # no centre on a wall


def test_fixangles_wrap_rule():
    th = np.array([[2 * np.pi - 0.1], [0.0], [2 * np.pi - 0.1]])
    out = fixangles(th)
    assert np.allclose(out[:, 0], [2 * np.pi - 0.1, 2 * np.pi, 2 * np.pi - 0.1])


def _small_inputs(nr=4, nt=16, w=3):
    return make_inputs(scaled_workload(w, nr, nt))


def test_annulus_area_is_exact():
    inp = _small_inputs(8, 32)
    m = inp.mesh
    total = 0.0
    for q in range(6):
        _, _, om, _, _ = so.ip_derivatives(m.el_theta, m.el_r, q)
        total += om.sum()
    assert total == pytest.approx(np.pi * (R_OUT ** 2 - R_IN ** 2), rel=1e-13)


def test_element_matrix_invariants():
    inp = _small_inputs()
    Ke, Ge, Me, fe = so.element_matrices(inp)
    assert np.allclose(Ke, np.transpose(Ke, (0, 2, 1)), rtol=0, atol=1e-13 * np.abs(Ke).max())
    ev = np.linalg.eigvalsh(Ke)
    scale = ev[:, -1:]
    assert np.all(ev > -1e-12 * scale)
    # two exact rigid translations and one near-null rigid rotation: rank 9
    assert np.all(ev[:, 1] < 1e-12 * scale[:, 0])
    assert np.all(ev[:, 3] > 1e-6 * scale[:, 0])
    tx = np.zeros(12); tx[0::2] = 1
    tz = np.zeros(12); tz[1::2] = 1
    assert np.abs(Ke @ tx).max() < 1e-11 * np.abs(Ke).max()
    assert np.abs(Ke @ tz).max() < 1e-11 * np.abs(Ke).max()
    # divergence of a translation is zero: G_e^T t = 0
    assert np.abs(np.einsum("eai,a->ei", Ge, tx)).max() < 1e-12 * np.abs(Ge).max()
    # M_e is SPD
    assert np.all(np.linalg.eigvalsh(Me) > 0)


def test_mass_matrix_closed_form_for_unit_viscosity():
    inp = _small_inputs(8, 32)
    inp = StokesInputs(inp.mesh, inp.workload, inp.T, inp.rho, np.ones_like(inp.eta_ip), inp.C6, inp.phi, inp.delta)
    _, _, Me, _ = so.element_matrices(inp)
    # 1^T M 1 summed over elements = area of the annulus when eta = 1
    assert Me.sum() == pytest.approx(np.pi * (R_OUT ** 2 - R_IN ** 2), rel=1e-13)
    KK, GG, MM, Fb = so.assembly_stokes_cylindric(inp)
    pc = so.lumped_preconditioner(MM)
    assert (1.0 / pc).sum() == pytest.approx(np.pi * (R_OUT ** 2 - R_IN ** 2), rel=1e-13)


def test_isotropic_tensor_matches_isotropic_assembly():
    inp = _small_inputs()
    iso = direct_rotation_2d(rotmatrix(np.cos(inp.phi), np.sin(inp.phi)), 4 / 3, 4 / 3, 1.0, -2 / 3)
    C6 = tuple(np.repeat(c[:, None], 6, axis=1) for c in iso)
    # rotation invariance of the isotropic tensor
    assert np.allclose(C6[0], 4 / 3) and np.allclose(C6[2], 1.0) and np.allclose(C6[3], -2 / 3)
    assert np.allclose(C6[4], 0, atol=1e-15) and np.allclose(C6[5], 0, atol=1e-15)
    a = StokesInputs(inp.mesh, inp.workload, inp.T, inp.rho, inp.eta_ip, C6, inp.phi, inp.delta)
    Ka, Ga, Ma, fa = so.element_matrices(a)
    Ki, Gi, Mi, fi = so.element_matrices(a, isotropic=True)
    assert np.allclose(Ka, Ki, rtol=0, atol=1e-12 * np.abs(Ki).max())


def test_direct_rotation_matches_fourth_order_tensor_rotation():
    a, c, e, h = 4 / 3, 1.1, 0.3, -0.5
    Cf = np.zeros((2, 2, 2, 2))
    Cf[0, 0, 0, 0], Cf[1, 1, 1, 1] = a, c
    Cf[0, 0, 1, 1] = Cf[1, 1, 0, 0] = h
    for idx in [(0, 1, 0, 1), (0, 1, 1, 0), (1, 0, 0, 1), (1, 0, 1, 0)]:
        Cf[idx] = e
    for phi in [0.0, 0.37, 1.2, 2.9]:
        # directRotation2D(rotmatrix(cos, sin)) rotates the (x, z) tensor by -phi
        R = np.array([[np.cos(-phi), -np.sin(-phi)], [np.sin(-phi), np.cos(-phi)]])
        Cr = np.einsum("ia,jb,kc,ld,abcd->ijkl", R, R, R, R, Cf)
        want = [Cr[0, 0, 0, 0], Cr[1, 1, 1, 1], Cr[0, 1, 0, 1], Cr[0, 0, 1, 1], Cr[0, 0, 0, 1], Cr[1, 1, 0, 1]]
        got = direct_rotation_2d(rotmatrix(np.cos(phi), np.sin(phi)), a, c, e, h)
        assert np.allclose(got, want, atol=1e-14)
        got_pi = direct_rotation_2d(rotmatrix(np.cos(phi + np.pi), np.sin(phi + np.pi)), a, c, e, h)
        assert np.allclose(got, got_pi, atol=1e-13)


def test_force_sums_vanish_by_symmetry():
    inp = _small_inputs(8, 32)
    inp = StokesInputs(inp.mesh, inp.workload, inp.T, np.ones_like(inp.rho), inp.eta_ip, inp.C6, inp.phi,
                       inp.delta)
    _, _, _, Fb = so.assembly_stokes_cylindric(inp)
    assert abs(Fb[0::2].sum()) < 1e-13 * np.abs(Fb).sum()
    assert abs(Fb[1::2].sum()) < 1e-13 * np.abs(Fb).sum()


def test_splitmix64_known_values():
    # SplitMix64 with state 0: first outputs 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4
    u = splitmix64_uniform(0, 2)
    assert u[0] == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53
    assert u[1] == (0x6E789E6AA1B965F4 >> 11) * 2.0 ** -53


def _dense_saddle_solution(sol):
    """Dense LAPACK solve of the free-dof saddle system the reference iterates on.

    rm0space! (IterativeSolvers.jl:172-186) removes the mean of the continuity
    residual at every step, so the reference converges to K u - G p = f,
    -G^T u = c 1, mean(p) = 0 with an unknown constant c (c is not zero on a
    discrete curved free-slip boundary, where G 1 != 0 on the free dofs).
    """
    K = sol.K.toarray()
    G = sol.G.toarray()
    f = sol.F
    fr = sol.ifree
    nP = G.shape[1]
    nf = fr.size
    A = np.zeros((nf + nP + 1, nf + nP + 1))
    A[:nf, :nf] = K[np.ix_(fr, fr)]
    A[:nf, nf:nf + nP] = -G[fr]
    A[nf:nf + nP, :nf] = -G[fr].T
    A[nf:nf + nP, -1] = -1.0
    A[-1, nf:nf + nP] = 1.0
    b = np.zeros(nf + nP + 1)
    b[:nf] = f[fr]
    x = sla.lstsq(A, b, cond=1e-14)[0]
    U = np.zeros(K.shape[0])
    U[fr] = x[:nf]
    return U, x[nf:nf + nP]


def test_schur_pcg_matches_dense_saddle_solve():
    inp = _small_inputs(4, 16, w=1)
    sol, Uraw = so.solve_stokes(inp, rtol=1e-13, itmax=400)
    assert sol.info.converged
    Ud, Pd = _dense_saddle_solution(sol)
    Ud, _ = so.updatevelocity2(Ud, inp.rho, inp.mesh)
    assert np.linalg.norm(Ud - sol.U) / np.linalg.norm(sol.U) < 1e-9
    P = sol.P - sol.P.mean()
    assert np.linalg.norm(Pd - P) / np.linalg.norm(P) < 1e-9


def test_full_residual_of_the_dense_saddle_solution_vanishes():
    """The judged residual ||[f + G p - K u ; -G^T u - c]|| / ||f|| (c the mean of -G^T u) is at
    rounding level for the dense LAPACK solution of K u - G p = f, -G^T u = c 1 (the system the
    reference iterates on), and a change of p shows in it: the signs of its terms are those of the
    saddle system, independently of the Schur PCG."""
    inp = _small_inputs(4, 16, w=1)
    sol, _ = so.solve_stokes(inp, rtol=1e-13, itmax=400)
    Ud, Pd = _dense_saddle_solution(sol)
    assert so.full_residual(sol.K, sol.G, sol.F, Ud, Pd, sol.ifree) < 1e-10
    bumped = Pd + 1e-3 * np.abs(Pd).max() * np.cos(np.arange(Pd.size))
    assert so.full_residual(sol.K, sol.G, sol.F, Ud, bumped, sol.ifree) > 1e-6


def test_updatevelocity2_removes_net_rotation():
    inp = _small_inputs(4, 16, w=1)
    sol, _ = so.solve_stokes(inp)
    m = inp.mesh
    assert abs(so.volume_integral(sol.U[0::2] * inp.rho, m)) < 1e-12 * so.volume_integral(
        np.abs(sol.U[0::2]) * inp.rho, m)
    # the Cartesian velocity is the node-wise rotation of the polar one
    assert np.allclose(np.hypot(sol.Ucart[0::2], sol.Ucart[1::2]), np.hypot(sol.U[0::2], sol.U[1::2]))


def test_reference_stopping_rule_and_itmin():
    inp = _small_inputs(4, 16, w=1)
    KK, GG, MM, Fb = so.assembly_stokes_cylindric(inp)
    K, G, F = so.prepare_matrices(KK, GG, Fb, so.rotation_matrix(inp.mesh.theta))
    _, ifree = inp.mesh.free_slip()
    _, _, info = so.stokes_pcg(K, MM, G, F, ifree, rtol=1.0)
    assert info.iterations == 10          # itmin = 10 (IterativeSolvers.jl:97,156)
    _, _, info = so.stokes_pcg(K, MM, G, F, ifree, rtol=1e-30, itmax=7)
    assert info.iterations == 7 and not info.converged   # silent cap (:96)


# ----------------------------------------------------------------------------- closed-form pins of the
# frame, the body force, the 1/eta mass and the placement of the tensor in D.  Each one is fixed by
# the mathematics (polar geometry, integrals of polynomials in r, the fourth-order tensor contraction),
# not by a second transcription of the oracle's formulas.

def _polar_to_cartesian(theta, r):
    """polar2cartesian (src/Grid/mesher.jl:36): x = r sin(theta), z = r cos(theta)."""
    return r * np.sin(theta), r * np.cos(theta)


def test_rotation_matrix_maps_node_frames_to_the_polar_unit_vectors():
    # rotation_matrix (StokesSolver.jl:203-245): U_cart = TT U with U = (u_theta, u_r) per node.
    # With x = r sin(theta), z = r cos(theta): r_hat = (x, z) / r and theta_hat = d(x, z)/dtheta / r.
    m = build_annulus(4, 16)
    TT = so.rotation_matrix(m.theta)
    x, z = _polar_to_cartesian(m.theta, m.r)
    radial = np.zeros(m.nU)
    radial[1::2] = 1.0
    u = TT @ radial
    assert np.allclose(u[0::2], x / m.r, atol=1e-15) and np.allclose(u[1::2], z / m.r, atol=1e-15)
    tangential = np.zeros(m.nU)
    tangential[0::2] = 1.0
    u = TT @ tangential
    assert np.allclose(u[0::2], z / m.r, atol=1e-15) and np.allclose(u[1::2], -x / m.r, atol=1e-15)
    # a rigid rotation u = omega x r = (z, -x) has u_theta = r and u_r = 0 in the node frames
    rot = np.zeros(m.nU)
    rot[0::2], rot[1::2] = z, -x
    back = TT.T @ rot
    assert np.allclose(back[0::2], m.r, atol=1e-14) and np.allclose(back[1::2], 0.0, atol=1e-14)


def _element_centroid_theta(m):
    return np.mod(m.el_theta.mean(axis=0), 2 * np.pi)


def test_body_force_is_radial_buoyancy_with_linear_density():
    # _fill_Fbblk! + _element_density (StokesSolver.jl:838-845, :1017-1022): f = g rho r_hat with rho
    # interpolated linearly from the vertex values.  rho = a + b r is held exactly by that
    # interpolation (elements are affine in (theta, r)), so over the half-annulus x > 0
    #   sum f_x = g int_0^pi int (a + b r) sin(theta) r dr dtheta = 2 g [a (ro^2 - ri^2)/2 + b (ro^3 - ri^3)/3]
    # and the same for f_z over z > 0; the other half-plane sums are the negatives.  A force along
    # theta_hat, or sin and cos exchanged, gives zero for these sums.
    inp = _small_inputs(8, 32)
    m = inp.mesh
    a, b, g = 0.7, 1.9, 2.5
    rho = a + b * m.r
    inp = StokesInputs(m, inp.workload, inp.T, rho, inp.eta_ip, inp.C6, inp.phi, inp.delta, g=g)
    _, _, _, fe = so.element_matrices(inp)
    want = 2 * g * (a * (R_OUT ** 2 - R_IN ** 2) / 2 + b * (R_OUT ** 3 - R_IN ** 3) / 3)
    th_c = _element_centroid_theta(m)
    fx, fz = fe[:, 0::2].sum(1), fe[:, 1::2].sum(1)
    right = th_c < np.pi                                   # x > 0
    top = (th_c < np.pi / 2) | (th_c > 3 * np.pi / 2)      # z > 0
    # the rule is exact for polynomials; sin and cos over a cell of width 2 pi / 32 leave ~1e-7
    assert fx[right].sum() == pytest.approx(want, rel=1e-6)
    assert fx[~right].sum() == pytest.approx(-want, rel=1e-6)
    assert fz[top].sum() == pytest.approx(want, rel=1e-6)
    assert fz[~top].sum() == pytest.approx(-want, rel=1e-6)
    # in the node frames the force has a positive radial sum and no net tangential part (the 2-triangle
    # split is not mirror-symmetric, so the tangential sum is small rather than zero: 1.6e-10 here)
    KK, GG, MM, Fb = so.assembly_stokes_cylindric(inp)
    F = so.rotation_matrix(m.theta).T @ Fb
    total = g * 2 * np.pi * (a * (R_OUT ** 2 - R_IN ** 2) / 2 + b * (R_OUT ** 3 - R_IN ** 3) / 3)
    assert F[1::2].sum() == pytest.approx(total, rel=2e-2)        # int g rho dOmega, up to cos(dtheta)
    assert abs(F[0::2].sum()) < 1e-6 * np.abs(F).sum()


@pytest.mark.parametrize("kind", ["constant", "radial"])
def test_pressure_mass_scales_with_inverse_viscosity(kind):
    # _fill_Mblk! (StokesSolver.jl:825-836): M_e = sum_q omega_q / eta_q NP NP^T, so 1^T M 1 = int 1/eta.
    # constant eta = 3.7: area / 3.7.  eta_q = 1 / (1 + r_q): int (1 + r) r dr dtheta, a polynomial the
    # 6-point rule integrates exactly (this also pins which point each eta.ip column belongs to).
    inp = _small_inputs(8, 32)
    m = inp.mesh
    x_ip, _ = ip_triangle6()
    if kind == "constant":
        eta = np.full((m.nel, 6), 3.7, order="F")
        want = np.pi * (R_OUT ** 2 - R_IN ** 2) / 3.7
    else:
        lam = np.stack([1 - x_ip[:, 0] - x_ip[:, 1], x_ip[:, 0], x_ip[:, 1]])   # barycentric (3, 6)
        rq = (m.el_r.T @ lam)                                                    # (nel, 6)
        eta = np.asfortranarray(1.0 / (1.0 + rq))
        want = 2 * np.pi * ((R_OUT ** 2 - R_IN ** 2) / 2 + (R_OUT ** 3 - R_IN ** 3) / 3)
    inp = StokesInputs(m, inp.workload, inp.T, inp.rho, eta, inp.C6, inp.phi, inp.delta)
    _, _, Me, _ = so.element_matrices(inp)
    assert Me.sum() == pytest.approx(want, rel=1e-13)


def _fabric_fourth_order(a, c, e, h):
    """Orthotropic 2-D viscous tensor in its own frame (indices 0 = x, 1 = z)."""
    Cf = np.zeros((2, 2, 2, 2))
    Cf[0, 0, 0, 0], Cf[1, 1, 1, 1] = a, c
    Cf[0, 0, 1, 1] = Cf[1, 1, 0, 0] = h
    for idx in [(0, 1, 0, 1), (0, 1, 1, 0), (1, 0, 0, 1), (1, 0, 1, 0)]:
        Cf[idx] = e
    return Cf


def _quadratic_field(rng):
    """u_x, u_z quadratic in (theta, r): held exactly by P2 on elements affine in (theta, r)."""
    cx, cz = rng.standard_normal(6), rng.standard_normal(6)

    def val(cf, t, r):
        return cf[0] + cf[1] * t + cf[2] * r + cf[3] * t * r + cf[4] * t * t + cf[5] * r * r

    def d_t(cf, t, r):
        return cf[1] + cf[3] * r + 2 * cf[4] * t

    def d_r(cf, t, r):
        return cf[2] + cf[3] * t + 2 * cf[5] * r

    return cx, cz, val, d_t, d_r


def _element_nodes_polar(m):
    """(theta, r) of the six element nodes from the unwrapped vertex angles, (6, nel)."""
    th, r = m.el_theta, m.el_r
    mid = [(0, 1), (1, 2), (2, 0)]
    tn = np.concatenate([th, np.stack([0.5 * (th[i] + th[j]) for i, j in mid])])
    rn = np.concatenate([r, np.stack([0.5 * (r[i] + r[j]) for i, j in mid])])
    return tn, rn


def _points_polar(m, x_ip):
    lam = np.stack([1 - x_ip[:, 0] - x_ip[:, 1], x_ip[:, 0], x_ip[:, 1]])     # (3, nq)
    return m.el_theta.T @ lam, m.el_r.T @ lam                                   # (nel, nq)


def _cartesian_gradient(cf, t, r, d_t, d_r):
    """Chain rule for x = r sin t, z = r cos t: d/dx = sin t d/dr + cos t / r d/dt,
    d/dz = cos t d/dr - sin t / r d/dt."""
    gt, gr = d_t(cf, t, r), d_r(cf, t, r)
    return np.sin(t) * gr + np.cos(t) / r * gt, np.cos(t) * gr - np.sin(t) / r * gt


def _point_weights(m, w_ip):
    """omega_q = r_q |det d(theta, r)/d(xi, eta)| w_q: the polar area element r dr dtheta."""
    th, r = m.el_theta, m.el_r
    det = (th[1] - th[0]) * (r[2] - r[0]) - (th[2] - th[0]) * (r[1] - r[0])
    return np.abs(det)[:, None] * w_ip[None, :]


@pytest.mark.parametrize("nip", [6, 7])
def test_velocity_block_is_the_viscous_energy_of_the_rotated_tensor(nip):
    # _fill_Kblk!{Anisotropic} (src/src/Stokes/StokesSolver.jl:744-808): u^T K_e u must equal
    #   sum_q omega_q eta_q eps_ij C_ijkl eps_kl
    # with C the fourth-order tensor whose Voigt entries (xx, zz, xz) are eta11 = C_xxxx,
    # eta33 = C_zzzz, eta55 = C_xzxz, eta13 = C_xxzz, eta15 = C_xxxz, eta35 = C_zzxz (the naming that
    # directRotation2D's outputs carry, pinned above).  The velocity is quadratic in (theta, r), so
    # P2 holds it exactly and its Cartesian strain rate follows from the chain rule at each point.
    # eta15 != eta35 for a rotated tensor with C11 != C33, so exchanging them in D fails here.
    rng = np.random.default_rng(7)
    inp = _small_inputs(4, 16)
    m = inp.mesh
    nel = m.nel
    x_ip, w_ip = ip_triangle7() if nip == 7 else ip_triangle6()
    Cf = _fabric_fourth_order(4 / 3, 0.6, 0.3, -0.5)
    phi = rng.uniform(0, np.pi, nel)
    R = np.stack([np.stack([np.cos(phi), -np.sin(phi)], -1), np.stack([np.sin(phi), np.cos(phi)], -1)], -2)
    C4 = np.einsum("eia,ejb,ekc,eld,abcd->eijkl", R, R, R, R, Cf)
    comps = [C4[:, 0, 0, 0, 0], C4[:, 1, 1, 1, 1], C4[:, 0, 1, 0, 1], C4[:, 0, 0, 1, 1], C4[:, 0, 0, 0, 1],
             C4[:, 1, 1, 0, 1]]
    assert np.abs(comps[4] - comps[5]).min() > 1e-3 or np.abs(comps[4] - comps[5]).max() > 0.1
    C6 = tuple(np.asfortranarray(np.repeat(c[:, None], nip, axis=1)) for c in comps)
    eta = np.asfortranarray(0.5 + rng.random((nel, nip)))
    inp = StokesInputs(m, inp.workload, inp.T, inp.rho, eta, C6, phi, inp.delta)
    Ke, Ge, _, _ = so.element_matrices(inp, nip=nip)

    cx, cz, val, d_t, d_r = _quadratic_field(rng)
    tn, rn = _element_nodes_polar(m)
    ue = np.zeros((nel, 12))
    ue[:, 0::2], ue[:, 1::2] = val(cx, tn, rn).T, val(cz, tn, rn).T
    tq, rq = _points_polar(m, x_ip)
    uxx, uxz = _cartesian_gradient(cx, tq, rq, d_t, d_r)
    uzx, uzz = _cartesian_gradient(cz, tq, rq, d_t, d_r)
    eps = np.empty((nel, nip, 2, 2))
    eps[..., 0, 0], eps[..., 1, 1] = uxx, uzz
    eps[..., 0, 1] = eps[..., 1, 0] = 0.5 * (uxz + uzx)
    dens = np.einsum("eqij,eijkl,eqkl->eq", eps, C4, eps)
    omega = rq * _point_weights(m, w_ip)
    want = (omega * eta * dens).sum(1)
    got = np.einsum("ea,eab,eb->e", ue, Ke, ue)
    assert np.allclose(got, want, rtol=1e-11, atol=1e-12 * np.abs(want).max())

    # _fill_Gblk! (:812-823): u^T G_e p = sum_q omega_q p_q div u, for p linear (P1) and u as above
    pv = rng.standard_normal((3, nel))
    lam = np.stack([1 - x_ip[:, 0] - x_ip[:, 1], x_ip[:, 0], x_ip[:, 1]])
    pq = pv.T @ lam
    want = (omega * pq * (uxx + uzz)).sum(1)
    got = np.einsum("ea,eai,ie->e", ue, Ge, pv)
    assert np.allclose(got, want, rtol=1e-11, atol=1e-12 * np.abs(want).max())


def test_prepare_matrices_act_in_the_node_frames():
    # _prepare_matrices (StokesSolver.jl:192-201) rotates K, G and F into the (u_theta, u_r) frames of
    # the nodes.  In those frames:
    #  * u = (r, 0) per node is the rigid rotation (z, -x): its energy u^T K u is at the level of K's
    #    near-null rotation mode, while a field of the same size that is not rigid is not;
    #  * u = (0, r) per node is the field (x, z), whose divergence is 2: sum_i (G^T u)_i = int div u = 2 area
    #    (P2 holds (x, z) only approximately on the curved cells, hence the tolerance);
    #  * the body force g rho r_hat of a uniform density has a positive radial sum ~ g rho area and a tangential
    #    sum near zero.
    inp = _small_inputs(8, 64)
    m = inp.mesh
    inp = StokesInputs(m, inp.workload, inp.T, np.ones_like(inp.rho), inp.eta_ip, inp.C6, inp.phi, inp.delta)
    KK, GG, MM, Fb = so.assembly_stokes_cylindric(inp)
    K, G, F = so.prepare_matrices(KK, GG, Fb, so.rotation_matrix(m.theta))
    area = np.pi * (R_OUT ** 2 - R_IN ** 2)
    rigid = np.zeros(m.nU)
    rigid[0::2] = m.r
    shear = np.zeros(m.nU)
    shear[0::2] = m.r * np.cos(2 * m.theta)
    e_rigid, e_shear = rigid @ (K @ rigid), shear @ (K @ shear)
    assert e_shear > 0 and abs(e_rigid) < 1e-6 * e_shear
    radial = np.zeros(m.nU)
    radial[1::2] = m.r
    assert (G.T @ radial).sum() == pytest.approx(2 * area, rel=1e-3)
    assert F[1::2].sum() == pytest.approx(inp.g * area, rel=1e-2)
    assert abs(F[0::2].sum()) < 1e-6 * np.abs(F).sum()


def test_volume_integral_of_polynomials_in_r_is_exact():
    # volume_integral (src/Stress/FSE.jl:412-464): int V dOmega with the polar weight r detJ w_q and P2
    # interpolation of the nodal values.  P2 holds 1, r and r^2 exactly on elements affine in (theta, r),
    # and the integrands r^k r are polynomials the 6-point rule integrates exactly.
    m = build_annulus(8, 32)
    for k in range(3):
        want = 2 * np.pi * (R_OUT ** (k + 2) - R_IN ** (k + 2)) / (k + 2)
        assert so.volume_integral(m.r ** k, m) == pytest.approx(want, rel=1e-13)


def test_incomplete_cholesky_reproduces_the_matrix_on_its_pattern():
    # IC(0) is defined by (L D L^T)_ij = A_ij on A's pattern; where the elimination order creates no
    # fill (a path graph in natural order) it is the exact Cholesky factorisation.
    inp = _small_inputs(8, 32)
    KK, GG, MM, Fb = so.assembly_stokes_cylindric(inp)
    M = so.symmetric_from_lower(MM).tocsr()
    order, colour = so.greedy_colour_order(M)
    Mc = M.tocoo()
    off = Mc.row != Mc.col
    assert np.all(colour[Mc.row[off]] != colour[Mc.col[off]])         # colours are independent sets
    L, D = so.incomplete_cholesky(M, order)
    assert np.all(D > 0)
    pos = np.empty(M.shape[0], dtype=np.int64)
    pos[order] = np.arange(M.shape[0])
    Lc = L.tocoo()
    assert np.all(pos[Lc.col] < pos[Lc.row])                          # strictly earlier in the order
    Lu = (L + sp.identity(M.shape[0])).tocsr()
    A = (Lu @ sp.diags(D) @ Lu.T).tocsr()
    on = np.asarray(A[Mc.row, Mc.col]).ravel()
    assert np.abs(on - Mc.data).max() < 1e-14 * np.abs(Mc.data).max()
    n = 9
    T = sp.diags([np.full(n - 1, -1.0), np.full(n, 2.5), np.full(n - 1, -1.0)], [-1, 0, 1]).tocsr()
    L, D = so.incomplete_cholesky(T, np.arange(n))
    Lu = (L + sp.identity(n)).toarray()
    assert np.allclose(Lu @ np.diag(D) @ Lu.T, T.toarray(), atol=1e-14)


def test_pardiso_variant_keeps_the_incoming_pressure():
    # StokesPcCG_pardiso (IterativeSolvers.jl:1-17) keeps P and adds GG * P to the right-hand side: from
    # the converged pressure the initial Schur residual vanishes (it does not for a pressure subtracted
    # instead of added, or one zeroed), and with the IC(0) preconditioner of MM the iteration converges
    # to the solution of the dense saddle solve.
    inp = _small_inputs(4, 16, w=1)
    sol, _ = so.solve_stokes(inp, rtol=1e-13, itmax=400)
    K, G, F, ifree, MM = sol.K, sol.G, sol.F, sol.ifree, sol.MM
    inner = so.inner_direct(K, ifree)
    _, _, cold = so.stokes_pcg(K, MM, G, F, ifree, rtol=1e-8, inner=inner)
    _, _, warm = so.stokes_pcg(K, MM, G, F, ifree, rtol=1e-8, inner=inner, P0=sol.P)
    assert warm.residuals[0] < 1e-9 * cold.residuals[0]
    U, P, info = so.stokes_pcg(K, MM, G, F, ifree, rtol=1e-12, itmax=400, inner=inner,
                               pc=so.ic_preconditioner(MM), P0=0.5 * sol.P)
    assert info.converged
    Ud, Pd = _dense_saddle_solution(sol)
    Ud, _ = so.updatevelocity2(Ud, inp.rho, inp.mesh)
    U, _ = so.updatevelocity2(U, inp.rho, inp.mesh)
    assert np.linalg.norm(U - Ud) / np.linalg.norm(Ud) < 1e-9
    assert np.linalg.norm((P - P.mean()) - Pd) / np.linalg.norm(Pd) < 1e-9


def test_clamped_tensor_is_the_all_ones_matrix_below_the_lower_bound():
    # The top-level _fill_Kblk! (StokesSolver.jl:756,770-804) clamps every eta C entry to [1e18, 1e25]
    # (applybounds, src/PiC/PiC.jl:834).  Non-dimensional viscosities (eta C <= 1e2) all become 1e18, so
    # D = 1e18 [[1,1,1],[1,1,1],[1,1,1]] and u^T K_e u = 1e18 sum_q omega_q (u_x,x + u_z,z + u_x,z + u_z,x)^2;
    # in the dimensional range the clamp is the identity.
    rng = np.random.default_rng(5)
    inp = _small_inputs(4, 16)
    m = inp.mesh
    x_ip, w_ip = ip_triangle6()
    Ke, _, _, _ = so.element_matrices(inp, clamp=True)
    cx, cz, val, d_t, d_r = _quadratic_field(rng)
    tn, rn = _element_nodes_polar(m)
    ue = np.zeros((m.nel, 12))
    ue[:, 0::2], ue[:, 1::2] = val(cx, tn, rn).T, val(cz, tn, rn).T
    tq, rq = _points_polar(m, x_ip)
    uxx, uxz = _cartesian_gradient(cx, tq, rq, d_t, d_r)
    uzx, uzz = _cartesian_gradient(cz, tq, rq, d_t, d_r)
    omega = rq * _point_weights(m, w_ip)
    want = 1e18 * (omega * (uxx + uzz + uxz + uzx) ** 2).sum(1)
    got = np.einsum("ea,eab,eb->e", ue, Ke, ue)
    assert np.allclose(got, want, rtol=1e-11, atol=1e-12 * np.abs(want).max())
    dim = StokesInputs(m, inp.workload, inp.T, inp.rho, inp.eta_ip * 1e21, inp.C6, inp.phi, inp.delta)
    C = np.stack([np.asarray(c) for c in inp.C6]) * dim.eta_ip[None]
    inside = (C > 1e18) & (C < 1e25)
    if inside.all():
        Kc, _, _, _ = so.element_matrices(dim, clamp=True)
        Kd, _, _, _ = so.element_matrices(dim)
        assert np.allclose(Kc, Kd, rtol=1e-14, atol=0)
    else:
        # negative eta13 entries hit the lower bound: the clamp changes K
        Kc, _, _, _ = so.element_matrices(dim, clamp=True)
        Kd, _, _, _ = so.element_matrices(dim)
        assert not np.allclose(Kc, Kd)
