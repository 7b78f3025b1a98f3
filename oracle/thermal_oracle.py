"""FP64 CPU oracle of the thermal diffusion step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module; the product path never calls it.

Restates ``solveDiffusion_threaded`` (src/ThermalDiffusion/AssemblerThermal.jl:1-62):

* ``thermal_matrices`` -- ``assemble_element!`` (:117-209) with the 7-point rule of
  ``diffusion_immutables(gr, 2, 3, 6, 7)`` (MAIN.jl:167-169): K_e = sum_q K dN^T dN w_q,
  M_e = sum_q rho_q Cp N^T N w_q, f_e = sum_q dQdT dt N w_q, w_q = r_q detJ w_ip, rho_q the
  P1 interpolation of rho at the vertices, derivatives by the polar double Jacobian;
  summed into global CSC matrices (``update_stiffness_matrix!``, :217-223).
* ``thermal_step`` -- ``_prepare_matrices!`` (src/ThermalDiffusion/ThermalSolver.jl:54-63,
  alpha = 1): Rhs = M T0 + F, A = M + dt K; ``_apply_bcs!`` (:38-41); the free block
  solved directly (SuperLU) in place of the reference's Krylov.gmres (Krylov.jl 0.7.13,
  an iterative solve of the same system, stopping near 1.5e-8 of |Rhs|).

Pinned (tests/test_thermal_oracle.py): sum of M_T = area for rho = Cp = 1; K_T 1 = 0;
u^T K_T u = K area for u = r (|grad r| = 1, held exactly by P2); the large-dt step equals
the analytic steady conduction profile ln(r/r_out)/ln(r_in/r_out) to discretisation error; a
step with insulated walls gains exactly dQdT dt area of heat (1^T M_T (T - T0)).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from dimspers_amd.quadrature import ip_triangle7, sf_dN_tri6, sf_N_tri3, sf_N_tri6


def thermal_matrices(mesh, rho, K, Cp, dQdT=0.0, dt=1.0):
    th, r = mesh.el_theta.T, mesh.el_r.T                      # (nel, 3)
    e2n = mesh.e2n                                            # (6, nel)
    nel, nn = e2n.shape[1], mesh.theta.size
    R21, R31 = r[:, 1] - r[:, 0], r[:, 2] - r[:, 0]
    Th31, Th21 = th[:, 2] - th[:, 0], th[:, 1] - th[:, 0]
    detJ = Th21 * R31 - R21 * Th31
    x, w = ip_triangle7()
    N = sf_N_tri6(x[:, 0], x[:, 1])                           # (7, 6)
    dN = sf_dN_tri6(x[:, 0], x[:, 1])                         # (7, 2, 6)
    N3 = sf_N_tri3(x[:, 0], x[:, 1])                          # (7, 3)
    rho_v = np.asarray(rho)[e2n[:3]].T                        # (nel, 3)
    Ke = np.zeros((nel, 6, 6))
    Me = np.zeros((nel, 6, 6))
    fe = np.zeros((nel, 6))
    for q in range(len(w)):
        thq, rq = th @ N3[q], r @ N3[q]
        rhoq = rho_v @ N3[q]
        sn, cs = np.sin(thq), np.cos(thq)
        cr, sr = cs / rq, sn / rq
        j00 = R31 * cr - Th31 * sn
        j01 = -R21 * cr + Th21 * sn
        j10 = -R31 * sr - Th31 * cs
        j11 = R21 * sr + Th21 * cs
        dx = (j00[:, None] * dN[q, 0] + j01[:, None] * dN[q, 1]) / detJ[:, None]
        dz = (j10[:, None] * dN[q, 0] + j11[:, None] * dN[q, 1]) / detJ[:, None]
        om = rq * detJ * w[q]
        Ke += K * (dx[:, :, None] * dx[:, None, :] + dz[:, :, None] * dz[:, None, :]) * om[:, None, None]
        Me += np.outer(N[q], N[q])[None] * (om * rhoq * Cp)[:, None, None]
        if dQdT != 0.0:
            fe += N[q][None] * (dQdT * dt * om)[:, None]
    rows = np.repeat(e2n.T[:, :, None], 6, axis=2).ravel()
    cols = np.repeat(e2n.T[:, None, :], 6, axis=1).ravel()
    KT = sp.csc_matrix((Ke.ravel(), (rows, cols)), shape=(nn, nn))
    MT = sp.csc_matrix((Me.ravel(), (rows, cols)), shape=(nn, nn))
    F = np.bincount(e2n.T.ravel(), weights=fe.ravel(), minlength=nn)
    return KT, MT, F


def thermal_step(mesh, T0, rho, K, Cp, dt, fixed, vfix, dQdT=0.0):
    """T after one step; fixed: 0-based nodes with values vfix."""
    KT, MT, F = thermal_matrices(mesh, rho, K, Cp, dQdT, dt)
    rhs = MT @ T0 + F
    A = (MT + dt * KT).tocsc()
    T = np.array(T0, dtype=np.float64, copy=True)
    T[fixed] = vfix
    xf = np.zeros_like(T)
    xf[fixed] = vfix
    rhs = rhs - A @ xf
    free = np.setdiff1d(np.arange(T.size), fixed)
    T[free] = spla.spsolve(A[free][:, free].tocsc(), rhs[free])
    return T
