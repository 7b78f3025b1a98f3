"""FP64 CPU oracle of the Stokes hot path -- TEST INFRASTRUCTURE ONLY.

This module is the oracle of the GPU library.  Only ``tests/``, the smoke
check and the CPU-baseline leg of ``bench.py`` may import it; the product
path never calls it.

It restates, in vectorised numpy/scipy, the reference Julia path:

* ``element_matrices``      -- src/Stokes/StokesSolver.jl:247-404 (blocked assembly,
  6-point rule), :522-541 (_fill_R_J!), :543-639 (_ip_loop!), :641-686
  (_derivative_weights!), src/src/Stokes/StokesSolver.jl:744-808 (unclamped
  anisotropic _fill_Kblk!), :812-845 (_fill_Gblk!, _fill_Mblk!, _fill_Fbblk!),
  :1017-1022 (_element_density).
* ``assembly_stokes_cylindric`` -- StokesSolver.jl:488-520 with the triplet
  layout of src/Algebra/Sparsity.jl:31-103 (lower-triangular K and M, full G)
  and ``accumarray`` (src/Utilities.jl:1-8).
* ``rotation_matrix``       -- StokesSolver.jl:203-245.
* ``prepare_matrices``      -- StokesSolver.jl:192-201.
* ``stokes_pcg``            -- src/IterativeSolvers.jl:94-168 with
  ``_preconditioner(:jacobi)`` = lumped inverse (:189-217, :238-240),
  ``rm0space!`` (:172-186), ``_updatesolution!`` (:246-259), Polak-Ribiere beta
  (:163) and the direct inner solve ``lu(K[ifree, ifree])``
  (src/PardisoSolvers.jl:247-278; SuperLU here in place of UMFPACK).
* ``volume_integral``/``updatevelocity2`` -- src/Stress/FSE.jl:412-464,
  StokesSolver.jl:170-184.

Pinning (tests/test_oracle_*.py): quadrature exactness and partition of
unity, exact annulus area, element-matrix symmetry / rank / translation
null spaces, isotropic-tensor equivalence with the isotropic _fill_Kblk!
(StokesSolver.jl:688-742), tensor rotation against an independent fourth-order
tensor rotation, a dense LAPACK solve of the whole saddle-point system, golden
fixtures, and closed forms for each piece a misreading could change:
``rotation_matrix`` maps the node frames onto (theta_hat, r_hat) of
x = r sin(theta), z = r cos(theta); the body force of a density linear in r
has the exact half-annulus sums of g rho r_hat; 1^T M 1 = int 1/eta for
constant and radially varying eta; u^T K_e u equals the viscous energy
sum_q omega_q eta_q eps:C:eps of a fourth-order tensor for a velocity P2
holds exactly (fixes where eta15 and eta35 sit in D), and u^T G_e p equals
sum_q omega_q p div u; ``prepare_matrices`` gives a rigid rotation near-zero
energy and the field (x, z) divergence 2; ``volume_integral`` integrates
1, r, r^2 exactly; ``full_residual`` vanishes on the dense LAPACK solution and
moves with its pressure.  The Polak-Ribiere beta is not pinned apart from the
Fletcher-Reeves form: with exact inner solves the two are equal.  The reference ships no tests or
vectors and Julia is absent, so nothing is compared against a run of the
reference itself.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from dimspers_amd.quadrature import ip_triangle6, ip_triangle7, sf_dN_tri6, sf_N_tri3, sf_N_tri6


# ----------------------------------------------------------------------------- element level
def geometry(el_theta, el_r):
    """_fill_R_J! (StokesSolver.jl:522-541): returns R21, R31, Th21, Th31, detJ."""
    R21 = el_r[1] - el_r[0]
    R31 = el_r[2] - el_r[0]
    Th21 = el_theta[1] - el_theta[0]
    Th31 = el_theta[2] - el_theta[0]
    detJ = Th21 * R31 - Th31 * R21
    return R21, R31, Th21, Th31, detJ


def _rule(nip):
    return ip_triangle7() if nip == 7 else ip_triangle6()


def ip_derivatives(el_theta, el_r, q, nip=6):
    """_derivative_weights! (StokesSolver.jl:641-686) at integration point q of the nip-point rule.

    Returns (dNdx, dNdz) of shape (nel, 6), omega (nel,), theta_q, r_q.
    """
    x_ip, w_ip = _rule(nip)
    N3 = sf_N_tri3(x_ip[q, 0], x_ip[q, 1])
    dNds = sf_dN_tri6(x_ip[q, 0], x_ip[q, 1])      # (2, 6)
    R21, R31, Th21, Th31, detJ = geometry(el_theta, el_r)
    th = N3 @ el_theta
    r = N3 @ el_r
    s, c = np.sin(th), np.cos(th)
    inv = 1.0 / r / detJ
    jx0 = (R31 * c - Th31 * s * r) * inv
    jx1 = (-R21 * c + Th21 * s * r) * inv
    jz0 = (-R31 * s - Th31 * c * r) * inv
    jz1 = (R21 * s + Th21 * c * r) * inv
    dNdx = jx0[:, None] * dNds[0][None, :] + jx1[:, None] * dNds[1][None, :]
    dNdz = jz0[:, None] * dNds[0][None, :] + jz1[:, None] * dNds[1][None, :]
    omega = r * detJ * w_ip[q]
    return dNdx, dNdz, omega, th, r


def element_matrices(inp, isotropic: bool = False, nip: int = 6, clamp: bool = False):
    """Per-element K_e (12x12), G_e (12x3), M_e (3x3), f_e (12) in the Cartesian frame.

    ``nip`` = 6 is the live blocked assembly's rule (StokesSolver.jl:262); 7 the threaded
    assembler's (src/Stokes/AssemblerStokes.jl:174), which reads column q of eta.ip and of the
    tensor at point q.  ``clamp`` applies applybounds(eta C, 1e25, 1e18) to every component of the
    anisotropic D as the top-level _fill_Kblk! does (StokesSolver.jl:756,770-804; src/PiC/PiC.jl:834).

    Local velocity dofs are x1 z1 x2 z2 ... x6 z6.  K_e = sum_q omega_q B^T D_q B
    with B_j = [[dN_j/dx, 0], [0, dN_j/dz], [dN_j/dz, dN_j/dx]] and
    D_q = eta_ip[e, q] * [[C11, C13, C15], [C13, C33, C35], [C15, C35, C55]]
    (the tensor naming of src/src/Stokes/StokesSolver.jl:749-754).
    ``isotropic`` uses D = eta * [[4/3, -2/3, 0], [-2/3, 4/3, 0], [0, 0, 1]]
    (StokesSolver.jl:688-742).
    """
    mesh = inp.mesh
    nel = mesh.nel
    x_ip, _ = _rule(nip)
    K = np.zeros((nel, 12, 12))
    G = np.zeros((nel, 12, 3))
    M = np.zeros((nel, 3, 3))
    f = np.zeros((nel, 12))
    rho_v = inp.rho[mesh.e2n[:3]]                   # (3, nel) vertex densities
    for q in range(nip):
        dNdx, dNdz, om, th, r = ip_derivatives(mesh.el_theta, mesh.el_r, q, nip)
        NP = sf_N_tri3(x_ip[q, 0], x_ip[q, 1])      # (3,)
        Nq = sf_N_tri6(x_ip[q, 0], x_ip[q, 1])      # (6,)
        eta = inp.eta_ip[:, q]
        B = np.zeros((nel, 3, 12))
        B[:, 0, 0::2] = dNdx
        B[:, 1, 1::2] = dNdz
        B[:, 2, 0::2] = dNdz
        B[:, 2, 1::2] = dNdx
        if isotropic:
            D = np.zeros((nel, 3, 3))
            D[:, 0, 0] = D[:, 1, 1] = 4.0 / 3.0
            D[:, 0, 1] = D[:, 1, 0] = -2.0 / 3.0
            D[:, 2, 2] = 1.0
        else:
            c11, c33, c55, c13, c15, c35 = (a[:, q] for a in inp.C6)
            D = np.stack([
                np.stack([c11, c13, c15], -1),
                np.stack([c13, c33, c35], -1),
                np.stack([c15, c35, c55], -1),
            ], -2)
        if clamp and not isotropic:
            D = np.maximum(np.minimum(D * eta[:, None, None], 1e25), 1e18) * om[:, None, None]
        else:
            D = D * (eta * om)[:, None, None]
        K += np.einsum("eia,eij,ejb->eab", B, D, B)
        # G_e[(x_j | z_j), i] += omega NP_i dN_j/d(x | z)   (StokesSolver.jl:812-823)
        G[:, 0::2, :] += (om[:, None, None] * dNdx[:, :, None]) * NP[None, None, :]
        G[:, 1::2, :] += (om[:, None, None] * dNdz[:, :, None]) * NP[None, None, :]
        # M_e += omega/eta NP NP^T                        (StokesSolver.jl:825-836)
        M += (om / eta)[:, None, None] * np.outer(NP, NP)[None]
        # f_e[x_j | z_j] += omega g rho_q N_j (sin | cos)   (StokesSolver.jl:838-845)
        rho_q = NP @ rho_v
        cte = (om * inp.g * rho_q)[:, None] * Nq[None, :]
        f[:, 0::2] += np.sin(th)[:, None] * cte
        f[:, 1::2] += np.cos(th)[:, None] * cte
    return K, G, M, f


def el2dof(e2n):
    """EL2DOF (Sparsity.jl:46-48), 0-based: x dof 2n, z dof 2n+1."""
    d = np.empty((12, e2n.shape[1]), dtype=np.int64)
    d[0::2] = 2 * e2n
    d[1::2] = 2 * e2n + 1
    return d


# ----------------------------------------------------------------------------- global level
def assembly_stokes_cylindric(inp, isotropic: bool = False, nip: int = 6, clamp: bool = False):
    """Global KK (lower triangle), GG, MM (lower triangle), Fb; Cartesian frame.

    Returns scipy CSC matrices (0-based) and Fb, like StokesSolver.jl:247-404.
    """
    mesh = inp.mesh
    nU, nP = mesh.nU, mesh.nP
    Ke, Ge, Me, fe = element_matrices(inp, isotropic, nip, clamp)
    dof = el2dof(mesh.e2n)                           # (12, nel)
    I = np.broadcast_to(dof[:, None, :], (12, 12, mesh.nel))
    J = np.broadcast_to(dof[None, :, :], (12, 12, mesh.nel))
    Kfull = sp.coo_matrix((np.transpose(Ke, (1, 2, 0)).ravel(), (I.ravel(), J.ravel())),
                          shape=(nU, nU)).tocsc()
    KK = sp.tril(Kfull, format="csc")
    pI = np.broadcast_to(dof[:, None, :], (12, 3, mesh.nel))
    pJ = np.broadcast_to(mesh.e2nP[None, :, :], (12, 3, mesh.nel))
    GG = sp.coo_matrix((np.transpose(Ge, (1, 2, 0)).ravel(), (pI.ravel(), pJ.ravel())),
                       shape=(nU, nP)).tocsc()
    mI = np.broadcast_to(mesh.e2nP[:, None, :], (3, 3, mesh.nel))
    mJ = np.broadcast_to(mesh.e2nP[None, :, :], (3, 3, mesh.nel))
    Mfull = sp.coo_matrix((np.transpose(Me, (1, 2, 0)).ravel(), (mI.ravel(), mJ.ravel())),
                          shape=(nP, nP)).tocsc()
    MM = sp.tril(Mfull, format="csc")
    Fb = np.zeros(nU)
    np.add.at(Fb, dof.ravel(), fe.T.ravel())
    for A in (KK, GG, MM):
        A.sum_duplicates()
        A.sort_indices()
    return KK, GG, MM, Fb


def rotation_matrix(theta):
    """Block-diagonal TT, per node [[cos, sin], [-sin, cos]] (StokesSolver.jl:203-245)."""
    n = theta.shape[0]
    c, s = np.cos(theta), np.sin(theta)
    rows = np.concatenate([2 * np.arange(n), 2 * np.arange(n) + 1, 2 * np.arange(n), 2 * np.arange(n) + 1])
    cols = np.concatenate([2 * np.arange(n), 2 * np.arange(n) + 1, 2 * np.arange(n) + 1, 2 * np.arange(n)])
    vals = np.concatenate([c, c, s, -s])
    return sp.csc_matrix((vals, (rows, cols)), shape=(2 * n, 2 * n))


def symmetric_from_lower(L):
    """``SparseMatrixCSC(Symmetric(L, :L))``: full matrix from the lower triangle."""
    L = sp.tril(L, format="csc")
    return (L + sp.tril(L, k=-1, format="csc").T).tocsc()


def prepare_matrices(KK, GG, Fb, TT):
    """_prepare_matrices (StokesSolver.jl:192-201): rotate to (u_theta, u_r) node frames."""
    K = symmetric_from_lower(KK)
    K.eliminate_zeros()
    GG = GG.copy()
    GG.eliminate_zeros()
    K = (TT.T @ K @ TT).tocsc()
    G = (TT.T @ GG).tocsc()
    F = TT.T @ Fb
    K = symmetric_from_lower(K)
    K.sort_indices()
    G.sort_indices()
    return K, G, F


def lumped_preconditioner(MM):
    """_preconditioner(:jacobi) -> lumpedpreconditioner (IterativeSolvers.jl:189-240)."""
    Mfull = symmetric_from_lower(MM)
    return 1.0 / np.asarray(Mfull.sum(axis=1)).ravel()


def rm0space(a):
    """rm0space! (IterativeSolvers.jl:172-186)."""
    a -= a.mean()
    return a


@dataclass
class PCGInfo:
    iterations: int = 0
    converged: bool = False
    residuals: list = field(default_factory=list)


def inner_direct(K, ifree):
    """``lu(K[ifree, ifree])`` (PardisoSolvers.jl:247-256) as a solve callable."""
    Kff = K[ifree][:, ifree].tocsc()
    lu = spla.splu(Kff, permc_spec="COLAMD")
    n = K.shape[0]

    def solve(y):
        z = np.zeros(n)
        z[ifree] = lu.solve(y[ifree])
        return z

    return solve


def null_vector(K, ifree, mesh, iters=3):
    """Lowest eigenvector of K[ifree, ifree] (the discrete rigid rotation) and its eigenvalue.

    Inverse iteration with the sparse LU, started from the nodal rigid rotation
    u_theta = r, u_r = 0.  With free slip on both walls this eigenvalue falls
    like h^6; the reference never forms the vector but its LU solve contains it.
    Not from the reference: used for the projected (null-space) formulation.
    """
    n = K.shape[0]
    Kff = K[ifree][:, ifree].tocsc()
    lu = spla.splu(Kff, permc_spec="COLAMD")
    z = np.zeros(n)
    z[0::2] = mesh.r
    z = z[ifree]
    z /= np.linalg.norm(z)
    v = z.copy()
    for _ in range(iters):
        v = lu.solve(v)
        v /= np.linalg.norm(v)
    if v @ z < 0:
        v = -v
    out = np.zeros(n)
    out[ifree] = v
    return out, float(out @ (K @ out))


def inner_projected(K, ifree, v):
    """Direct inner solve on the complement of the null vector v: z = Pi K^-1 Pi y."""
    Kff = K[ifree][:, ifree].tocsc()
    lu = spla.splu(Kff, permc_spec="COLAMD")
    n = K.shape[0]
    vf = v[ifree]

    def solve(y):
        yf = y[ifree] - (vf @ y[ifree]) * vf
        x = lu.solve(yf)
        x -= (vf @ x) * vf
        z = np.zeros(n)
        z[ifree] = x
        return z

    return solve


def stokes_pcg(K, MM, G, F, ifree, rtol=1e-8, itmax=200, itmin=10, inner=None, pc=None, P0=None):
    """StokesPcCG (IterativeSolvers.jl:94-168).  Returns U, P, PCGInfo.

    ``pc``: the lumped inverse (default), another vector (d = r * pc) or a callable d = pc(r).
    ``P0``: the incoming pressure of StokesPcCG_pardiso (IterativeSolvers.jl:1-17), which keeps P
    (its fill! is commented out) and adds GG * P to the right-hand side; None starts from zero.
    """
    nU, nP = G.shape
    U = np.zeros(nU)
    P = np.zeros(nP) if P0 is None else np.array(P0, dtype=np.float64)
    Rhs = F + G @ P
    solve = inner if inner is not None else inner_direct(K, ifree)
    U = solve(Rhs)
    GT = G.T.tocsr()
    r = -(GT @ U)
    rm0space(r)
    tol = np.linalg.norm(r) * rtol
    if pc is None:
        pc = lumped_preconditioner(MM)
    apply_pc = pc if callable(pc) else (lambda res, v=pc: res * v)
    d = apply_pc(r)
    q = d.copy()
    info = PCGInfo(residuals=[np.linalg.norm(r)])
    for it in range(1, itmax + 1):
        info.iterations = it
        rm0space(q)
        rd = r @ d
        y = G @ q
        z = solve(y)
        Sq = GT @ z
        qSq = q @ Sq
        rlast = r.copy()
        alpha = rd / qSq
        P += alpha * q
        U += alpha * z
        r -= alpha * Sq
        rm0space(r)
        nr = np.linalg.norm(r)
        info.residuals.append(nr)
        if nr < tol and it >= itmin:
            info.converged = True
            break
        d = apply_pc(r)
        beta = ((r - rlast) @ d) / rd
        q = d + beta * q
    return U, P, info


def greedy_colour_order(A):
    """Vertices grouped by a greedy first-fit colouring in natural order (colour 0 first)."""
    A = sp.csr_matrix(A)
    n = A.shape[0]
    colour = -np.ones(n, dtype=np.int64)
    for i in range(n):
        used = {colour[j] for j in A.indices[A.indptr[i]:A.indptr[i + 1]] if colour[j] >= 0}
        c = 0
        while c in used:
            c += 1
        colour[i] = c
    return np.argsort(colour, kind="stable"), colour


def incomplete_cholesky(A, order):
    """IC(0) of the symmetric matrix A in the elimination order ``order``: A ~ L D L^T with L unit
    lower triangular (in that order) on A's pattern.  A pivot that is not positive restarts the
    factorisation of A + alpha diag(A) with alpha = 1e-3, doubled on every restart.

    The pressure preconditioner of StokesPcCG_pardiso is ``lldl(MM)`` (IterativeSolvers.jl:32,
    :242-244) of LimitedLDLFactorizations 0.4.1, an incomplete LDL^T with a fill-reducing
    permutation and limited memory; its factors are parity unpinned (no Julia).  This incomplete
    factorisation of the same matrix is what the library computes; its defining property
    (L D L^T equals A on the pattern) pins it.  Returns (L, D) with L in the original numbering
    (a scipy CSR whose row i holds the entries L_ij of earlier vertices j) and D per vertex."""
    A = sp.csr_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    pos = np.empty(n, dtype=np.int64)
    pos[order] = np.arange(n)
    alpha = 0.0
    while True:
        rows = {}
        D = np.zeros(n)
        ok = True
        for i in order:
            cols = A.indices[A.indptr[i]:A.indptr[i + 1]]
            vals = A.data[A.indptr[i]:A.indptr[i + 1]]
            row = {}
            dii = 0.0
            earlier = sorted((pos[j], j, v) for j, v in zip(cols, vals) if pos[j] < pos[i])
            for j, v in zip(cols, vals):
                if j == i:
                    dii = v * (1.0 + alpha)
            for _, j, v in earlier:
                s_ = v - sum(row[k] * D[k] * lk for k, lk in rows[j].items() if k in row)
                row[j] = s_ / D[j]
                dii -= row[j] ** 2 * D[j]
            rows[i] = row
            D[i] = dii
            if not dii > 1e-12 * abs(A[i, i]):
                ok = False
                break
        if ok:
            break
        alpha = 2.0 * alpha if alpha > 0 else 1e-3
    I = [i for i in rows for _ in rows[i]]
    J = [j for i in rows for j in rows[i]]
    V = [rows[i][j] for i in rows for j in rows[i]]
    L = sp.csr_matrix((V, (I, J)), shape=(n, n))
    return L, D


def ic_preconditioner(MM):
    """d = (L D L^T)^-1 r with the IC(0) of the symmetric MM in the greedy colour order."""
    Mfull = symmetric_from_lower(MM)
    order, _ = greedy_colour_order(Mfull)
    L, D = incomplete_cholesky(Mfull, order)
    Lu = (L + sp.identity(L.shape[0])).tocsr()
    A = (Lu @ sp.diags(D) @ Lu.T).tocsc()
    lu = spla.splu(A)
    return lu.solve


def volume_integral(V, mesh):
    """volume_integral (FSE.jl:412-464): sum_e sum_q (N.V_e) r_q detJ w_q."""
    x_ip, w_ip = ip_triangle6()
    _, _, _, _, detJ = geometry(mesh.el_theta, mesh.el_r)
    Ve = V[mesh.e2n]                                  # (6, nel)
    total = np.zeros(mesh.nel)
    for q in range(6):
        N3 = sf_N_tri3(x_ip[q, 0], x_ip[q, 1])
        N = sf_N_tri6(x_ip[q, 0], x_ip[q, 1])
        r = N3 @ mesh.el_r
        total += (N @ Ve) * r * detJ * w_ip[q]
    return total.sum()


def updatevelocity2(U, rho, mesh):
    """Net-rotation removal (StokesSolver.jl:170-184); returns (U, Ucart)."""
    U = U.copy()
    ut = U[0::2]
    omega = volume_integral(ut * rho, mesh) / volume_integral(mesh.r * rho, mesh)
    U[0::2] -= omega * mesh.r
    TT = rotation_matrix(mesh.theta)
    return U, TT @ U


def full_residual(K, G, F, U, P, ifree, null=None):
    """||[f + G p - K u ; -G^T u - c]|| / ||[f ; 0]|| on free velocity dofs.

    c is the mean of -G^T u: the reference removes it from the continuity
    residual at every iteration (rm0space!, IterativeSolvers.jl:131,153), so
    its solutions satisfy continuity up to a constant.  With ``null`` (the
    near-null vector) the momentum residual is taken on its complement.
    """
    ru = (F + G @ P - K @ U)
    if null is not None:
        ru = ru - (null @ ru) * null
    ru = ru[ifree]
    rp = -(G.T @ U)
    rp = rp - rp.mean()
    return np.sqrt(ru @ ru + rp @ rp) / np.linalg.norm(F[ifree])


@dataclass
class OracleSolution:
    U: np.ndarray
    P: np.ndarray
    Ucart: np.ndarray
    info: PCGInfo
    K: object
    G: object
    MM: object
    F: np.ndarray
    ifree: np.ndarray


def solve_stokes(inp, rtol=1e-8, itmax=200, itmin=10, null_mode="literal", bcs=None):
    """solveStokes (StokesSolver.jl:7-54) with the :suitesparse inner solve.

    null_mode "literal" is the reference algorithm (lu(K[ifree, ifree]));
    "project" solves on the complement of the near-null vector of K.
    ``bcs`` = (fixed, ifree), 0-based (UBC.Omega, UBC.ifree); default free slip.
    """
    mesh = inp.mesh
    KK, GG, MM, Fb = assembly_stokes_cylindric(inp)
    TT = rotation_matrix(mesh.theta)
    K, G, F = prepare_matrices(KK, GG, Fb, TT)
    fixed, ifree = mesh.free_slip() if bcs is None else bcs
    inner = None
    if null_mode == "project":
        v, _ = null_vector(K, ifree, mesh)
        inner = inner_projected(K, ifree, v)
    U, P, info = stokes_pcg(K, MM, G, F, ifree, rtol=rtol, itmax=itmax, itmin=itmin, inner=inner)
    Urot, Ucart = updatevelocity2(U, inp.rho, mesh)
    return OracleSolution(Urot, P, Ucart, info, K, G, MM, F, ifree), U
