"""Host-side mirror of the reference Stokes interface, backed by the GPU library.

The reference call chain (src/Stokes/StokesSolver.jl:7-54) is
``solveStokes -> assembly_stokes_cylindric -> _prepare_matrices -> _apply_bcs ->
StokesPcCG -> updatevelocity2``.  The functions here keep those names, argument
meanings and return values; Julia's duck-typed containers become small Python
objects with the same fields (``rho.ρ`` -> ``rho.rho``, ``eta.ip``, the
six-field ``StiffnessTensor``, ``coordinates.θ/.r`` -> ``.theta/.r``,
``UBC.Ω/.ifree/.vfix``).  Sparse results are scipy CSC matrices with 0-based
indices.  Every call runs on the GPU through ``libdimspers_stokes.so``; there
is no CPU path.
"""
from __future__ import annotations

from dataclasses import dataclass
from types import SimpleNamespace

import numpy as np

from . import _lib
from .mesh import AnnulusMesh


@dataclass
class StiffnessTensor:
    """``StiffnessTensor`` (src/Rheology/anisotropy.jl:1-8): six nel x nip arrays."""

    η11: np.ndarray
    η33: np.ndarray
    η55: np.ndarray
    η13: np.ndarray
    η15: np.ndarray
    η35: np.ndarray

    def components(self):
        return (self.η11, self.η33, self.η55, self.η13, self.η15, self.η35)


def tensor_from_inputs(inp) -> StiffnessTensor:
    return StiffnessTensor(*inp.C6)


def boundary_conditions(mesh: AnnulusMesh):
    """``velocity_bcs`` (src/Setup/boundary_conditions.jl:15-37), 0-based."""
    fixed, ifree = mesh.free_slip()
    return SimpleNamespace(Ω=fixed, ifree=ifree, vfix=np.zeros(fixed.size))


_contexts: dict = {}


def context_for(gr: AnnulusMesh) -> _lib.Context:
    """One device context per mesh object (patterns and slots are built once)."""
    key = id(gr)
    ent = _contexts.get(key)
    if ent is None or ent[0] is not gr:
        ent = (gr, _lib.Context.from_mesh(gr))
        _contexts[key] = ent
    return ent[1]


def release_contexts():
    for _, ctx in _contexts.values():
        ctx.close()
    _contexts.clear()


def _fields(ctx, g, ρ, η, 𝓒):
    C6 = None if 𝓒 is None else 𝓒.components()
    ctx.set_fields(g, ρ.rho, η.ip, C6)


def assembly_stokes_cylindric(gr, coordinates, g, ρ, η, 𝓒, εII=None, T=None, P=None, PhaseID=None,
                              KKidx=None, GGidx=None, MMidx=None, isotropic=False):
    """GPU assembly with the reference's outputs (StokesSolver.jl:247-404).

    Returns (KK, GG, MM, Fb, η): KK and MM lower-triangular, all in the
    Cartesian frame, no boundary conditions applied.  ``coordinates`` must be
    the element vertex coordinates of ``gr`` (they are part of the mesh set-up
    on the device); ``εII, T, P, PhaseID`` and the triplet index lists are
    unused by the reference's live path and accepted for signature parity.
    """
    ctx = context_for(gr)
    if coordinates is not None:
        assert np.array_equal(np.asarray(coordinates.theta), gr.el_theta)
    _fields(ctx, g, ρ, η, None if isotropic else 𝓒)
    flags = _lib.ASM_FP64 | _lib.ASM_CARTESIAN | _lib.ASM_NO_BC | (_lib.ASM_ISOTROPIC if isotropic else 0)
    ctx.assemble(flags)
    KK = ctx.export_csc(_lib.MAT_K, lower_only=True)
    GG = ctx.export_csc(_lib.MAT_G)
    MM = ctx.export_csc(_lib.MAT_M, lower_only=True)
    Fb = ctx.export_rhs()
    return KK, GG, MM, Fb, η


def assemble_rotated(gr, g, ρ, η, 𝓒, precision=("fp64",), bc=True, multigrid=True):
    """Device-resident operators after ``_prepare_matrices`` (and ``_apply_bcs``).

    The rotation TT^T K TT (StokesSolver.jl:192-201) is applied element by
    element; with ``bc`` the radial dofs of both walls are fixed.
    """
    ctx = context_for(gr)
    _fields(ctx, g, ρ, η, 𝓒)
    flags = 0
    if "fp64" in precision:
        flags |= _lib.ASM_FP64
    if "fp32" in precision:
        flags |= _lib.ASM_FP32
    if not bc:
        flags |= _lib.ASM_NO_BC
    if not multigrid:
        flags |= _lib.ASM_NO_MG
    ctx.assemble(flags)
    return ctx


def StokesPcCG(ctx: _lib.Context, method="plain", rtol=1e-8, itmax=200, itmin=10, stop="schur", inner="mg",
               **kw):
    """Schur-complement PCG (src/IterativeSolvers.jl:94-168) on device-resident operators.

    ``method`` is "plain" (FP64, the reference algorithm), "mixed" (FP64
    refinement around FP32 inner work) or "learned".  Returns (U, P, report)
    with U in the (u_theta, u_r) node frame before net-rotation removal.
    """
    o = _lib.default_opts(
        method={"plain": _lib.SOLVE_PLAIN, "mixed": _lib.SOLVE_MIXED, "learned": _lib.SOLVE_LEARNED}[method],
        inner={"mg": _lib.INNER_MG, "jacobi": _lib.INNER_JACOBI}[inner],
        stop={"schur": _lib.STOP_SCHUR, "full": _lib.STOP_FULL}[stop],
        rtol=rtol, itmax=itmax, itmin=itmin, **kw)
    rep = ctx.solve(o)
    U, P = ctx.solution()
    return U, P, rep


def updatevelocity2(ctx: _lib.Context):
    """Net-rotation removal on the device solution (StokesSolver.jl:170-184).

    Returns (U, Ucart, omega)."""
    return ctx.updatevelocity()


def solveStokes(U, P, gr, g, ρ, η, 𝓒, UBC=None, solver="plain", **kw):
    """``solveStokes`` (StokesSolver.jl:7-54) on the GPU.

    Writes the solution into ``U`` and ``P`` and returns
    (Ucartesian, Upolar, U, Ucart, P, report) like the reference (the
    Cartesian/polar point lists are returned as (nnod, 2) arrays).
    """
    precision = ("fp64",) if solver == "plain" else ("fp64", "fp32")
    ctx = assemble_rotated(gr, g, ρ, η, 𝓒, precision=precision)
    Us, Ps, rep = StokesPcCG(ctx, method=solver, **kw)
    Urot, Ucart, _ = updatevelocity2(ctx)
    U[:] = Urot
    P[:] = Ps
    Upolar = Urot.reshape(-1, 2)
    Ucartesian = Ucart.reshape(-1, 2)
    return Ucartesian, Upolar, U, Ucart, P, rep
