"""Host-side mirror of the reference's fabric chain, backed by the GPU library.

MAIN.jl:184-185 and 266-269 run ``FSE, F = getFSE(F, FSE)`` (src/Stress/FSE.jl:131-355) and
``𝓒 = anisotropic_tensor(FSE.fse, D, Valη)`` (src/Rheology/anisotropy.jl:140-285) with
``D = getDEM(file, Δη, ϕ)`` (:60-93).  The names and argument meanings are kept; the pieces that
live in the external Fabrics.jl package are inputs here: the viscosity cut-off ``w`` =
``weakening(Δη, ϕ)`` and ``fabric_shape(a1, 1, parameterization)``, given as (a1 nodes, r1, r2)
samples.  The HDF5 read of ``DEMloader`` stays with the caller (the arrays "C", "ax", "vol").
Every call runs on the GPU through ``libdimspers_stokes.so``; there is no CPU path.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from .stokes import StiffnessTensor

HEALING = {"NoHealing": _lib.FSE_NO_HEALING, "FabricDestructionOnly": _lib.FSE_FABRIC_DESTRUCTION,
           "AnnealingFabricDestruction": _lib.FSE_ANNEALING_DESTRUCTION}


@dataclass
class FiniteStrain:
    """``FiniteStrain`` (FSE.jl:33-67): the ellipses of every point (nel x nip x 6: x1, x2, y1, y2,
    a1, a2) and the healing type its constructor picks from (ϵ, annealing_rate)."""

    fse: np.ndarray
    ϵ: float = 1e3
    annealing_rate: float = 0.0

    @classmethod
    def new(cls, nel, nip=6, ϵ=1e3, annealing_rate=0.0):
        fse = np.tile(np.array([1.0, 0.0, 0.0, 1.0, 1.0, 1.0]), (nel, nip, 1))
        return cls(fse, ϵ, annealing_rate)

    @property
    def healing(self) -> int:
        if self.annealing_rate != 0:
            if self.ϵ == 0:
                raise NotImplementedError("AnnealingOnly (FSE.jl:233-241) reads undefined names in the reference")
            return _lib.FSE_ANNEALING_DESTRUCTION
        return _lib.FSE_FABRIC_DESTRUCTION if self.ϵ != 0 else _lib.FSE_NO_HEALING


@dataclass
class DEM:
    """``DEM`` (anisotropy.jl:30-41) as the device holds it, plus the fabric-shape samples."""

    𝓒: np.ndarray
    a: np.ndarray
    ϕ: float
    w: float
    shape_a1: np.ndarray
    shape_r1: np.ndarray
    shape_r2: np.ndarray


def getDEM(tensor, axes, volume, w, shape_a1, shape_r1, shape_r2) -> DEM:
    """``getDEM`` on the arrays DEMloader reads; w = weakening(Δη, ϕ); fabric_shape sampled."""
    return DEM(np.asarray(tensor, dtype=np.float64), np.asarray(axes, dtype=np.float64), float(volume), float(w),
               *(np.asarray(v, dtype=np.float64) for v in (shape_a1, shape_r1, shape_r2)))


def getFSE(ctx: _lib.Context, F, FSE: FiniteStrain):
    """``FSE, F = getFSE(F, FSE)``: F is nel x nip x 2 x 2; returns the updated pair."""
    ell, F = ctx.fse(F, mode=FSE.healing, eps=FSE.ϵ)
    return FiniteStrain(ell, FSE.ϵ, FSE.annealing_rate), F


def anisotropic_tensor(ctx: _lib.Context, FSE: FiniteStrain | None, D: DEM, install=0) -> StiffnessTensor:
    """``𝓒 = anisotropic_tensor(FSE.fse, D, Valη)``: six nel x 7 arrays.  FSE = None uses the
    ellipses the last getFSE left on the device; ``install`` also writes the result into the
    context's assembly and/or stress fields (``_lib.INSTALL_*``)."""
    ctx.set_dem(D.𝓒, D.a, D.w)
    fse = None if FSE is None else FSE.fse
    comps = ctx.anisotropic_tensor(fse=fse, shape_table=(D.shape_a1, D.shape_r1, D.shape_r2), install=install)
    return StiffnessTensor(*comps)
