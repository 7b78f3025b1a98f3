"""Seeded synthetic inputs for the Stokes path (host-side setup).

The reference builds these fields in its driver (``MAIN.jl:97-146``) from a
temperature field, an Arrhenius viscosity and a DEM fabric lookup.  The five
workloads here replace that with closed-form generators so the GPU path and
the oracle see identical arrays:

* temperature at every P2 node
  ``T = (r_out - r)/(r_out - r_in) + 0.05 cos(4 theta) sin(pi (r - r_in)) + 1e-3 u``
  (noise in [0, amp) as in ``src/Setup/Setup.jl:87``);
* ``eta.ip[e, q] = exp(-ln(C) T_q)`` with ``T_q`` the P2 interpolant at
  integration point q (6-point rule, ``src/Stokes/StokesSolver.jl:262,608``);
* one fabric director angle ``phi_e`` per element and a normal/shear ratio
  ``delta_e``; the viscous tensor is the fabric-frame tensor
  (a = c = 4/3, h = -2/3, e = 1/delta) rotated by ``directRotation2D``
  (``src/Rheology/anisotropy.jl:287-292,503-536``) and copied to every ip;
* density ``rho = Ra T`` at every node with Ra = 1e5 and g = 1, so the body
  force is ``Ra T r_hat`` (``StokesSolver.jl:838-845``).

Random numbers come from SplitMix64, one stream per field:
seed = 0x5EED0000 + 16*workload + stream, u = (x >> 11) * 2**-53.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mesh import R_IN, R_OUT, AnnulusMesh, build_annulus, structured_ids
from .quadrature import ip_triangle6, sf_N_tri6

RA = 1.0e5
GRAV = 1.0

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """n uniforms in [0, 1) from SplitMix64 started at ``seed``."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def stream_seed(workload: int, stream: int) -> int:
    return 0x5EED0000 + 16 * workload + stream


@dataclass(frozen=True)
class Workload:
    """One of the five benchmark / parity workloads."""

    index: int
    ncell_r: int
    ncell_t: int
    contrast: float          # eta = exp(-ln(contrast) T)
    ratio: str               # "const10" | "uniform1_100" | "loguniform1_1e3"
    smooth_director: bool = False


WORKLOADS = {
    1: Workload(1, 32, 256, 1.0e3, "const10"),
    2: Workload(2, 128, 1024, 1.0e4, "const10"),
    3: Workload(3, 256, 2048, 1.0e4, "uniform1_100"),
    4: Workload(4, 256, 2048, 1.0e4, "uniform1_100"),
    5: Workload(5, 512, 4096, 1.0e6, "loguniform1_1e3", smooth_director=True),
}


def scaled_workload(index: int, ncell_r: int, ncell_t: int) -> Workload:
    """The field rules of workload ``index`` on a different mesh size (tests)."""
    w = WORKLOADS[index]
    return Workload(w.index, ncell_r, ncell_t, w.contrast, w.ratio, w.smooth_director)


def rotmatrix(x, y):
    """``rotmatrix`` (anisotropy.jl:287-292): normalised (cos a, sin a, -sin a)."""
    h = 1.0 / np.sqrt(y * y + x * x)
    return x * h, y * h, -y * h


def direct_rotation_2d(R, a, c, e, h):
    """Rotate a 2-D orthotropic viscous tensor (anisotropy.jl:503-536).

    ``R`` = (x1, x2, y1) with x1 = y2 = cos, x2 = sin, y1 = -sin; (a, c, e, h)
    are the fabric-frame C11, C33, C55, C13.  Returns the Cartesian
    (eta11, eta33, eta55, eta13, eta15, eta35).  The polynomial is the
    standard fourth-order-tensor rotation, so the products must match the
    reference's term for term.
    """
    x1, x2, y1 = R
    y2 = x1
    x1x1, x2x2, y1y1, y2y2 = x1 * x1, x2 * x2, y1 * y1, y2 * y2
    s = x1 * y2 + x2 * y1
    n11 = (c * x2x2 + h * x1x1 + 4.0 * e * x1x1) * x2x2 + (a * x1x1 + h * x2x2) * x1x1
    n33 = (c * y2y2 + h * y1y1 + 4.0 * e * y1y1) * y2y2 + (a * y1y1 + h * y2y2) * y1y1
    n55 = (c * x2 * y2 + h * x1 * y1) * x2 * y2 + s * s * e + (a * x1 * y1 + h * x2 * y2) * x1 * y1
    n13 = ((c * x2x2 + h * x1x1) * y2 + 4.0 * e * x1 * x2 * y1) * y2 + (a * x1x1 + h * x2x2) * y1y1
    n15 = ((c * x2x2 + h * x1x1) * y2 + 2.0 * s * e * x1) * x2 + (a * x1x1 + h * x2x2) * x1 * y1
    n35 = ((c * y2y2 + h * y1y1) * x2 + 2.0 * s * e * y1) * y2 + (a * y1y1 + h * y2y2) * x1 * y1
    return n11, n33, n55, n13, n15, n35


def fabric_tensor(phi: np.ndarray, delta: np.ndarray):
    """Cartesian tensor of a fabric with director ``phi`` and ratio ``delta``."""
    R = rotmatrix(np.cos(phi), np.sin(phi))
    return direct_rotation_2d(R, 4.0 / 3.0, 4.0 / 3.0, 1.0 / delta, -2.0 / 3.0)


@dataclass
class StokesInputs:
    """The arrays ``assembly_stokes_cylindric`` reads, in reference layout."""

    mesh: AnnulusMesh
    workload: Workload
    T: np.ndarray            # (nnod,)
    rho: np.ndarray          # (nnod,) rho.rho
    eta_ip: np.ndarray       # (nel, 6) Fortran order, eta.ip
    C6: tuple                # six (nel, 6) arrays: eta11, eta33, eta55, eta13, eta15, eta35
    phi: np.ndarray          # (nel,) director angle
    delta: np.ndarray        # (nel,) normal/shear ratio
    g: float = GRAV


def _smooth_director(mesh: AnnulusMesh, phi: np.ndarray, sigma_cells: float = 5.0):
    """Gaussian smoothing of (cos 2phi, sin 2phi) on the cell grid, periodic in theta."""
    ir, it, _ = structured_ids(mesh)
    nr, nt = mesh.ncell_r, mesh.ncell_t
    c2 = np.zeros((nr, nt))
    s2 = np.zeros((nr, nt))
    np.add.at(c2, (ir, it), np.cos(2 * phi))
    np.add.at(s2, (ir, it), np.sin(2 * phi))
    rad = int(np.ceil(3 * sigma_cells))
    k = np.exp(-0.5 * (np.arange(-rad, rad + 1) / sigma_cells) ** 2)
    k /= k.sum()

    def conv_t(f):  # periodic in theta
        out = np.zeros_like(f)
        for o, w in zip(range(-rad, rad + 1), k):
            out += w * np.roll(f, o, axis=1)
        return out

    def conv_r(f):  # zero-flux (reflect) in r
        pad = np.pad(f, ((rad, rad), (0, 0)), mode="reflect")
        out = np.zeros_like(f)
        for idx, w in enumerate(k):
            out += w * pad[idx : idx + nr]
        return out

    c2 = conv_r(conv_t(c2))
    s2 = conv_r(conv_t(s2))
    ang = 0.5 * np.arctan2(s2, c2)
    ang = np.mod(ang, np.pi)
    return ang[ir, it]


def make_inputs(workload: Workload | int, mesh: AnnulusMesh | None = None) -> StokesInputs:
    if isinstance(workload, int):
        workload = WORKLOADS[workload]
    if mesh is None:
        mesh = build_annulus(workload.ncell_r, workload.ncell_t)
    w = workload.index
    u_T = splitmix64_uniform(stream_seed(w, 0), mesh.nnod)
    th, r = mesh.theta, mesh.r
    T = (R_OUT - r) / (R_OUT - R_IN) + 0.05 * np.cos(4.0 * th) * np.sin(np.pi * (r - R_IN)) + 1.0e-3 * u_T

    x_ip, _ = ip_triangle6()
    N = sf_N_tri6(x_ip[:, 0], x_ip[:, 1])          # (6 ips, 6 nodes)
    Tq = (N @ T[mesh.e2n]).T                       # (nel, 6)
    eta_ip = np.asfortranarray(np.exp(-np.log(workload.contrast) * Tq))

    phi = np.pi * splitmix64_uniform(stream_seed(w, 1), mesh.nel)
    if workload.smooth_director:
        phi = _smooth_director(mesh, phi)
    u_d = splitmix64_uniform(stream_seed(w, 2), mesh.nel)
    if workload.ratio == "const10":
        delta = np.full(mesh.nel, 10.0)
    elif workload.ratio == "uniform1_100":
        delta = 1.0 + 99.0 * u_d
    elif workload.ratio == "loguniform1_1e3":
        delta = 10.0 ** (3.0 * u_d)
    else:
        raise ValueError(workload.ratio)
    comps = fabric_tensor(phi, delta)
    C6 = tuple(np.asfortranarray(np.repeat(cmp[:, None], 6, axis=1)) for cmp in comps)
    rho = RA * T
    return StokesInputs(mesh, workload, T, rho, eta_ip, C6, phi, delta)
