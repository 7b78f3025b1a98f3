"""CPU oracle of the learned local-Schur model -- TEST INFRASTRUCTURE ONLY.

The learned preconditioner has no counterpart in the reference; this module
restates the build's own definitions so the GPU kernels can be checked:

* ``local_schur``: S_e = G_e^T (K_e + mu N N^T)^-1 G_e with N the orthonormal
  Cartesian rigid motions of the element nodes and mu = trace(K_e)/12, in the
  canonical vertex order, divided by area / exp(mean log eta); K_e and G_e come
  from the FP64 element oracle (stokes_oracle.element_matrices).
* ``features``: (mean log eta, cos 2(phi - theta_c), sin 2(phi - theta_c),
  log delta, area / mean area of the element's cell -- elements split*q ..
  split*q + split - 1 share cell q); on the four-triangle ``Grid`` the director is
  measured from the corner-corner edge (less pi/2 on the radial edges of types 1 and
  3) with the sine's sign flipped on the mirrored types 2 and 3.
* ``forward`` / ``gradient``: the 5-64-64-64-6 GELU network with the BF16
  rounding points of the GPU kernels (``bf16=True``) or in exact FP64
  (``bf16=False``, pinned against central finite differences).
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

from oracle import stokes_oracle as so

OFF = {}
_o = 0
for _name, _n in (("W1", 5 * 64), ("b1", 64), ("W2", 64 * 64), ("b2", 64), ("W3", 64 * 64), ("b3", 64),
                  ("W4", 64 * 6), ("b4", 6)):
    OFF[_name] = (_o, _n)
    _o += _n
N_PARAM = _o
SHAPES = {"W1": (5, 64), "W2": (64, 64), "W3": (64, 64), "W4": (64, 6)}


def unpack(p):
    out = {}
    for k, (o, n) in OFF.items():
        v = np.asarray(p[o:o + n], dtype=np.float64)
        out[k] = v.reshape(SHAPES[k]) if k in SHAPES else v
    return out


def bf16(x):
    """Round to bfloat16 (nearest even) through float32, as the kernels do."""
    u = np.asarray(x, dtype=np.float32).copy().view(np.uint32)
    u = (u + np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))) & np.uint32(0xFFFF0000)
    return u.view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def forward(p, norm, X, bf=True, keep=False):
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W = unpack(p)
    nm = np.asarray(norm, dtype=np.float64)
    xn = (np.asarray(X, dtype=np.float64) - nm[0:5]) / nm[5:10]
    if bf:
        xn = np.asarray(xn, dtype=np.float32).astype(np.float64)
    h0 = r(xn)
    z1 = h0 @ r(W["W1"]) + W["b1"]
    h1 = r(gelu(z1))
    z2 = h1 @ r(W["W2"]) + W["b2"]
    h2 = r(gelu(z2))
    z3 = h2 @ r(W["W3"]) + W["b3"]
    h3 = r(gelu(z3))
    y = h3 @ r(W["W4"]) + W["b4"]
    if keep:
        return y, (h0, z1, h1, z2, h2, z3, h3)
    return y * nm[16:22] + nm[10:16]


def gradient(p, norm, X, T, bf=True):
    """(loss, gradient) of the mean squared error of the normalised outputs."""
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    W = unpack(p)
    nm = np.asarray(norm, dtype=np.float64)
    y, (h0, z1, h1, z2, h2, z3, h3) = forward(p, norm, X, bf=bf, keep=True)
    tn = (np.asarray(T, dtype=np.float64) - nm[10:16]) / nm[16:22]
    n = y.shape[0]
    d = y - tn
    loss = np.mean(d * d)
    dY = 2.0 * d / (6.0 * n)
    if bf:
        dY = np.asarray(dY, dtype=np.float32).astype(np.float64)
    dZ3 = (r(dY) @ r(W["W4"]).T) * gelu_d(z3)
    dZ2 = (r(dZ3) @ r(W["W3"]).T) * gelu_d(z2)
    dZ1 = (r(dZ2) @ r(W["W2"]).T) * gelu_d(z1)
    g = np.zeros(N_PARAM)
    for name, H, D in (("W1", h0, dZ1), ("W2", h1, dZ2), ("W3", h2, dZ3), ("W4", h3, dY)):
        o, k = OFF[name]
        g[o:o + k] = (H.T @ r(D)).ravel()
    for name, D in (("b1", dZ1), ("b2", dZ2), ("b3", dZ3), ("b4", dY)):
        o, k = OFF[name]
        g[o:o + k] = D.sum(0)
    return loss, g


CANON = ((0, 1, 2), (1, 2, 0))
# four triangles per cell, [corner k, corner k+1, centre]: corners swapped for types 2 and 3 (the
# mirror images of types 0 and 1 in r and in theta)
CANON4 = ((0, 1, 2), (0, 1, 2), (1, 0, 2), (1, 0, 2))


def local_schur(inp, phi, delta, mean_area=None):
    """Exact normalised local Schur blocks (nel x 6) and features (nel x 5)."""
    from dimspers_amd.inputs import StokesInputs, fabric_tensor

    m = inp.mesh
    comps = fabric_tensor(phi, delta)
    C6 = tuple(np.repeat(c[:, None], 6, axis=1) for c in comps)
    Ke, Ge, _, _ = so.element_matrices(StokesInputs(m, inp.workload, inp.T, inp.rho, inp.eta_ip, C6, phi, delta))
    th, r = m.el_theta, m.el_r
    detJ = so.geometry(th, r)[4]
    area = detJ * r.sum(0) / 6.0
    if mean_area is None:
        mean_area = area.mean()
    # node coordinates (vertices, then edge midpoints in (theta, r))
    tn = np.concatenate([th, 0.5 * (th + np.roll(th, -1, axis=0))])
    rn = np.concatenate([r, 0.5 * (r + np.roll(r, -1, axis=0))])
    xn, zn = rn * np.sin(tn), rn * np.cos(tn)
    N = np.zeros((m.nel, 12, 3))
    N[:, 0::2, 0] = 1.0
    N[:, 1::2, 1] = 1.0
    N[:, 0::2, 2] = (zn - zn.mean(0)).T
    N[:, 1::2, 2] = (-(xn - xn.mean(0))).T
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    mu = np.trace(Ke, axis1=1, axis2=2) / 12.0
    Kr = Ke + mu[:, None, None] * np.einsum("eik,ejk->eij", N, N)
    S = np.einsum("eai,eab->eib", Ge, np.linalg.solve(Kr, Ge))
    logeta = np.log(inp.eta_ip[:, :6]).mean(1)
    se = area / np.exp(logeta)
    out = np.zeros((m.nel, 6))
    split = getattr(m, "split", 2)
    canon = CANON4 if split == 4 else CANON
    for t in range(split):
        idx = np.arange(t, m.nel, split)
        pp = canon[t]
        Sc = S[idx][:, pp][:, :, pp] / se[idx, None, None]
        out[idx] = np.stack([Sc[:, 0, 0], Sc[:, 1, 0], Sc[:, 1, 1], Sc[:, 2, 0], Sc[:, 2, 1], Sc[:, 2, 2]], 1)
    thc = th.mean(0)
    psi, sgn = phi - thc, np.ones(m.nel)
    if split == 4:
        # director from the corner-corner edge (radial on types 1 and 3), mirrored types 2 and 3
        t = np.arange(m.nel) % 4
        psi = psi - 0.5 * np.pi * (t % 2)
        sgn = np.where(t >= 2, -1.0, 1.0)
    shape = area / area.reshape(-1, split).mean(1).repeat(split)
    X = np.stack([logeta, np.cos(2 * psi), sgn * np.sin(2 * psi), np.log(delta), shape], 1)
    return out, X, S


def log_cholesky(S6):
    """(log L00, L10, log L11, L20, L21, log L22) of the blocks given as S00 S10 S11 S20 S21 S22."""
    s00, s10, s11, s20, s21, s22 = S6.T
    l00 = np.sqrt(s00)
    l10 = s10 / l00
    l20 = s20 / l00
    l11 = np.sqrt(s11 - l10 ** 2)
    l21 = (s21 - l20 * l10) / l11
    l22 = np.sqrt(s22 - l20 ** 2 - l21 ** 2)
    return np.stack([np.log(l00), l10, np.log(l11), l20, l21, np.log(l22)], 1)
