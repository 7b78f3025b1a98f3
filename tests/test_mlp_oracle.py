"""Pins of the learned-model oracle (CPU only)."""
import numpy as np
import pytest

from dimspers_amd.inputs import make_inputs, scaled_workload
from oracle import mlp_oracle as mo


def _params(seed=0):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(mo.N_PARAM) * 0.2
    norm = np.concatenate([rng.standard_normal(5) * 0.1, 1 + rng.random(5), rng.standard_normal(6) * 0.1,
                           1 + rng.random(6)])
    return p, norm


def test_parameter_count_matches_the_architecture():
    assert mo.N_PARAM == 5 * 64 + 64 + 2 * (64 * 64 + 64) + 64 * 6 + 6 == 9094


def test_exact_gradient_matches_central_differences():
    p, norm = _params(1)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((40, 5))
    T = rng.standard_normal((40, 6))
    loss, g = mo.gradient(p, norm, X, T, bf=False)
    for k in list(rng.choice(mo.N_PARAM, 40, replace=False)) + [mo.OFF["b4"][0], mo.OFF["W1"][0]]:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (mo.gradient(pp, norm, X, T, bf=False)[0] - mo.gradient(pm, norm, X, T, bf=False)[0]) / (2 * h)
        assert fd == pytest.approx(g[k], rel=1e-5, abs=1e-9)


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-3])
    r = mo.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 1e-3) / 1e-3 < 2 ** -8


def test_local_schur_blocks_are_spd_and_close_to_the_pseudo_inverse_form():
    inp = make_inputs(scaled_workload(3, 4, 16))
    S6, X, S = mo.local_schur(inp, inp.phi, inp.delta)
    ev = np.linalg.eigvalsh(S)
    assert np.all(ev > 0)
    # against G^T K^+ G with the pseudo-inverse that treats the near-null element rotation
    # (eigenvalue ~1e-7 of the largest at this mesh) as null, like the rigid-motion term
    from dimspers_amd.inputs import StokesInputs, fabric_tensor
    from oracle import stokes_oracle as so
    comps = fabric_tensor(inp.phi, inp.delta)
    C6 = tuple(np.repeat(c[:, None], 6, axis=1) for c in comps)
    Ke, Ge, _, _ = so.element_matrices(StokesInputs(inp.mesh, inp.workload, inp.T, inp.rho, inp.eta_ip, C6,
                                                    inp.phi, inp.delta))
    Sp = np.einsum("eai,eab,ebj->eij", Ge, np.linalg.pinv(Ke, rcond=1e-6, hermitian=True), Ge)
    rel = np.abs(Sp - S).max(axis=(1, 2)) / np.abs(S).max(axis=(1, 2))
    assert rel.max() < 1e-4
    # log-Cholesky parameters reproduce the blocks
    L = mo.log_cholesky(S6)
    l00, l10, l11, l20, l21, l22 = np.exp(L[:, 0]), L[:, 1], np.exp(L[:, 2]), L[:, 3], L[:, 4], np.exp(L[:, 5])
    assert np.allclose(l00 ** 2, S6[:, 0]) and np.allclose(l20 ** 2 + l21 ** 2 + l22 ** 2, S6[:, 5])
    # features
    assert np.allclose(X[:, 3], np.log(inp.delta))
    assert np.allclose(X[:, 1] ** 2 + X[:, 2] ** 2, 1.0)
