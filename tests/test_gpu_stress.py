"""GPU parity of the stress post-processor and the deformation gradient (C ABI) against
the FP64 oracle, and the exact-gradient property at the full 256 x 2048 size."""
import numpy as np
import pytest

from dimspers_amd import _lib
from dimspers_amd.inputs import make_inputs, scaled_workload
from oracle import stress_oracle as so

pytestmark = pytest.mark.gpu

FIELDS = _lib.STRESS_FIELDS


def _ctx(inp):
    print(f"[case] stress mesh {inp.mesh.ncell_r}x{inp.mesh.ncell_t}", flush=True)
    ctx = _lib.Context.from_mesh(inp.mesh)
    ctx.set_fields(inp.g, inp.rho, inp.eta_ip, inp.C6)
    return ctx


def _fields7(inp, rng):
    """eta.ip and tensor arrays with 7 columns: seeded eta, the workload tensor per element."""
    nel = inp.mesh.nel
    eta = np.asfortranarray(np.exp(rng.normal(size=(nel, 7))))
    C7 = [np.asfortranarray(np.repeat(np.asarray(c)[:, :1], 7, axis=1)) for c in inp.C6]
    return eta, C7


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("mesh", [(8, 64, 3), (32, 256, 4), (7, 18, 5)])
@pytest.mark.parametrize("aniso", [False, True])
def test_stress_matches_oracle(mesh, aniso):
    nr, nt, w = mesh
    inp = make_inputs(scaled_workload(w, nr, nt))
    rng = np.random.default_rng(11)
    eta, C7 = _fields7(inp, rng)
    U = rng.normal(size=inp.mesh.nU)
    ctx = _ctx(inp)
    ctx.set_stress_fields(eta, C7 if aniso else None, nip=7)
    got = ctx.stress(U)
    want = so.stress(inp.mesh, U, eta, C7 if aniso else None)
    errs = {k: _rel(got[k], want[k]) for k in FIELDS}
    print(f"[result] stress {nr}x{nt} aniso={aniso}: max rel err {max(errs.values()):.2e}")
    for k in FIELDS:
        assert errs[k] < 1e-12, (k, errs[k])
    ctx.close()


def test_stress_on_the_device_velocity_after_a_solve():
    """Solve, remove the net rotation, then stress! on the device Cartesian velocity (NULL Ucart),
    as MAIN.jl:217-243 chains them."""
    inp = make_inputs(scaled_workload(3, 8, 64))
    ctx = _ctx(inp)
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
    rep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_MIXED, stop=_lib.STOP_FULL, rtol=1e-9))
    assert rep.converged
    _, Ucart, _ = ctx.updatevelocity()
    rng = np.random.default_rng(2)
    eta, C7 = _fields7(inp, rng)
    ctx.set_stress_fields(eta, C7)
    got = ctx.stress()                     # device velocity
    want = so.stress(inp.mesh, Ucart, eta, C7)
    for k in FIELDS:
        assert _rel(got[k], want[k]) < 1e-12, k
    ctx.close()


@pytest.mark.parametrize("mesh", [(8, 64, 3), (32, 256, 4), (7, 18, 5)])
def test_deformation_gradient_matches_oracle(mesh):
    nr, nt, w = mesh
    inp = make_inputs(scaled_workload(w, nr, nt))
    rng = np.random.default_rng(4)
    eta, _ = _fields7(inp, rng)
    ctx = _ctx(inp)
    ctx.set_stress_fields(eta, nip=7)
    U = rng.normal(size=inp.mesh.nU)
    F0 = rng.normal(size=(inp.mesh.nel, 7, 2, 2)) * 0.1 + np.eye(2)
    got = ctx.deformation_gradient(F0, 0.02, U)
    want = so.deformation_gradient(inp.mesh, U, F0, 0.02)
    err = _rel(got, want)
    print(f"[result] deformation gradient {nr}x{nt}: max rel err {err:.2e}")
    assert err < 1e-13
    # two steps in a row keep matching (F stays on the host between calls)
    got2 = ctx.deformation_gradient(got, 0.02, U)
    assert _rel(got2, so.deformation_gradient(inp.mesh, U, want, 0.02)) < 1e-13
    ctx.close()


def test_linear_velocity_gives_exact_gradients_at_full_size():
    """Size-independent property at 256 x 2048: a velocity linear in (theta, r) has the
    closed-form Cartesian gradient at every point away from the theta = 0 seam."""
    inp = make_inputs(scaled_workload(4, 256, 2048))
    m = inp.mesh
    a, b, c, d, e, f = 0.7, -1.3, 0.2, 2.1, 0.4, -0.5
    U = np.empty(m.nU)
    U[0::2] = a * m.theta + b * m.r + c
    U[1::2] = d * m.theta + e * m.r + f
    ctx = _lib.Context.from_mesh(m)
    print("[case] stress linear field 256x2048", flush=True)
    ctx.set_stress_fields(np.ones((m.nel, 7), order="F"), nip=7)
    got = ctx.stress(U)
    _, _, thq, rq = so.cartesian_gradients(m)
    ok = np.all((m.theta[m.e2n] > 0.5) & (m.theta[m.e2n] < 2 * np.pi - 0.5), axis=0)
    sn, cs = np.sin(thq), np.cos(thq)
    exz = 0.5 * ((b * cs - a * sn / rq) + (e * sn + d * cs / rq))
    for k, want in (("exx", b * sn + a * cs / rq), ("ezz", e * cs - d * sn / rq), ("exz", exz)):
        err = np.abs(got[k][ok] - want[ok]).max()
        print(f"[result] {k}: {err:.2e}")
        assert err < 1e-9 * np.abs(want).max()
    ctx.close()


def test_stress_error_paths():
    inp = make_inputs(scaled_workload(3, 4, 16))
    ctx = _lib.Context.from_mesh(inp.mesh)
    with pytest.raises(_lib.StkError) as e:
        ctx.s_nip = 7
        ctx.stress(np.zeros(inp.mesh.nU))
    assert e.value.code == _lib.STK_E_STATE
    with pytest.raises(_lib.StkError) as e:
        ctx.set_stress_fields(np.ones((inp.mesh.nel, 7), order="F"), nip=5)
    assert e.value.code == _lib.STK_E_SHAPE
    ctx.close()


@pytest.mark.parametrize("mesh", [(8, 64, 3), (32, 256, 4), (7, 18, 5)])
def test_stats_match_oracle(mesh):
    from oracle import stats_oracle
    nr, nt, w = mesh
    inp = make_inputs(scaled_workload(w, nr, nt))
    rng = np.random.default_rng(9)
    U = rng.normal(size=inp.mesh.nU)
    T = rng.uniform(size=inp.mesh.theta.size)
    ctx = _lib.Context.from_mesh(inp.mesh)
    print(f"[case] stats {nr}x{nt}", flush=True)
    got = ctx.stats(T, U)
    want = stats_oracle.write_stats(inp.mesh, U, T)
    for g, wv in zip(got, want):
        assert abs(g - wv) <= 1e-12 * abs(wv), (got, want)
    ctx.close()


def test_stats_closed_forms_at_full_size():
    """256 x 2048: a uniform unit velocity and T = 1 give Urms^2 = <T> = polygon area / annulus area."""
    inp = make_inputs(scaled_workload(4, 256, 2048))
    m = inp.mesh
    ctx = _lib.Context.from_mesh(m)
    print("[case] stats closed form 256x2048", flush=True)
    U = np.tile([1.0, 0.0], m.theta.size)
    urms, tav, _, _ = ctx.stats(np.ones(m.theta.size), U)
    ratio = 0.5 * m.ncell_t * np.sin(2 * np.pi / m.ncell_t) / np.pi
    assert abs(tav - ratio) < 1e-12
    assert abs(urms ** 2 - ratio) < 1e-12
    ctx.close()


def test_two_contexts_keep_their_own_stress_rules():
    """Context A uses the 6-point rule, context B the 7-point rule; each computes stress! and
    deformation_gradient! after the other's set-up, and both still match the oracle with their own
    rule (the rule travels with each launch, it is not shared by the contexts of a process)."""
    inp = make_inputs(scaled_workload(3, 8, 64))
    rng = np.random.default_rng(21)
    eta, C7 = _fields7(inp, rng)
    U = rng.normal(size=inp.mesh.nU)
    a, b = _ctx(inp), _ctx(inp)
    a.set_stress_fields(eta, C7, nip=6)
    b.set_stress_fields(eta, C7, nip=7)
    for ctx, nip in ((a, 6), (b, 7), (a, 6)):
        got = ctx.stress(U)
        want = so.stress(inp.mesh, U, eta, C7, nip=nip)
        assert got["exx"].shape == (inp.mesh.nel, nip)
        for k in FIELDS:
            assert _rel(got[k], want[k]) < 1e-12, (nip, k)
        F0 = np.tile(np.eye(2), (inp.mesh.nel, nip, 1, 1)) + 0.01 * rng.normal(size=(inp.mesh.nel, nip, 2, 2))
        gotF = ctx.deformation_gradient(F0, 0.02, U)
        assert _rel(gotF, so.deformation_gradient(inp.mesh, U, F0, 0.02, nip=nip)) < 1e-13, nip
    print("[result] two contexts with 6- and 7-point stress rules match the oracle", flush=True)
    a.close()
    b.close()
