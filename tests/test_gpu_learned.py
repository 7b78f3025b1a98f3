"""GPU tests of the learned local-Schur preconditioner against its CPU oracle."""
import numpy as np
import pytest

from dimspers_amd import _lib
from dimspers_amd.inputs import make_inputs, scaled_workload
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _ctx(inp, flags=_lib.ASM_FP64 | _lib.ASM_FP32):
    print(f"[case] mesh {inp.mesh.ncell_r}x{inp.mesh.ncell_t} workload {inp.workload.index}", flush=True)
    ctx = _lib.Context.from_mesh(inp.mesh)
    ctx.set_fields(inp.g, inp.rho, inp.eta_ip, inp.C6)
    ctx.set_fabric(inp.phi, inp.delta)
    ctx.assemble(flags)
    return ctx


def _random_model(seed):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(mo.N_PARAM) * 0.15).astype(np.float32)
    norm = np.concatenate([rng.standard_normal(5) * 0.1, 1 + rng.random(5), rng.standard_normal(6) * 0.1,
                           1 + rng.random(6)]).astype(np.float32)
    return p, norm


def test_local_schur_blocks_match_oracle():
    inp = make_inputs(scaled_workload(3, 4, 16))
    ctx = _ctx(inp)
    S, X = ctx.local_schur()
    So, Xo, _ = mo.local_schur(inp, inp.phi, inp.delta)
    assert np.abs(S - So).max() / np.abs(So).max() < 1e-10
    assert np.abs(X - Xo).max() < 1e-5
    ctx.close()


def test_mlp_forward_matches_oracle():
    inp = make_inputs(scaled_workload(3, 4, 16))
    ctx = _ctx(inp)
    p, norm = _random_model(5)
    ctx.mlp_params(p, norm)
    rng = np.random.default_rng(6)
    X = rng.standard_normal((101, 5)).astype(np.float32)     # not a multiple of the 32-sample tile
    Y = ctx.mlp_forward(X)
    Yo = mo.forward(p, norm, X.astype(np.float64), bf=True)
    err = np.abs(Y - Yo).max() / np.abs(Yo).max()
    print(f"[result] forward max rel err {err:.2e}")
    assert err < 5e-3
    # every output column individually (catches a transposed MFMA accumulator layout)
    for j in range(6):
        assert np.corrcoef(Y[:, j], Yo[:, j])[0, 1] > 0.9999
    ctx.close()


def test_mlp_gradient_matches_oracle():
    inp = make_inputs(scaled_workload(3, 4, 16))
    ctx = _ctx(inp)
    p, norm = _random_model(7)
    ctx.mlp_params(p, norm)
    rng = np.random.default_rng(8)
    X = rng.standard_normal((300, 5)).astype(np.float32)
    T = rng.standard_normal((300, 6)).astype(np.float32)
    loss, g = ctx.mlp_gradient(X, T)
    lo, go = mo.gradient(p, norm, X.astype(np.float64), T.astype(np.float64), bf=True)
    assert loss == pytest.approx(lo, rel=1e-4)
    for name, (o, k) in mo.OFF.items():
        a, b = g[o:o + k], go[o:o + k]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"[result] grad {name}: rel L2 err {rel:.2e}")
        assert rel < 2e-2, name
    ctx.close()


def test_training_reduces_the_loss():
    inp = make_inputs(scaled_workload(3, 8, 64))
    ctx = _ctx(inp)
    rep = ctx.mlp_train(_lib.default_train_opts(snapshots=100, per_snapshot=256, epochs=6, batch=512))
    print(f"[result] train: {rep.as_dict()}")
    assert rep.samples == 25600 and rep.steps == 300
    assert rep.final_loss < 0.3 * rep.initial_loss
    assert np.isfinite(rep.validation_loss) and rep.validation_loss < 0.3
    ctx.close()


def test_learned_solve_converges_and_reports_its_gain():
    inp = make_inputs(scaled_workload(3, 32, 256))
    ctx = _ctx(inp)
    rep_t = ctx.mlp_train(_lib.default_train_opts(snapshots=400))
    ctx.learned_setup()
    out = {}
    for name, m in (("mixed", _lib.SOLVE_MIXED), ("learned", _lib.SOLVE_LEARNED), ("plain", _lib.SOLVE_PLAIN)):
        r = ctx.solve(_lib.default_opts(method=m, stop=_lib.STOP_FULL, rtol=1e-8, itmax=400))
        U, P = ctx.solution()
        out[name] = (r, U, P)
        print(f"[result] {name}: {r.solve_ms:.1f} ms outer {r.outer_iterations} inner {r.inner_iterations} "
              f"sweeps {r.refinements} fallbacks {r.fallbacks} residual {r.full_residual:.2e}")
        assert r.converged and r.full_residual < 1e-8
    Upl, Ppl = out["plain"][1], out["plain"][2]
    for name in ("mixed", "learned"):
        U, P = out[name][1], out[name][2]
        eu = np.linalg.norm(U - Upl) / np.linalg.norm(Upl)
        ep = np.linalg.norm((P - P.mean()) - (Ppl - Ppl.mean())) / np.linalg.norm(Ppl - Ppl.mean())
        print(f"[result] {name} vs plain FP64: errU {eu:.2e} errP {ep:.2e}")
        assert eu < 1e-7 and ep < 1e-7
    assert out["learned"][0].fallbacks == 0
    ctx.close()


def test_learned_solve_on_the_stress_workload_fields():
    """The learned + mixed solve with the stress workload's fields (contrast 1e6, anisotropy ratio up
    to 1e3, smooth director) at 16x128, trained on those fields, against the plain FP64 solve."""
    inp = make_inputs(scaled_workload(5, 16, 128))
    ctx = _ctx(inp)
    ctx.mlp_train(_lib.default_train_opts(snapshots=200))
    ctx.learned_setup()
    out = {}
    for name, m in (("learned", _lib.SOLVE_LEARNED), ("plain", _lib.SOLVE_PLAIN)):
        r = ctx.solve(_lib.default_opts(method=m, stop=_lib.STOP_FULL, rtol=1e-9, itmax=800))
        U, P = ctx.solution()
        out[name] = (U, P)
        print(f"[result] {name}: outer {r.outer_iterations} inner {r.inner_iterations} sweeps {r.refinements} "
              f"fallbacks {r.fallbacks} residual {r.full_residual:.2e}")
        assert r.converged and r.full_residual < 1e-9
    (U, P), (Upl, Ppl) = out["learned"], out["plain"]
    eu = np.linalg.norm(U - Upl) / np.linalg.norm(Upl)
    ep = np.linalg.norm((P - P.mean()) - (Ppl - Ppl.mean())) / np.linalg.norm(Ppl - Ppl.mean())
    print(f"[result] learned vs plain FP64: errU {eu:.2e} errP {ep:.2e}")
    assert eu < 1e-7 and ep < 1e-7
    ctx.close()


def test_poisoned_model_falls_back_to_the_plain_preconditioner():
    inp = make_inputs(scaled_workload(3, 16, 128))
    ctx = _ctx(inp)
    ctx.mlp_train(_lib.default_train_opts(snapshots=50, per_snapshot=256, epochs=1, batch=1024))
    p, norm = ctx.mlp_params()
    rng = np.random.default_rng(9)
    o, k = mo.OFF["W4"]
    p[o:o + k] = rng.standard_normal(k).astype(np.float32)
    norm[16 + 1] = norm[16 + 3] = norm[16 + 4] = 50.0   # off-diagonal factors of +-250: misshapen blocks
    ctx.mlp_params(p, norm)
    ctx.learned_setup()
    r = ctx.solve(_lib.default_opts(method=_lib.SOLVE_LEARNED, stop=_lib.STOP_FULL, rtol=1e-8))
    print(f"[result] poisoned: outer {r.outer_iterations} sweeps {r.refinements} fallbacks {r.fallbacks} "
          f"residual {r.full_residual:.2e}")
    assert r.converged and r.fallbacks >= 1
    ctx.close()


def test_solver_options_give_the_same_solution():
    """inner_warm (warm-started inner solves), pc_sweeps (more Jacobi sweeps on the learned
    operator) and the fine-level smoother (Gauss-Seidel or Chebyshev) change the work, not the
    answer: all reach the full residual and agree with the plain FP64 solve to 1e-7."""
    inp = make_inputs(scaled_workload(3, 32, 256))
    ctx = _ctx(inp)
    ctx.mlp_train(_lib.default_train_opts(snapshots=200, per_snapshot=256, epochs=2))
    ctx.learned_setup()
    ctx.solve(_lib.default_opts(method=_lib.SOLVE_PLAIN, stop=_lib.STOP_FULL, rtol=1e-10, itmax=400))
    U0, P0 = ctx.solution()
    its = {}
    for kw in (dict(inner_warm=1), dict(pc_sweeps=8), dict(inner_warm=1, pc_sweeps=1),
               dict(fine_smoother=_lib.SMOOTH_CHEB), dict(fine_smoother=_lib.SMOOTH_GS)):
        rep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_LEARNED, stop=_lib.STOP_FULL, rtol=1e-9, **kw))
        assert rep.converged and rep.full_residual < 1e-9, kw
        U, P = ctx.solution()
        eu = np.linalg.norm(U - U0) / np.linalg.norm(U0)
        ep = np.linalg.norm((P - P.mean()) - (P0 - P0.mean())) / np.linalg.norm(P0 - P0.mean())
        print(f"[result] {kw}: outer {rep.outer_iterations} inner {rep.inner_iterations} errU {eu:.1e} errP {ep:.1e}")
        assert eu < 1e-7 and ep < 1e-7, kw
        its[str(kw)] = rep.inner_iterations
    # one forward and one backward Gauss-Seidel sweep smooth as well as Chebyshev(2) (three
    # products): about as many inner iterations, not more (a wrong sweep would still converge)
    gs, ch = its[str(dict(fine_smoother=_lib.SMOOTH_GS))], its[str(dict(fine_smoother=_lib.SMOOTH_CHEB))]
    assert gs <= 1.15 * ch + 2, (gs, ch)
    ctx.close()


def test_reassembly_gives_the_fresh_context_result():
    """A time step re-assembles on the same mesh with new fields: every per-assembly product
    (FP32 copies, R K, the coloured smoother values, the coarse inverse, the near-null vector,
    the learned operator) is rebuilt, so a re-assembled context solves bit for bit like a fresh
    one assembled with the new fields."""
    inp = make_inputs(scaled_workload(3, 32, 256))
    rng = np.random.default_rng(21)
    eta_b = np.asfortranarray(inp.eta_ip * np.exp(rng.normal(scale=0.5, size=inp.eta_ip.shape)))
    params = None
    results = []
    for fresh in (False, True):
        ctx = _lib.Context.from_mesh(inp.mesh)
        if not fresh:
            ctx.set_fields(inp.g, inp.rho, inp.eta_ip, inp.C6)
            ctx.set_fabric(inp.phi, inp.delta)
            ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
            ctx.mlp_train(_lib.default_train_opts(snapshots=200, per_snapshot=256, epochs=2))
            params = ctx.mlp_params()
            ctx.learned_setup()
            ctx.solve(_lib.default_opts(method=_lib.SOLVE_LEARNED, stop=_lib.STOP_FULL, rtol=1e-8))
        else:
            ctx.set_fabric(inp.phi, inp.delta)
            ctx.mlp_params(*params)
        ctx.set_fields(inp.g, inp.rho, eta_b, inp.C6)
        ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
        ctx.learned_setup()
        out = []
        for m in (_lib.SOLVE_LEARNED, _lib.SOLVE_MIXED):
            rep = ctx.solve(_lib.default_opts(method=m, stop=_lib.STOP_FULL, rtol=1e-8))
            assert rep.converged and rep.full_residual < 1e-8
            U, P = ctx.solution()
            out.append((rep.outer_iterations, rep.inner_iterations, U, P))
        results.append(out)
        ctx.close()
    for (o1, i1, U1, P1), (o2, i2, U2, P2) in zip(*results):
        print(f"[result] re-assembled {o1}/{i1}, fresh {o2}/{i2}")
        assert (o1, i1) == (o2, i2)
        assert np.array_equal(U1, U2) and np.array_equal(P1, P2)
