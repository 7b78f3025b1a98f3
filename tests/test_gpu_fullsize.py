"""Full-size solves (256 x 2048, 4.73M dofs, and 128 x 1024, 512 x 4096): the mixed-precision and the
learned + mixed paths meet the 1e-8 full-system residual and agree with a plain FP64 Schur PCG solved
to a 1e-11 residual within the parity bar (1e-7 relative L2 on velocity and pressure).  The oracle
cannot run at these sizes; the methods share only the assembled operators, so their agreement with the
tight FP64 solve is the size-independent check."""
import numpy as np
import pytest

from dimspers_amd import _lib
from dimspers_amd.inputs import make_inputs

pytestmark = pytest.mark.gpu


def _errs(U, P, U0, P0):
    eu = np.linalg.norm(U - U0) / np.linalg.norm(U0)
    ep = np.linalg.norm((P - P.mean()) - (P0 - P0.mean())) / np.linalg.norm(P0 - P0.mean())
    return eu, ep


def _tight_fp64(ctx):
    """Plain FP64 Schur PCG run to a 1e-11 full-system residual: the reference the fast solves are
    measured against, three orders tighter than the 1e-8 they are asked for.  The FP64 floor of this
    residual at 256 x 2048 is about 2e-12 (asked for 1e-12 or 1e-13 the solve stalls there), and the
    solution at that floor differs from this one by 5e-13 in velocity."""
    rep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_PLAIN, stop=_lib.STOP_FULL, rtol=1e-11, itmax=800))
    print(f"[result] tight FP64 reference: {rep.solve_ms:.1f} ms ({rep.outer_iterations}/{rep.inner_iterations}), "
          f"residual {rep.full_residual:.2e}", flush=True)
    assert rep.converged and rep.full_residual < 1e-11
    return ctx.solution()


def _context(inp):
    m = inp.mesh
    ctx = _lib.Context.from_mesh(m)
    ctx.set_fields(inp.g, inp.rho, inp.eta_ip, inp.C6)
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
    return ctx


def test_mixed_precision_at_full_size_against_a_tight_fp64_solve():
    """Workload 3 (256 x 2048, contrast 1e4, ratio U[1, 100] per element): FP32 assembly copies,
    SpMV and inner PCG under FP64 refinement reach the 1e-8 full residual; velocity and pressure
    errors against the tight FP64 solve are below 1e-7."""
    inp = make_inputs(3)
    print(f"[case] workload 3 mixed {inp.mesh.ncell_r}x{inp.mesh.ncell_t}", flush=True)
    ctx = _context(inp)
    rep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_MIXED, stop=_lib.STOP_FULL, rtol=1e-8))
    assert rep.converged and rep.full_residual < 1e-8
    U, P = ctx.solution()
    U0, P0 = _tight_fp64(ctx)
    eu, ep = _errs(U, P, U0, P0)
    print(f"[result] mixed {rep.solve_ms:.1f} ms ({rep.outer_iterations}/{rep.inner_iterations}, "
          f"{rep.refinements} sweeps), residual {rep.full_residual:.2e}: errU {eu:.2e} errP {ep:.2e}", flush=True)
    assert eu < 1e-7 and ep < 1e-7
    ctx.close()


def test_learned_mixed_at_full_size_against_a_tight_fp64_solve_and_its_gain():
    """Workload 4 (256 x 2048, learned): the learned + mixed solve the benchmark times and the mixed
    solve without the model both reach the 1e-8 full residual with errors below 1e-7 against the tight
    FP64 solve, and the model saves Schur iterations over the same solver without it."""
    inp = make_inputs(4)
    print(f"[case] workload 4 learned {inp.mesh.ncell_r}x{inp.mesh.ncell_t}", flush=True)
    ctx = _context(inp)
    ctx.set_fabric(inp.phi, inp.delta)
    ctx.mlp_train(_lib.default_train_opts())
    ctx.learned_setup()
    reps, sols = {}, {}
    for name, method in (("learned", _lib.SOLVE_LEARNED), ("mixed", _lib.SOLVE_MIXED)):
        rep = ctx.solve(_lib.default_opts(method=method, stop=_lib.STOP_FULL, rtol=1e-8))
        assert rep.converged and rep.full_residual < 1e-8
        reps[name], sols[name] = rep, ctx.solution()
    U0, P0 = _tight_fp64(ctx)
    for name in ("learned", "mixed"):
        rep = reps[name]
        eu, ep = _errs(*sols[name], U0, P0)
        print(f"[result] {name} {rep.solve_ms:.1f} ms ({rep.outer_iterations}/{rep.inner_iterations}, "
              f"{rep.refinements} sweeps, {rep.fallbacks} fallbacks), residual {rep.full_residual:.2e}: "
              f"errU {eu:.2e} errP {ep:.2e}", flush=True)
        assert eu < 1e-7 and ep < 1e-7
    assert reps["learned"].fallbacks == 0
    assert reps["learned"].outer_iterations < reps["mixed"].outer_iterations
    ctx.close()


def test_plain_fp64_drop_in_at_128x1024():
    """128 x 1024 (1.18M dofs, contrast 1e4, ratio 10): the plain FP64 Schur PCG (the reference's
    lumped-mass preconditioner) meets the 1e-8 full-system residual and agrees with the mixed solve."""
    inp = make_inputs(2)
    m = inp.mesh
    print(f"[case] plain drop-in {m.ncell_r}x{m.ncell_t}", flush=True)
    ctx = _lib.Context.from_mesh(m)
    ctx.set_fields(inp.g, inp.rho, inp.eta_ip, inp.C6)
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
    prep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_PLAIN, stop=_lib.STOP_FULL, rtol=1e-8, itmax=400))
    assert prep.converged and prep.full_residual < 1e-8
    U0, P0 = ctx.solution()
    rep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_MIXED, stop=_lib.STOP_FULL, rtol=1e-8))
    assert rep.converged and rep.full_residual < 1e-8
    U, P = ctx.solution()
    eu = np.linalg.norm(U - U0) / np.linalg.norm(U0)
    ep = np.linalg.norm((P - P.mean()) - (P0 - P0.mean())) / np.linalg.norm(P0 - P0.mean())
    print(f"[result] plain {prep.solve_ms:.1f} ms ({prep.outer_iterations}/{prep.inner_iterations}), "
          f"mixed {rep.solve_ms:.1f} ms ({rep.outer_iterations}/{rep.inner_iterations}): errU {eu:.2e} errP {ep:.2e}")
    assert eu < 1e-7 and ep < 1e-7
    ctx.close()


def test_learned_and_mixed_agree_at_the_largest_size():
    """512 x 4096 (18.9M dofs, viscosity contrast 1e6): learned + mixed, mixed without the model
    and plain FP64 all meet the 1e-8 full-system residual; the fast solves agree with plain FP64 to the
    parity bar."""
    inp = make_inputs(5)
    m = inp.mesh
    print(f"[case] largest solve {m.ncell_r}x{m.ncell_t}", flush=True)
    ctx = _lib.Context.from_mesh(m)
    ctx.set_fields(inp.g, inp.rho, inp.eta_ip, inp.C6)
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
    ctx.set_fabric(inp.phi, inp.delta)
    ctx.mlp_train(_lib.default_train_opts())
    ctx.learned_setup()
    sols = []
    for method in (_lib.SOLVE_LEARNED, _lib.SOLVE_MIXED):
        rep = ctx.solve(_lib.default_opts(method=method, stop=_lib.STOP_FULL, rtol=1e-8))
        assert rep.converged and rep.full_residual < 1e-8
        sols.append(ctx.solution())
        print(f"[result] method {method}: {rep.solve_ms:.1f} ms ({rep.outer_iterations}/{rep.inner_iterations}), "
              f"residual {rep.full_residual:.2e}")
    # inner solves to 1e-10, as bench.py's plain FP64 leg
    prep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_PLAIN, stop=_lib.STOP_FULL, rtol=1e-8, itmax=400,
                                       inner_rtol=1e-10))
    assert prep.converged and prep.full_residual < 1e-8
    U0, P0 = ctx.solution()
    print(f"[result] plain FP64: {prep.solve_ms:.1f} ms ({prep.outer_iterations}/{prep.inner_iterations}), "
          f"residual {prep.full_residual:.2e}")
    for name, (U, P) in zip(("learned", "mixed"), sols):
        eu = np.linalg.norm(U - U0) / np.linalg.norm(U0)
        ep = np.linalg.norm((P - P.mean()) - (P0 - P0.mean())) / np.linalg.norm(P0 - P0.mean())
        print(f"[result] {name} vs plain FP64: errU {eu:.2e} errP {ep:.2e}")
        assert eu < 1e-7 and ep < 1e-7
    ctx.close()
