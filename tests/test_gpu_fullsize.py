"""Full-size (256 x 2048, 4.73M dofs) solves: the learned + mixed path that the benchmark times
agrees with the plain FP64 Schur PCG, and both meet the 1e-8 full-system residual.  The oracle cannot
run at this size; the two GPU methods share only the assembled operators, so their agreement to
the parity bar (1e-7 relative L2) is the size-independent check."""
import numpy as np
import pytest

from dimspers_amd import _lib
from dimspers_amd.inputs import make_inputs

pytestmark = pytest.mark.gpu


def test_learned_mixed_matches_plain_fp64_at_full_size():
    inp = make_inputs(4)
    m = inp.mesh
    print(f"[case] full-size solve {m.ncell_r}x{m.ncell_t}", flush=True)
    ctx = _lib.Context.from_mesh(m)
    ctx.set_fields(inp.g, inp.rho, inp.eta_ip, inp.C6)
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
    ctx.set_fabric(inp.phi, inp.delta)
    ctx.mlp_train(_lib.default_train_opts())
    ctx.learned_setup()
    rep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_LEARNED, stop=_lib.STOP_FULL, rtol=1e-8))
    assert rep.converged and rep.full_residual < 1e-8
    U, P = ctx.solution()
    prep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_PLAIN, stop=_lib.STOP_FULL, rtol=1e-8, itmax=400))
    assert prep.converged and prep.full_residual < 1e-8
    U0, P0 = ctx.solution()
    eu = np.linalg.norm(U - U0) / np.linalg.norm(U0)
    ep = np.linalg.norm((P - P.mean()) - (P0 - P0.mean())) / np.linalg.norm(P0 - P0.mean())
    print(f"[result] learned {rep.solve_ms:.1f} ms ({rep.outer_iterations}/{rep.inner_iterations}), "
          f"plain {prep.solve_ms:.1f} ms: errU {eu:.2e} errP {ep:.2e}")
    assert eu < 1e-7 and ep < 1e-7
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
