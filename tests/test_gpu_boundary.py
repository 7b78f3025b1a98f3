"""GPU tests of the drop-in boundary: fixed sets other than free slip.

The reference offers three velocity boundary conditions (src/Setup/boundary_conditions.jl:15-66):
free slip on both walls (the default of MAIN.jl), free slip with the tangential dof of one node
pinned, and no slip.  Only free slip leaves the rigid rotation as a near-null mode of K; with a
fixed tangential dof K[ifree, ifree] is well conditioned and the literal reference algorithm (the
LU inner solve) is the oracle.
"""
import numpy as np
import pytest

from dimspers_amd import _lib
from dimspers_amd.inputs import make_inputs, scaled_workload
from oracle import stokes_oracle as so

pytestmark = pytest.mark.gpu


def _errs(U, P, Uo, Po):
    eu = np.linalg.norm(U - Uo) / np.linalg.norm(Uo)
    ep = np.linalg.norm((P - P.mean()) - (Po - Po.mean())) / np.linalg.norm(Po - Po.mean())
    return eu, ep


@pytest.mark.parametrize("kind", ["no_slip", "pinned"])
def test_other_velocity_bcs_match_the_literal_oracle(kind):
    inp = make_inputs(scaled_workload(1, 8, 32))
    m = inp.mesh
    bcs = m.no_slip() if kind == "no_slip" else m.free_slip_pinned(m.ncell_r + 1)
    print(f"[case] {kind} 8x32 workload 1, {bcs[0].size} fixed dofs", flush=True)
    sol, Uo = so.solve_stokes(inp, rtol=1e-13, itmax=400, bcs=bcs)    # Uo: before updatevelocity2
    Po = sol.P
    ctx = _lib.Context.from_mesh(m, bcs[0])
    ctx.set_fields(inp.g, inp.rho, inp.eta_ip, inp.C6)
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
    # the multigrid is built for free slip only: the methods that need it are refused, with the reason
    for method in (_lib.SOLVE_PLAIN, _lib.SOLVE_MIXED):
        with pytest.raises(_lib.StkError) as e:
            ctx.solve(_lib.default_opts(method=method, inner=_lib.INNER_MG, stop=_lib.STOP_FULL))
        assert e.value.code == _lib.STK_E_NOMESH and "free slip" in str(e.value)
    # no near-null rotation: deflating it is refused
    with pytest.raises(_lib.StkError) as e:
        ctx.solve(_lib.default_opts(method=_lib.SOLVE_PLAIN, inner=_lib.INNER_JACOBI, null_mode=_lib.NULL_DEFLATE))
    assert e.value.code == _lib.STK_E_ARG
    rep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_PLAIN, inner=_lib.INNER_JACOBI, stop=_lib.STOP_FULL,
                                      rtol=1e-11, itmax=400, inner_itmax=50000))
    U, P = ctx.solution()
    eu, ep = _errs(U, P, Uo, Po)
    print(f"[result] {kind}: outer {rep.outer_iterations} inner {rep.inner_iterations} "
          f"residual {rep.full_residual:.2e} errU {eu:.2e} errP {ep:.2e}", flush=True)
    assert rep.converged and rep.full_residual < 1e-11
    assert eu < 1e-8 and ep < 1e-8
    ctx.close()
