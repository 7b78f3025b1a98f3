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
from oracle import mlp_oracle as mo
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


def _oracle_operators(inp):
    KK, GG, MM, Fb = so.assembly_stokes_cylindric(inp)
    K, G, F = so.prepare_matrices(KK, GG, Fb, so.rotation_matrix(inp.mesh.theta))
    return K, MM, G, F


def test_caller_operators_equal_the_assembled_ones_and_are_checked():
    """stk_set_operators with the oracle's operators leaves the same FP64 operators as the
    library's own assembly (K[ifree, ifree], G on free rows, full M) and solves to the same answer;
    the lower-triangle form of K is accepted, an entry outside the mesh pattern is refused
    (STK_E_SHAPE), and the multigrid methods are refused with the reason (STK_E_NOMESH)."""
    import scipy.sparse as sp
    inp = make_inputs(scaled_workload(3, 8, 32))
    m = inp.mesh
    print("[case] caller operators 8x32 workload 3", flush=True)
    K, MM, G, F = _oracle_operators(inp)
    own = _lib.Context.from_mesh(m)
    own.set_fields(inp.g, inp.rho, inp.eta_ip, inp.C6)
    own.assemble(_lib.ASM_FP64)
    opts = _lib.default_opts(method=_lib.SOLVE_PLAIN, inner=_lib.INNER_JACOBI, stop=_lib.STOP_FULL, rtol=1e-10,
                             itmax=400, inner_itmax=50000)
    rep_own = own.solve(opts)
    Uown, Pown = own.solution()
    for k_lower in (False, True):
        ctx = _lib.Context.from_mesh(m)
        ctx.set_operators(sp.tril(K) if k_lower else K, MM, G, F, K_lower=k_lower)
        for which in (_lib.MAT_K, _lib.MAT_G, _lib.MAT_GT, _lib.MAT_M):
            a, b = ctx.export_csc(which), own.export_csc(which)
            assert abs(a - b).max() <= 1e-12 * abs(b).max(), which
        rep = ctx.solve(opts)
        U, P = ctx.solution()
        eu, ep = _errs(U, P, Uown, Pown)
        print(f"[result] K_lower={k_lower}: outer {rep.outer_iterations} (own {rep_own.outer_iterations}) "
              f"errU {eu:.2e} errP {ep:.2e}", flush=True)
        assert rep.converged and eu < 1e-9 and ep < 1e-9
        with pytest.raises(_lib.StkError) as e:
            ctx.solve(_lib.default_opts(method=_lib.SOLVE_MIXED, stop=_lib.STOP_FULL))
        assert e.value.code == _lib.STK_E_NOMESH and "stk_set_operators" in str(e.value)
        ctx.close()
    # an entry between two nodes that share no element
    bad = K.tolil()
    bad[0, m.nU - 1] = bad[m.nU - 1, 0] = 1.0
    ctx = _lib.Context.from_mesh(m)
    with pytest.raises(_lib.StkError) as e:
        ctx.set_operators(bad.tocsc(), MM, G, F)
    assert e.value.code == _lib.STK_E_SHAPE and "outside" in str(e.value)
    ctx.close()
    own.close()


@pytest.mark.parametrize("method", [_lib.SOLVE_PLAIN, _lib.SOLVE_MIXED])
def test_warm_start_from_the_solution_needs_no_iterations(method):
    """A solve with warm_start = 1 starts from the context's solution: from the converged one it
    meets the judged residual at once (no Schur iteration), and from a perturbed pressure it
    converges back to the same solution."""
    inp = make_inputs(scaled_workload(3, 16, 128))
    print(f"[case] warm start 16x128 method {method}", flush=True)
    ctx = _lib.Context.from_mesh(inp.mesh)
    ctx.set_fields(inp.g, inp.rho, inp.eta_ip, inp.C6)
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
    rep = ctx.solve(_lib.default_opts(method=method, stop=_lib.STOP_FULL, rtol=1e-9))
    assert rep.converged and rep.outer_iterations > 0
    U0, P0 = ctx.solution()
    again = ctx.solve(_lib.default_opts(method=method, stop=_lib.STOP_FULL, rtol=1e-9, warm_start=1))
    assert again.converged and again.outer_iterations == 0
    U1, P1 = ctx.solution()
    assert np.array_equal(P1, P0) and np.abs(U1 - U0).max() <= 1e-14 * np.abs(U0).max()
    rng = np.random.default_rng(3)
    ctx.set_solution(U0, P0 * (1 + 0.1 * rng.standard_normal(P0.size)))
    rep2 = ctx.solve(_lib.default_opts(method=method, stop=_lib.STOP_FULL, rtol=1e-9, warm_start=1))
    U2, P2 = ctx.solution()
    eu, ep = _errs(U2, P2, U0, P0)
    print(f"[result] warm from a perturbed pressure: outer {rep2.outer_iterations} (cold {rep.outer_iterations}) "
          f"errU {eu:.2e} errP {ep:.2e}", flush=True)
    assert rep2.converged and eu < 1e-7 and ep < 1e-7
    ctx.close()


@pytest.mark.parametrize("method", [_lib.SOLVE_PLAIN, _lib.SOLVE_MIXED, _lib.SOLVE_LEARNED])
@pytest.mark.parametrize("stop", [_lib.STOP_SCHUR, _lib.STOP_FULL])
def test_zero_forcing_gives_the_zero_solution(method, stop):
    """g = 0: the right-hand side vanishes.  StokesPcCG (src/IterativeSolvers.jl:114-137) would
    divide 0 by 0 there (r = 0, so r.d = q.Sq = 0 and alpha is NaN); the drop-in reports a
    converged solve without a Schur iteration and returns exactly U = 0, P = 0."""
    inp = make_inputs(scaled_workload(3, 16, 128))
    print(f"[case] zero forcing 16x128 method {method} stop {stop}", flush=True)
    ctx = _lib.Context.from_mesh(inp.mesh)
    ctx.set_fields(0.0, inp.rho, inp.eta_ip, inp.C6)
    ctx.set_fabric(inp.phi, inp.delta)
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
    if method == _lib.SOLVE_LEARNED:
        rng = np.random.default_rng(5)
        p = (rng.standard_normal(mo.N_PARAM) * 0.15).astype(np.float32)
        norm = np.concatenate([np.zeros(5), np.ones(5), np.zeros(6), np.ones(6)]).astype(np.float32)
        ctx.mlp_params(p, norm)
        ctx.learned_setup()
    rep = ctx.solve(_lib.default_opts(method=method, stop=stop, rtol=1e-8))
    U, P = ctx.solution()
    print(f"[result] outer {rep.outer_iterations} converged {rep.converged} "
          f"max|U| {np.abs(U).max():.1e} max|P| {np.abs(P).max():.1e}", flush=True)
    assert rep.converged and rep.outer_iterations == 0 and rep.fallbacks == 0
    assert np.all(np.isfinite(U)) and np.all(np.isfinite(P))
    assert not U.any() and not P.any()
    ctx.close()
