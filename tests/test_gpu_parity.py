"""GPU parity tests: the HIP path through the C ABI against the FP64 oracle.

Bars: assembled operators within 1e-12 (relative to the largest entry) of the
oracle, FP32 operators within 2e-7, SpMV within 1e-13 / 1e-6, solutions within
1e-7 relative L2 on velocity and pressure.  Velocity comparisons are made in
the node (u_theta, u_r) frame; pressures after mean removal.
"""
import numpy as np
import pytest

from dimspers_amd import _lib
from dimspers_amd.inputs import StokesInputs, make_inputs, scaled_workload
from oracle import stokes_oracle as so

pytestmark = pytest.mark.gpu


def _ctx(inp, flags=_lib.ASM_FP64, C=True):
    print(f"[case] mesh {inp.mesh.ncell_r}x{inp.mesh.ncell_t} workload {inp.workload.index} flags {flags}", flush=True)
    ctx = _lib.Context.from_mesh(inp.mesh)
    ctx.set_fields(inp.g, inp.rho, inp.eta_ip, inp.C6 if C else None)
    ctx.assemble(flags)
    return ctx


def _rel(a, b):
    a = a.toarray() if hasattr(a, "toarray") else a
    b = b.toarray() if hasattr(b, "toarray") else b
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def small():
    return make_inputs(scaled_workload(3, 4, 16))


@pytest.fixture(scope="module")
def medium():
    return make_inputs(scaled_workload(1, 8, 64))


@pytest.mark.parametrize("iso", [False, True])
def test_assembly_cartesian_matches_oracle(small, iso):
    flags = _lib.ASM_FP64 | _lib.ASM_CARTESIAN | _lib.ASM_NO_BC | (_lib.ASM_ISOTROPIC if iso else 0)
    ctx = _ctx(small, flags)
    KK, GG, MM, Fb = so.assembly_stokes_cylindric(small, isotropic=iso)
    assert _rel(ctx.export_csc(_lib.MAT_K, lower_only=True), KK) < 1e-12
    assert _rel(ctx.export_csc(_lib.MAT_G), GG) < 1e-12
    assert _rel(ctx.export_csc(_lib.MAT_M, lower_only=True), MM) < 1e-12
    assert np.abs(ctx.export_rhs() - Fb).max() / np.abs(Fb).max() < 1e-12
    # the exported lower triangle has the reference's structure (every element pair present)
    assert ctx.export_csc(_lib.MAT_K, lower_only=True).nnz == (184 * 4 * 16 + 32 * 16 + 2 * small.mesh.nnod) // 2
    ctx.close()


def test_assembly_rotated_matches_prepare_matrices(medium):
    ctx = _ctx(medium, _lib.ASM_FP64 | _lib.ASM_NO_BC)
    KK, GG, MM, Fb = so.assembly_stokes_cylindric(medium)
    K, G, F = so.prepare_matrices(KK, GG, Fb, so.rotation_matrix(medium.mesh.theta))
    assert _rel(ctx.export_csc(_lib.MAT_K), K) < 1e-12
    assert _rel(ctx.export_csc(_lib.MAT_G), G) < 1e-12
    assert _rel(ctx.export_csc(_lib.MAT_GT), G.T) < 1e-12
    assert np.abs(ctx.export_rhs() - F).max() / np.abs(F).max() < 1e-12
    ctx.close()


def test_boundary_conditions_and_fp32(medium):
    ctx = _ctx(medium, _lib.ASM_FP64 | _lib.ASM_FP32)
    KK, GG, MM, Fb = so.assembly_stokes_cylindric(medium)
    K, G, F = so.prepare_matrices(KK, GG, Fb, so.rotation_matrix(medium.mesh.theta))
    fixed, ifree = medium.mesh.free_slip()
    Kg = ctx.export_csc(_lib.MAT_K).tocsr()
    assert _rel(Kg[ifree][:, ifree], K[ifree][:, ifree]) < 1e-12
    Kd = Kg[fixed].toarray()
    assert np.array_equal(Kd[np.arange(fixed.size), fixed], np.ones(fixed.size))
    assert np.count_nonzero(Kd) == fixed.size          # identity rows for the fixed dofs
    Gg = ctx.export_csc(_lib.MAT_G).tocsr()
    assert _rel(Gg[ifree], G[ifree]) < 1e-12 and Gg[fixed].count_nonzero() == 0
    # FP32 copies are the FP64 operators rounded
    rng = np.random.default_rng(3)
    x = rng.standard_normal(ctx.nU)
    y64 = ctx.spmv(_lib.MAT_K, x, 64)
    y32 = ctx.spmv(_lib.MAT_K, x, 32)
    assert np.abs(y32 - y64).max() / np.abs(y64).max() < 2e-6
    ctx.close()


def test_spmv_matches_scipy(medium):
    ctx = _ctx(medium, _lib.ASM_FP64 | _lib.ASM_NO_BC)
    rng = np.random.default_rng(4)
    for which, n in ((_lib.MAT_K, ctx.nU), (_lib.MAT_G, ctx.nP), (_lib.MAT_GT, ctx.nU), (_lib.MAT_M, ctx.nP)):
        A = ctx.export_csc(which)
        x = rng.standard_normal(n)
        y = ctx.spmv(which, x, 64)
        ref = A @ x
        assert np.abs(y - ref).max() / np.abs(ref).max() < 1e-13, which
    ctx.close()


@pytest.mark.parametrize("mesh", [(8, 64, 1), (64, 512, 3)])
def test_restriction_times_k_is_the_product(mesh):
    """R K, assembled element by element (rotated P2 -> P1 interpolation blocks, constrained
    K), equals the product of the exported restriction and K to rounding; R itself holds the
    identity at vertices and half rotations (row norm 1/2) at edge midpoints, with the dofs of
    fixed nodes masked.  The V-cycle forms its coarse right-hand side as R b - (R K) x."""
    nr, nt, w = mesh
    inp = make_inputs(scaled_workload(w, nr, nt))
    m = inp.mesh
    ctx = _ctx(inp, _lib.ASM_FP64 | _lib.ASM_FP32)
    K = ctx.export_csc(_lib.MAT_K)
    R = ctx.export_csc(_lib.MAT_R)
    RK = ctx.export_csc(_lib.MAT_RK)
    assert R.shape == RK.shape == (2 * m.nvert, m.nU)
    fixed, _ = m.free_slip()
    Rd = R.toarray() if m.nU < 20000 else None
    if Rd is not None:
        free = np.ones(m.nU, dtype=bool)
        free[fixed] = False
        assert np.abs(Rd[:, ~free]).max() == 0.0              # fixed fine dofs: zero columns
        assert np.abs(Rd[~free[:2 * m.nvert], :]).max() == 0.0  # fixed vertex dofs: zero rows
        v = np.flatnonzero(free[:2 * m.nvert])
        assert np.array_equal(Rd[v, v], np.ones(v.size))      # vertex rows: identity block
        mid = Rd[:, 2 * m.nvert:].reshape(m.nvert, 2, -1, 2)   # (vertex, c, midpoint, c')
        norms = np.sqrt((mid ** 2).sum(axis=3))
        full = (np.abs(mid).sum(axis=(1, 3)) > 0) & np.all(free[:2 * m.nvert].reshape(-1, 2), axis=1)[:, None] \
            & np.all(free[2 * m.nvert:].reshape(-1, 2), axis=1)[None, :]
        assert full.sum() > 0
        assert np.allclose(norms.transpose(0, 2, 1)[full], 0.5, atol=1e-15)
    ref = (R @ K).tocsc()
    err = abs(RK - ref).max() / abs(ref).max()
    print(f"[result] R K vs R @ K {nr}x{nt}: {err:.2e}")
    assert err < 1e-13
    rng = np.random.default_rng(12)
    x = rng.standard_normal(m.nU)
    want = R @ (K @ x)
    assert np.abs(ctx.spmv(_lib.MAT_RK, x, 64) - want).max() < 1e-12 * np.abs(want).max()
    assert np.abs(ctx.spmv(_lib.MAT_RK, x, 32) - want).max() < 1e-5 * np.abs(want).max()
    assert np.abs(ctx.spmv(_lib.MAT_R, x, 64) - R @ x).max() < 1e-13 * np.abs(R @ x).max()
    ctx.close()


def test_assembly_is_bitwise_deterministic(medium):
    ctx = _ctx(medium, _lib.ASM_FP64)
    a = ctx.export_csc(_lib.MAT_K).data.copy()
    f = ctx.export_rhs()
    ctx.assemble(_lib.ASM_FP64)
    assert np.array_equal(a, ctx.export_csc(_lib.MAT_K).data)
    assert np.array_equal(f, ctx.export_rhs())
    r1 = ctx.solve(_lib.default_opts())
    U1, P1 = ctx.solution()
    r2 = ctx.solve(_lib.default_opts())
    U2, P2 = ctx.solution()
    assert np.array_equal(U1, U2) and np.array_equal(P1, P2)
    assert r1.outer_iterations == r2.outer_iterations
    ctx.close()


def _oracle_projected(inp, nip=6):
    KK, GG, MM, Fb = so.assembly_stokes_cylindric(inp, nip=nip)
    K, G, F = so.prepare_matrices(KK, GG, Fb, so.rotation_matrix(inp.mesh.theta))
    _, ifree = inp.mesh.free_slip()
    v, lam = so.null_vector(K, ifree, inp.mesh)
    U, P, info = so.stokes_pcg(K, MM, G, F, ifree, rtol=1e-12, itmax=400, inner=so.inner_projected(K, ifree, v))
    return U, P, v, (K, G, MM, F, ifree)


def _errs(U, P, Uo, Po):
    eu = np.linalg.norm(U - Uo) / np.linalg.norm(Uo)
    ep = np.linalg.norm((P - P.mean()) - (Po - Po.mean())) / np.linalg.norm(Po - Po.mean())
    return eu, ep


@pytest.mark.parametrize("mesh", [(4, 16, 3), (8, 64, 1), (32, 256, 1), (5, 24, 3), (7, 18, 1), (8, 64, 5)])
@pytest.mark.parametrize("method", ["plain", "mixed"])
def test_solve_matches_oracle_projected(mesh, method):
    """(5, 24) and (7, 18): odd radial counts and azimuthal counts that are not powers of two, so
    partial 64-row slices at every level and no h-coarsening of the P1 level.  (8, 64) with the
    fields of the stress workload: viscosity contrast 1e6, log-uniform anisotropy ratio up to 1e3
    and a smooth director field."""
    nr, nt, w = mesh
    inp = make_inputs(scaled_workload(w, nr, nt))
    ctx = _ctx(inp, _lib.ASM_FP64 | _lib.ASM_FP32)
    Uo, Po, v, _ = _oracle_projected(inp)
    m = _lib.SOLVE_PLAIN if method == "plain" else _lib.SOLVE_MIXED
    rep = ctx.solve(_lib.default_opts(method=m, stop=_lib.STOP_FULL, rtol=1e-10, itmax=400))
    assert rep.converged and rep.full_residual < 1e-10
    U, P = ctx.solution()
    eu, ep = _errs(U, P, Uo, Po)
    print(f"[result] {method} {nr}x{nt}: outer {rep.outer_iterations} errU {eu:.2e} errP {ep:.2e}")
    assert eu < 1e-7 and ep < 1e-7
    ctx.close()


def test_inner_solvers_without_multigrid_match_oracle():
    """The plain method with the Jacobi-preconditioned inner CG (STK_INNER_JACOBI), both when the
    multigrid is not built (STK_ASM_NO_MG) and when it is built but not used, gives the oracle's
    projected solution; a multigrid solve asked of an assembly without the hierarchy is refused
    (STK_E_NOMESH), and so are the FP32 methods with Jacobi inner solves (STK_E_ARG: at their
    loose inner tolerance only the multigrid preconditioner makes progress)."""
    inp = make_inputs(scaled_workload(3, 4, 16))
    Uo, Po, _, _ = _oracle_projected(inp)
    method, m = "plain", _lib.SOLVE_PLAIN
    for flags in (_lib.ASM_FP64 | _lib.ASM_FP32 | _lib.ASM_NO_MG, _lib.ASM_FP64 | _lib.ASM_FP32):
        ctx = _ctx(inp, flags)
        inner = _lib.INNER_JACOBI
        if flags & _lib.ASM_NO_MG:
            with pytest.raises(_lib.StkError) as e:
                ctx.solve(_lib.default_opts(method=m, inner=_lib.INNER_MG, stop=_lib.STOP_FULL, rtol=1e-10))
            assert e.value.code == _lib.STK_E_NOMESH
        with pytest.raises(_lib.StkError) as e:
            ctx.solve(_lib.default_opts(method=_lib.SOLVE_MIXED, inner=inner, stop=_lib.STOP_FULL, rtol=1e-8))
        assert e.value.code == _lib.STK_E_ARG
        rep = ctx.solve(_lib.default_opts(method=m, inner=inner, stop=_lib.STOP_FULL, rtol=1e-10, itmax=400,
                                          inner_itmax=20000))
        assert rep.converged and rep.full_residual < 1e-10
        U, P = ctx.solution()
        eu, ep = _errs(U, P, Uo, Po)
        print(f"[result] {method} flags {flags} inner {inner}: outer {rep.outer_iterations} "
              f"inner {rep.inner_iterations} errU {eu:.2e} errP {ep:.2e}")
        assert eu < 1e-7 and ep < 1e-7
        ctx.close()


def test_solve_byte_count_matches_the_kernel_formulas():
    """stk_report.bytes of a plain solve with the Jacobi inner CG: each inner iteration is one FP64
    K p with its two dots (the K SpMV bytes of stk_time_kernel plus p and v read once more) and
    three element-wise passes over 15 FP64 vectors; the inner iterations dominate, so bytes per
    inner iteration sit just above that sum.  Almost every launch states its byte count."""
    inp = make_inputs(scaled_workload(3, 16, 128))
    ctx = _ctx(inp, _lib.ASM_FP64 | _lib.ASM_FP32)
    rep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_PLAIN, inner=_lib.INNER_JACOBI, stop=_lib.STOP_SCHUR,
                                      rtol=1e-6, itmax=400, inner_rtol=1e-8, inner_itmax=100000))
    _, kb = ctx.time_kernel(_lib.KERNEL_SPMV_K64, reps=1)
    nU = inp.mesh.nU
    per_it = kb + 2 * nU * 8 + 15 * nU * 8
    ratio = rep.bytes / rep.inner_iterations / per_it
    print(f"[result] bytes {rep.bytes:.4e} inner {rep.inner_iterations} launches {rep.launches} "
          f"floor {rep.launches_floor} ratio {ratio:.4f}")
    assert 1.0 <= ratio < 1.1
    assert rep.launches > 4 * rep.inner_iterations and rep.launches_floor < 0.02 * rep.launches
    # an FP32 multigrid solve counts its bytes too
    rep2 = ctx.solve(_lib.default_opts(method=_lib.SOLVE_MIXED, stop=_lib.STOP_FULL, rtol=1e-8))
    assert rep2.bytes > rep2.inner_iterations * kb / 2 and rep2.launches_floor < 0.1 * rep2.launches
    ctx.close()


def test_fp16_smoother_matches_fp32_smoother():
    """The FP16 scaled copy of K only changes the fine-level smoother: the mixed solve
    reaches the same full residual, the same solution (1e-7 of the plain FP64 one) and
    about the same number of Schur iterations as with the FP32 smoother."""
    inp = make_inputs(scaled_workload(3, 64, 512))
    out = {}
    for name, flags in (("fp16", _lib.ASM_FP16_SMOOTHER), ("fp32", 0)):
        ctx = _ctx(inp, _lib.ASM_FP64 | _lib.ASM_FP32 | flags)
        rep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_MIXED, stop=_lib.STOP_FULL, rtol=1e-8))
        assert rep.converged and rep.full_residual < 1e-8 and rep.fallbacks == 0
        U, P = ctx.solution()
        if name == "fp16":
            prep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_PLAIN, stop=_lib.STOP_FULL, rtol=1e-10, itmax=400))
            assert prep.converged
            out["plain"] = ctx.solution()
        out[name] = (U, P, rep)
        print(f"[result] {name} smoother: outer {rep.outer_iterations} inner {rep.inner_iterations} "
              f"residual {rep.full_residual:.2e}")
        ctx.close()
    Upl, Ppl = out["plain"]
    for name in ("fp16", "fp32"):
        eu, ep = _errs(out[name][0], out[name][1], Upl, Ppl)
        assert eu < 1e-7 and ep < 1e-7, (name, eu, ep)
    assert out["fp16"][2].outer_iterations <= 1.5 * out["fp32"][2].outer_iterations + 2


@pytest.mark.parametrize("mesh", [(4, 16, 3), (32, 256, 1)])
def test_deflated_solve_matches_reference_lu(mesh):
    """STK_NULL_DEFLATE reproduces lu(K[ifree, ifree]) of the reference (IterativeSolvers.jl:94-168)."""
    nr, nt, w = mesh
    inp = make_inputs(scaled_workload(w, nr, nt))
    ctx = _ctx(inp, _lib.ASM_FP64)
    sol, Uraw = so.solve_stokes(inp, rtol=1e-12, itmax=400)
    rep = ctx.solve(_lib.default_opts(null_mode=_lib.NULL_DEFLATE, stop=_lib.STOP_FULL, rtol=1e-11, itmax=400))
    U, P = ctx.solution()
    eu, ep = _errs(U, P, Uraw, sol.P)
    Uv, Uc, omega = ctx.updatevelocity()
    eu2 = np.linalg.norm(Uv - sol.U) / np.linalg.norm(sol.U)
    ec = np.linalg.norm(Uc - sol.Ucart) / np.linalg.norm(sol.Ucart)
    print(f"[result] deflate {nr}x{nt}: raw errU {eu:.2e} errP {ep:.2e} after updatevelocity2 {eu2:.2e} cart {ec:.2e}")
    # solveStokes returns U after updatevelocity2 (StokesSolver.jl:50): that is the compared result.
    # The raw U carries the near-null rotation with amplitude ~ (v.f)/lambda, which follows the
    # rounding of lambda and is removed by the net-rotation step.
    assert ep < 1e-7 and eu2 < 1e-7 and ec < 1e-7
    ctx.close()


def test_reference_stopping_rule_iteration_count(medium):
    """STK_STOP_SCHUR with rtol 1e-8 / itmin 10 stops where the reference's loop stops."""
    ctx = _ctx(medium, _lib.ASM_FP64)
    _, _, _, (K, G, MM, F, ifree) = _oracle_projected(medium)
    v, _ = so.null_vector(K, ifree, medium.mesh)
    _, _, info = so.stokes_pcg(K, MM, G, F, ifree, rtol=1e-8, inner=so.inner_projected(K, ifree, v))
    rep = ctx.solve(_lib.default_opts())
    assert rep.converged
    assert abs(rep.outer_iterations - info.iterations) <= 1
    assert rep.schur_residual < 1e-8
    ctx.close()


def test_updatevelocity2_matches_oracle(medium):
    ctx = _ctx(medium, _lib.ASM_FP64)
    ctx.solve(_lib.default_opts(stop=_lib.STOP_FULL, rtol=1e-11, itmax=400))
    U, P = ctx.solution()
    Ur, Uc = so.updatevelocity2(U, medium.rho, medium.mesh)
    Ug, Ucg, omega = ctx.updatevelocity()
    assert np.abs(Ug - Ur).max() / np.abs(Ur).max() < 1e-12
    assert np.abs(Ucg - Uc).max() / np.abs(Uc).max() < 1e-12
    ctx.close()


def test_error_paths(small):
    ctx = _lib.Context.from_mesh(small.mesh)
    with pytest.raises(_lib.StkError) as e:
        ctx.solve(_lib.default_opts())
    assert e.value.code == _lib.STK_E_STATE
    ctx.set_fields(small.g, small.rho, small.eta_ip, small.C6)
    ctx.assemble(_lib.ASM_FP64)
    with pytest.raises(_lib.StkError) as e:
        ctx.solve(_lib.default_opts(method=7))
    assert e.value.code == _lib.STK_E_ARG
    with pytest.raises(_lib.StkError) as e:
        ctx.solve(_lib.default_opts(method=_lib.SOLVE_MIXED))   # no FP32 operators assembled
    assert e.value.code == _lib.STK_E_STATE
    # iteration cap: reported, not silent
    rep = ctx.solve(_lib.default_opts(itmax=3, itmin=1, rtol=1e-14), allow_notconv=True)
    assert rep.converged == 0 and rep.outer_iterations == 3
    with pytest.raises(_lib.StkError) as e:
        ctx.solve(_lib.default_opts(itmax=3, itmin=1, rtol=1e-14))
    assert e.value.code == _lib.STK_E_NOTCONV
    # a non-finite viscosity is caught by the solve
    eta = small.eta_ip.copy()
    eta[5, 2] = np.nan
    ctx.set_fields(small.g, small.rho, eta, small.C6)
    with pytest.raises(_lib.StkError) as e:
        ctx.assemble(_lib.ASM_FP64)      # the multigrid set-up meets the NaN first
        ctx.solve(_lib.default_opts())
    assert e.value.code == _lib.STK_E_NAN
    ctx.close()


def test_forced_failures_are_reported(small):
    """The failure cases a caller can force: the singular system without boundary conditions is
    refused; a NaN in the right-hand side (density) stops the solve with STK_E_NAN; a negative
    viscosity (K negative definite) is refused by the multigrid set-up or the solve, never solved
    silently (the coarsest multigrid factorisation reports it).  The reference runs all three to
    its iteration cap without a flag."""
    ctx = _lib.Context.from_mesh(small.mesh)
    ctx.set_fields(small.g, small.rho, small.eta_ip, small.C6)
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_NO_BC)
    with pytest.raises(_lib.StkError) as e:
        ctx.solve(_lib.default_opts())
    assert e.value.code == _lib.STK_E_STATE
    rho = small.rho.copy()
    rho[7] = np.nan
    ctx.set_fields(small.g, rho, small.eta_ip, small.C6)
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
    for method in (_lib.SOLVE_PLAIN, _lib.SOLVE_MIXED):
        with pytest.raises(_lib.StkError) as e:
            ctx.solve(_lib.default_opts(method=method))
        assert e.value.code == _lib.STK_E_NAN
    ctx.set_fields(small.g, small.rho, -small.eta_ip, small.C6)
    with pytest.raises(_lib.StkError) as e:
        ctx.assemble(_lib.ASM_FP64)
        ctx.solve(_lib.default_opts())
    print(f"[result] negative viscosity refused with code {e.value.code}: {e.value}")
    assert e.value.code == _lib.STK_E_SHAPE
    ctx.close()


def test_fp32_only_assembly(medium):
    """STK_ASM_FP32 alone assembles the operators in FP32 (the FP32 element kernels, Galerkin
    levels and set-up): K and G equal the FP64 assembly to FP32 arithmetic (measured 2.5e-6 of
    the largest entry: sums over six points and two rotations); the solves that need
    FP64 operators (plain, and the refinement of mixed) are refused, not run on missing data."""
    ref = _ctx(medium, _lib.ASM_FP64)
    K64, G64 = ref.export_csc(_lib.MAT_K), ref.export_csc(_lib.MAT_G)
    ref.close()
    ctx = _ctx(medium, _lib.ASM_FP32)
    K32, G32 = ctx.export_csc(_lib.MAT_K), ctx.export_csc(_lib.MAT_G)
    ek, eg = _rel(K32, K64), _rel(G32, G64)
    print(f"[result] FP32-only assembly vs FP64: K {ek:.2e} G {eg:.2e}")
    assert ek < 1e-5 and eg < 1e-5
    for method in (_lib.SOLVE_PLAIN, _lib.SOLVE_MIXED):
        with pytest.raises(_lib.StkError) as e:
            ctx.solve(_lib.default_opts(method=method, stop=_lib.STOP_FULL))
        assert e.value.code == _lib.STK_E_STATE
    ctx.close()


def test_bad_mesh_is_rejected():
    m = make_inputs(scaled_workload(1, 4, 16)).mesh
    e2n = m.e2n.copy()
    e2n[0, 0] = m.nnod + 5
    with pytest.raises(_lib.StkError) as e:
        _lib.Context(m.nel, m.nnod, m.nvert, e2n + 1, m.el_theta, m.el_r, m.theta, m.r, m.free_slip()[0] + 1)
    assert e.value.code == _lib.STK_E_SHAPE


@pytest.mark.parametrize("size", [(4, 16), (32, 256)])
def test_colourings_are_valid(size):
    """check_colored_mesh (src/Grid/coloring.jl:97-111) for the assembly's element colouring: no
    two elements of a colour share a node; and for the smoother's node colouring: no two nodes
    coupled in K share a colour."""
    m = make_inputs(scaled_workload(1, *size)).mesh
    print(f"[case] colourings {size}", flush=True)
    ctx = _lib.Context.from_mesh(m)
    ec, nec = ctx.colouring(_lib.COLOUR_ELEMENTS)
    assert ec.min() == 0 and ec.max() == nec - 1
    for k in range(nec):
        nodes = np.asarray(m.e2n)[:, ec == k].ravel()
        assert np.unique(nodes).size == nodes.size, k
    inp = make_inputs(scaled_workload(1, *size))
    ctx.set_fields(inp.g, inp.rho, inp.eta_ip, inp.C6)
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_NO_BC)
    K = ctx.export_csc(_lib.MAT_K).tocoo()
    nc, nnc = ctx.colouring(_lib.COLOUR_NODES)
    i, j = K.row // 2, K.col // 2
    off = i != j
    assert np.all(nc[i[off]] != nc[j[off]])
    assert nnc == nc.max() + 1
    print(f"[result] {nec} element colours, {nnc} node colours, {off.sum()} coupled node pairs checked")
    ctx.close()


def _fields7(inp):
    """eta.ip at the 7 points of ip_triangle7 (the workload's law at each point, as make_inputs does at
    the 6 points) and the tensor in 7 columns."""
    from dimspers_amd.quadrature import ip_triangle7, sf_N_tri6
    m = inp.mesh
    x7, _ = ip_triangle7()
    N = sf_N_tri6(x7[:, 0], x7[:, 1])                       # (7 points, 6 nodes)
    Tq = (N @ inp.T[m.e2n]).T                               # (nel, 7)
    eta7 = np.asfortranarray(np.exp(-np.log(inp.workload.contrast) * Tq))
    C7 = tuple(np.asfortranarray(np.repeat(np.asarray(c)[:, :1], 7, axis=1)) for c in inp.C6)
    return StokesInputs(m, inp.workload, inp.T, inp.rho, eta7, C7, inp.phi, inp.delta)


@pytest.mark.parametrize("variant", ["7point", "clamp", "clamp_dimensional"])
def test_assembly_variants_match_oracle(small, variant):
    """The 7-point rule of the threaded assembler (AssemblerStokes.jl:174) and the clamp of the top-level
    _fill_Kblk! (StokesSolver.jl:756): Cartesian operators against the oracle within 1e-12; the clamp on
    non-dimensional and on dimensional (eta ~ 1e21) viscosities."""
    inp, nip, clamp = small, 6, False
    if variant == "7point":
        inp, nip = _fields7(small), 7
    elif variant.startswith("clamp"):
        clamp = True
        if variant == "clamp_dimensional":
            inp = StokesInputs(small.mesh, small.workload, small.T, small.rho, small.eta_ip * 1e21, small.C6,
                               small.phi, small.delta)
    flags = _lib.ASM_FP64 | _lib.ASM_CARTESIAN | _lib.ASM_NO_BC
    flags |= _lib.ASM_7POINT if nip == 7 else 0
    flags |= _lib.ASM_CLAMP if clamp else 0
    ctx = _ctx(inp, flags)
    KK, GG, MM, Fb = so.assembly_stokes_cylindric(inp, nip=nip, clamp=clamp)
    errs = (_rel(ctx.export_csc(_lib.MAT_K, lower_only=True), KK), _rel(ctx.export_csc(_lib.MAT_G), GG),
            _rel(ctx.export_csc(_lib.MAT_M, lower_only=True), MM),
            np.abs(ctx.export_rhs() - Fb).max() / np.abs(Fb).max())
    print(f"[result] assembly {variant}: K {errs[0]:.1e} G {errs[1]:.1e} M {errs[2]:.1e} F {errs[3]:.1e}", flush=True)
    assert max(errs) < 1e-12
    ctx.close()


def test_seven_point_operators_solve_like_the_oracle():
    """The 7-point rule through the whole path (rotation, boundary conditions, the multigrid built
    from the 7-point element matrices, the mixed method) at 8 x 64 against the projected oracle."""
    inp = _fields7(make_inputs(scaled_workload(1, 8, 64)))
    ctx = _ctx(inp, _lib.ASM_FP64 | _lib.ASM_FP32 | _lib.ASM_7POINT)
    Uo, Po, _, _ = _oracle_projected(inp, nip=7)
    rep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_MIXED, stop=_lib.STOP_FULL, rtol=1e-10, itmax=400))
    U, P = ctx.solution()
    eu, ep = _errs(U, P, Uo, Po)
    print(f"[result] 7-point mixed: outer {rep.outer_iterations} inner {rep.inner_iterations} errU {eu:.2e} "
          f"errP {ep:.2e}", flush=True)
    assert rep.converged and eu < 1e-7 and ep < 1e-7
    # the 6-point assembly of the same fields is a different operator
    ctx6 = _ctx(inp, _lib.ASM_FP64)
    assert _rel(ctx6.export_csc(_lib.MAT_K), ctx.export_csc(_lib.MAT_K)) > 1e-8
    ctx.close()
    ctx6.close()
