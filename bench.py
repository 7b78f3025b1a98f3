"""Stokes solve benchmark on one MI355X.

Workload: the learned / mixed-precision configuration at 256 x 2048 cells
(1,048,576 quadratic triangles, 4,202,496 velocity + 526,336 pressure dofs):
viscosity contrast 1e4, anisotropy ratio uniform in [1, 100], seeded fields.
One step = one Schur-complement solve, from zero, of the assembled system to a
1e-8 relative full-system residual (FP64, evaluated on the device).  The
operators are assembled and resident in HBM before the timed region; assembly
is timed separately.  The plain FP64 method is timed on the same system as the
comparison point, and the errors of the fast solve are measured against it.

Prints one JSON line (see the repository documentation for the fields).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# FP64 flop per element of the assembly kernel, counted from its loops: per integration point
# (6) the derivative weights (~50), D B (108 FMA), the lower triangle of B^T (D B) (234 FMA), G (36 FMA),
# M (~9) and F (~12 FMA); then the node-frame rotation of the 21 K blocks (~336 FMA) and the P1
# projection of the element matrix for the multigrid (~650 FMA).
ASM_FLOP_PER_ELEMENT = 6 * (50 + 2 * (108 + 234 + 36 + 12) + 9) + 2 * (336 + 650)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


# Inner iterations of the plain FP64 method with the point-Jacobi inner CG (the algorithm of the CPU
# restatement) to the 1e-8 full residual, per workload index, measured on the GPU by
# `STK_TRACE=1 python bench.py --jacobi-count --workload W` (FP64 throughout, so the count is the
# algorithm's).
JACOBI_INNER_ITERATIONS = {4: 1_020_144}


def cpu_baseline(inp, workload, budget_s):
    """The C++/OpenMP CPU restatement (oracle/cpu_stokes.cpp: element assembly, node frames, free
    slip, StokesPcCG with the lumped Schur preconditioner and a projected point-Jacobi CG inner
    solve) on all host cores, on the benchmarked inputs: the whole assembly, then ``budget_s`` of
    the solve.  The solve time is the measured time per inner iteration times the iteration
    count of the same algorithm at this size."""
    from oracle import cpu_port

    rep, _, _ = cpu_port.run(inp, budget_s=budget_s, inner_rtol=1e-10)
    n_it = JACOBI_INNER_ITERATIONS.get(workload)
    value = None if n_it is None else (rep["assembly_s"] + n_it * rep["s_per_inner_iteration"]) * 1000.0
    return rep, n_it, value


def superlu_aside(workload):
    """The FP64 numpy/SciPy oracle (StokesPcCG with the SuperLU inner solve, one core) on the
    same workload definition at 64 x 512 cells."""
    from dimspers_amd.inputs import make_inputs, scaled_workload
    from oracle import stokes_oracle as so

    inp = make_inputs(scaled_workload(workload, 64, 512))
    t0 = time.perf_counter()
    sol, _ = so.solve_stokes(inp, rtol=1e-8)
    dt = time.perf_counter() - t0
    return {"s": round(dt, 2), "cells": "64x512", "dofs": inp.mesh.nU + inp.mesh.nP,
            "outer_iterations": sol.info.iterations, "cores": 1,
            "note": "numpy/SciPy oracle, SuperLU inner solve; not scaled"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", type=int, default=4)
    ap.add_argument("--method", default="learned", choices=["mixed", "learned"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--kernels", action="store_true",
                    help="only time the dominant kernel (50 launches) and the copy reference (for counter runs)")
    ap.add_argument("--rates", action="store_true",
                    help="time every kernel family of the solve with its algorithmic bytes (GB/s, fraction of 8 TB/s)")
    ap.add_argument("--diagnostics", action="store_true",
                    help="only time the SpMV, one V-cycle and one inner CG iteration")
    ap.add_argument("--jacobi-count", action="store_true",
                    help="only count the inner iterations of the plain FP64 method with the point-Jacobi inner CG "
                         "(the CPU restatement's algorithm) to the 1e-8 full residual")
    ap.add_argument("--cpu-budget", type=float, default=20.0,
                    help="seconds of CPU-restatement solve time in the CPU baseline sample")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the path is not sharded across GPUs.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from dimspers_amd import _lib
    from dimspers_amd.inputs import WORKLOADS, make_inputs

    w = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    inp = make_inputs(w)
    m = inp.mesh
    log(f"workload {w.index}: {w.ncell_r}x{w.ncell_t} cells, nel {m.nel}, nU {m.nU}, nP {m.nP}, "
        f"inputs {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    ctx = _lib.Context.from_mesh(m)
    log(f"context (patterns, colouring, multigrid structure) {time.perf_counter() - t0:.1f}s, "
        f"{ctx.ncolours} colours, {ctx.mg_levels} P1 levels")
    ctx.set_fields(inp.g, inp.rho, inp.eta_ip, inp.C6)
    t0 = time.perf_counter()
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
    # first assembly of the context: also builds the multigrid operators, the coarsest inverse and
    # the near-null vector; a time step's re-assembly repeats the same work (timed below)
    first_setup_ms = (time.perf_counter() - t0) * 1000.0
    log(f"assembly + solver set-up {first_setup_ms / 1000.0:.2f}s")
    asm_ms, asm_bytes = ctx.time_kernel(_lib.KERNEL_ASSEMBLY, reps=3)
    log(f"element assembly kernels {asm_ms:.2f} ms ({asm_bytes / asm_ms / 1e6:.0f} GB/s compulsory)")

    if args.rates:
        res = {"element_assembly_fp64": {"avg_us": round(asm_ms * 1e3, 1), "bytes": asm_bytes,
                                         "GB_per_s": round(asm_bytes / asm_ms / 1e6, 1),
                                         "note": "k_element and the gathers, compulsory bytes"}}
        for name, kid, reps in (("spmv_k32", _lib.KERNEL_SPMV_K32, 50), ("spmv_k64", _lib.KERNEL_SPMV_K64, 50),
                                ("cheb_k32", _lib.KERNEL_CHEB_K32, 50), ("cheb_k64", _lib.KERNEL_CHEB_K64, 50),
                                ("cg_product_kp32", _lib.KERNEL_KP32, 50),
                                ("gs_backward_sweep32_per_launch", _lib.KERNEL_GS_BWD32, 50),
                                ("g32", _lib.KERNEL_G32, 50), ("gt32", _lib.KERNEL_GT32, 50),
                                ("coarsest_dense32", _lib.KERNEL_COARSE32, 50),
                                ("restriction32_level0", _lib.KERNEL_RESTRICT32, 50),
                                ("cg_update_and_norm32", _lib.KERNEL_CGUPD32, 50), ("copy_256MiB", _lib.KERNEL_COPY, 20)):
            kms, kb = ctx.time_kernel(kid, reps=reps)
            res[name] = {"avg_us": round(kms * 1e3, 2), "bytes": kb, "GB_per_s": round(kb / kms / 1e6, 1),
                         "frac_of_8TBps": round(kb / kms / 1e6 / 8000.0, 3)}
            log(f"{name}: {kms * 1e3:.1f} us, {kb / kms / 1e6:.0f} GB/s")
        ctx.close()
        print(json.dumps({"rates": res}), flush=True)
        return

    if args.jacobi_count:
        # a solve of several minutes: STK_TRACE=1 in the environment (read once, at the library's first
        # trace check) prints one progress line per inner solve on stderr
        if os.environ.get("STK_TRACE", "0") in ("", "0"):
            log("STK_TRACE=1 in the environment prints progress during this long solve")
        jopts = _lib.default_opts(method=_lib.SOLVE_PLAIN, inner=_lib.INNER_JACOBI, stop=_lib.STOP_FULL,
                                  rtol=1e-8, itmax=400, inner_rtol=1e-10, inner_itmax=10_000_000)
        rep = ctx.solve(jopts)
        ctx.close()
        print(json.dumps({"jacobi_count": {"workload": w.index, "outer_iterations": rep.outer_iterations,
                                           "inner_iterations": int(rep.inner_iterations),
                                           "refinements": rep.refinements, "converged": bool(rep.converged),
                                           "full_residual": rep.full_residual, "solve_ms": round(rep.solve_ms, 1)}}),
              flush=True)
        return

    if args.kernels or args.diagnostics:
        res = {}
        kernels = ((("gs_backward_sweep32_per_launch", _lib.KERNEL_GS_BWD32, 50), ("kp32", _lib.KERNEL_KP32, 50),
                    ("copy_256MiB", _lib.KERNEL_COPY, 20)) if args.kernels else
                   (("spmv_k32", _lib.KERNEL_SPMV_K32, 50), ("vcycle32", _lib.KERNEL_VCYCLE32, 20),
                    ("inner_cg_iteration32", _lib.KERNEL_INNER_IT32, 5)))
        for name, kid, reps in kernels:
            kms, kb = ctx.time_kernel(kid, reps=reps)
            res[name] = {"avg_us": round(kms * 1e3, 2), "bytes": kb, "GB/s": round(kb / kms / 1e6, 1) if kb else None}
            log(f"{name}: {kms * 1e3:.1f} us" + (f", {kb / kms / 1e6:.0f} GB/s" if kb else ""))
        ctx.close()
        print(json.dumps({"kernels": res}), flush=True)
        return

    method = _lib.SOLVE_LEARNED if args.method == "learned" else _lib.SOLVE_MIXED
    fast_opts = _lib.default_opts(method=method, stop=_lib.STOP_FULL, rtol=1e-8)
    mixed_opts = _lib.default_opts(method=_lib.SOLVE_MIXED, stop=_lib.STOP_FULL, rtol=1e-8)
    # plain FP64 at its fastest setting that keeps its outer convergence (inner 1e-10:
    # same 53 Schur iterations and residual as 1e-11 at 256 x 2048, 10% less time)
    plain_opts = _lib.default_opts(method=_lib.SOLVE_PLAIN, stop=_lib.STOP_FULL, rtol=1e-8, itmax=400,
                                   inner_rtol=1e-10)

    train = None
    setup_ms = 0.0
    if method == _lib.SOLVE_LEARNED:
        ctx.set_fabric(inp.phi, inp.delta)
        trep = ctx.mlp_train(_lib.default_train_opts())
        train = {"ms": round(trep.ms, 1), "samples": trep.samples, "steps": trep.steps,
                 "initial_loss": trep.initial_loss, "final_loss": trep.final_loss,
                 "validation_loss": trep.validation_loss}
        log(f"MLP training on the GPU: {trep.ms:.0f} ms, {trep.samples} samples, {trep.steps} steps, "
            f"loss {trep.initial_loss:.3f} -> {trep.final_loss:.4f} (validation {trep.validation_loss:.4f})")
        ctx.learned_setup()
        t0 = time.perf_counter()
        for _ in range(3):
            ctx.learned_setup()
        setup_ms = (time.perf_counter() - t0) * 1000.0 / 3
        log(f"learned preconditioner set-up (features, MLP inference, assembly of S): {setup_ms:.2f} ms")

    for k in range(args.warmup):
        rep = ctx.solve(fast_opts)
        log(f"warmup {k}: {rep.solve_ms:.1f} ms, {rep.outer_iterations} outer, {rep.inner_iterations} inner, "
            f"residual {rep.full_residual:.2e}")
    times, reps = [], []
    t_all = time.perf_counter()
    for k in range(args.steps):
        rep = ctx.solve(fast_opts)
        times.append(rep.solve_ms)
        reps.append(rep)
        log(f"step {k}: {rep.solve_ms:.1f} ms, {rep.outer_iterations} outer, {rep.inner_iterations} inner, "
            f"{rep.refinements} sweeps, residual {rep.full_residual:.2e}")
    wall = (time.perf_counter() - t_all) * 1000.0 / args.steps + setup_ms
    U_fast, P_fast = ctx.solution()
    ms = float(np.mean(times)) + setup_ms

    mixed = None
    mixed_rep = plain_rep = None
    if method == _lib.SOLVE_LEARNED:
        mrep = mixed_rep = ctx.solve(mixed_opts)
        mixed = {"ms": round(mrep.solve_ms, 2), "outer": mrep.outer_iterations, "inner": int(mrep.inner_iterations),
                 "residual": mrep.full_residual}
        log(f"mixed without the MLP: {mrep.solve_ms:.1f} ms, {mrep.outer_iterations} outer, "
            f"{mrep.inner_iterations} inner, residual {mrep.full_residual:.2e}")

    plain = None
    if not args.no_plain:
        prep = plain_rep = ctx.solve(plain_opts)
        U_pl, P_pl = ctx.solution()
        eu = float(np.linalg.norm(U_fast - U_pl) / np.linalg.norm(U_pl))
        ep = float(np.linalg.norm((P_fast - P_fast.mean()) - (P_pl - P_pl.mean())) / np.linalg.norm(P_pl - P_pl.mean()))
        plain = dict(ms=prep.solve_ms, outer=prep.outer_iterations, inner=int(prep.inner_iterations),
                     residual=prep.full_residual, errU=eu, errP=ep)
        log(f"plain FP64: {prep.solve_ms:.1f} ms, {prep.outer_iterations} outer, {prep.inner_iterations} inner, "
            f"residual {prep.full_residual:.2e}; fast-vs-plain errU {eu:.2e} errP {ep:.2e}")
        # errors of the fast solve against a plain FP64 solve three orders tighter (1e-11 full residual;
        # untimed): the accuracy the 1e-7 bar is measured against
        trep = ctx.solve(_lib.default_opts(method=_lib.SOLVE_PLAIN, stop=_lib.STOP_FULL, rtol=1e-11, itmax=800))
        U_t, P_t = ctx.solution()
        plain["vs_tight_fp64"] = dict(
            rtol=1e-11, residual=trep.full_residual,
            errU=float(np.linalg.norm(U_fast - U_t) / np.linalg.norm(U_t)),
            errP=float(np.linalg.norm((P_fast - P_fast.mean()) - (P_t - P_t.mean())) / np.linalg.norm(P_t - P_t.mean())))
        log(f"fast solve vs a 1e-11 FP64 solve: errU {plain['vs_tight_fp64']['errU']:.2e} "
            f"errP {plain['vs_tight_fp64']['errP']:.2e}")

    # per time step the reference re-assembles and refactorises: time one re-assembly with the
    # solver set-up it triggers (multigrid operators, coarsest inverse, near-null vector)
    t0 = time.perf_counter()
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
    resetup_ms = (time.perf_counter() - t0) * 1000.0
    log(f"re-assembly with the solver set-up: {resetup_ms:.1f} ms")

    # stress post-processor on the same mesh (7-point rule, anisotropic law)
    eta7 = np.asfortranarray(np.repeat(np.asarray(inp.eta_ip)[:, :1], 7, axis=1))
    C7 = [np.asfortranarray(np.repeat(np.asarray(c)[:, :1], 7, axis=1)) for c in inp.C6]
    ctx.set_stress_fields(eta7, C7, nip=7)
    st_ms, st_bytes = ctx.time_kernel(_lib.KERNEL_STRESS, reps=20)
    fg_ms, fg_bytes = ctx.time_kernel(_lib.KERNEL_DEFGRAD, reps=20)
    log(f"stress!: {st_ms * 1e3:.1f} us ({st_bytes / st_ms / 1e6:.0f} GB/s compulsory); "
        f"deformation_gradient!: {fg_ms * 1e3:.1f} us ({fg_bytes / fg_ms / 1e6:.0f} GB/s)")

    # fabric chain on the same mesh: getFSE on 7 points per element, then anisotropic_tensor on the
    # DEM table DEM_1e-3_vol30 (the reference's newDEM data, kept as a test fixture)
    fse_ms, fse_bytes = ctx.time_kernel(_lib.KERNEL_FSE, reps=20)
    dem = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "dem_tables.npz"))
    ctx.set_dem(dem["DEM_1em3_vol30_C"], dem["DEM_1em3_vol30_ax"], 0.03)
    ang = np.linspace(0.0, np.pi, 7 * m.nel).reshape(m.nel, 7)
    Fsyn = np.zeros((m.nel, 7, 2, 2))
    Fsyn[..., 0, 0], Fsyn[..., 1, 0] = 3.0 * np.cos(ang), 3.0 * np.sin(ang)
    Fsyn[..., 0, 1], Fsyn[..., 1, 1] = -np.sin(ang) / 3.0, np.cos(ang) / 3.0
    ctx.fse(Fsyn)
    del Fsyn, ang
    an_ms, an_bytes = ctx.time_kernel(_lib.KERNEL_ANISO, reps=20)
    log(f"getFSE: {fse_ms * 1e3:.1f} us ({fse_bytes / fse_ms / 1e6:.0f} GB/s); "
        f"anisotropic_tensor: {an_ms * 1e3:.1f} us ({an_bytes / an_ms / 1e6:.0f} GB/s)")

    # one thermal diffusion step on the same mesh (heated bottom, non-dimensional K = Cp = rho = 1)
    nn = m.theta.size
    top, bot = np.flatnonzero(m.r == m.r.max()), np.flatnonzero(m.r == m.r.min())
    tfix = np.concatenate([top, bot]) + 1
    tval = np.concatenate([np.zeros(top.size), np.ones(bot.size)])
    ctx.thermal_solve(inp.T, np.ones(nn), 1.0, 1.0, 1e-4, tfix, tval)
    _, trep = ctx.thermal_solve(inp.T, np.ones(nn), 1.0, 1.0, 1e-4, tfix, tval)
    log(f"thermal step (assembly + BCs + Jacobi-CG to 1.5e-8): {trep.solve_ms:.2f} ms, {trep.outer_iterations} its")

    # dominant kernel of the solve by total time: the backward Gauss-Seidel colour kernel
    # (seven launches per V-cycle, timed over whole sweeps, reported per launch)
    kms, kbytes = ctx.time_kernel(_lib.KERNEL_GS_BWD32, reps=50)
    achieved = kbytes / (kms * 1e-3) / 1e9
    pms, pbytes = ctx.time_kernel(_lib.KERNEL_KP32, reps=50)
    sms, sbytes = ctx.time_kernel(_lib.KERNEL_SPMV_K32, reps=50)
    log(f"dominant kernel (backward Gauss-Seidel colour launch, FP32): {kms * 1e3:.1f} us, {achieved:.0f} GB/s; "
        f"CG product K p {pms * 1e3:.1f} us, {pbytes / pms / 1e6:.0f} GB/s; "
        f"plain FP32 SpMV {sms * 1e3:.1f} us, {sbytes / sms / 1e6:.0f} GB/s")

    # one time step of each method as a time loop runs it: re-assembly with the solver set-up it
    # triggers, then the solve.  The fast methods assemble FP64 + FP32 (FP32 multigrid, coarsest
    # inverse, near-null vector); the plain method alone assembles FP64 only (FP64 multigrid and
    # near-null vector).  MLP training is excluded: the network is trained once and reused.
    per_step = {"learned_or_mixed_ms": round(resetup_ms + ms, 1), "reassembly_and_setup_fast_ms": round(resetup_ms, 1)}
    if mixed is not None:
        per_step["mixed_ms"] = round(resetup_ms + mixed["ms"], 1)
    if not args.no_plain:
        t0 = time.perf_counter()
        ctx.assemble(_lib.ASM_FP64)
        asm64_ms = (time.perf_counter() - t0) * 1000.0
        prep2 = ctx.solve(plain_opts)
        per_step["plain_ms"] = round(asm64_ms + prep2.solve_ms, 1)
        per_step["reassembly_and_setup_fp64_ms"] = round(asm64_ms, 1)
        per_step["speedup_vs_plain"] = round(per_step["plain_ms"] / per_step["learned_or_mixed_ms"], 2)
        log(f"per time step: fast {per_step['learned_or_mixed_ms']:.1f} ms, plain {per_step['plain_ms']:.1f} ms "
            f"(FP64-only re-assembly + set-up {asm64_ms:.1f} ms)")

    cpu = None
    aside = None
    if not args.no_cpu:
        crep, n_it, cval = cpu_baseline(inp, w.index, args.cpu_budget)
        log(f"CPU restatement: {crep['threads']} threads, assembly {crep['assembly_s']:.2f} s, "
            f"{crep['inner_iterations']} inner iterations in {crep['solve_s']:.1f} s "
            f"({crep['s_per_inner_iteration'] * 1e3:.2f} ms each)")
        cpu = {"value": None if cval is None else round(cval, 1), "unit": "ms", "cores": crep["threads"],
               "kind": "port",
               "sample": (f"C++/OpenMP FP64 restatement of assembly_stokes_cylindric + _prepare_matrices + StokesPcCG "
                          f"(lumped Schur PCG, projected point-Jacobi CG inner solve to 1e-10) on the benchmarked "
                          f"{w.ncell_r}x{w.ncell_t} inputs, {crep['threads']} OpenMP threads of {os.cpu_count()} "
                          f"host CPUs: whole assembly {crep['assembly_s']:.2f} s, then {crep['solve_s']:.1f} s of "
                          f"the solve ({crep['inner_iterations']} inner iterations, "
                          f"{crep['s_per_inner_iteration'] * 1e3:.2f} ms each); value = assembly + that time per "
                          f"iteration x " + (f"{n_it} inner iterations (the count of the same algorithm to the "
                                              f"1e-8 full residual at this size, measured by bench.py --jacobi-count)"
                                              if n_it is not None else "the iteration count, not measured for "
                                                                      "this workload: value left null")),
               "assembly_s": crep["assembly_s"], "s_per_inner_iteration": crep["s_per_inner_iteration"],
               "sample_inner_iterations": crep["inner_iterations"], "total_inner_iterations": n_it}
        aside = superlu_aside(w.index)
        log(f"SciPy/SuperLU oracle at 64x512 (aside): {aside['s']:.2f} s")

    def bandwidth(rep_):
        # the solve's own algorithmic bytes (kernel operands, as the library counts them) over its time
        gbs = rep_.bytes / (rep_.solve_ms * 1e-3) / 1e9
        return {"bytes": rep_.bytes, "GB_per_s": round(gbs, 1), "frac_of_8TBps": round(gbs / PEAK_HBM_GBS, 4),
                "launches": int(rep_.launches), "launches_at_8B_floor": int(rep_.launches_floor)}

    solve_bw = {args.method: bandwidth(reps[-1])}
    if mixed_rep is not None:
        solve_bw["mixed"] = bandwidth(mixed_rep)
    if plain_rep is not None:
        solve_bw["plain_fp64"] = bandwidth(plain_rep)
    log("solve bandwidth: " + ", ".join(f"{k} {v['GB_per_s']:.0f} GB/s ({v['launches']} launches)"
                                         for k, v in solve_bw.items()))

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_gs_backward_k32_256x2048.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("bytes_per_launch")

    ndof = m.nU + m.nP
    out = {
        "metric": "Stokes solve time to 1e-8 residual (learned+mixed vs plain FP64 PCG), 1 GPU",
        "value": round(ms, 2),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 inner / fp64 refinement",
        "data": "synthetic (seeded fields of the workload definition)",
        "config": {"workload": f"annulus {w.ncell_r}x{w.ncell_t} cells, {m.nel} P2-P1 triangles, "
                               f"contrast {w.contrast:g}, anisotropy {w.ratio}", "method": args.method,
                   "nU": m.nU, "nP": m.nP},
        "dofs_per_s": round(ndof / (ms * 1e-3)),
        "outer_iterations": reps[-1].outer_iterations,
        "inner_iterations": int(reps[-1].inner_iterations),
        "refinements": reps[-1].refinements,
        "fallbacks": reps[-1].fallbacks,
        "full_residual": reps[-1].full_residual,
        "schur_residual_last_sweep": reps[-1].schur_residual,
        "assembly_ms": round(asm_ms, 3),
        "assembly_and_solver_setup_ms": round(resetup_ms, 1),
        "assembly": {"ms": round(asm_ms, 3), "compulsory_GB_per_s": round(asm_bytes / asm_ms / 1e6, 1),
                     "GFLOP_per_s_fp64": round(ASM_FLOP_PER_ELEMENT * m.nel / asm_ms / 1e6, 1),
                     "flop_per_element": ASM_FLOP_PER_ELEMENT},
        "learned_setup_ms": round(setup_ms, 3),
        "mlp_training": train,
        "mixed_without_mlp": mixed,
        "plain_fp64": plain,
        "speedup_vs_plain": round(plain["ms"] / ms, 2) if plain else None,
        "value_per_step": per_step["learned_or_mixed_ms"],
        "per_time_step": per_step,
        "solve_bandwidth": solve_bw,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "traffic_source": "carried over from profiles/traffic_gs_backward_k32_256x2048.json "
                                       "(rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes over bench.py --kernels, "
                                       "scaled by a 256 MiB copy); not measured in this run",
                     "kernel": "k_gs_colour<256,false,2> (backward multicolour block Gauss-Seidel sweep on K, FP32; "
                               "seven colour launches per sweep, figures per launch)",
                     "avg_us": round(kms * 1e3, 2), "bytes_per_launch": kbytes},
        "cg_product_kp": {"kernel": "k_spmv_sell22<256,float,float,EpStoreDot2<float>,XPlain<float>>",
                          "avg_us": round(pms * 1e3, 2), "bytes_per_launch": pbytes,
                          "GB_per_s": round(pbytes / pms / 1e6, 1)},
        "cpu_baseline": cpu,
        "cpu_aside_superlu_64x512": aside,
        "thermal_step": {"ms": round(trep.solve_ms, 2), "cg_iterations": trep.outer_iterations,
                         "dt": 1e-4, "relative_residual": trep.schur_residual},
        "stress_postprocessor": {
            "stress_us": round(st_ms * 1e3, 1), "stress_GB_per_s": round(st_bytes / st_ms / 1e6, 1),
            "deformation_gradient_us": round(fg_ms * 1e3, 1),
            "deformation_gradient_GB_per_s": round(fg_bytes / fg_ms / 1e6, 1)},
        "fabric": {
            "fse_us": round(fse_ms * 1e3, 1), "fse_GB_per_s": round(fse_bytes / fse_ms / 1e6, 1),
            "anisotropic_tensor_us": round(an_ms * 1e3, 1),
            "anisotropic_tensor_GB_per_s": round(an_bytes / an_ms / 1e6, 1), "points": 7 * m.nel},
    }
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
