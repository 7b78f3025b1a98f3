"""The C ABI driven from C (tests/c_abi/abi_driver.c), as a Julia ccall binding would drive it:
the header compiles as C11 and the driver links against the library (CPU), and on the GPU the
C driver's solve equals the ctypes path's bit for bit."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from dimspers_amd import _lib
from dimspers_amd.inputs import make_inputs, scaled_workload
from oracle import cpu_port

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "c_abi" / "abi_driver.c"
LIBDIR = ROOT / "dimspers_amd"


def _compile(out: Path) -> Path:
    exe = out / "abi_driver"
    cmd = ["gcc", "-std=c11", "-O2", "-Wall", "-Werror", "-I", str(ROOT / "include"), str(SRC), "-o", str(exe),
           "-L", str(LIBDIR), "-ldimspers_stokes", f"-Wl,-rpath,{LIBDIR}", "-Wl,-rpath-link,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    return exe


def test_header_compiles_as_c_and_the_driver_links(tmp_path):
    if not (LIBDIR / "libdimspers_stokes.so").exists():
        pytest.skip("library not built")
    assert _compile(tmp_path).exists()


def _write_inputs(inp, d: Path):
    cpu_port.write_inputs(inp, d)


@pytest.mark.gpu
@pytest.mark.parametrize("method", [_lib.SOLVE_MIXED, _lib.SOLVE_PLAIN])
def test_c_driver_matches_the_ctypes_path(tmp_path, method):
    inp = make_inputs(scaled_workload(3, 16, 128))
    print(f"[case] C ABI driver 16x128 method {method}", flush=True)
    exe = _compile(tmp_path)
    _write_inputs(inp, tmp_path)
    res = subprocess.run([str(exe), str(tmp_path), str(method)], capture_output=True, text=True, timeout=120,
                         env=dict(os.environ))
    print(res.stdout.strip(), res.stderr.strip()[-2000:])
    assert res.returncode == 0
    conv, outer, inner, resid, rc_bad, msg = (tmp_path / "report.txt").read_text().split()
    assert int(conv) == 1 and float(resid) < 1e-8
    assert int(rc_bad) == _lib.STK_E_ARG and msg == "message"
    U_c = np.fromfile(tmp_path / "U.bin", dtype=np.float64)
    P_c = np.fromfile(tmp_path / "P.bin", dtype=np.float64)
    ctx = _lib.Context.from_mesh(inp.mesh)
    ctx.set_fields(inp.g, inp.rho, inp.eta_ip, inp.C6)
    ctx.assemble(_lib.ASM_FP64 | _lib.ASM_FP32)
    rep = ctx.solve(_lib.default_opts(method=method, stop=_lib.STOP_FULL, rtol=1e-8))
    U, P = ctx.solution()
    ctx.close()
    assert rep.outer_iterations == int(outer) and rep.inner_iterations == int(inner)
    assert np.array_equal(U, U_c) and np.array_equal(P, P_c)


def _write_operators(d: Path, K, MM, G, F, P0, k_lower, warm, ppc, stop, null_mode, itmax, rtol, inner_rtol):
    import scipy.sparse as sp

    def put(name, A):
        A = sp.csc_matrix(A)
        A.sort_indices()
        (d / f"{name}colptr.bin").write_bytes(np.asarray(A.indptr + 1, dtype=np.int64).tobytes())
        (d / f"{name}rowval.bin").write_bytes(np.asarray(A.indices + 1, dtype=np.int64).tobytes())
        (d / f"{name}nzval.bin").write_bytes(np.asarray(A.data, dtype=np.float64).tobytes())
        return A.nnz
    Ks = sp.tril(K, format="csc") if k_lower else K
    nK, nM, nG = put("K", Ks), put("M", MM), put("G", G)
    (d / "Rhs.bin").write_bytes(np.asarray(F, dtype=np.float64).tobytes())
    if P0 is not None:
        (d / "P0.bin").write_bytes(np.asarray(P0, dtype=np.float64).tobytes())
    (d / "ops.txt").write_text(f"{nK} {nM} {nG} {int(k_lower)} 1 {int(warm)} {ppc} {stop} {null_mode} {itmax} "
                               f"{rtol!r} {inner_rtol!r}\n")


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["StokesPcCG", "StokesPcCG_pardiso"])
def test_c_driver_solves_operators_the_caller_assembled(tmp_path, variant):
    """StokesPcCG(U, P, KK, MM, GG, Rhs, ifree) on operators the caller built (here: the oracle's
    assembly, _prepare_matrices and MM), handed over through stk_set_operators from C.

    StokesPcCG (IterativeSolvers.jl:94-168): zero start, lumped preconditioner, Schur rtol 1e-8 -> the
    oracle's literal algorithm (LU inner solve); the null-space deflation reproduces the LU solve of
    K[ifree, ifree], whose near-null rotation FP64 still resolves at this size.
    StokesPcCG_pardiso (:1-92): the incoming pressure kept (Rhs .+= GG*P), the incomplete Cholesky of
    MM, rtol 1e-6 -> the oracle's same iteration from the same P0 with IC(0) of MM: the same iteration
    count (within one) and the same iterate."""
    from oracle import stokes_oracle as so
    inp = make_inputs(scaled_workload(1, 8, 32))
    m = inp.mesh
    print(f"[case] C ABI driver, caller operators 8x32, {variant}", flush=True)
    KK, GG, MM, Fb = so.assembly_stokes_cylindric(inp)
    K, G, F = so.prepare_matrices(KK, GG, Fb, so.rotation_matrix(m.theta))
    _, ifree = m.free_slip()
    inner = so.inner_direct(K, ifree)
    if variant == "StokesPcCG":
        Uo, Po, info = so.stokes_pcg(K, MM, G, F, ifree, rtol=1e-8, inner=inner)
        P0, warm, ppc, rtol = None, 0, _lib.PPC_LUMPED, 1e-8
    else:
        Ul, Pl, _ = so.stokes_pcg(K, MM, G, F, ifree, rtol=1e-12, itmax=400, inner=inner)
        P0 = 0.9 * Pl + 0.05 * np.abs(Pl).max() * np.cos(3 * m.theta[:m.nvert])   # a previous step's pressure
        Uo, Po, info = so.stokes_pcg(K, MM, G, F, ifree, rtol=1e-6, inner=inner, pc=so.ic_preconditioner(MM), P0=P0)
        warm, ppc, rtol = 1, _lib.PPC_IC, 1e-6
    assert info.converged
    exe = _compile(tmp_path)
    _write_inputs(inp, tmp_path)
    _write_operators(tmp_path, K, MM, G, F, P0, k_lower=(variant == "StokesPcCG"), warm=warm, ppc=ppc,
                     stop=_lib.STOP_SCHUR, null_mode=_lib.NULL_DEFLATE, itmax=200, rtol=rtol, inner_rtol=1e-13)
    res = subprocess.run([str(exe), str(tmp_path), str(_lib.SOLVE_PLAIN), "operators"], capture_output=True,
                         text=True, timeout=120, env=dict(os.environ))
    print(res.stdout.strip(), res.stderr.strip()[-2000:])
    assert res.returncode == 0
    conv, outer, inner_its, resid, rc_bad, msg = (tmp_path / "report.txt").read_text().split()
    assert int(conv) == 1 and int(rc_bad) == _lib.STK_E_ARG
    U = np.fromfile(tmp_path / "U.bin", dtype=np.float64)
    P = np.fromfile(tmp_path / "P.bin", dtype=np.float64)
    # solveStokes removes the net rotation right after the solve (updatevelocity2,
    # StokesSolver.jl:48-50); before it, the LU answer carries an ill-determined multiple of the
    # near-null rotation of K[ifree, ifree]
    U, _ = so.updatevelocity2(U, inp.rho, m)
    Uo, _ = so.updatevelocity2(Uo, inp.rho, m)
    eu = np.linalg.norm(U - Uo) / np.linalg.norm(Uo)
    ep = np.linalg.norm((P - P.mean()) - (Po - Po.mean())) / np.linalg.norm(Po - Po.mean())
    print(f"[result] {variant}: outer {outer} (oracle {info.iterations}), errU {eu:.2e} errP {ep:.2e}", flush=True)
    assert abs(int(outer) - info.iterations) <= 1
    assert eu < 1e-7 and ep < 1e-7
