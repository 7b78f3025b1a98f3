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
    m = inp.mesh
    fixed, _ = m.free_slip()
    arrays = {
        "e2n": np.ascontiguousarray(np.asarray(m.e2n + 1, dtype=np.int64).T),
        "el_theta": np.ascontiguousarray(np.asarray(m.el_theta, dtype=np.float64).T),
        "el_r": np.ascontiguousarray(np.asarray(m.el_r, dtype=np.float64).T),
        "node_theta": np.ascontiguousarray(m.theta, dtype=np.float64),
        "node_r": np.ascontiguousarray(m.r, dtype=np.float64),
        "fixed": np.ascontiguousarray(fixed + 1, dtype=np.int64),
        "rho": np.ascontiguousarray(inp.rho, dtype=np.float64),
        "eta": np.asfortranarray(np.asarray(inp.eta_ip)[:, :6], dtype=np.float64),
    }
    for k, c in enumerate(inp.C6):
        arrays[f"C{k}"] = np.asfortranarray(np.asarray(c)[:, :6], dtype=np.float64)
    for name, a in arrays.items():
        (d / f"{name}.bin").write_bytes(a.tobytes(order="A"))
    (d / "meta.txt").write_text(f"{m.nel} {m.nnod} {m.nvert} {fixed.size} {inp.g!r}\n")


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
