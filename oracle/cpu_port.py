"""Build and run the C++/OpenMP CPU restatement (cpu_stokes.cpp) -- TEST INFRASTRUCTURE ONLY.

The executable is the CPU baseline that bench.py times beside the GPU library; tests check that
it solves the oracle's system.  Nothing in the library calls it.
"""
from __future__ import annotations

import json
import os
import subprocess
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SOURCE = HERE / "cpu_stokes.cpp"
BINARY = HERE / "_bin" / "cpu_stokes"
# AVX2/FMA code: runs on any x86-64 host of the last decade (the build host's -march=native may not)
CXXFLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-std=c++17", "-Wall"]


def build(force: bool = False) -> Path:
    """Compile cpu_stokes.cpp with g++ into oracle/_bin/ (when missing or older than the source)."""
    if not force and BINARY.exists() and BINARY.stat().st_mtime >= SOURCE.stat().st_mtime:
        return BINARY
    BINARY.parent.mkdir(exist_ok=True)
    tmp = BINARY.with_suffix(".tmp")
    subprocess.run(["g++", *CXXFLAGS, "-o", str(tmp), str(SOURCE)], check=True)
    os.replace(tmp, BINARY)
    return BINARY


def write_inputs(inp, d: Path, fixed=None) -> None:
    """The input directory of cpu_stokes and tests/c_abi/abi_driver.c: meta.txt and raw arrays in
    the Julia layouts (1-based e2n as 6 x nel, el_* as 3 x nel, eta and the tensor as nel x 6)."""
    m = inp.mesh
    if fixed is None:
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


def run(inp, budget_s: float, inner_rtol: float = 1e-10, threads: int | None = None, keep=None):
    """Assemble and solve ``inp`` with the CPU restatement; the solve stops after ``budget_s``.

    Returns (report dict, U, P): U in node frames before the net-rotation removal.
    ``keep``: a directory to use instead of a temporary one.
    """
    exe = build()
    env = dict(os.environ)
    if threads is not None:
        env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory(prefix="cpu_stokes_") as tmp:
        d = Path(keep) if keep is not None else Path(tmp)
        write_inputs(inp, d)
        res = subprocess.run([str(exe), str(d), repr(float(budget_s)), repr(float(inner_rtol))],
                             capture_output=True, text=True, env=env, check=False)
        if res.returncode != 0:
            raise RuntimeError(f"cpu_stokes failed ({res.returncode}): {res.stderr[-2000:]}")
        rep = json.loads(res.stdout.strip().splitlines()[-1])
        U = np.fromfile(d / "cpu_U.bin", dtype=np.float64)
        P = np.fromfile(d / "cpu_P.bin", dtype=np.float64)
    return rep, U, P
