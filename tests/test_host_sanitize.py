"""The host set-up code (dimspers_amd/csrc/setup.cpp: patterns, colourings, slot maps, structure
detection, multigrid structure and transfers, Gauss-Seidel layout, gather lists, pressure IC, CSC
import) built with g++ under AddressSanitizer and UndefinedBehaviorSanitizer against a host-memory
stand-in of the HIP runtime (tests/host_sanitize/), and run on the CPU on both structured meshes,
a fixed set without the multigrid, and the oracle's operators.  Any sanitizer report fails the run
(-fno-sanitize-recover=all); the driver also checks the invariants of each step."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from dimspers_amd.inputs import make_inputs, scaled_workload
from dimspers_amd.mesh import build_annulus, build_grid
from oracle import cpu_port
from oracle import stokes_oracle as so

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    exe = tmp_path_factory.mktemp("asan") / "setup_driver"
    cmd = ["g++", "-std=c++17", "-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-Wall", "-Wno-unknown-pragmas", "-I", str(ROOT / "tests" / "host_sanitize"),
           "-I", str(ROOT / "dimspers_amd" / "csrc"), "-o", str(exe),
           str(ROOT / "tests" / "host_sanitize" / "setup_driver.cpp"), str(ROOT / "dimspers_amd" / "csrc" / "setup.cpp")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]
    return exe


def _run(driver, d, *extra):
    res = subprocess.run([str(driver), str(d), *extra], capture_output=True, text=True, timeout=600,
                         env={"ASAN_OPTIONS": "detect_leaks=1", "UBSAN_OPTIONS": "print_stacktrace=1"})
    assert res.returncode == 0, res.stdout + res.stderr[-4000:]
    fields = dict(kv.split("=") for kv in res.stdout.split()[1:])
    return {k: int(v) for k, v in fields.items()}


@pytest.mark.parametrize("kind,size,levels", [("split1", (32, 256), 2), ("grid", (32, 256), 3),
                                              ("split1", (7, 18), 1), ("grid", (5, 24), 1)])
def test_setup_is_clean_under_sanitizers(driver, tmp_path, kind, size, levels):
    """(7, 18) and (5, 24): odd radial counts and azimuthal counts that are not powers of two (partial
    64-row slices, the P1 level only)."""
    mesh = build_annulus(*size) if kind == "split1" else build_grid(*size)
    cpu_port.write_inputs(make_inputs(scaled_workload(3, *size), mesh=mesh), tmp_path)
    out = _run(driver, tmp_path)
    assert out["split"] == (2 if kind == "split1" else 4) and out["levels"] == levels


def test_setup_without_multigrid_and_operator_import_are_clean(driver, tmp_path):
    """No-slip fixes tangential dofs: no multigrid; the oracle's prepared K (lower triangle), MM and
    GG go through the CSC import, and an entry outside the mesh's pattern is refused."""
    inp = make_inputs(scaled_workload(3, 8, 64))
    fixed, ifree = inp.mesh.no_slip()
    cpu_port.write_inputs(inp, tmp_path, fixed=fixed)
    KK, GG, MM, Fb = so.assembly_stokes_cylindric(inp)
    K, G, F = so.prepare_matrices(KK, GG, Fb, so.rotation_matrix(inp.mesh.theta))

    def put(name, A):
        A = sp.csc_matrix(A)
        A.sort_indices()
        (tmp_path / f"{name}colptr.bin").write_bytes(np.asarray(A.indptr + 1, dtype=np.int64).tobytes())
        (tmp_path / f"{name}rowval.bin").write_bytes(np.asarray(A.indices + 1, dtype=np.int64).tobytes())
        (tmp_path / f"{name}nzval.bin").write_bytes(np.asarray(A.data, dtype=np.float64).tobytes())
        return A.nnz
    nK, nM, nG = put("K", sp.tril(K, format="csc")), put("M", MM), put("G", G)
    (tmp_path / "Rhs.bin").write_bytes(np.asarray(F, dtype=np.float64).tobytes())
    (tmp_path / "ops.txt").write_text(f"{nK} {nM} {nG} 1 1\n")
    out = _run(driver, tmp_path, "operators")
    assert out["split"] == 2 and out["levels"] == 0 and out["kvals"] > 0
