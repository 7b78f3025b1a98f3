"""C ABI checks that need no GPU: the library loads and exports every declared symbol."""
import ctypes
import re
from pathlib import Path

from dimspers_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(stk_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = declared_functions()
    assert len(names) >= 13
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_the_header():
    assert set(_lib.EXPORTED) == set(declared_functions())


def test_struct_layouts(tmp_path):
    # every field of the ctypes mirrors sits where gcc puts it from the header
    import subprocess
    structs = {"stk_opts": _lib.StkOpts, "stk_report": _lib.StkReport, "stk_train_opts": _lib.StkTrainOpts,
               "stk_train_report": _lib.StkTrainReport}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dimspers_stokes.h"', 'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    root = Path(__file__).resolve().parent.parent
    res = subprocess.run(["gcc", "-std=c11", "-I", str(root / "include"), str(src), "-o", str(exe)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        cname, field, val = line.split()
        got[(cname, field)] = int(val)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
    o = _lib.default_opts()
    assert o.itmax == 200 and o.itmin == 10 and o.rtol == 1e-8
    assert o.warm_start == 0 and o.pressure_pc == _lib.PPC_LUMPED


def test_last_error_without_context():
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.stk_create(ctypes.byref(h), -1, 0, 0, None, None, None, None, None, None, 0)
    assert rc == _lib.STK_E_SHAPE
    assert b"mesh sizes" in lib.stk_last_error(None)
