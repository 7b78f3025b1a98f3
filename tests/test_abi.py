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


def test_struct_layouts():
    # int32 fields first, doubles 8-byte aligned
    assert ctypes.sizeof(_lib.StkOpts) == 10 * 4 + 5 * 8 + 2 * 4
    assert ctypes.sizeof(_lib.StkReport) == 8 + 8 + 8 + 3 * 8
    o = _lib.default_opts()
    assert o.itmax == 200 and o.itmin == 10 and o.rtol == 1e-8


def test_last_error_without_context():
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.stk_create(ctypes.byref(h), -1, 0, 0, None, None, None, None, None, None, 0)
    assert rc == _lib.STK_E_SHAPE
    assert b"mesh sizes" in lib.stk_last_error(None)
