"""Write tests/golden/dem_tables.npz from the reference's DEM tables (newDEM/*.h5).

Run in the container that holds the reference tree:  python tests/golden/make_dem_tables.py

``DEMloader`` (src/Rheology/anisotropy.jl:48-54) reads three datasets: "C" (n x 9 tensor),
"ax" (n x 3 semi-axes) and "vol" (the volume fraction).  No HDF5 library is installed, so the
script reads the contiguous-storage layout messages of the files (version 3, class 1: an
8-byte address and an 8-byte size) and takes the datasets by size: 9n, 3n and 1 doubles.  The
bytes are in Julia's column-major order (HDF5.jl writes the dimensions reversed).  The result is
checked against the structure MAIN.jl relies on: unit middle axes, blocks of equal a1, and a
volume equal to the one in the file name.
"""
from __future__ import annotations

import re
import struct
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/newDEM")
OUT = Path(__file__).with_name("dem_tables.npz")


def contiguous_datasets(blob: bytes):
    """(address, size) of every contiguous-layout message whose extent lies inside the file."""
    found = []
    for m in re.finditer(rb"\x03\x01", blob):
        i = m.end()
        if i + 16 > len(blob):
            continue
        addr, size = struct.unpack_from("<QQ", blob, i)
        if size and size % 8 == 0 and addr + size <= len(blob) and addr >= 512:
            found.append((addr, size))
    return sorted(set(found))


def read_table(path: Path):
    blob = path.read_bytes()
    assert blob[:8] == b"\x89HDF\r\n\x1a\n", path
    sets = {size // 8: np.frombuffer(blob, "<f8", size // 8, addr) for addr, size in contiguous_datasets(blob)}
    n9 = [k for k in sets if k % 9 == 0 and k // 9 * 3 in sets]
    assert len(n9) == 1 and 1 in sets, (path, sorted(sets))
    n = n9[0] // 9
    tensor = sets[9 * n].reshape(9, n).T.copy()
    axes = sets[3 * n].reshape(3, n).T.copy()
    vol = float(sets[1][0])
    assert np.all(axes[:, 1] == 1.0)
    blk = int(np.argmax(axes[:, 0] != axes[0, 0]))
    assert blk > 1 and n % blk == 0
    assert np.all(axes[:, 0].reshape(-1, blk) == axes[::blk, 0][:, None])
    assert abs(vol - int(re.search(r"vol(\d+)", path.name).group(1)) / 100) < 1e-12
    return tensor, axes, vol


def main():
    arrays = {}
    for p in sorted(SRC.glob("DEM_*.h5")):
        tensor, axes, vol = read_table(p)
        key = p.stem.replace("-", "m")
        arrays[f"{key}_C"] = tensor
        arrays[f"{key}_ax"] = axes
        arrays[f"{key}_vol"] = np.array(vol)
        print(p.name, tensor.shape, axes.shape, vol)
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
