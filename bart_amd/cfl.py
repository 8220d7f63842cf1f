"""BART array files from Python: ``readcfl`` / ``writecfl`` for ``name.hdr`` +
``name.cfl`` pairs and ``.ra`` files (python/cfl.py and misc/io.c:380-468,
:960-1040 of BART).

Arrays are returned with BART's dims in order (dim 0 first) in Fortran
(column-major) order, as BART's own reader does; trailing singleton dims
are dropped.
"""
import struct

import numpy as np

RA_MAGIC = 0x7961727261776172
RA_TYPE_COMPLEX = 4


def _read_header(name):
    with open(name + ".hdr", "r") as f:
        lines = f.read().split("\n")
    for i, line in enumerate(lines):
        if line.strip() == "# Dimensions":
            return [int(v) for v in lines[i + 1].split()]
    raise ValueError(f"{name}.hdr: no '# Dimensions' section")


def _trim(dims):
    dims = list(dims)
    while len(dims) > 1 and dims[-1] == 1:
        dims.pop()
    return dims


def readcfl(name):
    """Reads ``name.hdr``/``name.cfl`` or ``name`` ending in ``.ra``."""
    if name.endswith(".ra"):
        with open(name, "rb") as f:
            magic, flags, eltype, elbyte, size, ndims = struct.unpack("<6Q", f.read(48))
            if magic != RA_MAGIC or flags & 1 or eltype != RA_TYPE_COMPLEX or elbyte != 8:
                raise ValueError(f"{name}: not a little-endian complex float ra file")
            dims = list(struct.unpack(f"<{ndims}Q", f.read(8 * ndims)))
            n = int(np.prod(dims)) if dims else 1
            if size != 8 * n:
                raise ValueError(f"{name}: size mismatch")
            data = np.frombuffer(f.read(8 * n), dtype=np.complex64)
    else:
        dims = _read_header(name)
        n = int(np.prod(dims))
        data = np.fromfile(name + ".cfl", dtype=np.complex64, count=n)
        if data.size != n:
            raise ValueError(f"{name}.cfl: short file")
    dims = _trim(dims)
    return data.reshape(dims, order="F")


def writecfl(name, array, ndims=None):
    """Writes a complex array (dim 0 first, any memory order)."""
    a = np.asarray(array, dtype=np.complex64)
    dims = list(a.shape) if a.ndim else [1]
    if ndims is not None:
        dims = dims + [1] * (ndims - len(dims))
    flat = np.asfortranarray(a).reshape(-1, order="F")
    if name.endswith(".ra"):
        with open(name, "wb") as f:
            f.write(struct.pack("<6Q", RA_MAGIC, 0, RA_TYPE_COMPLEX, 8, 8 * flat.size, len(dims)))
            f.write(struct.pack(f"<{len(dims)}Q", *dims))
            f.write(flat.tobytes())
        return
    with open(name + ".hdr", "w") as f:
        f.write("# Dimensions\n")
        f.write(" ".join(str(d) for d in dims) + " \n")
    flat.tofile(name + ".cfl")
