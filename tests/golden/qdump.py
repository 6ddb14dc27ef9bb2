"""Reader for the self-describing binary written by oracle/ref/qnb_dump.f90.

record := int32 namelen, name, int32 kind (1=int32, 2=float64), int32 ndim,
          int32 dims[ndim], data in Fortran (column-major) order.
Arrays are returned as numpy arrays with the Fortran shape; 1-element arrays are
returned as Python scalars.
"""
import numpy as np


def read_dump(path):
    out = {}
    with open(path, "rb") as f:
        buf = f.read()
    off = 0
    n = len(buf)
    while off < n:
        (namelen,) = np.frombuffer(buf, "<i4", 1, off)
        off += 4
        name = buf[off:off + namelen].decode()
        off += namelen
        kind, ndim = np.frombuffer(buf, "<i4", 2, off)
        off += 8
        dims = tuple(int(v) for v in np.frombuffer(buf, "<i4", ndim, off))
        off += 4 * ndim
        cnt = int(np.prod(dims)) if dims else 1
        dt = "<i4" if kind == 1 else "<f8"
        arr = np.frombuffer(buf, dt, cnt, off).copy()
        off += cnt * (4 if kind == 1 else 8)
        arr = arr.reshape(dims, order="F")
        if arr.size == 1 and ndim == 1 and dims == (1,):
            arr = arr.item()
        out[name] = arr
    return out
