"""PLY ingestion (include/phaser_ply.h, in libphaser_host.so) from Python.

read_ply(path) -> dict with "xyzi" (n, 4) float32, "extra" (n, 2) float32
(reflectivity, ambient; zero where the file lacks the field), "range" (n,)
float64 or None, and has_reflectivity / has_ambient flags.  The field names
default to the reference's flags (ply-helper.cc:8-22)."""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libphaser_host.so")
_lib = None


class Fields(C.Structure):
    _fields_ = [("intensity", C.c_char_p), ("reflectivity", C.c_char_p),
                ("ambient", C.c_char_p), ("range", C.c_char_p)]


class Cloud(C.Structure):
    _fields_ = [("n", C.c_int64), ("xyzi", C.POINTER(C.c_float)),
                ("extra", C.POINTER(C.c_float)), ("range", C.POINTER(C.c_double)),
                ("has_reflectivity", C.c_int32), ("has_ambient", C.c_int32),
                ("has_range", C.c_int32), ("reserved", C.c_int32)]


EXPORTS = {
    "phz_ply_default_fields": (None, [C.POINTER(Fields)]),
    "phz_ply_read": (C.c_int, [C.c_char_p, C.POINTER(Fields), C.POINTER(Cloud)]),
    "phz_ply_free": (None, [C.POINTER(Cloud)]),
    "phz_ply_last_error": (C.c_char_p, []),
    "phz_ply_write_xyzi": (C.c_int, [C.c_char_p, C.c_int64, C.c_void_p]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            f = getattr(_lib, name)
            f.restype = res
            f.argtypes = args
    return _lib


class PlyError(RuntimeError):
    pass


def read_ply(path, intensity=None, reflectivity=None, ambient=None, range_field=None):
    f = Fields()
    lib().phz_ply_default_fields(C.byref(f))
    for name, v in (("intensity", intensity), ("reflectivity", reflectivity),
                    ("ambient", ambient), ("range", range_field)):
        if v is not None:
            setattr(f, name, v.encode())
    c = Cloud()
    if lib().phz_ply_read(str(path).encode(), C.byref(f), C.byref(c)) != 0:
        raise PlyError(lib().phz_ply_last_error().decode())
    try:
        n = c.n
        out = {
            "xyzi": np.ctypeslib.as_array(c.xyzi, (n, 4)).copy(),
            "extra": np.ctypeslib.as_array(c.extra, (n, 2)).copy(),
            "range": np.ctypeslib.as_array(c.range, (n,)).copy() if c.has_range else None,
            "has_reflectivity": bool(c.has_reflectivity),
            "has_ambient": bool(c.has_ambient),
        }
    finally:
        lib().phz_ply_free(C.byref(c))
    return out


def write_ply_xyzi(path, xyzi):
    """The ASCII PLY the command-line driver writes (phz_ply_write_xyzi)."""
    a = np.ascontiguousarray(xyzi, np.float32)
    assert a.ndim == 2 and a.shape[1] == 4
    if lib().phz_ply_write_xyzi(str(path).encode(), a.shape[0], a.ctypes.data) != 0:
        raise PlyError(lib().phz_ply_last_error().decode())


_TYPES = {"char": "i1", "uchar": "u1", "short": "i2", "ushort": "u2", "int": "i4",
          "uint": "u4", "float": "f4", "double": "f8"}


def write_ply(path, columns, fmt="binary_little_endian", faces=None):
    """Writes a vertex element with the given {name: (ply type, values)}
    columns (and optionally a face element of index lists), for tests and for
    handing synthetic clouds to the C++ driver."""
    names = list(columns)
    n = len(next(iter(columns.values()))[1])
    head = ["ply", f"format {fmt} 1.0", "comment written by phaser_amd.ply",
            f"element vertex {n}"]
    head += [f"property {columns[k][0]} {k}" for k in names]
    if faces is not None:
        head += [f"element face {len(faces)}", "property list uchar int vertex_indices"]
    head.append("end_header")
    with open(path, "wb") as fh:
        fh.write(("\n".join(head) + "\n").encode())
        if fmt == "ascii":
            for i in range(n):
                fh.write((" ".join(repr(np.asarray(columns[k][1])[i].item()) for k in names)
                          + "\n").encode())
            for f in faces or []:
                fh.write((" ".join(str(v) for v in [len(f), *f]) + "\n").encode())
            return
        order = "<" if fmt == "binary_little_endian" else ">"
        dt = np.dtype([(k, order + _TYPES[columns[k][0]]) for k in names])
        rec = np.zeros(n, dt)
        for k in names:
            rec[k] = columns[k][1]
        fh.write(rec.tobytes())
        for f in faces or []:
            fh.write(np.asarray([len(f)], order + "u1").tobytes())
            fh.write(np.asarray(f, order + "i4").tobytes())
