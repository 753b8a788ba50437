"""ctypes loader for oracle/librefcpu.so (CPU baseline / checker only)."""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "librefcpu.so"
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        p, i64, i32 = C.c_void_p, C.c_int64, C.c_int32
        lib.refcpu_threads.restype = i32
        lib.refcpu_csr_spmv_c128.argtypes = [i64, p, p, p, p, p]
        lib.refcpu_csr_spmv_f64.argtypes = [i64, p, p, p, p, p]
        lib.refcpu_sector.restype = i64
        lib.refcpu_sector.argtypes = [i32, p, i32, i32, p, p, i32, p, p, p, p, p, i32, p, p, p, p, C.POINTER(p)]
        lib.refcpu_free.argtypes = [p]
        lib.refcpu_nbody_count.restype = i64
        lib.refcpu_nbody_count.argtypes = [p, i64, i32, i32, p, p, i32, p]
        lib.refcpu_nbody_fill.restype = i64
        lib.refcpu_nbody_fill.argtypes = [p, i64, i32, i32, p, p, p, p, i32, p, p, p, p]
        _lib = lib
    return _lib


def threads():
    return int(load().refcpu_threads())


def spmv(indptr, indices, data, x, y):
    lib = load()
    n = indptr.size - 1
    if np.iscomplexobj(data):
        lib.refcpu_csr_spmv_c128(n, indptr.ctypes.data, indices.ctypes.data, data.ctypes.data,
                                 x.ctypes.data, y.ctypes.data)
    else:
        lib.refcpu_csr_spmv_f64(n, indptr.ctypes.data, indices.ctypes.data, data.ctypes.data,
                                x.ctypes.data, y.ctypes.data)
    return y
