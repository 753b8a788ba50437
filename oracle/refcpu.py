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


def sector(loc_dims, glob_diags=None, glob_sectors=None, link_diags=None, link_sectors=None, pairs=None,
           nb_diags=None, nb_sectors=None, nb_sites=None):
    """refcpu_sector: U(1) global, link-pair and additive n-body constraints;
    returns the uint16 configuration table (lexicographic)."""
    lib = load()
    loc = np.ascontiguousarray(loc_dims, dtype=np.int32)
    n_sites = loc.size
    d_max = int(loc.max())

    def f64(a, shape):
        return np.zeros(shape) if a is None else np.ascontiguousarray(a, dtype=np.float64)

    n_glob = 0 if glob_diags is None else glob_diags.shape[0]
    glob = f64(glob_diags, (1,))
    gsec = f64(glob_sectors, (1,))
    pl = [] if pairs is None else [(d, int(i), int(j)) for d, ps in enumerate(pairs) for (i, j) in ps]
    pdir = np.array([p[0] for p in pl] or [0], dtype=np.int32)
    p0 = np.array([p[1] for p in pl] or [0], dtype=np.int32)
    p1 = np.array([p[2] for p in pl] or [0], dtype=np.int32)
    link = f64(link_diags, (1,))
    lsec = f64(link_sectors, (1,))
    n_nb = 0 if nb_diags is None else nb_diags.shape[0]
    nbd = f64(nb_diags, (1,))
    nbsec = f64(nb_sectors, (1,))
    offs = [0]
    flat = []
    for s in (nb_sites or []):
        flat += [int(x) for x in s]
        offs.append(len(flat))
    nboff = np.array(offs, dtype=np.int32)
    nbs = np.array(flat or [0], dtype=np.int32)
    out = C.c_void_p()
    n = lib.refcpu_sector(n_sites, loc.ctypes.data, d_max, n_glob, glob.ctypes.data, gsec.ctypes.data, len(pl),
                          pdir.ctypes.data, p0.ctypes.data, p1.ctypes.data, link.ctypes.data, lsec.ctypes.data,
                          n_nb, nboff.ctypes.data, nbs.ctypes.data, nbd.ctypes.data, nbsec.ctypes.data,
                          C.byref(out))
    buf = (C.c_uint16 * (n * n_sites)).from_address(out.value) if n else None
    cfg = np.frombuffer(buf, dtype=np.uint16).reshape(n, n_sites).copy() if n else np.zeros((0, n_sites), np.uint16)
    lib.refcpu_free(out)
    return cfg


def nbody(configs, op_stack, sites, op_tol=1e-10):
    """refcpu_nbody_count / _fill for one placement of n >= 2 factors
    (factor f acts on site sites[f] with op_stack[f, sites[f]]); returns
    (rows, cols, vals, misses) in the reference's emission order."""
    lib = load()
    cfg = np.ascontiguousarray(configs, dtype=np.uint16)
    n, n_sites = cfg.shape
    n_ops = len(sites)
    d_max = op_stack.shape[-1]
    counts = np.zeros((n_ops, d_max), dtype=np.int32)
    kets = np.zeros((n_ops, d_max, d_max), dtype=np.int32)
    vals = np.zeros((n_ops, d_max, d_max), dtype=np.complex128)
    for f, s in enumerate(sites):
        O = op_stack[f, s]
        for a in range(d_max):
            bs = [b for b in range(d_max) if abs(O[a, b]) > op_tol]
            counts[f, a] = len(bs)
            for t, b in enumerate(bs):
                kets[f, a, t] = b
                vals[f, a, t] = O[a, b]
    st = np.ascontiguousarray(sites, dtype=np.int32)
    offsets = np.zeros(n, dtype=np.int64)
    tot = lib.refcpu_nbody_count(cfg.ctypes.data, n, n_sites, n_ops, st.ctypes.data, counts.ctypes.data, d_max,
                                 offsets.ctypes.data)
    rows = np.zeros(tot, dtype=np.int32)
    cols = np.zeros(tot, dtype=np.int32)
    out = np.zeros(tot, dtype=np.complex128)
    miss = lib.refcpu_nbody_fill(cfg.ctypes.data, n, n_sites, n_ops, st.ctypes.data, counts.ctypes.data,
                                 kets.ctypes.data, vals.ctypes.data, d_max, offsets.ctypes.data, rows.ctypes.data,
                                 cols.ctypes.data, out.ctypes.data)
    return rows, cols, out, int(miss)
