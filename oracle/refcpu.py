"""ctypes loader for oracle/librefcpu.so (CPU baseline / checker only)."""
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "librefcpu.so"
LIB_NATIVE = HERE / "librefcpu_native.so"
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def use_native():
    """Builds librefcpu_native.so (-march=native) on this host and makes the
    following calls use it; returns False (keeping the portable build) when
    the host has no compiler.  Call before the first use of the library."""
    global LIB, _lib
    try:
        subprocess.run(["make", "-s", "-B", "-C", str(HERE), "native"], check=True, capture_output=True)
    except (OSError, subprocess.CalledProcessError):
        return False
    LIB, _lib = LIB_NATIVE, None
    return True


def set_threads(n):
    """OpenMP team size of the following calls."""
    load().refcpu_set_threads(int(n))


def host_cpu():
    """(model name, sockets, physical cores, logical CPUs) of this host."""
    model, phys = "unknown", set()
    try:
        cur = {}
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if cur:
                    phys.add((cur.get("physical id", "0"), cur.get("core id", cur.get("processor"))))
                cur = {}
                continue
            k, v = (t.strip() for t in line.split(":", 1))
            cur[k] = v
            if k == "model name":
                model = v
        if cur:
            phys.add((cur.get("physical id", "0"), cur.get("core id", cur.get("processor"))))
    except OSError:
        pass
    sockets = len({p[0] for p in phys}) or 1
    return model, sockets, len(phys) or os.cpu_count(), os.cpu_count()


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        p, i64, i32 = C.c_void_p, C.c_int64, C.c_int32
        lib.refcpu_threads.restype = i32
        lib.refcpu_set_threads.argtypes = [i32]
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


def qed_build_and_e0(lvals, has_obc, g, m, log=None):
    """The reference's build + lowest-eigenvalue path for QED with matter on
    the host cores, through this C/OpenMP restatement of its kernels:
    sector enumeration (symmetry_sector.py:785-967), two-pass triplets with
    binary search per term placement in the reference's term order
    (sym_qmb_operations.py:95-602, QED_model.py:235-394), SciPy coalescing of
    the triplets into CSR (qmb_hamiltonian.py:173-225; done per row block),
    and eigsh(which="SA", k=1) on a LinearOperator over the C SpMV
    (qmb_hamiltonian.py:260-332, ham_format="linear").  Returns (E0, N, nnz,
    timings)."""
    from time import perf_counter

    import scipy.sparse as sp
    from scipy.sparse.linalg import LinearOperator, eigsh

    from . import lattice as lat
    from .model import OP_TOL, QEDOracle

    say = log or (lambda msg: None)
    T = {}
    t0 = perf_counter()
    orc = QEDOracle(1, False, lvals, has_obc, enumerate_sector=False)
    dg = lambda name: np.real(np.einsum("sii->si", orc.ops[name]))
    cfg = sector(orc.loc_dims, dg("N")[None], np.array([orc.n_sites // 2], float), orc.link_diags,
                 np.zeros(orc.dim), orc.pairs)
    N = cfg.shape[0]
    T["sector_s"] = perf_counter() - t0
    say(f"sector N={N} in {T['sector_s']:.1f}s")
    c = orc.couplings(g, m)
    blocks = []

    def local(name, strength, mask=None):
        for s_ in range(orc.n_sites):
            if mask is not None and not mask[s_]:
                continue
            v = orc.ops[name][s_][cfg[:, s_], cfg[:, s_]]
            keep = np.nonzero(np.abs(v) >= OP_TOL)[0].astype(np.int32)
            blocks.append((keep, keep, v[keep] * strength))

    def nbody_list(names, site_lists, strength, dagger):
        stack = np.stack([orc.ops[n] for n in names])
        for sites in site_lists:
            r, cc, v, miss = nbody(cfg, stack, sites)
            if miss:
                raise RuntimeError("lookup miss")
            blocks.append((r, cc, v * strength))
            if dagger:
                blocks.append((cc, r, np.conj(v * strength)))
        say(f"term {names[0]}.. on {len(site_lists)} placements done ({perf_counter() - t0:.1f}s)")

    t0 = perf_counter()
    local("E2", c["E"])
    plaqs = [p for p in (lat.plaquette_sites(lat.zig_zag(orc.lvals, i), orc.lvals, ["x", "y"], orc.has_obc)
                         for i in range(orc.n_sites)) if p]
    nbody_list(["C_px,py", "C_py,mx", "C_my,px", "C_mx,my"], plaqs, c["B"], True)
    for stag in ("even", "odd"):
        mask = lat.staggered_mask(orc.lvals, stag)
        local("N", c[f"m_{stag}"], [mask[lat.zig_zag(orc.lvals, i)] for i in range(orc.n_sites)])
    for d in orc.dirs:
        for stag in ("even", "odd"):
            mask = lat.staggered_mask(orc.lvals, stag)
            bonds = []
            for i in range(orc.n_sites):
                co = lat.zig_zag(orc.lvals, i)
                b = lat.neighbor_sites(co, orc.lvals, d, orc.has_obc)
                if b is not None and mask[co]:
                    bonds.append(b)
            nbody_list([f"Q_p{d}_dag", f"Q_m{d}"], bonds, c[f"t{d}_{stag}"], True)
    T["triplets_s"] = perf_counter() - t0
    raw = sum(b[0].size for b in blocks)
    say(f"{raw} triplets in {T['triplets_s']:.1f}s")
    # coalesce per row block (same result as one SciPy call over all triplets)
    t0 = perf_counter()
    n_chunks = 8
    edges = np.linspace(0, N, n_chunks + 1).astype(np.int64)
    parts = []
    for k in range(n_chunks):
        lo, hi = edges[k], edges[k + 1]
        rs, cs, vs = [], [], []
        for r, cc, v in blocks:
            sel = np.nonzero((r >= lo) & (r < hi))[0]
            rs.append(r[sel] - lo); cs.append(cc[sel]); vs.append(v[sel])
        part = sp.coo_matrix((np.concatenate(vs), (np.concatenate(rs), np.concatenate(cs))),
                             shape=(hi - lo, N)).tocsr()
        part.sum_duplicates()
        part.sort_indices()
        parts.append(part)
        say(f"CSR rows {lo}..{hi} done ({perf_counter() - t0:.1f}s)")
    blocks.clear()
    H = sp.vstack(parts, format="csr")
    T["csr_s"] = perf_counter() - t0
    nnz = H.nnz
    indptr = H.indptr.astype(np.int64)
    indices = H.indices.astype(np.int32)
    data = H.data.astype(np.complex128)
    del H, parts
    say(f"nnz={nnz}; eigsh")
    t0 = perf_counter()
    y = np.empty(N, dtype=np.complex128)

    def mv(x):
        x = np.ascontiguousarray(x, dtype=np.complex128).ravel()
        spmv(indptr, indices, data, x, y)
        return y.copy()

    op = LinearOperator((N, N), matvec=mv, dtype=np.complex128)
    v0 = np.random.default_rng(0).random(N).astype(np.complex128)
    w = eigsh(op, k=1, which="SA", v0=v0 / np.linalg.norm(v0), return_eigenvectors=False)
    T["eigsh_s"] = perf_counter() - t0
    say(f"E0={w[0]} in {T['eigsh_s']:.1f}s")
    return float(np.min(w)), N, nnz, T
