"""Builds and runs the Gram-product benchmark harness (gram_bench.hip) on one
GPU: the Krylov solvers' VALU Gram kernel against the MFMA formulations, at
the column lengths of the QED 4x3 sector (28,904,615) and the Z3 4x4
zero-flux sector (14,348,907), for every basis width and block width the
solvers use.  Prints one markdown table row per (kernel, shape): time per
launch (HIP events, mean of 5), algorithmic bytes (nv + nw) * n * 16, GB/s and
the fraction of the 8 TB/s HBM peak, and the relative difference of G from
the VALU kernel's.

    python ed-lgt_amd/perf/gram_bench.py [--n N ...]
"""
import argparse
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "libgram_bench.so"
KINDS = {0: "VALU gram_partial (solver)", 1: "MFMA, operands from global", 3: "MFMA, LDS-staged operands",
         2: "VALU, register accumulation"}


def build():
    cmd = ["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", str(HERE / "gram_bench.hip"),
           "-o", str(LIB)]
    subprocess.run(cmd, check=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, nargs="*", default=[28904615, 14348907])
    ap.add_argument("--no-build", action="store_true")
    args = ap.parse_args()
    if not args.no_build:
        build()
    import torch  # noqa: F401  (one HIP runtime in the process)
    lib = C.CDLL(str(LIB))
    lib.probe_layout.argtypes = [C.c_void_p]
    lib.probe_gram.argtypes = [C.c_int, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                               C.POINTER(C.c_float)]
    out = np.zeros((64, 4))
    assert lib.probe_layout(out.ctypes.data) == 0
    A = np.array([[1000.0 * i + k for k in range(4)] for i in range(16)])
    B = np.array([[0.5 * (k + 1) * (j + 1) for j in range(16)] for k in range(4)])
    D = A @ B
    assert all(out[l, q] == D[4 * q + l // 16, l % 16] for l in range(64) for q in range(4)), "MFMA layout"
    print("| n | nw | nv | kernel | ms | GB/s | frac | rel diff |")
    print("|---|---|---|---|---|---|---|---|")
    for n in args.n:
        for nw in (1, 2, 4):
            for nv in (4, 8, 16, 32, 64):
                ref = None
                for kind in (0, 1, 3, 2):
                    if kind == 2 and nv > 16:
                        continue
                    g = np.zeros((nv, nw), dtype=np.complex128)
                    ms = C.c_float()
                    st = lib.probe_gram(kind, n, nv, nw, 1024, 5, g.ctypes.data, C.byref(ms))
                    assert st == 0, st
                    ref = g if ref is None else ref
                    diff = float(np.max(np.abs(g - ref)) / np.max(np.abs(ref)))
                    gbs = (nv + nw) * n * 16 / (ms.value * 1e-3) / 1e9
                    print(f"| {n} | {nw} | {nv} | {KINDS[kind]} | {ms.value:.3f} | {gbs:.0f} | {gbs / 8000:.3f} | "
                          f"{diff:.1e} |", flush=True)


if __name__ == "__main__":
    main()
