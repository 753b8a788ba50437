"""Writes tests/golden/qed4x3_pbc_e0.json: the lowest eigenvalue of 2D U(1)
QED on the 4x3 periodic lattice with staggered fermions (spin-1 links,
g = 1, m = 0.5; 28,904,615 states), computed on the host cores by the
oracle's C/OpenMP restatement of the reference's build path
(oracle/refcpu.qed_build_and_e0: sector enumeration, two-pass triplets,
SciPy coalescing, eigsh(k=1, which="SA") on a LinearOperator).

    python tests/golden/make_qed4x3_e0.py [out.json]   # from the repository root; ~40 GB RAM

The GPU test of this workload compares its device solves with this value.
"""
import json
import os
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import refcpu  # noqa: E402


def main():
    refcpu.build()
    e0, n, nnz, timings = refcpu.qed_build_and_e0([4, 3], [False, False], 1.0, 0.5, log=print)
    out = {
        "workload": "QED 4x3 PBC, staggered fermions, spin-1 links, g=1, m=0.5",
        "command": "python tests/golden/make_qed4x3_e0.py",
        "path": "oracle/refcpu.qed_build_and_e0 (host C/OpenMP + SciPy eigsh)",
        "threads": refcpu.threads(),
        "sector_dim": n,
        "nnz": nnz,
        "E0": e0,
        "timings_s": timings,
    }
    path = Path(sys.argv[1]) if len(sys.argv) > 1 else Path(__file__).resolve().parent / "qed4x3_pbc_e0.json"
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
