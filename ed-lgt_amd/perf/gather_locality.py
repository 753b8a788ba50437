"""Gather locality of the Hamiltonian's column pattern: how many distinct
128-byte x lines one gather instruction touches, for three ways of assigning
stored entries to the 64 lanes of a wave:

  stream  the CSR-stream kernel's order (lane = consecutive stored entries),
  sell    sliced ELL with one row per lane (lane = row, instruction = the
          k-th entry of each of 64 consecutive rows),
  term    one row per lane, instruction = one term of the Hamiltonian (entries
          grouped by the key difference between column and row configuration,
          which identifies the placement and its local transition).

Prints entries per line request (higher = more coalescing) for complex128 x
(8 elements per 128-byte line).  Runs the engine's sector and CSR build on
the GPU, then analyses a window of rows on the host.

    python ed-lgt_amd/perf/gather_locality.py --lvals 4 3 --pbc   # QED 4x3 PBC, g=1, m=0.5
"""
import argparse
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))


def lines_per_group(group, lines, n_lines):
    return np.unique(group * n_lines + lines).size


def analyse(rp, col, keys, r0, r1, per_line=8, lanes=64):
    p0, p1 = int(rp[r0]), int(rp[r1])
    c = col[p0:p1].astype(np.int64)
    lens = np.diff(rp[r0:r1 + 1]).astype(np.int64)
    row = np.repeat(np.arange(r0, r1, dtype=np.int64), lens)
    lines = c // per_line
    nl = int(lines.max()) + 1
    nnz = c.size
    out = {}
    out["stream"] = nnz / lines_per_group(np.arange(nnz) // lanes, lines, nl)
    k = np.arange(nnz) - (np.repeat(rp[r0:r1], lens) - p0)
    out["sell"] = nnz / lines_per_group(((row - r0) // lanes) * 4096 + k, lines, nl)
    kd = keys[c] - keys[row]
    _, kdi = np.unique(kd, return_inverse=True)
    out["term"] = nnz / lines_per_group(((row - r0) // lanes) * (int(kdi.max()) + 1) + kdi.ravel(), lines, nl)
    return nnz, out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--lvals", type=int, nargs="+", default=[4, 3])
    ap.add_argument("--pbc", action="store_true")
    ap.add_argument("--rows", type=int, default=2_000_000, help="rows analysed (a window in the middle)")
    args = ap.parse_args()
    from edlgt_amd import QED_Model
    mod = QED_Model(spin=1, pure_theory=False, lvals=args.lvals, has_obc=[not args.pbc] * len(args.lvals))
    H = mod.build_Hamiltonian(1.0, 0.5)
    rp, col, _ = H.to_host()
    cfg = mod.sector_configs.astype(np.int64)
    dims = np.asarray(mod.loc_dims, np.int64)
    keys = np.zeros(cfg.shape[0], np.int64)
    for s in range(cfg.shape[1]):  # mixed radix, site 0 most significant
        keys = keys * dims[s] + cfg[:, s]
    N = rp.size - 1
    r0 = max(0, N // 2 - args.rows // 2)
    r1 = min(N, r0 + args.rows)
    nnz, out = analyse(rp, col, keys, r0, r1)
    print(f"QED {args.lvals} {'PBC' if args.pbc else 'OBC'}: N={N} rows [{r0}, {r1}) nnz={nnz}")
    print("| order | gathered entries per 128-byte line request |")
    print("|---|---|")
    for k, v in out.items():
        print(f"| {k} | {v:.2f} |")


if __name__ == "__main__":
    main()
