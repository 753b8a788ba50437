"""Copies the arrays of the reference's SU(2) golden dynamics file
(validation/su2/gold_dynamics_case01_dense.npz: L=6 PBC, j=1/2, N_tot=6, g=1,
m=5, state "V", t = 0..0.5 step 0.01; fidelity F, electric energy E2 and the
half-chain entanglement entropy S, dense propagation) into a fixture here.
Loaded with allow_pickle=False: plain float arrays only."""
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/validation/su2/gold_dynamics_case01_dense.npz")
DST = Path(__file__).with_name("su2_dynamics_L6_pbc.npz")

if __name__ == "__main__":
    g = np.load(SRC, allow_pickle=False)
    arrays = {k: np.asarray(g[k], dtype=np.float64) for k in ("time_steps", "F", "E2", "S")}
    np.savez(DST, **arrays)
    print({k: v.shape for k, v in arrays.items()})
