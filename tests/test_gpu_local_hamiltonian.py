"""GPU: SU2_Model.build_local_Hamiltonian (SU2_model.py:626-699), the energy
density around a site of the 1D chain with matter, against the oracle's
assembly of the same terms from its own operators (E_square and the
staggered mass N-1 at R0; [Qpx_dag, Qmx] times t/2 and [Qpx, Qmx_dag] times
-t/2 on the bonds (R0, R0+1) and (R0-1, R0)): pattern bit-exact, values at
1e-14; summed over R0 it equals the model's Hamiltonian (periodic and open
chains); building it leaves the model's own program and couplings alone.
Between two momentum sectors (set_momentum_pair) the rectangular block
against the oracle's projectors."""
import numpy as np
import pytest

from oracle.model import SU2Oracle, Triplets, nbody_triplets

pytestmark = pytest.mark.gpu


def _oracle_local(orc, g, m, R0, coeffs):
    n = orc.n_sites
    T = Triplets(orc.N)
    d = orc.ops["E_square"].shape[1]

    def place(names, sites, strength):
        stack = np.zeros((len(sites), n, d, d), dtype=np.complex128)
        for f, (nm, s) in enumerate(zip(names, sites)):
            stack[f, s] = orc.ops[nm][s]
        r, c, v = nbody_triplets(stack, sites, orc.configs, orc.keys, orc.loc_dims)
        T.add(r, c, v * strength)

    place(["E_square"], [R0], coeffs["E"])
    place(["N-1"], [R0], (-1) ** R0 * coeffs["m"])
    half = 0.5 * coeffs["tx_even"]
    bonds = [(R0, (R0 + 1) % n), ((R0 - 1) % n, R0)]
    if orc.has_obc[0]:  # no bond across the open ends
        bonds = [b for b in bonds if b[1] == b[0] + 1]
    for names, st in ((["Qpx_dag", "Qmx"], half), (["Qpx", "Qmx_dag"], -half)):
        for b in bonds:
            place(names, list(b), st)
    return T.csr()


@pytest.mark.parametrize("L,obc", [(6, False), (8, True)])
def test_local_hamiltonian(gpu, L, obc):
    from edlgt_amd import SU2_Model
    print(f"local Hamiltonian L={L} obc={obc}", flush=True)
    orc = SU2Oracle(False, [L], [obc], sectors=[L])
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[L], lvals=[L], has_obc=[obc])
    H = mod.build_Hamiltonian(1.3, 0.7).to_scipy()
    coeffs = dict(mod.coeffs)
    prog = [dict(p) for p in mod._placements]
    total = None
    for R0 in range(L):
        h = mod.build_local_Hamiltonian(1.3, 0.7, R0).to_scipy()
        want = _oracle_local(orc, 1.3, 0.7, R0, mod.coeffs)
        assert np.array_equal(h.indptr, want.indptr) and np.array_equal(h.indices, want.indices), R0
        assert np.max(np.abs(h.data - want.data), initial=0.0) <= 1e-14 * np.max(np.abs(want.data)), R0
        assert abs(h - h.conj().T).max() <= 1e-14 * np.max(np.abs(h.data)), R0
        assert mod.Hlocal.Ham.n == mod.sector_dim
        total = h if total is None else total + h
    diff = (total - H)
    assert abs(diff).max() <= 1e-13 * np.max(np.abs(H.data))
    assert mod.coeffs == coeffs and [dict(p) for p in mod._placements] == prog
    with pytest.raises(ValueError):
        mod.build_local_Hamiltonian(1.3, 0.7, L)


@pytest.mark.parametrize("k1,k2", [(0, 0), (0, 1), (1, 2)])
def test_local_hamiltonian_between_momentum_sectors(gpu, k1, k2):
    """set_momentum_pair / check_momentum_pair (quantum_model.py:278-398) and
    the pair-mode block P_k1^dagger H_R0 P_k2 against the oracle's projectors
    (oracle/momentum.py) at 1e-13; <psi_k1|H_R0|psi_k2> from the stored
    triplets (mixed_exp_val_data, qmb_state.py:657-679), as the reference's
    convolution matrix takes it."""
    from edlgt_amd import SU2_Model
    from edlgt_amd.models import mixed_exp_val_data
    from oracle.momentum import momentum_basis
    print(f"local Hamiltonian between k={k1} and k={k2}", flush=True)
    orc = SU2Oracle(False, [6], [False], sectors=[6])
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[6], lvals=[6], has_obc=[False])
    mod.build_Hamiltonian(1.0, 5.0)
    mod.set_momentum_pair([k1], [k2], [2], False)
    mod.check_momentum_pair()
    assert mod.momentum_basis is None
    A = mod.build_local_Hamiltonian(1.0, 5.0, 0).toarray()
    BL = momentum_basis(orc.configs, orc.keys, orc.loc_dims, [6], [2], [k1])
    BR = momentum_basis(orc.configs, orc.keys, orc.loc_dims, [6], [2], [k2])
    want = (BL.conj().T @ _oracle_local(orc, 1.0, 5.0, 0, mod.coeffs) @ BR).toarray()
    assert A.shape == want.shape
    assert np.max(np.abs(A - want)) <= 1e-13 * np.max(np.abs(want))
    rng = np.random.default_rng(k1 + 3 * k2)
    p1 = rng.normal(size=A.shape[0]) + 1j * rng.normal(size=A.shape[0])
    p2 = rng.normal(size=A.shape[1]) + 1j * rng.normal(size=A.shape[1])
    H = mod.Hlocal
    got = mixed_exp_val_data(p1, p2, H.row_list, H.col_list, H.value_list)
    assert abs(got - np.vdot(p1, want @ p2)) <= 1e-12 * max(1.0, abs(got))
