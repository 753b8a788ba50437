"""GPU: the two largest workloads at their full sizes.

QED 4x3: 2D U(1) QED on the 4x3 periodic lattice with staggered fermions,
spin-1 links, g = 1, m = 0.5 (QED_model.py:23-233; 28,904,615 states).  The
sector size and stored-entry count, the CSR structure (monotone row pointer,
sorted columns, structural symmetry and H[r,c] = conj(H[c,r]) on sampled
rows), and the lowest eigenvalue through every operator form (FP64, mixed
FP32/FP64, dictionary-coded, matrix-free) against each other and against the
host-path value in tests/golden/qed4x3_pbc_e0.json (oracle/refcpu restatement
of the reference's build + eigsh; command recorded in the fixture).  The
reference's own KATs pin per-site energies at atol 1e-10
(validation/qed/test_qed_2d_matter.py:43-74); here the totals are compared.

SU(2) 20 sites: 1D SU(2) spin-1/2 with matter, OBC, N_tot = 20, g = m = 1
(SU2_model.py:31-171; 377,379,369 states, 6,957,066,525 stored entries, so
64-bit entry offsets).  A few Krylov steps of exp(-iH dt) from the bare
vacuum must conserve the norm and <H> (exact invariants of a unitary
generated by H).
"""
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIXTURE_PATH = Path(__file__).resolve().parent / "golden" / "qed4x3_pbc_e0.json"
N_QED4X3 = 28_904_615
NNZ_QED4X3 = 578_948_151


@pytest.fixture(scope="module")
def qed4x3(gpu):
    global FIXTURE
    FIXTURE = json.loads(FIXTURE_PATH.read_text())
    import torch
    from edlgt_amd import QED_Model
    torch.cuda.empty_cache()
    print("QED 4x3 PBC: building sector and CSR", flush=True)
    mod = QED_Model(spin=1, pure_theory=False, lvals=[4, 3], has_obc=[False, False])
    mod.matrix_free = False
    H = mod.build_Hamiltonian(1.0, 0.5)
    yield mod, H
    del mod.H, H
    torch.cuda.empty_cache()


def test_qed4x3_sector_and_csr_structure(qed4x3):
    mod, H = qed4x3
    assert mod.sector_dim == N_QED4X3 == FIXTURE["sector_dim"]
    assert H.nnz == NNZ_QED4X3 == FIXTURE["nnz"]
    rp, col, val = H.to_host()
    assert rp[0] == 0 and rp[-1] == H.nnz
    lens = np.diff(rp)
    assert lens.min() >= 1 and lens.max() <= 64
    assert col.min() >= 0 and col.max() < N_QED4X3
    rng = np.random.default_rng(5)
    rows = np.sort(rng.choice(N_QED4X3, 10_000, replace=False))
    # every row's first and last state too
    rows = np.unique(np.concatenate([rows, [0, N_QED4X3 - 1]]))
    starts, ends = rp[rows], rp[rows + 1]
    pos = np.concatenate([np.arange(s, e) for s, e in zip(starts, ends)])
    r_ent = np.repeat(rows, ends - starts)
    c_ent, v_ent = col[pos], val[pos]
    # columns strictly ascending within each row
    same_row = r_ent[1:] == r_ent[:-1]
    assert np.all(np.diff(c_ent)[same_row] > 0)
    # the transposed entry exists once and holds the conjugate value
    cs, ce = rp[c_ent], rp[c_ent + 1]
    width = int((ce - cs).max())
    idx = cs[:, None] + np.arange(width)[None, :]
    valid = idx < ce[:, None]
    idx = np.where(valid, idx, 0)
    hit = (col[idx] == r_ent[:, None]) & valid
    assert np.all(hit.sum(axis=1) == 1), "pattern not structurally symmetric"
    j = idx[np.arange(idx.shape[0]), hit.argmax(axis=1)]
    scale = float(np.abs(val).max())
    assert np.max(np.abs(val[j] - np.conj(v_ent))) <= 1e-14 * scale


def test_qed4x3_lowest_eigenvalue_every_form(qed4x3):
    import torch
    from edlgt_amd import PREC_MIXED, QED_Model
    mod, H = qed4x3
    e_ref = FIXTURE["E0"]
    print("QED 4x3: FP64 Lanczos", flush=True)
    mod.diagonalize_Hamiltonian(1, "sparse")
    e64 = float(mod.H.energies[0])
    assert abs(e64 - e_ref) <= 1e-10, (e64, e_ref)
    assert abs(mod.H.Nenergies[0] - e_ref / 12) <= 1e-11
    print("QED 4x3: mixed Lanczos", flush=True)
    mod.diagonalize_Hamiltonian(1, "sparse", precision=PREC_MIXED)
    emix = float(mod.H.energies[0])
    assert abs(emix - e64) <= 1e-10, (emix, e64)
    print("QED 4x3: real-gauge FP64 and mixed Lanczos", flush=True)
    mod.diagonalize_Hamiltonian(1, "sparse", real_gauge=True)
    assert mod.H.real_gauge
    erg = float(mod.H.energies[0])
    assert abs(erg - e64) <= 1e-10, (erg, e64)
    psi = mod.H.Npsi[0].psi_device
    assert float(torch.linalg.vector_norm(H.matvec(psi) - erg * psi)) <= 1e-8
    mod.diagonalize_Hamiltonian(1, "sparse", real_gauge=True, precision=PREC_MIXED)
    assert abs(float(mod.H.energies[0]) - e64) <= 1e-10
    del psi
    mod._rg = None
    torch.cuda.empty_cache()
    print("QED 4x3: coded Lanczos", flush=True)
    assert H.compress() > 0
    ev, _, _, _ = H.lanczos(1, tol=1e-12, want_vectors=False)
    assert abs(float(ev[0]) - e64) <= 1e-10, (ev[0], e64)
    del mod.H.Npsi
    torch.cuda.empty_cache()
    print("QED 4x3: matrix-free Lanczos", flush=True)
    mf = QED_Model(spin=1, pure_theory=False, lvals=[4, 3], has_obc=[False, False])
    mf.matrix_free = True
    Hf = mf.build_Hamiltonian(1.0, 0.5)
    ev, _, _, _ = Hf.lanczos(1, tol=1e-12, want_vectors=False)
    assert abs(float(ev[0]) - e64) <= 1e-10, (ev[0], e64)
    del mf, Hf
    torch.cuda.empty_cache()


def test_su2_20_sites_krylov_invariants(gpu):
    import torch
    from edlgt_amd import SU2_Model
    torch.cuda.empty_cache()
    print("SU(2) 20 sites: sector", flush=True)
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[20], lvals=[20], has_obc=[True])
    assert mod.sector_dim == 377_379_369
    print("SU(2) 20 sites: Hamiltonian", flush=True)
    mod.matrix_free = False
    H = mod.build_Hamiltonian(1.0, 1.0)
    assert H.nnz == 6_957_066_525 and H.nnz > 2**31
    # coded values (5 bytes per entry) so that the Krylov basis fits beside it
    assert H.compress() > 0
    H.release_values()
    psi0 = mod.get_qmb_state_from_configs([mod.overlap_QMB_state("V")])
    psi = psi0
    y = torch.empty_like(psi)

    def energy(v):
        H.matvec(v, y)
        return complex(torch.vdot(v, y).item())

    e0 = energy(psi)
    assert abs(e0.imag) <= 1e-12 * max(1.0, abs(e0.real))
    dt = 0.05
    for step in range(4):
        print(f"SU(2) 20 sites: Krylov step {step + 1}", flush=True)
        mod.time_evolution_Hamiltonian(psi, [0.0, dt], m=20, reorth=False, store_states=False)
        psi = mod.H.psi_final
        nrm = float(torch.linalg.vector_norm(psi).item())
        assert abs(nrm - 1.0) <= 1e-10, (step, nrm)
        e = energy(psi)
        assert abs(e.real - e0.real) <= 1e-10 * max(1.0, abs(e0.real)), (step, e, e0)
        assert abs(e.imag) <= 1e-10 * max(1.0, abs(e0.real))
    assert float(mod.H.krylov_err.max()) <= 1e-10
    # the state moved: the bare vacuum is not stationary
    assert abs(complex(torch.vdot(psi0, psi).item())) ** 2 < 1 - 1e-6
    mod.H.release_krylov()
    del mod, H, psi, psi0, y
    torch.cuda.empty_cache()
