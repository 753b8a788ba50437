"""CPU: the real gauge of QED and SU(2) with staggered matter.

D = diag(i^(sum_j a_j N_j)), a_j = (x_j + z_j) mod 2, turns the complex QED
Hamiltonian with matter (hopping -i/2 along x, -i/2 / +i/2 along z, real along
y: QED_model.py:719-776) into a real matrix D^H H D with the same spectrum.
Pinned here against the definition, with the oracle's complex H:
  * D^H H D has exactly zero imaginary parts on lattices whose periodic axes
    along x and z have even extents (1D, 2D, 3D, open and periodic), and keeps
    imaginary entries on a lattice periodic along x with an odd extent;
  * the engine's program for it (QED_Model._real_gauge_placements: each
    placement's strength times a fixed power of i), assembled here on the host
    with the oracle's triplet routine, equals D^H H D entry for entry, and the
    unmodified program equals H (which pins this host assembly).
The same holds for SU(2) with matter (hopping -2 sqrt(2) i along x, real
along y and z: SU2_model.py:555-624) with a_j = x_j mod 2 and N_tot.
The engine's model is built on the host with the sector taken from the oracle
(the device enumeration is pinned bit-exactly elsewhere); the configuration
printout, which reads the same host-side model, is checked here too."""
import numpy as np
import pytest

import edlgt_amd.models as models
from oracle import lattice as lat
from oracle.model import QEDOracle, SU2Oracle, Triplets, nbody_triplets


def _phase_exponents(orc, a, name="N"):
    nd = np.rint(np.real(np.einsum("sii->si", orc.ops[name]))).astype(np.int64)
    s = np.zeros(orc.N, dtype=np.int64)
    for j in range(orc.n_sites):
        s += a[j] * nd[j][orc.configs[:, j]]
    return s % 4


def _weights(lvals, with_z=True):
    a = []
    for i in range(int(np.prod(lvals))):
        co = lat.zig_zag(lvals, i)
        a.append((co[0] + (co[2] if len(co) > 2 and with_z else 0)) % 2)
    return np.array(a, dtype=np.int64)


def _gauged(H, s):
    ph = 1j ** s
    C = H.tocoo()
    G = C.copy()
    G.data = np.conj(ph[C.row]) * C.data * ph[C.col]
    return G.tocsr()


CASES = [([6], [True]), ([4], [False]), ([2, 2], [False, False]), ([3, 2], [True, True]),
         ([4, 2], [True, False]), ([2, 2, 2], [True, True, True])]


@pytest.mark.parametrize("lvals,obc", CASES)
def test_gauge_makes_the_hamiltonian_real(lvals, obc):
    orc = QEDOracle(1, False, lvals, obc)
    H = orc.hamiltonian(1.3, 0.7)
    assert np.max(np.abs(H.data.imag)) > 0.1  # complex before
    G = _gauged(H, _phase_exponents(orc, _weights(lvals)))
    assert np.max(np.abs(G.data.imag)) == 0.0
    # same spectrum (unitary similarity)
    if orc.N <= 600:
        e1 = np.linalg.eigvalsh(H.toarray())
        e2 = np.linalg.eigvalsh(G.real.toarray())
        assert np.max(np.abs(e1 - e2)) < 1e-12


def test_odd_periodic_extent_has_no_real_gauge():
    """3x2 periodic: the wrap-around x bond joins two sites with the same a_j,
    so its hopping entries stay imaginary."""
    orc = QEDOracle(1, False, [3, 2], [False, False])
    G = _gauged(orc.hamiltonian(1.3, 0.7), _phase_exponents(orc, _weights([3, 2])))
    assert np.max(np.abs(G.data.imag)) == 0.5


class _HostSector:
    """Stands in for the device sector: the oracle's table (bit-exact with the
    device enumeration in the GPU parity tests)."""

    def __init__(self, orc):
        self.n_states = orc.N


def _engine_model(monkeypatch, orc, lvals, obc):
    monkeypatch.setattr(models, "SectorBasis", lambda *a, **k: _HostSector(orc))
    mod = models.QED_Model(spin=1, pure_theory=False, lvals=lvals, has_obc=obc)
    mod._build_program = lambda real: mod._keep_program()  # keep the program on the host
    mod.build_Hamiltonian(1.3, 0.7)
    return mod


def _assemble(placements, tables, orc):
    """The engine's placement semantics on the host: entries of the ordered
    operator product times the strength; dagger placements add the conjugate
    transpose; single-factor placements keep |O_aa| >= 1e-10 (oracle routine)."""
    T = Triplets(orc.N)
    d = tables[0].shape[0]
    for P in placements:
        stack = np.zeros((len(P["sites"]), orc.n_sites, d, d), dtype=np.complex128)
        for f, (s, t) in enumerate(zip(P["sites"], P["tables"])):
            stack[f, s] = tables[t]
        r, c, v = nbody_triplets(stack, P["sites"], orc.configs, orc.keys, orc.loc_dims)
        v = v * P["strength"]
        if P.get("dagger"):
            T.add(c, r, np.conj(v))
        else:
            T.add(r, c, v)
    return T.csr()


@pytest.mark.parametrize("lvals,obc", [([6], [True]), ([4], [False]), ([2, 2], [False, False]),
                                       ([3, 2], [True, True]), ([2, 2, 2], [True, True, True])])
def test_engine_program_is_the_gauged_hamiltonian(monkeypatch, lvals, obc):
    orc = QEDOracle(1, False, lvals, obc)
    mod = _engine_model(monkeypatch, orc, lvals, obc)
    a, name = mod._gauge_weights()
    assert name == "N" and np.array_equal(a, _weights(lvals))
    H = orc.hamiltonian(1.3, 0.7)
    plain = _assemble(mod._placements, mod._tables, orc)
    assert np.array_equal(plain.indptr, H.indptr) and np.array_equal(plain.indices, H.indices)
    assert np.max(np.abs(plain.data - H.data)) <= 1e-14
    rg = mod._real_gauge_placements()
    assert rg is not None and len(rg) == len(mod._placements)
    G = _assemble(rg, mod._tables, orc)
    want = _gauged(H, _phase_exponents(orc, a))
    assert np.array_equal(G.indptr, want.indptr) and np.array_equal(G.indices, want.indices)
    assert np.max(np.abs(G.data.imag)) == 0.0
    assert np.max(np.abs(G.data - want.data)) <= 1e-14


def test_no_gauge_for_the_pure_theory_and_odd_periodic_program(monkeypatch):
    """The pure theory has no matter phase (already real: the engine builds it
    in float64); on the odd periodic lattice the per-placement program exists but
    its assembly keeps imaginary entries (the device build refuses it)."""
    orc = QEDOracle(1, False, [3, 2], [False, False])
    mod = _engine_model(monkeypatch, orc, [3, 2], [False, False])
    G = _assemble(mod._real_gauge_placements(), mod._tables, orc)
    assert np.max(np.abs(G.data.imag)) == 0.5
    pure = object.__new__(models.QED_Model)
    pure.spin, pure.pure_theory = 1, True
    assert pure._gauge_weights() is None


SU2_CASES = [([6], [True], 0.5), ([6], [False], 0.5), ([2, 2], [False, False], 0.5), ([4], [True], 1.0),
             ([2, 2, 2], [True, True, True], 0.5)]


def _su2(lvals, obc, spin):
    n = int(np.prod(lvals))
    return SU2Oracle(False, lvals, obc, sectors=[n], spin=spin, generic=spin > 0.5)


@pytest.mark.parametrize("lvals,obc,spin", SU2_CASES)
def test_su2_gauge_makes_the_hamiltonian_real(lvals, obc, spin):
    orc = _su2(lvals, obc, spin)
    H = orc.hamiltonian(1.3, 0.7)
    assert np.max(np.abs(H.data.imag)) > 1.0
    G = _gauged(H, _phase_exponents(orc, _weights(lvals, with_z=False), "N_tot"))
    assert np.max(np.abs(G.data.imag)) == 0.0


def test_su2_odd_periodic_chain_has_no_real_gauge():
    orc = SU2Oracle(False, [5], [False], sectors=[4])
    G = _gauged(orc.hamiltonian(1.3, 0.7), _phase_exponents(orc, _weights([5], False), "N_tot"))
    assert np.max(np.abs(G.data.imag)) > 1.0


@pytest.mark.parametrize("lvals,obc,spin", SU2_CASES[:4])
def test_su2_engine_program_is_the_gauged_hamiltonian(monkeypatch, lvals, obc, spin):
    orc = _su2(lvals, obc, spin)
    monkeypatch.setattr(models, "SectorBasis", lambda *a, **k: _HostSector(orc))
    n = int(np.prod(lvals))
    mod = models.SU2_Model(spin=spin, pure_theory=False, sectors=[n], lvals=lvals, has_obc=obc)
    mod._build_program = lambda real: mod._keep_program()
    mod.build_Hamiltonian(1.3, 0.7)
    a, name = mod._gauge_weights()
    assert name == "N_tot" and np.array_equal(a, _weights(lvals, with_z=False))
    H = orc.hamiltonian(1.3, 0.7)
    plain = _assemble(mod._placements, mod._tables, orc)
    assert np.array_equal(plain.indptr, H.indptr) and np.array_equal(plain.indices, H.indices)
    assert np.max(np.abs(plain.data - H.data)) <= 1e-14 * np.max(np.abs(H.data))
    G = _assemble(mod._real_gauge_placements(), mod._tables, orc)
    want = _gauged(H, _phase_exponents(orc, a, name))
    assert np.array_equal(G.indptr, want.indptr) and np.array_equal(G.indices, want.indices)
    assert np.max(np.abs(G.data.imag)) == 0.0
    assert np.max(np.abs(G.data - want.data)) <= 1e-14 * np.max(np.abs(H.data))


def test_print_state_config(monkeypatch, caplog):
    """print_state_config (QED_model.py:901-946): one line per site with the
    local index and its gauge-invariant state (matter, -x -y, +x +y links),
    the rows of the oracle's local basis for the site's label."""
    import logging
    orc = QEDOracle(1, False, [2, 2], [False, False])
    mod = _engine_model(monkeypatch, orc, [2, 2], [False, False])
    cfg = orc.configs[5]
    with caplog.at_level(logging.INFO, logger=models.logger.name):
        mod.print_state_config(cfg, amplitude=0.5)
    lines = [r.getMessage() for r in caplog.records]
    assert any("|psi|^2=0.25000000" in l for l in lines)
    assert any(l.strip().startswith("[") and l.split() == ["[", "M", "-x", "-y", "+x", "+y", "]"] for l in lines)
    rows = [l for l in lines if l.startswith("SITE")]
    assert len(rows) == 4
    for i, l in enumerate(rows):
        vals = [int(v) for v in l.split("[")[1].rstrip(" ]").split()]
        assert vals == [int(round(v)) for v in orc.gauge_states[orc.labels[i]][cfg[i]]]
