"""CPU oracle for the exact-diagonalization hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped engine (``ed-lgt_amd/``)
imports, links or calls anything in this directory.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use
it, and only as the checker (or as the timed CPU baseline), never as the thing
measured or shipped.

The oracle is an independent CPU restatement of the reference ``edlgt``
(gcataldi96/ed-lgt 0.2.3) algorithm for the path

    local dressed-site basis/operators -> per-site projection
    -> symmetry-sector enumeration (sorted basis)
    -> n-body triplet assembly with binary search -> canonical CSR
    -> Lanczos / eigsh and Krylov exp(-iHt)

Each function cites the reference file:line it restates.  Every function is
pinned by at least one of: the reference's own known-answer energies
(``validation/qed/*.py``, atol 1e-10 on per-site energies), brute-force
full-space Kronecker Hamiltonians projected onto the sector (the reference's
own technique, ``validation/qed/test_qed_1d_integrated.py:13-152``),
exhaustive sector enumeration, closed forms, and independent transfer-matrix
sector counts.  See ``tests/test_oracle_*.py`` for the pins.

Modules:
  lattice   -- geometry (site order, neighbours, plaquettes, link pairs, masks)
  qed_ops   -- U(1) rishon / dressed-site operators and gauge-invariant bases
  zn_ops    -- Z_n pure-gauge dressed-site operators and bases
  model     -- per-site projection, sector enumeration, triplets, CSR, eigsh,
               expm_multiply, and the QED / Z_n model drivers
  brute     -- full-space Kronecker Hamiltonians for tiny lattices
  refcpu.c  -- C/OpenMP restatement of the reference's Numba kernels
               (sector, two-pass triplets with binary search, COO->CSR,
               row-parallel SpMV), used as the timed CPU baseline
"""
