"""GPU: the allocation-free marginals entry point (edlgt_basis_marginals_noalloc,
partial sums in the library's static per-device scratch array) against the
allocating edlgt_basis_marginals, bit for bit, on two calls in a row -- the
second on a smaller sector -- for complex and real vectors."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _psi(torch, n, cplx, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    if cplx:
        v = torch.randn(n, dtype=torch.complex128, device="cuda", generator=g)
    else:
        v = torch.randn(n, dtype=torch.float64, device="cuda", generator=g)
    return (v / torch.linalg.vector_norm(v)).contiguous()


def _both(torch, L, basis, psi, d_max, cplx):
    lib = L.load()
    vt = L.VALUE_C128 if cplx else L.VALUE_F64
    ref = np.zeros(basis.n_sites * d_max)
    L.check(lib.edlgt_basis_marginals(basis.handle, L.ptr(psi), vt, d_max, ref.ctypes.data,
                                      L.stream_handle()))
    out = torch.full((basis.n_sites * d_max,), float("nan"), dtype=torch.float64, device="cuda")
    L.check(lib.edlgt_basis_marginals_noalloc(basis.handle, L.ptr(psi), vt, d_max, L.ptr(out),
                                              L.stream_handle()))
    torch.cuda.synchronize()
    return ref, out.cpu().numpy()


@pytest.mark.parametrize("cplx", [True, False])
def test_noalloc_marginals_equal_allocating(gpu, cplx):
    import torch
    from edlgt_amd import QED_Model, _lib as L
    # first call: QED 4x2 PBC with matter (140,647 states, 8 sites x 35 states)
    big = QED_Model(spin=1, pure_theory=False, lvals=[4, 2], has_obc=[False, False])
    # second call: QED 1D L=10 OBC (197 states), a smaller sector and histogram
    small = QED_Model(spin=1, pure_theory=False, lvals=[10], has_obc=[True])
    for step, mod in enumerate((big, small)):
        d_max = int(mod.loc_dims.max())
        psi = _psi(torch, mod.sector_dim, cplx, seed=11 + step)
        print(f"noalloc marginals: step {step} N={mod.sector_dim} d_max={d_max} cplx={cplx}", flush=True)
        ref, got = _both(torch, L, mod.basis, psi, d_max, cplx)
        assert np.array_equal(ref, got), np.max(np.abs(ref - got))
        # each site's marginal is a distribution of |psi|^2
        sums = got.reshape(mod.n_sites, d_max).sum(axis=1)
        assert np.allclose(sums, 1.0, atol=1e-12)


def test_noalloc_marginals_rejects_oversized_histogram(gpu):
    import torch
    from edlgt_amd import QED_Model, _lib as L
    mod = QED_Model(spin=1, pure_theory=False, lvals=[10], has_obc=[True])
    psi = _psi(torch, mod.sector_dim, True, seed=3)
    out = torch.zeros(10 * 4096, dtype=torch.float64, device="cuda")
    st = L.load().edlgt_basis_marginals_noalloc(mod.basis.handle, L.ptr(psi), L.VALUE_C128, 4096,
                                                L.ptr(out), L.stream_handle())
    assert st == L.ERR_INVALID_ARG
    assert "4096" in L.last_error()
