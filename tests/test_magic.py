"""CPU: the stabilizer Renyi-2 entropy restated from the reference
(oracle/magic.py), pinned by closed forms: a basis state and the uniform
superposition over a full product space are stabilizer states (SRE2 = 0); a
qubit cos(t)|0> + sin(t)|1> has M2 = (a^2 + b^2)^2 + 12 a^2 b^2 (a = cos^2 t,
b = sin^2 t), which is 1 at t = pi/4 and 3/4 for the magic state t = pi/8;
a product state's SRE2 is the sum over its factors; the support keeps the
smallest set of largest weights holding 1 - delta."""
import itertools

import numpy as np
import pytest

from oracle import magic


def _product_space(dims):
    return np.array(list(itertools.product(*[range(d) for d in dims])), dtype=np.int64)


def _qubit_product(thetas):
    psi = np.array([1.0 + 0j])
    for t in thetas:
        psi = np.kron(psi, np.array([np.cos(t), np.sin(t)]))
    return psi


def test_stabilizer_states_have_zero_sre():
    dims = [3, 2, 3]
    cfg = _product_space(dims)
    e = np.zeros(cfg.shape[0], complex)
    e[7] = 1.0
    assert abs(magic.sre2(e, dims, cfg)[0]) < 1e-14
    u = np.full(cfg.shape[0], 1 / np.sqrt(cfg.shape[0]), complex)
    assert abs(magic.sre2(u, dims, cfg, delta=1e-6)[0]) < 1e-13


@pytest.mark.parametrize("t", [np.pi / 4, np.pi / 8, 0.3])
def test_single_qubit_closed_form(t):
    a, b = np.cos(t) ** 2, np.sin(t) ** 2
    m2 = (a * a + b * b) ** 2 + 12 * a * a * b * b
    got = magic.sre2(np.array([np.cos(t), np.sin(t)], complex), [2], _product_space([2]), delta=1e-9)
    assert abs(got[0] + np.log(m2)) < 1e-14
    if t == np.pi / 8:
        assert abs(m2 - 0.75) < 1e-15


def test_product_state_is_additive():
    ts = [np.pi / 8, 0.3, np.pi / 4, 0.1]
    single = [magic.sre2(np.array([np.cos(t), np.sin(t)], complex), [2], _product_space([2]), delta=1e-12)[0]
              for t in ts]
    got = magic.sre2(_qubit_product(ts), [2] * 4, _product_space([2] * 4), delta=1e-12)
    assert abs(got[0] - sum(single)) < 1e-13
    assert got[2] == 16 and got[1] == 16


def test_support_is_the_smallest_heavy_set():
    psi = np.sqrt(np.array([0.5, 0.3, 0.15, 0.05]) + 0j)
    cfg = _product_space([4])
    idx, c, _, keys, disc = magic.extract_support(psi, [4], cfg, delta=0.1)
    assert list(idx) == [0, 1, 2] and abs(disc - 0.05) < 1e-15
    idx, *_ = magic.extract_support(psi, [4], cfg, delta=0.2)
    assert list(idx) == [0, 1]
    with pytest.raises(ValueError):
        magic.extract_support(psi, [4], cfg, delta=0.0)
