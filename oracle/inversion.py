"""Spatial inversion of the 1D SU(2) chain as a signed permutation of the
sector states (oracle restatement; test infrastructure).

Follows edlgt/symmetries/inversion_sym.py: site-centred inversion about
j0 = L/2 - 1, j -> (2 j0 - j) mod L (parity_perm_site, :18-43), or
bond-centred about (j0, j0 + 1), j -> (2 j0 + 1 - j) mod L (parity_perm_bond,
:46-74); the image configuration puts loc_perm[c_j] on site perm[j] and
carries prod_j loc_phase[c_j] (parity_image_config, :77-113); column r of P
holds that sign in the image's row (build_parity_operator, :116-161).  The
local maps are SU2_Model.local_parity_labels (models/SU2_model.py:708-741).
"""
import numpy as np
import scipy.sparse as sp

from .model import encode

# SU2_model.py:733-740: (loc_perm, loc_phase) of the six dressed states
SU2_PARITY_SITE = (np.array([0, 1, 3, 2, 4, 5]), np.array([1, -1, 1, -1, 1, -1]))
SU2_PARITY_BOND = (np.array([4, 5, 3, 2, 0, 1]), np.array([1, -1, 1, 1, 1, -1]))


def parity_operator(configs, keys, loc_dims, wrt_site, loc_maps=None):
    """P (N x N CSR, one +-1 per column) of the inversion about site L/2 - 1
    (wrt_site) or about the bond (L/2 - 1, L/2)."""
    loc_perm, loc_phase = loc_maps or (SU2_PARITY_SITE if wrt_site else SU2_PARITY_BOND)
    N, L = configs.shape
    center2 = 2 * (L // 2 - 1) + (0 if wrt_site else 1)
    perm = (center2 - np.arange(L)) % L
    c = configs.astype(np.int64)
    new = np.empty_like(c)
    new[:, perm] = np.asarray(loc_perm)[c]
    sign = np.prod(np.asarray(loc_phase, dtype=float)[c], axis=1)
    k = encode(new, loc_dims)
    row = np.searchsorted(keys, k)
    if np.any(row >= N) or np.any(keys[np.minimum(row, N - 1)] != k):
        raise ValueError("an image configuration is not in the sector")
    return sp.csr_matrix((sign, (row, np.arange(N))), shape=(N, N))
