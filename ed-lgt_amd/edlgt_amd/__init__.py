"""edlgt_amd -- MI355X-native engine for edlgt's exact-diagonalization hot path.

Drop-in model API (QED_Model, QuantumModel, QMB_hamiltonian, QMB_state) over
hand-written gfx950 HIP kernels reached through the C ABI of libedlgt_hip
(include/edlgt_hip.h).  Import the package with ``ed-lgt_amd`` on sys.path.
"""
from . import _lib
from ._lib import EdlgtError, NotConvergedError, PREC_FP64, PREC_MIXED
from .device import CSRMatrix, SectorBasis, Server
from .lattice import Lattice
from .models import QED_Model, QMB_hamiltonian, QMB_state, QuantumModel, SU2_Model, Zn_Model

__all__ = [
    "QED_Model", "SU2_Model", "Zn_Model", "QuantumModel", "QMB_hamiltonian", "QMB_state", "SectorBasis", "CSRMatrix", "Server",
    "Lattice", "EdlgtError", "NotConvergedError", "PREC_FP64", "PREC_MIXED",
]
