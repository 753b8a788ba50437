"""ctypes binding of libedlgt_hip (C ABI in include/edlgt_hip.h).

The shared library is built in-tree by ``ed-lgt_amd/build.py``.  Loading
fails loudly when it is missing: there is no CPU fallback for any
hot-path operation.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libedlgt_hip.so"

OK = 0
ERR_INVALID_ARG = -1
ERR_KEY_NOT_FOUND = -2
ERR_INDEX_OVERFLOW = -3
ERR_OUT_OF_MEMORY = -4
ERR_HIP = -5
ERR_NOT_CONVERGED = -6
ERR_NOT_REAL = -7

VALUE_F64, VALUE_C128, VALUE_F32, VALUE_C64 = 0, 1, 2, 3
PREC_FP64, PREC_MIXED = 0, 1
SYM_U1, SYM_Z2 = 0, 1


def SYM_MOD(n):
    """Additive modulo-n n-body constraint type (Z_n flux sectors)."""
    return 16 + int(n)


class EdlgtError(RuntimeError):
    """Device/runtime failure inside libedlgt_hip."""


class NotConvergedError(EdlgtError):
    pass


# Every symbol include/edlgt_hip.h declares, with its ctypes signature.
_p = C.c_void_p
_i32, _i64, _f64 = C.c_int32, C.c_int64, C.c_double
SIGNATURES = {
    "edlgt_last_error": (_i32, [C.c_char_p, _i64]),
    "edlgt_version": (_i32, []),
    "edlgt_device_info": (_i32, [_p, _p, _p]),
    "edlgt_sector_build": (_i32, [_i32, _p, _i32, _i32, _i32, _p, _p, _i32, _p, _p, _i32, _p, _p, _p,
                                  _i32, _i32, _p, _p, _p, _p, _p, _p]),
    "edlgt_basis_info": (_i32, [_p, _p, _p]),
    "edlgt_basis_copy_configs": (_i32, [_p, _p]),
    "edlgt_basis_key_words": (_i32, [_p, _p]),
    "edlgt_basis_copy_keys": (_i32, [_p, _p]),
    "edlgt_basis_lookup": (_i32, [_p, _p, _i64, _p]),
    "edlgt_basis_device_ptrs": (_i32, [_p, _p, _p, _p]),
    "edlgt_basis_destroy": (_i32, [_p]),
    "edlgt_hamiltonian_build": (_i32, [_p, _i32, _i32, _p, _i32, _p, _p, _p, _p, _p, _i32, _p, _p]),
    "edlgt_hamiltonian_build_matfree": (_i32, [_p, _i32, _i32, _p, _i32, _p, _p, _p, _p, _p, _i32, _p, _p]),
    "edlgt_hamiltonian_count": (_i32, [_p, _i32, _i32, _p, _i32, _p, _p, _p, _p, _p, _p, _p]),
    "edlgt_csr_is_matrix_free": (_i32, [_p, _p]),
    "edlgt_csr_info": (_i32, [_p, _p, _p, _p]),
    "edlgt_csr_copy": (_i32, [_p, _p, _p, _p]),
    "edlgt_csr_device_ptrs": (_i32, [_p, _p, _p, _p]),
    "edlgt_csr_dense_eigh": (_i32, [_p, _p, _p, _p]),
    "edlgt_csr_create": (_i32, [C.c_int64, C.c_int64, _p, _p, _p, C.c_int32, _p, _p]),
    "edlgt_basis_sre2": (_i32, [_p, _p, C.c_double, _p, _p, _p, _p, _p]),
    "edlgt_ensemble_density": (_i32, [_p, C.c_int64, C.c_int64, C.c_int64, _p, _p, _p, C.c_int64, _p, _p]),
    "edlgt_csr_make_low_precision": (_i32, [_p, _p]),
    "edlgt_csr_destroy": (_i32, [_p]),
    "edlgt_spmv": (_i32, [_p, _i32, _p, _p, _i32, _p]),
    "edlgt_spmm": (_i32, [_p, _i32, _i32, _p, _p, _i32, _p]),
    "edlgt_csr_autotune": (_i32, [_p, _i32, _i32, _i32, _p, _p, _p, _p, _p, _p]),
    "edlgt_csr_set_launch": (_i32, [_p, _i32, _i64]),
    "edlgt_hamiltonian_build_terms": (_i32, [_p, _i32, _i32, _p, _i32, _p, _p, _p, _p, _p, _p, _i32, _i32, _p, _p]),
    "edlgt_csr_copy_terms": (_i32, [_p, _p, _p]),
    "edlgt_csr_compress": (_i32, [_p, _p, _p, _p]),
    "edlgt_csr_release_values": (_i32, [_p]),
    "edlgt_spmm_coef": (_i32, [_p, _i32, _p, _p, _i32, _p, _p]),
    "edlgt_server_create": (_i32, [_p, _i32, _i32, _i32, _f64, _i32, _i64, _i32, _p, _p]),
    "edlgt_server_submit": (_i32, [_p, _i64, _p, C.c_uint64, _p]),
    "edlgt_server_info": (_i32, [_p, _p, _p, _p, _p, _p]),
    "edlgt_server_step": (_i32, [_p, _i32, _p, _p, _p, _p, _p, _p]),
    "edlgt_server_trim": (_i32, [_p]),
    "edlgt_server_destroy": (_i32, [_p]),
    "edlgt_lanczos": (_i32, [_p, _i32, _f64, _i32, _i64, _i32, _i32, _p, _p, _p, _p, _p, _p]),
    "edlgt_basis_marginals": (_i32, [_p, _p, _i32, _i32, _p, _p]),
    "edlgt_basis_marginals_noalloc": (_i32, [_p, _p, _i32, _i32, _p, _p]),
    "edlgt_basis_quarter_phase": (_i32, [_p, _p, _i32, _i32, _p, _p, _p]),
    "edlgt_basis_quarter_phase_c": (_i32, [_p, _p, _i32, _i32, _i32, _p, _p, _p]),
    "edlgt_basis_signed_image": (_i32, [_p, _p, _p, _p, _i32, _p, _p, _p]),
    "edlgt_schmidt_gram": (_i32, [_p, _i32, _i64, _p, _i32, _p, _p, _i64, _p]),
    "edlgt_hermitian_eigvals": (_i32, [_p, _i64, _i32, _p, _p, _i64, _p]),
    "edlgt_expval_terms": (_i32, [_p, _i32, _i32, _p, _i32, _p, _p, _p, _p, _p, _i32, _p, _p]),
    "edlgt_krylov_evolve": (_i32, [_p, _p, _i32, _p, _f64, _i64, _i32, _i32, _i32, _p, _p, _p, _p]),
    "edlgt_momentum_build": (_i32, [_p, _i32, _p, _p, _p, _p, _p]),
    "edlgt_momentum_build_tc": (_i32, [_p, _i32, _p, _p, _i32, _p, _p]),
    "edlgt_momentum_info": (_i32, [_p, _p, _p, _p, _p]),
    "edlgt_momentum_copy": (_i32, [_p, _p, _p, _p]),
    "edlgt_momentum_apply": (_i32, [_p, _i32, _p, _p, _p]),
    "edlgt_csr_momentum": (_i32, [_p, _p, _p, _p]),
    "edlgt_momentum_destroy": (_i32, [_p]),
    "edlgt_krylov_create": (_i32, [_p, _p, _i32, _i32, _i32, _i32, _p, _p]),
    "edlgt_krylov_run": (_i32, [_p, _p, _f64, _i64, _p, _p, _p]),
    "edlgt_krylov_destroy": (_i32, [_p]),
    "edlgt_random_uniform": (_i32, [_p, _i64, _i32, _p, _p]),
}

_lib = None


def load(path=None):
    """Opens libedlgt_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise ImportError(
            f"libedlgt_hip.so not found at {p}; build it with `python ed-lgt_amd/build.py` "
            "(the engine has no CPU fallback)"
        )
    # torch ships its own libamdhip64 with the same soname: loading it first makes
    # the dynamic linker bind this library to that one runtime.  The other order
    # maps two HIP runtimes into the process, and torch then finds no device and
    # cannot share streams or pointers with the engine.
    import torch  # noqa: F401
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def last_error():
    buf = C.create_string_buffer(4096)
    load().edlgt_last_error(buf, 4096)
    return buf.value.decode(errors="replace")


def check(status):
    """Maps a status code to the exception types the reference raises."""
    if status == OK:
        return
    msg = last_error()
    if status in (ERR_INVALID_ARG, ERR_KEY_NOT_FOUND, ERR_NOT_REAL):
        raise ValueError(msg)
    if status == ERR_INDEX_OVERFLOW:
        raise OverflowError(msg)
    if status == ERR_OUT_OF_MEMORY:
        raise MemoryError(msg)
    if status == ERR_NOT_CONVERGED:
        raise NotConvergedError(msg)
    raise EdlgtError(f"status {status}: {msg}")


def ptr(a):
    """Address of a contiguous numpy array, a torch tensor, or None."""
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        if not a.flags.c_contiguous:
            raise ValueError("array must be C-contiguous")
        return a.ctypes.data
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr()
    raise TypeError(type(a))


def stream_handle(stream=None):
    """hipStream_t of a torch stream (None -> the current torch stream)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream
