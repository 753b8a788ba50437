"""The C-ABI library loads and exports every symbol include/edlgt_hip.h declares
(no compute calls: runs without a GPU)."""
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "edlgt_hip.h").read_text()
    return sorted(set(re.findall(r"^int\s+(edlgt_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "edlgt_sector_build" in syms and "edlgt_lanczos" in syms
    assert len(syms) >= 18


def test_library_exports_every_declared_symbol():
    import ctypes
    from edlgt_amd import _lib
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for s in declared_symbols():
        assert hasattr(lib, s), s


def test_binding_covers_header():
    from edlgt_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared_symbols()
    _lib.load()


def test_last_error_roundtrip_without_gpu():
    # argument validation fails before any device call
    import ctypes
    from edlgt_amd import _lib
    lib = _lib.load()
    h = ctypes.c_void_p()
    st = lib.edlgt_sector_build(0, None, 1, 0, 0, None, None, 0, None, None, 0, None, None, None,
                                0, 0, None, None, None, None, None, ctypes.byref(h))
    assert st == _lib.ERR_INVALID_ARG
    assert "null" in _lib.last_error() or "n_sites" in _lib.last_error()


def test_one_hip_runtime_whatever_the_import_order():
    """Loading the engine before torch still maps a single HIP runtime (two
    runtimes in one process leave torch without a device)."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from edlgt_amd import _lib\n_lib.load()\nimport torch\n"
            "print(len({l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}))\n"
            % str(ROOT / "ed-lgt_amd"))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip().splitlines()[-1] == "1"


def test_random_uniform_argument_checks_without_gpu():
    # value type, negative length and null pointers fail before any device call
    import numpy as np
    from edlgt_amd import _lib
    lib = _lib.load()
    state = np.zeros(4, dtype=np.uint64)
    assert lib.edlgt_random_uniform(None, 0, _lib.VALUE_F64, None, None) == _lib.OK
    assert lib.edlgt_random_uniform(None, 5, _lib.VALUE_F64, state.ctypes.data, None) == _lib.ERR_INVALID_ARG
    assert lib.edlgt_random_uniform(None, -1, _lib.VALUE_F64, state.ctypes.data, None) == _lib.ERR_INVALID_ARG
    buf = np.zeros(4)
    assert lib.edlgt_random_uniform(buf.ctypes.data, 4, _lib.VALUE_F32, state.ctypes.data,
                                    None) == _lib.ERR_INVALID_ARG
    assert "value_type" in _lib.last_error()


def test_storage_plan_thresholds():
    """Stored while the CSR takes at most half the free memory, coded while
    the full build fits in 90 % and the 2-byte-coded form in half, else
    matrix-free (host-side rule; no device call)."""
    from edlgt_amd import _lib as L
    from edlgt_amd.device import storage_plan
    n, nnz = 1000, 20000
    full = nnz * 20 + 8 * (n + 1)
    assert storage_plan(n, nnz, L.VALUE_C128, 2 * full) == "stored"
    assert storage_plan(n, nnz, L.VALUE_C128, 2 * full - 2) == "coded"
    assert storage_plan(n, nnz, L.VALUE_C128, int(full / 0.9) + 1) == "coded"
    assert storage_plan(n, nnz, L.VALUE_C128, full) == "free"
    # real values: 12 bytes per entry with values
    assert storage_plan(n, nnz, L.VALUE_F64, 2 * (nnz * 12 + 8 * (n + 1))) == "stored"
