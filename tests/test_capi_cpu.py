"""CPU-side checks of the C ABI: the library loads without a GPU and exports
every entry point include/edb.h declares; argument validation that happens
before any device call."""
import ctypes
import os
import re

import pytest

from entropydecreaseboltzmann_amd import _lib


def declared_symbols():
    txt = open(_lib.HEADER_PATH).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(edb_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert len(syms) >= 30
    assert "edb_qcalc_calQ" in syms and "edb_solver_run" in syms


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_bindings_cover_header():
    assert set(declared_symbols()) == set(_lib._SIGS), set(declared_symbols()) ^ set(_lib._SIGS)


def test_invalid_arguments_rejected_before_device_use():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.edb_qcalc_create(2, 0, 8.0, ctypes.byref(h)) == _lib.EDB_EINVAL
    assert b"N must be" in lib.edb_last_error()
    assert lib.edb_qcalc_create(9, 3, 8.0, ctypes.byref(h)) == _lib.EDB_EINVAL
    assert lib.edb_qcalc_create(9, 0, -1.0, ctypes.byref(h)) == _lib.EDB_EINVAL
    p = _lib.SolverParams()
    p.N, p.batch, p.dt, p.precision, p.scheme = 9, 0, 0.01, 1, 3
    assert lib.edb_solver_create(ctypes.byref(p), None, ctypes.byref(h)) == _lib.EDB_EINVAL
    p.batch, p.scheme = 4, 9
    assert lib.edb_solver_create(ctypes.byref(p), None, ctypes.byref(h)) == _lib.EDB_EINVAL
    p.scheme, p.beta = 4, 0.0  # CFE needs beta > 0
    assert lib.edb_solver_create(ctypes.byref(p), None, ctypes.byref(h)) == _lib.EDB_EINVAL
    assert lib.edb_qcalc_create(65, 0, 8.0, ctypes.byref(h)) == _lib.EDB_EINVAL  # plain paths stop at 64
    p.scheme, p.beta, p.N = 3, 0.0, 40
    p.precision = _lib.PREC_FP32_BATCHED  # the batched kernels stop at N = 33
    assert lib.edb_solver_create(ctypes.byref(p), None, ctypes.byref(h)) == _lib.EDB_EINVAL
    assert b"batched precision" in lib.edb_last_error()
    p.N, p.precision = 9, 4  # no such precision
    assert lib.edb_solver_create(ctypes.byref(p), None, ctypes.byref(h)) == _lib.EDB_EINVAL
    st = _lib.EfixState()
    assert lib.edb_efix_init(ctypes.byref(st), 0, 1e-10) == _lib.EDB_EINVAL
    assert lib.edb_efix_init(ctypes.byref(st), 10, 1e-10) == _lib.EDB_OK
    assert st.average == 1.0 and st.average_flag == 0  # EntropyFix.h:20-22


def test_server_refuses_non_finite_initial_data():
    """A request whose initial data holds a NaN or an infinity is refused at
    submit with EDB_EINVAL, before any device work (a server with no groups
    holds no device memory)."""
    import numpy as np
    import entropydecreaseboltzmann_amd as edb
    srv = edb.Server(L=8.0, dt=0.01)
    for bad in (np.nan, np.inf, -np.inf):
        F0 = np.full((9, 9, 9), 1e-3)
        F0[4, 2, 7] = bad
        with pytest.raises(edb.EdbError) as e:
            srv.submit(1, 9, 0, 1.0, F0)
        assert e.value.status == _lib.EDB_EINVAL
        assert b"not finite" in _lib.load().edb_last_error()
    with pytest.raises(edb.EdbError):
        srv.submit(1, 9, 0, float("nan"), np.full((9, 9, 9), 1e-3))  # t_end


def test_server_refuses_too_many_steps():
    """t_end / dt beyond 1e8 steps is refused at submit with EDB_EINVAL, before
    the host loop that counts the drivers' steps (which would stall once t
    stops changing near 2^53 * dt) and before any device work."""
    import numpy as np
    import entropydecreaseboltzmann_amd as edb
    srv = edb.Server(L=8.0, dt=0.01)
    F0 = np.full((9, 9, 9), 1e-3)
    for t_end in (1.0e7, 1.0e300):
        with pytest.raises(edb.EdbError) as e:
            srv.submit(1, 9, 0, t_end, F0)
        assert e.value.status == _lib.EDB_EINVAL
        assert b"steps" in _lib.load().edb_last_error()
    assert srv.stats()["submitted"] == 0


def test_fft_rejects_bad_lengths():
    lib = _lib.load()
    buf = (ctypes.c_double * 2)()
    assert lib.edb_fft3d_fft(0, 4, 4, buf, buf) == _lib.EDB_EINVAL
    assert lib.edb_fft3d_fft(65, 4, 4, buf, buf) == _lib.EDB_EINVAL


def test_package_imports_without_gpu():
    import entropydecreaseboltzmann_amd as edb
    assert edb.nmodes(24) == 23 and edb.nmodes(25) == 25


def test_cli_usage_without_arguments():
    """testBKW3D checks its arguments before touching the GPU (the reference's
    main reads argv unchecked)."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "entropydecreaseboltzmann_amd", "bin", "testBKW3D")
    assert os.path.exists(exe), "build() builds the command-line drivers"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage: testBKW3D N dt TEnd" in r.stderr
    r = subprocess.run([exe, "2", "0.01", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
