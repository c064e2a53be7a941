"""ctypes binding of libedb.so (the C ABI declared in include/edb.h).

The library is built in-tree by ``entropydecreaseboltzmann_amd/csrc/Makefile``
(``__graft_entry__.build()``).  Loading fails loudly when it is missing: there
is no CPU fallback for any entry point.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libedb.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "edb.h")

EDB_OK = 0
EDB_EINVAL = 1
EDB_ENOMEM = 2
EDB_EHIP = 3
EDB_ENOTREADY = 4
EDB_EUNSUPPORTED = 5
EDB_EBUSY = 6

PREC_FP64_PLAIN = 0
PREC_FP32_BATCHED = 1
PREC_FP32_PLAIN = 2
PREC_FP64_COMPRESSED = 3

SCHEMES = {"FEuler": 0, "RK2": 1, "SSP3": 2, "RK4": 3, "CFE": 4, "DUFORT": 5}
FIX_NONE, FIX_DIAGNOSTIC, FIX_APPLIED = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_vp = ctypes.c_void_p


class EfixState(ctypes.Structure):
    _fields_ = [("Nt", ctypes.c_long), ("eps", ctypes.c_double),
                ("average", ctypes.c_double), ("average_flag", ctypes.c_int)]


class SolverParams(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("alpha", ctypes.c_int), ("L", ctypes.c_double),
                ("filter", ctypes.c_int), ("batch", ctypes.c_int),
                ("precision", ctypes.c_int), ("scheme", ctypes.c_int),
                ("dt", ctypes.c_double), ("beta", ctypes.c_double),
                ("fix_mode", ctypes.c_int), ("hermitian", ctypes.c_int),
                ("history", ctypes.c_int), ("conserve", ctypes.c_int)]


class ServerParams(ctypes.Structure):
    _fields_ = [("L", ctypes.c_double), ("filter", ctypes.c_int), ("dt", ctypes.c_double),
                ("fix_mode", ctypes.c_int), ("max_slots", ctypes.c_int), ("max_chunk", ctypes.c_int),
                ("hbm_budget", ctypes.c_long), ("conserve", ctypes.c_int),
                ("max_pending", ctypes.c_int)]


_SIGS = {
    "edb_last_error": (ctypes.c_char_p, []),
    "edb_version": (ctypes.c_int, [_ip, _ip]),
    "edb_device_count": (ctypes.c_int, [_ip]),
    "edb_sync": (ctypes.c_int, []),
    "edb_qcalc_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.POINTER(_vp)]),
    "edb_qcalc_close_filter": (ctypes.c_int, [_vp]),
    "edb_qcalc_genB": (ctypes.c_int, [_vp]),
    "edb_qcalc_genB_batched": (ctypes.c_int, [_vp]),
    "edb_qcalc_calQ": (ctypes.c_int, [_vp, _dp, _dp]),
    "edb_qcalc_calQ_device": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "edb_qcalc_calQ_batched": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "edb_qcalc_calQ_batched_device": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                     _vp, _vp, _vp]),
    "edb_qcalc_calQ_batched_device_noalloc": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                             _vp, _vp, _vp]),
    "edb_qcalc_copy_B": (ctypes.c_int, [_vp, _dp]),
    "edb_qcalc_tune_plain": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "edb_qcalc_plain_schedule": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip]),
    "edb_qcalc_info": (ctypes.c_int, [_vp, _ip, _ip, _lp, _lp]),
    "edb_qcalc_time_batched": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    "edb_qcalc_time_plain": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "edb_qcalc_tune": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "edb_qcalc_autotune": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip]),
    "edb_qcalc_set_conservative": (ctypes.c_int, [_vp, ctypes.c_int]),
    "edb_qcalc_genB_plain32": (ctypes.c_int, [_vp]),
    "edb_qcalc_tune_batched": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "edb_qcalc_batched_schedule": (ctypes.c_int, [_vp, _ip, _ip]),
    "edb_qcalc_calQ_plain32": (ctypes.c_int, [_vp, _dp, _dp]),
    "edb_qcalc_genB_compressed": (ctypes.c_int, [_vp]),
    "edb_qcalc_compressed_bytes": (ctypes.c_int, [_vp, _lp]),
    "edb_qcalc_calQ_compressed": (ctypes.c_int, [_vp, _dp, _dp]),
    "edb_qcalc_calQ_compressed_device": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "edb_qcalc_time_compressed": (ctypes.c_int, [_vp, ctypes.c_int, _dp]),
    "edb_solver_mean_history": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    "edb_solver_get_driver_real": (ctypes.c_int, [_vp, _dp]),
    "edb_qcalc_destroy": (ctypes.c_int, [_vp]),
    "edb_fft3d_fft": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "edb_fft3d_ifft": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]),
    "edb_fft3d_batched_device": (ctypes.c_int, [ctypes.c_int] * 6 + [_vp, _vp, _vp]),
    "edb_efix_init": (ctypes.c_int, [ctypes.POINTER(EfixState), ctypes.c_long, ctypes.c_double]),
    "edb_efix_set_average": (ctypes.c_int, [ctypes.POINTER(EfixState), ctypes.c_long, _dp]),
    "edb_efix_entropy_cal": (ctypes.c_int, [ctypes.POINTER(EfixState), ctypes.c_long, _dp, _dp]),
    "edb_efix_fixsol": (ctypes.c_int, [ctypes.POINTER(EfixState), ctypes.c_long, _dp, ctypes.c_double]),
    "edb_solver_create": (ctypes.c_int, [ctypes.POINTER(SolverParams), _vp, ctypes.POINTER(_vp)]),
    "edb_solver_set_initial_real": (ctypes.c_int, [_vp, ctypes.c_double, _dp]),
    "edb_solver_set_initial_spectral": (ctypes.c_int, [_vp, ctypes.c_double, _dp]),
    "edb_solver_step": (ctypes.c_int, [_vp, ctypes.c_int]),
    "edb_solver_run": (ctypes.c_int, [_vp, ctypes.c_double, _ip]),
    "edb_solver_use_graph": (ctypes.c_int, [_vp, ctypes.c_int]),
    "edb_solver_qkernel_time": (ctypes.c_int, [_vp, _dp, _lp]),
    "edb_solver_enable_qtiming": (ctypes.c_int, [_vp, ctypes.c_int]),
    "edb_solver_time": (ctypes.c_int, [_vp, _dp, _ip]),
    "edb_solver_get_spectral": (ctypes.c_int, [_vp, _dp]),
    "edb_solver_get_real": (ctypes.c_int, [_vp, _dp]),
    "edb_solver_entropy_history": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, _ip]),
    "edb_solver_moments": (ctypes.c_int, [_vp, _dp]),
    "edb_solver_bkw_error": (ctypes.c_int, [_vp, ctypes.c_double, _dp]),
    "edb_solver_destroy": (ctypes.c_int, [_vp]),
    "edb_server_create": (ctypes.c_int, [ctypes.POINTER(ServerParams), ctypes.POINTER(_vp)]),
    "edb_server_submit": (ctypes.c_int, [_vp, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]),
    "edb_server_poll": (ctypes.c_int, [_vp, _ip]),
    "edb_server_fetch": (ctypes.c_int, [_vp, _lp, _ip, _ip, _dp, _dp, ctypes.c_long]),
    "edb_server_stats": (ctypes.c_int, [_vp, _lp, _lp, _lp, _lp, _lp]),
    "edb_server_destroy": (ctypes.c_int, [_vp]),
}

_lib = None


def load():
    """Load libedb.so and set every signature; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class EdbError(RuntimeError):
    def __init__(self, status, where):
        msg = load().edb_last_error().decode(errors="replace")
        super().__init__(f"{where}: status {status}: {msg}")
        self.status = status


def check(status, where):
    if status != EDB_OK:
        raise EdbError(status, where)
