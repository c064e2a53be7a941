"""MI355X-native CalQ collision-operator path (Fourier-Galerkin Boltzmann).

Python face of libedb.so (include/edb.h) with the reference's names:

* ``QCalculator(N, alpha, L)`` -- ``genB``, ``calQ``, ``closeFilter``
  (CalQ/QCalculator.h:17-62), plus ``genB_batched`` / ``calQ_batched`` for the
  FP32 batched kernel.
* ``fft3d.fft`` / ``fft3d.ifft`` -- CalQ/fft3d.h:25-48.
* ``EntropyFix(Nt, eps)`` -- ``EntropyCal``, ``setAverage``, ``FixSol``
  (CalQ/EntropyFix.h:13-46).
* ``Solver`` -- the BKW3D drivers (CalQ/BKW.h:22-633) over a batch.

Every call runs on the GPU through the C ABI; importing works without a GPU,
calling does not.  Arrays are numpy complex128 in FFTW order (spectral) or
float64 (real space), shaped ``(N, N, N)`` or ``(batch, N, N, N)``.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (FIX_APPLIED, FIX_DIAGNOSTIC, FIX_NONE, PREC_FP32_BATCHED, PREC_FP32_PLAIN,
                   PREC_FP64_COMPRESSED, PREC_FP64_PLAIN, SCHEMES, EdbError, check, load)

__all__ = ["QCalculator", "fft3d", "EntropyFix", "Solver", "Server", "EdbError", "load",
           "device_count", "nmodes", "SCHEMES", "FIX_NONE", "FIX_DIAGNOSTIC",
           "FIX_APPLIED", "PREC_FP64_PLAIN", "PREC_FP32_BATCHED", "PREC_FP32_PLAIN",
           "PREC_FP64_COMPRESSED"]

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _c128(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    if shape is not None and a.shape != tuple(shape):
        a = a.reshape(shape)
    return a


def device_count():
    n = ctypes.c_int(0)
    check(load().edb_device_count(ctypes.byref(n)), "edb_device_count")
    return n.value


def nmodes(N):
    """Galerkin modes per dimension: 2n+1 with n = (N-1)//2."""
    return 2 * ((N - 1) // 2) + 1


class QCalculator:
    """QCalculator<c_type, r_type> (QCalculator.h:17-62) on the GPU."""

    def __init__(self, N, alpha=0, L=6.62132):
        self.N, self.alpha, self.L = int(N), int(alpha), float(L)
        self._h = ctypes.c_void_p()
        check(load().edb_qcalc_create(self.N, self.alpha, self.L, ctypes.byref(self._h)),
              "edb_qcalc_create")
        self.M = nmodes(self.N)

    def closeFilter(self):
        check(load().edb_qcalc_close_filter(self._h), "closeFilter")

    def genB(self):
        check(load().edb_qcalc_genB(self._h), "genB")

    def genB_batched(self):
        check(load().edb_qcalc_genB_batched(self._h), "genB_batched")

    def calQ(self, fin):
        N = self.N
        f = _c128(fin, (N, N, N))
        q = np.empty_like(f)
        check(load().edb_qcalc_calQ(self._h, _ptr(q.view(np.float64)), _ptr(f.view(np.float64))),
              "calQ")
        return q

    def genB_plain32(self):
        check(load().edb_qcalc_genB_plain32(self._h), "genB_plain32")

    def calQ_plain32(self, fin):
        """calQ with the FP32 copy of beta (FP64 products and sums)."""
        N = self.N
        f = _c128(fin, (N, N, N))
        q = np.empty_like(f)
        check(load().edb_qcalc_calQ_plain32(self._h, _ptr(q.view(np.float64)), _ptr(f.view(np.float64))),
              "calQ_plain32")
        return q

    def genB_compressed(self):
        """Compressed beta (F tabulated over |l+m|^2, |l-m|^2, plus per-mode
        sigma and Bmm) for the plain FP64 path; no M^6 array."""
        check(load().edb_qcalc_genB_compressed(self._h), "genB_compressed")

    def compressed_bytes(self):
        b = ctypes.c_long()
        check(load().edb_qcalc_compressed_bytes(self._h, ctypes.byref(b)), "compressed_bytes")
        return b.value

    def calQ_compressed(self, fin):
        """calQ with beta formed on the fly from the compressed tables: every
        beta equals genB's; the sum runs in another fixed order, so the result
        equals calQ to FP64 rounding."""
        N = self.N
        f = _c128(fin, (N, N, N))
        q = np.empty_like(f)
        check(load().edb_qcalc_calQ_compressed(self._h, _ptr(q.view(np.float64)), _ptr(f.view(np.float64))),
              "calQ_compressed")
        return q

    def time_compressed(self, reps=5):
        ms = ctypes.c_double()
        check(load().edb_qcalc_time_compressed(self._h, int(reps), ctypes.byref(ms)), "time_compressed")
        return ms.value

    def calQ_batched(self, fin, hermitian=False):
        N = self.N
        f = _c128(fin)
        b = f.size // (N ** 3)
        f = f.reshape(b, N, N, N)
        q = np.empty_like(f)
        check(load().edb_qcalc_calQ_batched(self._h, b, int(bool(hermitian)),
                                            _ptr(q.view(np.float64)), _ptr(f.view(np.float64))),
              "calQ_batched")
        return q

    def copy_B(self):
        M = self.M
        B = np.empty(M ** 6, dtype=np.float64)
        check(load().edb_qcalc_copy_B(self._h, _ptr(B)), "copy_B")
        return B.reshape(M ** 3, M ** 3)

    def info(self):
        N, M = ctypes.c_int(), ctypes.c_int()
        b64, b32 = ctypes.c_long(), ctypes.c_long()
        check(load().edb_qcalc_info(self._h, ctypes.byref(N), ctypes.byref(M), ctypes.byref(b64),
                                    ctypes.byref(b32)), "info")
        return {"N": N.value, "M": M.value, "bytes_plain64": b64.value, "bytes_pair32": b32.value}

    def time_batched(self, batch, hermitian=True, reps=5):
        ms = ctypes.c_double()
        check(load().edb_qcalc_time_batched(self._h, int(batch), int(bool(hermitian)), int(reps),
                                            ctypes.byref(ms)), "time_batched")
        return ms.value

    def time_plain(self, reps=5):
        ms = ctypes.c_double()
        check(load().edb_qcalc_time_plain(self._h, int(reps), ctypes.byref(ms)), "time_plain")
        return ms.value

    def tune(self, batched_waves=0, plain_slabs=0):
        check(load().edb_qcalc_tune(self._h, int(batched_waves), int(plain_slabs)), "tune")

    def autotune(self, batch=256):
        bw, ps = ctypes.c_int(), ctypes.c_int()
        check(load().edb_qcalc_autotune(self._h, int(batch), ctypes.byref(bw), ctypes.byref(ps)),
              "autotune")
        return {"batched_waves": bw.value, "plain_slabs": ps.value}

    def tune_plain(self, kernel=-1, slabs=0, fp32=False):
        """Stored-beta stream variant (0: per-term gather of f_[k-l]; 1-3: f_[k-l]
        from an LDS window, beta in groups of 4/2/8; -1 keeps it) and l-slab
        count (0 keeps it) of the plain kernel."""
        check(load().edb_qcalc_tune_plain(self._h, int(bool(fp32)), int(kernel), int(slabs)), "tune_plain")

    def plain_schedule(self, fp32=False):
        k, n = ctypes.c_int(), ctypes.c_int()
        check(load().edb_qcalc_plain_schedule(self._h, int(bool(fp32)), ctypes.byref(k), ctypes.byref(n)),
              "plain_schedule")
        return k.value, n.value

    def tune_batched(self, schedule, count=0):
        """schedule 0: stream-K rows (count waves); 1: quads (count workgroups)."""
        check(load().edb_qcalc_tune_batched(self._h, int(schedule), int(count)), "tune_batched")

    def batched_schedule(self):
        sc, n = ctypes.c_int(), ctypes.c_int()
        check(load().edb_qcalc_batched_schedule(self._h, ctypes.byref(sc), ctypes.byref(n)),
              "batched_schedule")
        return sc.value, n.value

    def set_conservative(self, on=True):
        """Moment correction of every Q this calculator returns (zero discrete
        mass, momentum and energy moments); no reference counterpart."""
        check(load().edb_qcalc_set_conservative(self._h, int(bool(on))), "set_conservative")

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            load().edb_qcalc_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class fft3d:
    """fft3d (fft3d.h:18-48): FFTW_FORWARD unnormalised / FFTW_BACKWARD x 1/N."""

    @staticmethod
    def fft(re):
        a = _c128(re)
        Nx, Ny, Nz = a.shape
        out = np.empty_like(a)
        check(load().edb_fft3d_fft(Nx, Ny, Nz, _ptr(a.view(np.float64)), _ptr(out.view(np.float64))),
              "fft3d.fft")
        return out

    @staticmethod
    def ifft(fre):
        a = _c128(fre)
        Nx, Ny, Nz = a.shape
        out = np.empty_like(a)
        check(load().edb_fft3d_ifft(Nx, Ny, Nz, _ptr(out.view(np.float64)), _ptr(a.view(np.float64))),
              "fft3d.ifft")
        return out


class EntropyFix:
    """EntropyFix<i_type, r_type> (EntropyFix.h:13-46); the state lives on the host."""

    def __init__(self, Nt, eps=1.0e-10):
        self.st = _lib.EfixState()
        check(load().edb_efix_init(ctypes.byref(self.st), int(Nt), float(eps)), "EntropyFix")

    @property
    def Averagef(self):
        return self.st.average

    def setAverage(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).ravel()
        check(load().edb_efix_set_average(ctypes.byref(self.st), f.size, _ptr(f)), "setAverage")

    def EntropyCal(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).ravel()
        H = ctypes.c_double()
        check(load().edb_efix_entropy_cal(ctypes.byref(self.st), f.size, _ptr(f), ctypes.byref(H)),
              "EntropyCal")
        return H.value

    def FixSol(self, f, aim):
        """Returns the fixed copy of f (the reference overwrites its argument)."""
        f = np.array(f, dtype=np.float64, copy=True).ravel()
        check(load().edb_efix_fixsol(ctypes.byref(self.st), f.size, _ptr(f), float(aim)), "FixSol")
        return f


class Solver:
    """The BKW3D explicit drivers (BKW.h:191-633) over a batch of instances."""

    def __init__(self, N, alpha=0, L=8.0, filter=False, batch=1, precision=PREC_FP64_PLAIN,
                 scheme="RK4", dt=0.01, beta=0.0, fix_mode=FIX_NONE, hermitian=True, history=0,
                 qcalc=None, conserve=False):
        p = _lib.SolverParams()
        p.N, p.alpha, p.L, p.filter = int(N), int(alpha), float(L), int(bool(filter))
        p.batch, p.precision = int(batch), int(precision)
        p.scheme = SCHEMES[scheme] if isinstance(scheme, str) else int(scheme)
        p.dt, p.beta, p.fix_mode = float(dt), float(beta), int(fix_mode)
        p.hermitian, p.history = int(bool(hermitian)), int(history)
        p.conserve = int(bool(conserve))
        self.p = p
        self.N, self.batch = p.N, p.batch
        self._h = ctypes.c_void_p()
        shared = qcalc.handle if qcalc is not None else None
        check(load().edb_solver_create(ctypes.byref(p), shared, ctypes.byref(self._h)),
              "edb_solver_create")

    def set_initial_real(self, F, T0=0.0):
        F = np.ascontiguousarray(F, dtype=np.float64).reshape(self.batch, -1)
        check(load().edb_solver_set_initial_real(self._h, float(T0), _ptr(F)), "set_initial_real")

    def set_initial_spectral(self, Fhat, T0=0.0):
        F = _c128(Fhat).reshape(self.batch, -1)
        check(load().edb_solver_set_initial_spectral(self._h, float(T0), _ptr(F.view(np.float64))),
              "set_initial_spectral")

    def step(self, n=1):
        check(load().edb_solver_step(self._h, int(n)), "step")

    def run(self, TEnd):
        n = ctypes.c_int()
        check(load().edb_solver_run(self._h, float(TEnd), ctypes.byref(n)), "run")
        return n.value

    def use_graph(self, enable=True):
        check(load().edb_solver_use_graph(self._h, int(bool(enable))), "use_graph")

    def enable_qtiming(self, enable=True):
        check(load().edb_solver_enable_qtiming(self._h, int(bool(enable))), "enable_qtiming")

    def qkernel_time(self):
        ms, n = ctypes.c_double(), ctypes.c_long()
        check(load().edb_solver_qkernel_time(self._h, ctypes.byref(ms), ctypes.byref(n)),
              "qkernel_time")
        return ms.value, n.value

    def time(self):
        t, n = ctypes.c_double(), ctypes.c_int()
        check(load().edb_solver_time(self._h, ctypes.byref(t), ctypes.byref(n)), "time")
        return t.value, n.value

    def get_spectral(self):
        N = self.N
        out = np.empty((self.batch, N, N, N), dtype=np.complex128)
        check(load().edb_solver_get_spectral(self._h, _ptr(out.view(np.float64))), "get_spectral")
        return out

    def get_real(self):
        N = self.N
        out = np.empty((self.batch, N, N, N), dtype=np.float64)
        check(load().edb_solver_get_real(self._h, _ptr(out)), "get_real")
        return out

    def entropy_history(self, steps):
        H = np.empty((steps, self.batch))
        Hf = np.empty((steps, self.batch))
        fx = np.empty((steps, self.batch), dtype=np.int32)
        check(load().edb_solver_entropy_history(self._h, int(steps), _ptr(H), _ptr(Hf),
                                                fx.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
              "entropy_history")
        return H, Hf, fx

    def mean_history(self, steps):
        m = np.empty((steps, self.batch))
        mf = np.empty((steps, self.batch))
        check(load().edb_solver_mean_history(self._h, int(steps), _ptr(m), _ptr(mf)), "mean_history")
        return m, mf

    def get_driver_real(self):
        """The driver's real-space F after the last step (fixed values where the fix ran)."""
        N = self.N
        out = np.empty((self.batch, N, N, N), dtype=np.float64)
        check(load().edb_solver_get_driver_real(self._h, _ptr(out)), "get_driver_real")
        return out

    def moments(self):
        out = np.empty((self.batch, 5))
        check(load().edb_solver_moments(self._h, _ptr(out)), "moments")
        return out

    def bkw_error(self, K):
        out = np.empty((self.batch, 2))
        check(load().edb_solver_bkw_error(self._h, float(K), _ptr(out)), "bkw_error")
        return out

    def close(self):
        if self._h:
            load().edb_solver_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous batching of independent RK4 solves (include/edb.h, edb_server_*)."""

    def __init__(self, L=8.0, filter=False, dt=0.01, fix_mode=FIX_APPLIED, max_slots=0, max_chunk=0,
                 hbm_budget=0, conserve=False, max_pending=0):
        p = _lib.ServerParams()
        p.L, p.filter, p.dt, p.fix_mode = float(L), int(bool(filter)), float(dt), int(fix_mode)
        p.max_slots, p.max_chunk, p.hbm_budget = int(max_slots), int(max_chunk), int(hbm_budget)
        p.conserve = int(bool(conserve))
        p.max_pending = int(max_pending)
        self.p = p
        self._h = ctypes.c_void_p()
        check(load().edb_server_create(ctypes.byref(p), ctypes.byref(self._h)), "edb_server_create")

    def submit(self, rid, N, alpha, t_end, F0):
        """Returns the status (0 ok, EDB_ENOMEM never fits, EDB_EBUSY not now)."""
        F = np.ascontiguousarray(F0, dtype=np.float64).ravel()
        if F.size != N ** 3:
            raise ValueError("F0 must have N^3 values")
        st = load().edb_server_submit(self._h, int(rid), int(N), int(alpha), float(t_end), _ptr(F))
        if st not in (_lib.EDB_OK, _lib.EDB_ENOMEM, _lib.EDB_EBUSY):
            check(st, "server.submit")
        return st

    def poll(self):
        n = ctypes.c_int()
        check(load().edb_server_poll(self._h, ctypes.byref(n)), "server.poll")
        return n.value

    def fetch(self, max_N=33):
        rid, N, steps = ctypes.c_long(), ctypes.c_int(), ctypes.c_int()
        t = ctypes.c_double()
        buf = np.empty(max_N ** 3)
        check(load().edb_server_fetch(self._h, ctypes.byref(rid), ctypes.byref(N), ctypes.byref(steps),
                                      ctypes.byref(t), _ptr(buf), buf.size), "server.fetch")
        n = N.value
        return rid.value, n, steps.value, t.value, buf[: n ** 3].reshape(n, n, n).copy()

    def stats(self):
        v = [ctypes.c_long() for _ in range(5)]
        check(load().edb_server_stats(self._h, *[ctypes.byref(x) for x in v]), "server.stats")
        return dict(zip(["submitted", "finished", "active", "waiting", "hbm_used"], [x.value for x in v]))

    def close(self):
        if self._h:
            load().edb_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
