"""CPU oracle for the CalQ path -- TEST INFRASTRUCTURE ONLY.

ctypes face of ``oracle/liboracle.so`` (a plain-C restatement of the reference,
see edb_oracle.c) and, when it was built in this checkout, of
``oracle/_ref/libedbref.so`` (the reference's own QCalculator.h / EntropyFix.h
compiled from their sources).  Only tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg import this module, and only to check results against it
(or to time the CPU baseline);
the product (entropydecreaseboltzmann_amd) never does.
"""
import contextlib
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libedbref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_o = None


class Efix(ctypes.Structure):
    _fields_ = [("Nt", ctypes.c_long), ("eps", ctypes.c_double),
                ("average", ctypes.c_double), ("average_flag", ctypes.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _o
    if _o is None:
        if not os.path.exists(LIB):
            build()
        _o = _bind(ctypes.CDLL(LIB))
    return _o


@contextlib.contextmanager
def using(path):
    """Route every function of this module through another build of the
    oracle's source (tests that plant a mistake in a copy of it and check that
    the pins catch it)."""
    global _o
    saved = lib()
    _o = _bind(ctypes.CDLL(path))
    try:
        yield
    finally:
        _o = saved


def _bind(o):
    """Declares the argument and result types of the oracle's entry points."""
    D, I = ctypes.c_double, ctypes.c_int
    o.edbo_nmodes.restype = I
    o.edbo_lambda.restype = D
    o.edbo_funF.restype = D
    o.edbo_funF.argtypes = [I, D, D]
    o.edbo_filtering.restype = D
    o.edbo_filtering.argtypes = [I, I]
    o.edbo_rval.restype = D
    o.edbo_rval.argtypes = [I, I, D]
    o.edbo_wrap.restype = I
    o.edbo_wrap.argtypes = [I, I, _ip]
    o.edbo_beta.restype = D
    o.edbo_beta.argtypes = [I, I, D, I, I, I, I, I, I, I]
    o.edbo_genB.argtypes = [I, I, D, I, _dp]
    o.edbo_calQ.argtypes = [I, _dp, _dp, _dp, I]
    o.edbo_calQ_direct.argtypes = [I, I, D, I, _dp, _dp]
    o.edbo_calQ_modes.argtypes = [I, I, D, I, _dp, I, ctypes.POINTER(ctypes.c_long), I, _dp]
    o.edbo_dft3.argtypes = [I, I, _dp, _dp]
    o.edbo_fft.argtypes = [I, _dp, _dp]
    o.edbo_ifft.argtypes = [I, _dp, _dp]
    P = ctypes.POINTER(Efix)
    o.edbo_efix_init.argtypes = [P, ctypes.c_long, D]
    o.edbo_set_average.argtypes = [P, ctypes.c_long, _dp]
    o.edbo_entropy_cal.restype = D
    o.edbo_entropy_cal.argtypes = [P, ctypes.c_long, _dp]
    o.edbo_fixsol.argtypes = [P, ctypes.c_long, _dp, D]
    o.edbo_flogf.restype = D
    o.edbo_flogf.argtypes = [D]
    o.edbo_cal_entropy.restype = D
    o.edbo_cal_entropy.argtypes = [I, D, _dp]
    o.edbo_bkw.restype = D
    o.edbo_bkw.argtypes = [D, D, D, D]
    o.edbo_grid_bkw.argtypes = [I, D, D, _dp]
    o.edbo_run.restype = I
    o.edbo_run.argtypes = [I, _dp, D, I, D, D, D, D, I, _dp, I, _dp, _dp, _dp, I]
    o.edbo_run_ex.restype = I
    o.edbo_run_ex.argtypes = [I, _dp, D, I, D, D, D, D, I, I, _dp, I, _dp, _dp, _dp, I]
    o.edbo_moment_correct.argtypes = [I, D, _dp]
    return o


def _p(a):
    return a.ctypes.data_as(_dp)


def nmodes(N):
    return lib().edbo_nmodes(N)


def genB(N, alpha=0, L=6.62132, filter=True):
    M = nmodes(N)
    B = np.empty(M ** 6)
    lib().edbo_genB(N, alpha, L, int(bool(filter)), _p(B))
    return B.reshape(M ** 3, M ** 3)


def calQ(N, B, fin, nthreads=8):
    f = np.ascontiguousarray(fin, dtype=np.complex128).reshape(N, N, N)
    q = np.empty_like(f)
    lib().edbo_calQ(N, _p(np.ascontiguousarray(B)), _p(f.view(np.float64)), _p(q.view(np.float64)),
                    int(nthreads))
    return q


def calQ_direct(N, fin, alpha=0, L=6.62132, filter=True):
    f = np.ascontiguousarray(fin, dtype=np.complex128).reshape(N, N, N)
    q = np.empty_like(f)
    lib().edbo_calQ_direct(N, alpha, L, int(bool(filter)), _p(f.view(np.float64)),
                           _p(q.view(np.float64)))
    return q


def calQ_modes(N, fin, kcs, alpha=0, L=6.62132, filter=True, absval=False):
    """Q (or, with absval, the sum of |terms|) at centred mode indices kcs,
    beta formed per term: no M^6 array."""
    f = np.ascontiguousarray(fin, dtype=np.complex128).reshape(N, N, N)
    k = np.ascontiguousarray(kcs, dtype=np.int64)
    out = np.empty(k.size, dtype=np.complex128)
    lib().edbo_calQ_modes(N, alpha, L, int(bool(filter)), _p(f.view(np.float64)), int(k.size),
                          k.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), int(bool(absval)),
                          _p(out.view(np.float64)))
    return out


def fft(re):
    a = np.ascontiguousarray(re, dtype=np.complex128)
    N = a.shape[0]
    out = np.empty_like(a)
    lib().edbo_fft(N, _p(a.view(np.float64)), _p(out.view(np.float64)))
    return out


def ifft(fre):
    a = np.ascontiguousarray(fre, dtype=np.complex128)
    N = a.shape[0]
    out = np.empty_like(a)
    lib().edbo_ifft(N, _p(out.view(np.float64)), _p(a.view(np.float64)))
    return out


class EntropyFix:
    def __init__(self, Nt, eps=1.0e-10):
        self.e = Efix()
        lib().edbo_efix_init(ctypes.byref(self.e), Nt, eps)

    def setAverage(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        lib().edbo_set_average(ctypes.byref(self.e), f.size, _p(f))

    def EntropyCal(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        return lib().edbo_entropy_cal(ctypes.byref(self.e), f.size, _p(f))

    def FixSol(self, f, aim):
        f = np.array(f, dtype=np.float64, copy=True).ravel()
        lib().edbo_fixsol(ctypes.byref(self.e), f.size, _p(f), aim)
        return f


def cal_entropy(N, R, F):
    Fc = np.ascontiguousarray(F, dtype=np.complex128)
    return lib().edbo_cal_entropy(N, R, _p(Fc.view(np.float64)))


def grid_bkw(N, R, K):
    F = np.empty((N, N, N), dtype=np.complex128)
    lib().edbo_grid_bkw(N, R, K, _p(F.view(np.float64)))
    return F.real.copy()


SCHEMES = {"FEuler": 0, "RK2": 1, "SSP3": 2, "RK4": 3, "CFE": 4, "DUFORT": 5}


def moment_correct(N, L, qhat):
    """Conservative moment correction of a spectrum (FFTW order); returns a copy."""
    q = np.array(qhat, dtype=np.complex128, copy=True).reshape(N, N, N)
    lib().edbo_moment_correct(N, L, _p(q.view(np.float64)))
    return q


def run(N, B, R, scheme, T0, TEnd, dt, fhat, beta=0.0, fix_mode=0, max_steps=100000, nthreads=8,
        conserve=False):
    """BKW3D driver restated on a spectral state; returns (fhat, t, H, Hfix)."""
    f = np.array(fhat, dtype=np.complex128, copy=True).reshape(N, N, N)
    t = np.zeros(max_steps)
    H = np.zeros(max_steps)
    Hf = np.zeros(max_steps)
    sc = SCHEMES[scheme] if isinstance(scheme, str) else scheme
    n = lib().edbo_run_ex(N, _p(np.ascontiguousarray(B)), R, sc, T0, TEnd, dt, beta, fix_mode,
                          int(bool(conserve)), _p(f.view(np.float64)), max_steps, _p(t), _p(H),
                          _p(Hf), nthreads)
    k = min(n, max_steps)
    return f, t[:k], H[:k], Hf[:k]


# --------------------------------------------------------------- reference

_r = None


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    """The reference's own headers compiled into oracle/_ref (may be absent)."""
    global _r
    if _r is None:
        r = ctypes.CDLL(REF_LIB)
        I = ctypes.c_int
        r.edbr_genB.argtypes = [I, I, I, _dp]
        r.edbr_calQ.argtypes = [I, I, I, _dp, _dp]
        r.edbr_calQ_reps.argtypes = [I, I, I, _dp, _dp, I]
        r.edbr_entropy_sequence.argtypes = [ctypes.c_long, _dp, _dp, _dp]
        r.edbr_fixsol_ptr.argtypes = [I, _dp, ctypes.c_double, _dp]
        _r = r
    return _r
