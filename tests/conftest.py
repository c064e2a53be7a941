import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


@pytest.fixture(scope="session")
def golden():
    out = {}
    for name in ("genB", "calQ", "efix"):
        with np.load(os.path.join(GOLDEN, f"golden_{name}.npz"), allow_pickle=False) as z:
            out[name] = {k: z[k] for k in z.files}
    return out


@pytest.fixture(scope="session")
def edb():
    import entropydecreaseboltzmann_amd as m
    m.load()
    if m.device_count() < 1:
        pytest.fail("GPU test run without a visible GPU")
    return m


def rand_spectrum(N, seed, hermitian=False):
    rng = np.random.default_rng(seed)
    if hermitian:
        f = rng.standard_normal((N, N, N))
        return np.fft.fftn(f)
    return rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))


def hermitize(fh):
    """Exact Hermitian part of a spectrum, (f_k + conj f_[-k]) / 2: fftn of real
    data is Hermitian only to rounding, and the Hermitian-half kernel mirrors
    Q_[-k] = conj Q_k exactly, so at modes where |Q| is at the rounding level of
    the large ones the comparison needs the exact symmetry."""
    rev = np.roll(np.flip(fh, axis=(-3, -2, -1)), 1, axis=(-3, -2, -1))
    return 0.5 * (fh + np.conj(rev))


def maxwellian_mixture(N, L, seed, ncomp=None):
    """Mixture of 2-4 isotropic Maxwellians (means U[-2,2]^3, T U[0.5,1.5],
    Dirichlet(1) weights), sampled on x_i = -L + i*2L/N and scaled to mass 1."""
    rng = np.random.default_rng(seed)
    k = ncomp or int(rng.integers(2, 5))
    w = rng.exponential(1.0, k)
    w /= w.sum()
    h = 2 * L / N
    x = -L + h * np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    f = np.zeros((N, N, N))
    for i in range(k):
        u = rng.uniform(-2, 2, 3)
        T = rng.uniform(0.5, 1.5)
        r2 = (X - u[0]) ** 2 + (Y - u[1]) ** 2 + (Z - u[2]) ** 2
        f += w[i] * np.exp(-r2 / (2 * T)) / (2 * np.pi * T) ** 1.5
    return f / (f.sum() * h ** 3)


def bimaxwellian_bump(N, L, seed):
    """0.9 x bi-Maxwellian (T_par along z, T_perp; both U[0.3, 2]) + 0.1 x a
    Maxwellian (T = 0.5) centred at |u| ~ U[2, 3] in a random direction, each
    part normalised to its mass on x_i = -L + i*2L/N (total mass 1)."""
    rng = np.random.default_rng(seed)
    tpar, tperp = rng.uniform(0.3, 2.0, 2)
    r = rng.uniform(2.0, 3.0)
    d = rng.standard_normal(3)
    u = r * d / np.linalg.norm(d)
    h = 2 * L / N
    x = -L + h * np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    bi = np.exp(-(X * X + Y * Y) / (2 * tperp) - Z * Z / (2 * tpar))
    bump = np.exp(-((X - u[0]) ** 2 + (Y - u[1]) ** 2 + (Z - u[2]) ** 2) / (2 * 0.5))
    return 0.9 * bi / (bi.sum() * h ** 3) + 0.1 * bump / (bump.sum() * h ** 3)


def centred_of_fftw(N):
    """Centred mode index of each FFTW-order grid point, flattened (-1 on the
    Nyquist planes of even N)."""
    n, M = (N - 1) // 2, 2 * ((N - 1) // 2) + 1
    w = np.array([(f if f <= n else f - N) for f in range(N)]) + n
    valid = (w >= 0) & (w < M)
    return np.where(valid[:, None, None] & valid[None, :, None] & valid[None, None, :],
                    (w[:, None, None] * M + w[None, :, None]) * M + w[None, None, :], -1).ravel()


class HipMemory:
    """Device buffers for the device-pointer entry points, through the same HIP
    runtime libedb.so links (not torch's bundled copy, whose pointers belong
    to another runtime instance)."""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.bufs = []

    def alloc(self, nbytes):
        p = self.ct.c_void_p()
        if self.hip.hipMalloc(self.ct.byref(p), self.ct.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
        self.bufs.append(p)
        return p

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        p = self.alloc(a.nbytes)
        self.copy_in(p, a)
        return p

    def copy_in(self, p, a):
        a = np.ascontiguousarray(a)
        if self.hip.hipMemcpy(p, a.ctypes.data_as(self.ct.c_void_p), self.ct.c_size_t(a.nbytes), 1) != 0:
            raise RuntimeError("hipMemcpy to the device failed")

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype=dtype)
        if self.hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
        if self.hip.hipMemcpy(out.ctypes.data_as(self.ct.c_void_p), p, self.ct.c_size_t(out.nbytes), 2) != 0:
            raise RuntimeError("hipMemcpy to the host failed")
        return out

    def fill(self, p, nbytes, byte):
        if self.hip.hipMemset(p, int(byte), self.ct.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMemset failed")

    def free(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []
