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
