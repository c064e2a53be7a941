"""The oracle's conservative moment correction, pinned by what the mathematics
fixes: the corrected spectrum has zero discrete mass, momentum and energy
moments computed in physical space (numpy's inverse FFT and grid sums, not
the correction's own Fourier weights); for odd N it equals the textbook
constrained-L2 projection of the grid values, Q - C^T (C C^T)^-1 C Q; for even
N the change lies in the span of the band-limited moment functions and leaves
the Nyquist planes at zero; it is idempotent; and an RK4 trajectory through
the oracle driver with the correction conserves energy where the plain one
drifts."""
import numpy as np
import pytest

import oracle
from conftest import maxwellian_mixture, rand_spectrum


def grid(N, L):
    h = 2 * L / N
    x = -L + h * np.arange(N)
    return h, np.meshgrid(x, x, x, indexing="ij")


def moment_functions(N, L):
    h, (X, Y, Z) = grid(N, L)
    return h, [np.ones_like(X), X, Y, Z, X * X + Y * Y + Z * Z]


def moments(N, L, qhat):
    """h^3 sum_x phi(x) Q(x) with Q = ifftn(qhat) (complex for general input)."""
    h, phis = moment_functions(N, L)
    g = np.fft.ifftn(qhat.reshape(N, N, N))
    return np.array([(g * p).sum() * h ** 3 for p in phis])


def galerkin(N, qhat):
    q = qhat.reshape(N, N, N).copy()
    if N % 2 == 0:
        q[N // 2, :, :] = 0
        q[:, N // 2, :] = 0
        q[:, :, N // 2] = 0
    return q


def moment_scale(N, L, qhat):
    """Size of the moments' terms: h^3 sum_x |phi(x)| |Q(x)|."""
    h, phis = moment_functions(N, L)
    g = np.abs(np.fft.ifftn(qhat.reshape(N, N, N)))
    return max((g * np.abs(p)).sum() * h ** 3 for p in phis)


@pytest.mark.parametrize("N,L", [(5, 8.0), (8, 8.0), (9, 6.62132), (12, 8.0), (16, 8.0)])
@pytest.mark.parametrize("hermitian", [True, False])
def test_corrected_moments_vanish(N, L, hermitian):
    q = galerkin(N, rand_spectrum(N, 3 * N, hermitian=hermitian))
    q[0, 0, 0] = 0  # as every collision spectrum
    assert np.abs(moments(N, L, q)).max() > 1e-3 * moment_scale(N, L, q)
    qc = oracle.moment_correct(N, L, q)
    assert np.abs(moments(N, L, qc)).max() <= 1e-14 * moment_scale(N, L, q)


@pytest.mark.parametrize("N", [3, 5, 7, 9])
def test_odd_N_equals_grid_projection(N):
    L = 8.0
    h, phis = moment_functions(N, L)
    q = rand_spectrum(N, 50 + N, hermitian=True)
    Qx = np.fft.ifftn(q).reshape(-1)
    C = np.stack([p.reshape(-1) for p in phis])
    proj = Qx - C.T @ np.linalg.solve(C @ C.T, C @ Qx)
    ref = np.fft.fftn(proj.reshape(N, N, N))
    got = oracle.moment_correct(N, L, q)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(q).max()


@pytest.mark.parametrize("N", [4, 8, 10])
def test_even_N_change_in_band_limited_span(N):
    L = 8.0
    h, phis = moment_functions(N, L)
    q = galerkin(N, rand_spectrum(N, 70 + N, hermitian=True))
    got = oracle.moment_correct(N, L, q)
    delta = (q - got).reshape(-1)
    # the Nyquist planes stay outside the Galerkin set
    assert np.all(galerkin(N, got) == got)
    P = np.stack([galerkin(N, np.fft.fftn(p)).reshape(-1) for p in phis], axis=1)
    coef, *_ = np.linalg.lstsq(P, delta, rcond=None)
    assert np.abs(P @ coef - delta).max() <= 1e-12 * np.abs(delta).max()


def test_idempotent_and_identity_on_conserving_input():
    N, L = 9, 8.0
    q = rand_spectrum(N, 5, hermitian=True)
    once = oracle.moment_correct(N, L, q)
    twice = oracle.moment_correct(N, L, once)
    assert np.abs(twice - once).max() <= 1e-13 * np.abs(q).max()
    assert not np.allclose(once, q)


@pytest.mark.slow
def test_rk4_with_correction_conserves_energy():
    N, L, dt = 9, 8.0, 0.05
    F0 = maxwellian_mixture(N, L, 11)
    B = oracle.genB(N, 0, L, 0)
    h, phis = moment_functions(N, L)

    def mom(fhat):
        return moments(N, L, fhat).real

    m0 = mom(np.fft.fftn(F0))
    drift = {}
    for conserve in (False, True):
        fo, *_ = oracle.run(N, B, L, "RK4", 0.0, 0.5 - 1e-12, dt, np.fft.fftn(F0), fix_mode=0,
                            conserve=conserve)
        m1 = mom(fo)
        drift[conserve] = np.abs(m1 - m0) / np.abs(m0[[0, 0, 0, 0, 4]]).max()
    assert drift[True].max() <= 1e-13
    assert drift[False][4] > 1e3 * max(drift[True][4], 1e-16)  # the uncorrected energy drifts
