"""Compressed beta (edb_qcalc_genB_compressed / calQ_compressed): beta formed
on the fly from F tabulated over (|l+m|^2, |l-m|^2) and the per-mode (sigma,
Bmm) pairs, in place of the M^6 array of QCalculator::genB
(QCalculator.h:158-229).

* where both fit, the result equals the stored-beta FP64 kernel to FP64
  rounding (every beta is the same double; the sums run in another order),
  with the modes of a wave taken as consecutive k3 (up to N = 28) or from one
  symmetry orbit (from N = 29);
* at N = 48 and 64, where the FP64 M^6 array needs 86 GB and 500 GB, selected
  modes are checked against the oracle's per-mode sum (beta from the
  reference's per-entry formula, no M^6 array) to 1e-12 of the summed |terms|,
  Q_0 is exactly 0 and Q is Hermitian for real data."""
import numpy as np
import pytest

import oracle
from conftest import maxwellian_mixture, rand_spectrum

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,alpha,fil", [(9, 0, 0), (9, 1, 1), (10, 0, 1), (16, 1, 0), (24, 0, 0),
                                         (25, 1, 1), (29, 1, 1), (32, 0, 0)])
def test_compressed_matches_stored(edb, N, alpha, fil):
    q = edb.QCalculator(N, alpha, 8.0)
    if not fil:
        q.closeFilter()
    q.genB()
    q.genB_compressed()
    for seed in (1, 2):
        f = rand_spectrum(N, 100 * N + seed)
        qs = q.calQ(f)
        qc = q.calQ_compressed(f)
        assert np.abs(qc - qs).max() <= 1e-12 * np.abs(qs).max()
        # fixed summation order: a second evaluation is bitwise identical
        np.testing.assert_array_equal(q.calQ_compressed(f), qc)
        # modes without a term (the Nyquist planes of even N, k = 0) are exactly 0 in both
        np.testing.assert_array_equal(qc == 0, qs == 0)
    if N >= 16:
        assert q.compressed_bytes() < q.info()["bytes_plain64"] / 50


def _centred_to_fftw(N, kcs):
    n, M = (N - 1) // 2, oracle.nmodes(N)
    c1, c2, c3 = kcs // (M * M), (kcs // M) % M, kcs % M
    return (((c1 - n) % N) * N + (c2 - n) % N) * N + (c3 - n) % N


def _negate(N, kcs):
    M = oracle.nmodes(N)
    c1, c2, c3 = kcs // (M * M), (kcs // M) % M, kcs % M
    return ((M - 1 - c1) * M + (M - 1 - c2)) * M + (M - 1 - c3)


@pytest.mark.parametrize("N,alpha,fil", [(48, 0, 0), (48, 1, 1), (64, 0, 0), (64, 1, 0)])
def test_compressed_large_grid_matches_oracle_modes(edb, N, alpha, fil):
    L = 8.0
    M = oracle.nmodes(N)
    q = edb.QCalculator(N, alpha, L)
    if not fil:
        q.closeFilter()
    assert q.info()["bytes_plain64"] > 80e9  # the stored FP64 array would not be practical
    q.genB_compressed()
    f = np.fft.fftn(maxwellian_mixture(N, L, 7 + N))
    Q = q.calQ_compressed(f).ravel()
    rng = np.random.default_rng(N + alpha)
    kcs = np.concatenate([[M ** 3 // 2 + 1], rng.choice(M ** 3, 5, replace=False)])
    kcs = np.concatenate([kcs, _negate(N, kcs)])
    ref = oracle.calQ_modes(N, f, kcs, alpha, L, fil)
    scale = oracle.calQ_modes(N, f, kcs, alpha, L, fil, absval=True).real
    got = Q[_centred_to_fftw(N, kcs)]
    assert np.all(np.abs(got - ref) <= 1e-12 * scale), np.abs(got - ref) / scale
    # k = 0: every term has m = -l, where beta is exactly 0
    assert Q[0] == 0
    # Hermitian symmetry of Q for real data, over the whole grid
    Qg = Q.reshape(N, N, N)
    Qm = np.conj(np.roll(np.flip(Qg), 1, axis=(0, 1, 2)))  # conj Q at -k (mod N)
    assert np.abs(Qg - Qm).max() <= 1e-12 * scale.max()


def test_solver_compressed_matches_plain(edb):
    N, L, dt = 9, 8.0, 0.05
    F0 = np.stack([maxwellian_mixture(N, L, 31 + i) for i in range(2)])
    out = {}
    for prec in (edb.PREC_FP64_PLAIN, edb.PREC_FP64_COMPRESSED):
        s = edb.Solver(N, 0, L, False, 2, prec, "RK4", dt, 0.0, edb.FIX_APPLIED, history=8)
        s.set_initial_real(F0)
        s.step(6)
        out[prec] = s.get_spectral()
    a, b = out[edb.PREC_FP64_COMPRESSED], out[edb.PREC_FP64_PLAIN]
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_time_compressed(edb):
    q = edb.QCalculator(24, 0, 8.0)
    q.closeFilter()
    q.genB_compressed()
    assert q.time_compressed(reps=2) > 0
    with pytest.raises(edb.EdbError):
        q.calQ(rand_spectrum(24, 1))  # the stored array was never built


def test_grid_limits(edb):
    from entropydecreaseboltzmann_amd import _lib
    with pytest.raises(edb.EdbError) as e:
        edb.QCalculator(65, 0, 8.0)
    assert e.value.status == _lib.EDB_EINVAL
    q = edb.QCalculator(40, 0, 8.0)  # plain paths only past N = 33
    with pytest.raises(edb.EdbError) as e:
        q.genB_batched()
    assert e.value.status == _lib.EDB_EUNSUPPORTED
    q.closeFilter()
    q.genB_plain32()  # the stored FP32 array still fits at N = 40 (14 GB)
    f = rand_spectrum(40, 5)
    q.genB_compressed()
    a, b = q.calQ_compressed(f), q.calQ_plain32(f)
    assert np.abs(a - b).max() <= 1e-6 * np.abs(a).max()  # FP32 beta against FP64 beta
    with pytest.raises(edb.EdbError) as e:
        edb.Solver(40, 0, 8.0, False, 64, edb.PREC_FP32_BATCHED, "RK4", 0.01, 0.0, edb.FIX_NONE)
    assert e.value.status == _lib.EDB_EINVAL


def test_failed_genB_leaves_nothing_built(edb):
    """The stored FP64 array at N = 64 (500 GB) cannot be allocated: genB
    reports EDB_ENOMEM on every call (a retry does not find a half-built array
    and report success), calQ refuses to run, and the handle still builds and
    runs the compressed beta."""
    from entropydecreaseboltzmann_amd import _lib
    q = edb.QCalculator(64, 0, 8.0)
    assert q.info()["bytes_plain64"] > 400e9
    for _ in range(2):
        with pytest.raises(edb.EdbError) as e:
            q.genB()
        assert e.value.status == _lib.EDB_ENOMEM
    f = rand_spectrum(64, 3)
    with pytest.raises(edb.EdbError):
        q.calQ(f)
    q.genB_compressed()
    assert np.isfinite(q.calQ_compressed(f)).all()


def test_solver_past_n33(edb):
    """An FP64 RK4 solve at N = 40 (generic DFT passes, compressed beta) against
    the same solve with the stored FP32 beta, and its mass conservation."""
    N, L, dt = 40, 8.0, 0.01
    F0 = maxwellian_mixture(N, L, 404)[None]
    out = {}
    for prec in (edb.PREC_FP64_COMPRESSED, edb.PREC_FP32_PLAIN):
        s = edb.Solver(N, 0, L, False, 1, prec, "RK4", dt, 0.0, edb.FIX_NONE)
        s.set_initial_real(F0)
        m0 = s.moments()
        s.step(2)
        out[prec] = s.get_real()
        m1 = s.moments()
        assert abs(m1[0, 0] - m0[0, 0]) <= 1e-12 * abs(m0[0, 0])
    a, b = out[edb.PREC_FP64_COMPRESSED], out[edb.PREC_FP32_PLAIN]
    assert np.abs(a - F0).max() > 1e-6 * np.abs(F0).max()  # the state moved
    assert np.abs(a - b).max() <= 1e-6 * np.abs(a).max()
