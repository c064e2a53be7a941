"""edb_qcalc_calQ_batched_device_noalloc (scratch in a file-scope device
array, no allocation) against edb_qcalc_calQ_batched_device on the same
device buffers: two calls in a row, the second with a smaller batch, on the
stream-K rows schedule (N = 24) and the workgroup-quad schedule (N = 16), with and
without the Hermitian half.
With the persistent count the scratch caps, results agree to the FP64 merge
order (1e-12 of max |Q|); with a count both paths share, bit for bit.  The
first result is also checked against the oracle at sampled modes (FP32 bar
3e-6 S_k, S_k = sum_l |beta||f_l||f_[k-l]|)."""
import ctypes

import numpy as np
import pytest

import oracle
from conftest import HipMemory, centred_of_fftw, hermitize, maxwellian_mixture

pytestmark = pytest.mark.gpu


def kernel_layout(fs, N, ld):
    """float2 [M^3 centred][ld] of FFTW-order spectra fs[batch][N][N][N]."""
    cen = centred_of_fftw(N)
    M = oracle.nmodes(N)
    fftw_of = np.empty(M ** 3, dtype=np.int64)
    fftw_of[cen[cen >= 0]] = np.flatnonzero(cen >= 0)
    b = fs.shape[0]
    out = np.zeros((M ** 3, ld), dtype=np.complex64)
    out[:, :b] = fs.reshape(b, -1)[:, fftw_of].T.astype(np.complex64)
    return out, fftw_of


def run_pair(edb, hm, q, N, batch, hermitian, seed):
    L = 8.0
    M = oracle.nmodes(N)
    ld = ((batch + 63) // 64) * 64
    fs = hermitize(np.stack([np.fft.fftn(maxwellian_mixture(N, L, seed + i)) for i in range(batch)]))
    fin, fftw_of = kernel_layout(fs, N, ld)
    d_in = hm.upload(fin)
    nbytes = M ** 3 * ld * 16
    d_a, d_b = hm.alloc(nbytes), hm.alloc(nbytes)
    hm.fill(d_a, nbytes, 0xFF)  # NaN patterns: every output entry must be written
    hm.fill(d_b, nbytes, 0xFF)
    lib = edb.load()
    print(f"noalloc: N={N} batch={batch} hermitian={hermitian} schedule={q.batched_schedule()}", flush=True)
    assert lib.edb_qcalc_calQ_batched_device(q.handle, batch, ld, int(hermitian), d_a, d_in, None) == 0
    st = lib.edb_qcalc_calQ_batched_device_noalloc(q.handle, batch, ld, int(hermitian), d_b, d_in, None)
    assert st == 0, lib.edb_last_error()
    a = hm.download(d_a, (M ** 3, ld), np.complex128)[:, :batch]
    b = hm.download(d_b, (M ** 3, ld), np.complex128)[:, :batch]
    return fs, fftw_of, a, b


@pytest.mark.parametrize("N,sched", [(24, 0), (16, 1)])
def test_noalloc_matches_batched_device(edb, N, sched):
    L = 8.0
    q = edb.QCalculator(N, 0, L)
    q.closeFilter()
    q.genB_batched()
    q.tune_batched(sched, 0)
    hm = HipMemory()
    try:
        for k, batch in enumerate((256, 70)):  # second call: smaller input
            fs, fftw_of, a, b = run_pair(edb, hm, q, N, batch, True, 60 * N + 1000 * k)
            assert np.isfinite(a).all() and np.isfinite(b).all()
            assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max(), (batch, np.abs(a - b).max())
            if k == 0:  # the existing entry point's operator, pinned to the oracle
                M = oracle.nmodes(N)
                rng = np.random.default_rng(N)
                kcs = rng.choice(M ** 3, 24, replace=False)
                for i in (0, batch - 1):
                    ref = oracle.calQ_modes(N, fs[i], kcs, 0, L, False)
                    S = oracle.calQ_modes(N, fs[i], kcs, 0, L, False, absval=True).real
                    assert np.all(np.abs(b[kcs, i] - ref) <= 3e-6 * S), i
        # a persistent count both paths use: bit for bit, twice, smaller second
        q.tune_batched(sched, 64)
        for k, batch in enumerate((130, 5)):
            _, _, a, b = run_pair(edb, hm, q, N, batch, True, 70 * N + 1000 * k)
            assert np.array_equal(a, b), batch
        # every k-row evaluated (no Hermitian mirror): the closed-form prefix
        # table then counts all rows / quads
        _, _, a, b = run_pair(edb, hm, q, N, 64, False, 80 * N)
        assert np.isfinite(a).all() and np.array_equal(a, b)
        # beyond the scratch: refused before any launch (the sizes are checked on the host)
        lib = edb.load()
        tiny = hm.alloc(256)
        big = 1 << 26  # prefix table alone > 200 MB at both sizes and schedules
        st = lib.edb_qcalc_calQ_batched_device_noalloc(q.handle, big, big, 1, tiny, tiny, None)
        assert st == 1 and b"MiB scratch" in lib.edb_last_error()
    finally:
        assert hm.hip.hipDeviceSynchronize() == 0
        hm.free()
