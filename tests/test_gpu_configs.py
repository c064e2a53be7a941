"""GPU parity at the benchmark configurations' own sizes.

* 256 mixtures at N = 24 sharing one FP32 collision kernel, RK4 dt = 0.01 with
  the applied entropy fix, every step replayed from a HIP graph on the
  autotuned schedule (BKW.h:478-516), against the FP64 oracle driver.
* The plain FP64 Q at N = 24 against the oracle (QCalculator.h:231-257).
* The batched Q of 64 bi-Maxwellian-plus-bump instances at N = 32 with the
  hard-sphere kernel (alpha = 1) against the oracle's per-mode sums.
* The FP64 moment and BKW-error reductions against numpy sums of the real
  state (testfft.cpp:11-16 for the BKW profile).
* The accuracy target at N = 24: FP32-beta error against the exact BKW
  solution within twice the FP64 run's at every step.
* Two solvers sharing one collision kernel, stepped back to back without a
  synchronisation between them, against solvers with kernels of their own.

Bars as in test_gpu_parity: FP64 paths 1e-12 relative; FP32 batched within
3e-6 of S_k = sum_l |beta||f_l||f_[k-l]| per mode, and 2e-6 relative along
an RK4 trajectory.
"""
import numpy as np
import pytest

import oracle
from conftest import bimaxwellian_bump, centred_of_fftw, hermitize, maxwellian_mixture, rand_spectrum

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_bench_workload_256_n24_graph_rk4_matches_oracle(edb):
    """The benchmark step itself: 256 x N=24, FP32 batched, RK4 dt = 0.01,
    entropy fix applied, graph replay, autotuned work split.  Instances 0, 63,
    64 and 255 (both ends of the first and last 64-instance groups) against the
    oracle's RK4 driver over 3 steps; mass of every instance to 1e-12."""
    N, L, dt, batch, steps = 24, 8.0, 0.01, 256, 3
    F0 = np.stack([maxwellian_mixture(N, L, 7000 + i) for i in range(batch)])
    q = edb.QCalculator(N, 0, L)
    q.closeFilter()
    q.genB_batched()
    q.autotune(batch)
    print("schedule", q.batched_schedule(), flush=True)
    s = edb.Solver(N, 0, L, False, batch, edb.PREC_FP32_BATCHED, "RK4", dt, 0.0, edb.FIX_APPLIED,
                   hermitian=True, history=8, qcalc=q)
    s.use_graph(True)
    s.set_initial_real(F0)
    m0 = s.moments()
    s.step(steps)
    m1 = s.moments()
    assert np.abs(m1[:, 0] - m0[:, 0]).max() <= 1e-12 * np.abs(m0[:, 0]).max()
    got = s.get_spectral()
    assert np.isfinite(got).all()
    t, n = s.time()
    assert n == steps
    B = oracle.genB(N, 0, L, 0)
    for i in (0, 63, 64, 255):
        fo, tt, H, Hf = oracle.run(N, B, L, "RK4", 0.0, t - 1e-12, dt, np.fft.fftn(F0[i]), fix_mode=2)
        assert len(tt) == steps
        assert rel(got[i], fo) < 2e-6, i


def test_plain_fp64_calQ_n24_matches_oracle(edb):
    """The FP64 stream of the single-instance run at N = 24 (M = 23 modes),
    BKW data and a random spectrum, against the oracle's calQ."""
    N, L = 24, 8.0
    q = edb.QCalculator(N, 0, L)
    q.closeFilter()
    q.genB()
    B = oracle.genB(N, 0, L, 0)
    for f in (np.fft.fftn(oracle.grid_bkw(N, L, 0.7)), rand_spectrum(N, 2424)):
        ref = oracle.calQ(N, B, f)
        got = q.calQ(f)
        assert rel(got, ref) < 1e-12
        assert got[0, 0, 0] == 0.0


def test_batched_n32_hard_spheres_64_matches_oracle(edb):
    """The N = 32, alpha = 1 batch of 64 bi-Maxwellian + bump instances (FP32
    paired kernel, default schedule) against the oracle's per-mode sums at
    512 modes of four instances: the 16 largest |Q| and 496 random ones."""
    N, L, batch = 32, 8.0, 64
    fs = hermitize(np.stack([np.fft.fftn(bimaxwellian_bump(N, L, 3200 + i)) for i in range(batch)]))
    q = edb.QCalculator(N, 1, L)
    q.closeFilter()
    q.genB_batched()
    got = q.calQ_batched(fs, hermitian=True)
    cen = centred_of_fftw(N)
    modes = np.flatnonzero(cen >= 0)
    rng = np.random.default_rng(32)
    for i in (0, 31, 32, 63):
        g = got[i].ravel()
        big = modes[np.argsort(np.abs(g[modes]))[-16:]]
        pick = np.unique(np.concatenate([big, rng.choice(modes, 496, replace=False)]))
        ref = oracle.calQ_modes(N, fs[i], cen[pick], 1, L, False)
        S = oracle.calQ_modes(N, fs[i], cen[pick], 1, L, False, absval=True).real
        err = np.abs(g[pick] - ref)
        assert np.all(err <= 3e-6 * S), (i, (err / S).max())
        assert got[i][0, 0, 0] == 0.0


@pytest.mark.parametrize("N", [16, 24])
def test_moments_and_bkw_error_match_numpy(edb, N):
    """edb_solver_moments (h^3 sums of f, v f, |v|^2 f) and edb_solver_bkw_error
    (L-infinity and h^3 L1 of F - f_BKW(K)) against numpy on get_real(), for BKW
    data after two RK4 steps and for mixtures."""
    L, dt = 8.0, 0.01
    h = 2 * L / N
    x = -L + h * np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    K0 = 0.7
    F0 = np.stack([oracle.grid_bkw(N, L, K0).real, maxwellian_mixture(N, L, 16 + N),
                   bimaxwellian_bump(N, L, 17 + N)])
    s = edb.Solver(N, 0, L, False, 3, edb.PREC_FP64_PLAIN, "RK4", dt, 0.0, edb.FIX_NONE)
    s.set_initial_real(F0)
    s.step(2)
    F = s.get_real()
    mom = s.moments()
    V2 = X * X + Y * Y + Z * Z
    for i in range(3):
        f = F[i]
        ref = np.array([f.sum(), (X * f).sum(), (Y * f).sum(), (Z * f).sum(), (V2 * f).sum()]) * h ** 3
        a = np.abs(f)
        scale = np.array([a.sum(), (np.abs(X) * a).sum(), (np.abs(Y) * a).sum(), (np.abs(Z) * a).sum(),
                          (V2 * a).sum()]) * h ** 3
        assert np.all(np.abs(mom[i] - ref) <= 1e-13 * scale), (i, mom[i], ref)
    K = 1 - (1 - K0) * np.exp(-2 * dt / 6)
    err = s.bkw_error(K)
    ex = oracle.grid_bkw(N, L, K).real
    # the device evaluates the profile with its own exp/pow: a few ulp of |f_BKW|
    tol = 1e-14 * np.abs(ex).max()
    for i in range(3):
        d = np.abs(F[i] - ex)
        assert abs(err[i, 0] - d.max()) <= tol, (i, err[i, 0], d.max())
        assert abs(err[i, 1] - d.sum() * h ** 3) <= tol * N ** 3 * h ** 3 + 1e-12 * d.sum() * h ** 3, i
    # a perturbation of known size at one grid point moves L-infinity to it
    s2 = edb.Solver(N, 0, L, False, 1, edb.PREC_FP64_PLAIN, "RK4", dt, 0.0, edb.FIX_NONE)
    G = oracle.grid_bkw(N, L, K0).real.copy()
    G[3, 5, 7] += 0.25
    s2.set_initial_real(G[None])
    e2 = s2.bkw_error(K0)
    assert abs(e2[0, 0] - 0.25) <= 1e-13
    assert abs(e2[0, 1] - 0.25 * h ** 3) <= 1e-13 * N ** 3 * h ** 3


def test_fp32_error_vs_bkw_within_twice_fp64(edb):
    """The accuracy target on the single-instance configuration: N = 24, L = 8,
    filter off, BKW data with K(0) = 0.5 and K(t) = 1 - e^{-t/6} / 2
    (testfft.cpp:11-16 profile), RK4 dt = 0.01 (BKW.h:452-519), 100 steps.
    At every step the FP32-beta runs (plain stream and batched kernel) have an
    L-infinity error against the exact solution within twice the FP64 plain
    run's, they stay within 1e-5 of its state, and all three conserve mass to
    1e-12."""
    N, L, dt, n = 24, 8.0, 0.01, 100
    K_of = lambda t: 1.0 - 0.5 * np.exp(-t / 6.0)
    F0 = oracle.grid_bkw(N, L, K_of(0.0)).real
    runs = {}
    for name, prec in (("fp64", edb.PREC_FP64_PLAIN), ("fp32_plain", edb.PREC_FP32_PLAIN),
                       ("fp32_batched", edb.PREC_FP32_BATCHED)):
        s = edb.Solver(N, 0, L, False, 1, prec, "RK4", dt, 0.0, edb.FIX_NONE)
        s.set_initial_real(F0[None])
        m0 = s.moments()[0, 0]
        errs = []
        for k in range(1, n + 1):
            s.step(1)
            errs.append(s.bkw_error(K_of(k * dt))[0, 0])
        assert abs(s.moments()[0, 0] - m0) <= 1e-12 * abs(m0), name
        runs[name] = (np.array(errs), s.get_real()[0])
        s.close()
    e64, F64 = runs["fp64"]
    assert np.all(np.isfinite(e64)) and e64.max() < 0.05 * F0.max()
    for name in ("fp32_plain", "fp32_batched"):
        e32, F32 = runs[name]
        assert np.all(e32 <= 2.0 * e64), (name, (e32 / e64).max())
        assert np.abs(F32 - F64).max() <= 1e-5 * np.abs(F64).max(), name


@pytest.mark.parametrize("precision", ["batched", "plain"])
def test_solvers_sharing_a_qcalc_step_concurrently(edb, precision):
    """Two solvers sharing one collision kernel, stepped back to back on their
    own streams with no synchronisation in between (their Q evaluations overlap
    on the device), give the results of solvers with kernels of their own."""
    N, L, dt = 16, 8.0, 0.01
    prec = edb.PREC_FP32_BATCHED if precision == "batched" else edb.PREC_FP64_PLAIN
    batch = 70 if precision == "batched" else 2
    Fa = np.stack([maxwellian_mixture(N, L, 800 + i) for i in range(batch)])
    Fb = np.stack([maxwellian_mixture(N, L, 900 + i) for i in range(batch)])
    q = edb.QCalculator(N, 0, L)
    q.closeFilter()
    if prec == edb.PREC_FP32_BATCHED:
        q.genB_batched()
    else:
        q.genB()
    shared = [edb.Solver(N, 0, L, False, batch, prec, "RK4", dt, 0.0, edb.FIX_APPLIED, qcalc=q)
              for _ in range(2)]
    own = [edb.Solver(N, 0, L, False, batch, prec, "RK4", dt, 0.0, edb.FIX_APPLIED) for _ in range(2)]
    for pair in (shared, own):
        pair[0].set_initial_real(Fa)
        pair[1].set_initial_real(Fb)
    for _ in range(3):
        shared[0].step(2)
        shared[1].step(2)
    for s in own:
        s.step(6)
    for a, b in zip(shared, own):
        assert np.array_equal(a.get_spectral(), b.get_spectral())
