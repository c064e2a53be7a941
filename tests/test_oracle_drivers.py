"""Pins of the oracle's per-step tail: BKW3D::flogf / calEntropy
(BKW.h:140-161), the drivers' entropy-fix gate and aim (BKW.h:225-235,
497-510), and the stage structure of all six drivers including CFE
(BKW.h:277-311) and DUFORT (BKW.h:539-631).

BKW.h needs fftw3.h, which this image lacks, so the drivers cannot be built
from the reference; they are pinned by closed forms instead:

* calEntropy: exact values of flogf on hand-built samples (the |f| < 1e-6 cut
  and the NaN of log at f <= -1e-6), the (2R)^3/N^3 scale on a constant
  field, and the entropy integral of a resolved Maxwellian,
  rho log rho - (3/2) rho (1 + log 2 pi T) - rho.
* The drivers, on a collision kernel built by hand so that Q = -lambda f on
  every mode but 0 (only the l = 0 row of B is non-zero): the state stays
  f_bar + c(t) g, each scheme multiplies c by a closed-form factor per step
  (the Taylor polynomials of e^{-z} for FEuler/RK2/SSP3/RK4, the CFE factor
  1 - lambda (1 - e^{-beta dt}) / beta, and the DUFORT start-up and
  DuFort-Frankel recurrences), and the entropy of f_bar + c g is a function of
  c alone that numpy evaluates from the pinned calEntropy.  With lambda < 0
  the entropy rises every step, which exercises the fix: not at the drivers'
  first step, only when the rise exceeds 1e-10, and then to the previous
  step's entropy (aim entropy0 / (2R)^3 in EntropyCal's units).

test_planted_mistakes_are_caught builds copies of the oracle source with one
plausible mistake each and checks that these pins fail on every copy.
"""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SRC = os.path.join(ROOT, "oracle", "edb_oracle.c")

N, R = 7, 4.0          # odd grid (M = N), velocity half-width
VOL = (2 * R) ** 3
AMP = 3.0              # f0 = AMP (1 + g): calEntropy ~ +150, above the drivers' entropy0 = 1 start


# ------------------------------------------------------------ calEntropy

def check_flogf_values():
    """flogf (BKW.h:153-161): 0 for |f| < 1e-6, f log f - f otherwise."""
    fl = oracle.lib().edbo_flogf
    assert fl(1.0) == -1.0
    assert abs(fl(math.e)) <= 1e-15
    assert fl(0.5e-6) == 0.0 and fl(-0.5e-6) == 0.0 and fl(0.0) == 0.0
    assert fl(0.999999e-6) == 0.0
    assert fl(1e-6) == pytest.approx(1e-6 * (math.log(1e-6) - 1.0), rel=1e-15)  # the cut is strict
    assert fl(2.0) == pytest.approx(2.0 * math.log(2.0) - 2.0, rel=1e-15)
    assert math.isnan(fl(-1e-6)) and math.isnan(fl(-0.3))  # log of a negative sample


def check_cal_entropy_scale():
    """calEntropy (BKW.h:140-152) of a constant field is (2R)^3 (c log c - c);
    the imaginary parts of the complex storage are ignored."""
    for n, r, c in ((5, 3.0, 0.7), (4, 8.0, 2.5)):
        F = np.full((n, n, n), c, dtype=np.complex128) + 5.0j
        assert oracle.cal_entropy(n, r, F) == pytest.approx((2 * r) ** 3 * (c * math.log(c) - c), rel=1e-14)
    F = np.full((5, 5, 5), 0.7, dtype=np.complex128)
    F[1, 2, 3] = 0.3e-6  # below the cut: contributes nothing
    H = oracle.cal_entropy(5, 3.0, F)
    assert H == pytest.approx(6.0 ** 3 / 125 * 124 * (0.7 * math.log(0.7) - 0.7), rel=1e-14)
    F[1, 2, 3] = -2e-6
    assert math.isnan(oracle.cal_entropy(5, 3.0, F))


def check_cal_entropy_maxwellian():
    """(2R)^3/N^3 sum flogf(f) -> int (f log f - f) dv for a resolved
    Maxwellian: rho log rho - 3/2 rho (1 + log 2 pi T) - rho.  At rho = 1e6 the
    samples under the 1e-6 cut carry < 1e-9 of it; at rho = 1, < 1e-3."""
    n, L = 32, 8.0
    h = 2 * L / n
    x = -L + h * np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    for rho, T, u, tol in ((1e6, 0.9, (0.3, -0.2, 0.1), 1e-9), (1.0, 1.3, (0.0, 0.4, 0.0), 1e-3)):
        f = rho * np.exp(-((X - u[0]) ** 2 + (Y - u[1]) ** 2 + (Z - u[2]) ** 2) / (2 * T)) / (2 * np.pi * T) ** 1.5
        H = oracle.cal_entropy(n, L, f.astype(np.complex128))
        exact = rho * math.log(rho) - 1.5 * rho * (1 + math.log(2 * math.pi * T)) - rho
        assert abs(H - exact) <= tol * abs(exact), (rho, H, exact)


# ------------------------------------------------- drivers on Q = -lambda f

def base_state():
    """f0 = AMP (1 + g) on the N^3 grid with a band-limited g (modes |k| <= 2);
    returns f0 and AMP g.  f_bar + c AMP g stays >= 0.4 AMP for c <= 1.5:
    every sample is above both entropy cuts (1e-6 in calEntropy, 1e-10 in
    EntropyCal)."""
    h = 2 * R / N
    x = -R + h * np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    w = 2 * np.pi / (2 * R)
    g = 0.3 * np.cos(w * X) + 0.2 * np.sin(2 * w * Y) + 0.1 * np.cos(w * (X + Z))
    return AMP * (1.0 + g), AMP * g


def relax_kernel(lam, fhat0):
    """B with Q_k = -lam f_k for k != 0 and Q_0 = 0: only the l = 0 row is
    non-zero (the term beta(0, k) f_0 f_k), beta(0, k) = -lam / f_0."""
    M = oracle.nmodes(N)
    n = (N - 1) // 2
    c0 = (n * M + n) * M + n
    B = np.zeros((M ** 3, M ** 3))
    B[c0, :] = -lam / fhat0[0, 0, 0].real
    B[c0, c0] = 0.0
    return B


def entropy_of(c):
    """calEntropy of f_bar + c g (numpy, from the definition pinned above)."""
    f0, g = base_state()
    f = AMP + c * g
    return VOL / N ** 3 * np.sum(f * np.log(f) - f)


def amplification(scheme, z, lam=None, dt=None, beta=None):
    if scheme == "FEuler":
        return 1 - z
    if scheme == "RK2":
        return 1 - z + z * z / 2
    if scheme == "SSP3":
        return 1 - z + z * z / 2 - z ** 3 / 6
    if scheme == "RK4":
        return 1 - z + z * z / 2 - z ** 3 / 6 + z ** 4 / 24
    if scheme == "CFE":
        return 1 - lam * (1 - math.exp(-beta * dt)) / beta
    raise ValueError(scheme)


def driver_times(T0, TEnd, dt, startup=0):
    t, out = T0, []
    for _ in range(startup):
        t += dt / 8
        out.append(t)
    while True:
        t += dt
        out.append(t)
        if not t < TEnd:
            return out


def run_relax(scheme, lam, dt, TEnd, fix_mode, beta=0.0):
    f0, g = base_state()
    fh0 = np.fft.fftn(f0)
    B = relax_kernel(lam, fh0)
    fo, t, H, Hf = oracle.run(N, B, R, scheme, 0.0, TEnd, dt, fh0, beta=beta, fix_mode=fix_mode)
    return fh0, fo, t, H, Hf


def coefficient(fhat):
    """(mean, c) of the real state f = mean + c g, and the residual off that line."""
    f0, g = base_state()
    f = np.fft.ifftn(fhat).real
    m = f.mean()
    c = np.dot((f - m).ravel(), g.ravel()) / np.dot(g.ravel(), g.ravel())
    return m, c, np.abs(f - m - c * g).max()


def check_amplification(scheme):
    """n steps multiply every mode but 0 by a(z)^n and keep mode 0; the step
    times follow the drivers' do { t += dt } while (t < TEnd)."""
    lam, dt, TEnd, beta = 0.8, 0.05, 0.3, 2.0
    fh0, fo, t, H, Hf = run_relax(scheme, lam, dt, TEnd, 0, beta=beta)
    assert list(t) == driver_times(0.0, TEnd, dt)
    a = amplification(scheme, lam * dt, lam, dt, beta)
    want = fh0 * a ** len(t)
    want[0, 0, 0] = fh0[0, 0, 0]
    assert np.abs(fo - want).max() <= 1e-13 * np.abs(fh0).max(), scheme
    np.testing.assert_array_equal(H, Hf)  # fix_mode 0: no fix


def check_all_amplifications():
    for scheme in ("FEuler", "RK2", "SSP3", "RK4", "CFE"):
        check_amplification(scheme)


def check_cfe_order():
    """CFE's step factor is first order in dt against e^{-lambda t}, and for
    beta -> 0 it becomes forward Euler."""
    lam, TEnd, beta = 0.8, 0.4, 2.0
    errs = []
    for dt in (0.02, 0.01):
        fh0, fo, t, *_ = run_relax("CFE", lam, dt, TEnd - 1e-9, 0, beta=beta)
        exact = math.exp(-lam * t[-1])
        errs.append(abs(fo[1, 0, 0] / fh0[1, 0, 0] - exact))
    assert 1.7 < errs[0] / errs[1] < 2.3, errs
    fh0, fo_c, *_ = run_relax("CFE", lam, 0.05, 0.2, 0, beta=1e-9)
    _, fo_e, *_ = run_relax("FEuler", lam, 0.05, 0.2, 0)
    assert np.abs(fo_c - fo_e).max() <= 1e-8 * np.abs(fh0).max()


def check_dufort():
    """DUFORT (BKW.h:539-631): F_old = initial state; 8 RK2 steps that advance t
    by dt/8 but use the full dt in their stages (factor 1 - z + z^2/2 each);
    then (1 + b) F^{n+1} = (1 + b) F^n + dt Q(F^n) - b F_old with b = beta dt,
    F_old the state before the previous update (the initial one for the first),
    on every mode.  Mode 0 has Q_0 = 0, so its mass moves by -b F_old / (1 + b)
    per step; the test kernel's Q_k = beta(0, k) f_0 f_k then carries the
    factor d = f_0 / f_0(0), so z = lambda dt d at each DuFort-Frankel step."""
    lam, dt, TEnd, beta = 0.8, 0.05, 0.2, 2.0
    fh0, fo, t, H, Hf = run_relax("DUFORT", lam, dt, TEnd, 0, beta=beta)
    want_t = driver_times(0.0, TEnd, dt, startup=8)
    assert len(t) == len(want_t) and list(t) == want_t
    b = beta * dt
    n_df = len(t) - 8
    d_old, d = 1.0, 1.0  # mode 0: conserved by the start-up, then the recurrence with z = 0
    dc = [d]
    for _ in range(n_df):
        d_old, d = d, ((1 + b) * d - b * d_old) / (1 + b)
        dc.append(d)
    z = lam * dt
    a2 = 1 - z + z * z / 2
    old, cur = 1.0, a2 ** 8
    for n in range(n_df):
        old, cur = cur, ((1 + b - z * dc[n]) * cur - b * old) / (1 + b)
    want = fh0 * cur
    want[0, 0, 0] = fh0[0, 0, 0] * dc[-1]
    assert np.abs(fo - want).max() <= 1e-13 * np.abs(fh0).max()


def fix_tol():
    # FixSol's early exit leaves EntropyCal within eps = 1e-10 below the aim,
    # its beta_R exit within |dH/dbeta| * 1e-10: (2R)^3 x 2e-10 in calEntropy units
    return VOL * 2e-10


def check_fix_gate_and_aim_applied():
    """lambda < 0 (entropy rises every step), fix applied: RK4 and SSP3 do not
    fix their first step and fix every later one; FEuler (whose counter starts
    at 1) the same; the fixed entropy is the previous step's, and the fixed
    state is f_bar + c g again (a blend toward the average)."""
    lam, dt, TEnd = -0.5, 0.05, 0.3
    for scheme in ("RK4", "SSP3", "FEuler"):
        fh0, fo, t, H, Hf = run_relax(scheme, lam, dt, TEnd, 2)
        a = amplification(scheme, lam * dt)
        assert H[0] == pytest.approx(entropy_of(a), rel=1e-12)
        assert Hf[0] == H[0]  # the first step is never fixed
        for k in range(1, len(t)):
            assert H[k] - Hf[k - 1] > 1e-10
            assert Hf[k] < H[k]
            assert Hf[k - 1] - fix_tol() <= Hf[k] <= Hf[k - 1] + 1e-12 * abs(Hf[k - 1]), (scheme, k)
        m, c, res = coefficient(fo)
        assert res <= 1e-12 * AMP and abs(m - AMP) <= 1e-12 * AMP
        assert entropy_of(c) == pytest.approx(Hf[-1], rel=1e-12)
        assert 1.0 < c < a  # pulled back toward f_bar, never past the first step's rise


def check_fix_diagnostic_leaves_state():
    """Diagnostic fix (the reference's drivers): the spectral state grows by
    a(z) per step untouched, the recorded entropy is that of the unfixed state,
    and the fixed entropy stays at the first step's."""
    lam, dt, TEnd = -0.5, 0.05, 0.3
    fh0, fo, t, H, Hf = run_relax("RK4", lam, dt, TEnd, 1)
    a = amplification("RK4", lam * dt)
    want = fh0 * a ** len(t)
    want[0, 0, 0] = fh0[0, 0, 0]
    assert np.abs(fo - want).max() <= 1e-13 * np.abs(fh0).max()
    for k in range(len(t)):
        assert H[k] == pytest.approx(entropy_of(a ** (k + 1)), rel=1e-12)
    assert Hf[0] == H[0]
    for k in range(1, len(t)):
        assert Hf[k - 1] - fix_tol() <= Hf[k] <= Hf[k - 1] + 1e-12 * abs(Hf[k - 1]), k
    # RK2 and DUFORT never fix in the reference (their fix is `if(0)`)
    for scheme in ("RK2", "DUFORT"):
        _, _, _, H2, Hf2 = run_relax(scheme, lam, dt, 0.1, 1, beta=2.0)
        np.testing.assert_array_equal(H2, Hf2)


def check_fix_threshold():
    """The gate is a rise of more than 1e-10 over the previous step: with
    lambda set so that each RK4 step raises H by ~3e-11 no fix ever runs; with
    ~3e-10 every step after the first is fixed."""
    dt, TEnd = 0.05, 0.3
    dHdc = (entropy_of(1 + 1e-6) - entropy_of(1 - 1e-6)) / 2e-6
    for rise, fixes in ((3e-11, False), (3e-10, True)):
        lam = -rise / (dHdc * dt)
        fh0, fo, t, H, Hf = run_relax("RK4", lam, dt, TEnd, 2)
        fixed = Hf != H
        assert not fixed[0]
        assert bool(fixed[1:].all()) == fixes and bool(fixed[1:].any()) == fixes, (rise, fixed)


PINS = [check_flogf_values, check_cal_entropy_scale, check_cal_entropy_maxwellian, check_all_amplifications,
        check_cfe_order, check_dufort, check_fix_gate_and_aim_applied, check_fix_diagnostic_leaves_state,
        check_fix_threshold]


@pytest.mark.parametrize("pin", PINS, ids=[p.__name__ for p in PINS])
def test_oracle_pin(pin):
    pin()


# Each entry: a plausible mistake as a one-place edit of the oracle source.
PLANTED = {
    "flogf cut at 1e-10": ("if (fabs(f) < 1.0e-6) return 0.;", "if (fabs(f) < 1.0e-10) return 0.;"),
    "flogf without -f": ("return f * log(f) - f;", "return f * log(f);"),
    "calEntropy over N^2": ("entropy *= pow(2 * R, 3.) / (double)N3;", "entropy *= pow(2 * R, 3.) / (double)(N * N);"),
    "fix gate at 0": ("entropy - entropy0 > 1.0e-10", "entropy - entropy0 > 0.0"),
    "fix gate at 1e-9": ("entropy - entropy0 > 1.0e-10", "entropy - entropy0 > 1.0e-9"),
    "fix aim without /(2R)^3": ("entropy0 / p->vol);", "entropy0);"),
    "RK4 fixes its first step": ("""        Fhat[l] = Fhat[l] + dt / 6. * (Q[l] + 2. * Q1[l] + 2. * Q2[l] + Q3[l]);
      entropy0 = post_step(&p, t, step > 0, entropy0);""", """        Fhat[l] = Fhat[l] + dt / 6. * (Q[l] + 2. * Q1[l] + 2. * Q2[l] + Q3[l]);
      entropy0 = post_step(&p, t, step >= 0, entropy0);"""),
    "FEuler fixes from its first step": ("entropy0 = post_step(&p, t, step > 1, entropy0);",
                                         "entropy0 = post_step(&p, t, step > 0, entropy0);"),
    "gate against the unfixed entropy": ("  p->rec++;\n  return entropy;", "  p->rec++;\n  return h_before;"),
    "RK4 stage weight": ("F2[l] = Fhat[l] + 0.5 * dt * Q1[l];", "F2[l] = Fhat[l] + 0.25 * dt * Q1[l];"),
    "SSP3 coefficient": ("F2[l] = 0.75 * Fhat[l] + 0.25 * F1[l] + 0.25 * dt * Q[l];",
                         "F2[l] = 0.7 * Fhat[l] + 0.3 * F1[l] + 0.25 * dt * Q[l];"),
    "CFE as forward Euler": ("double s = (1. - exp(-beta_cfe * dt)) / beta_cfe;", "double s = dt;"),
    "DUFORT start-up stages at dt/8": ("F1[l] = Fhat[l] + 0.5 * dt * Q[l];\n      drv_calQ(N, B, R, conserve, F1, Q, nthreads);\n      for (long l = 0; l < n2; ++l) Fhat[l] = Fhat[l] + dt * Q[l];",
                                       "F1[l] = Fhat[l] + 0.5 * tstep * Q[l];\n      drv_calQ(N, B, R, conserve, F1, Q, nthreads);\n      for (long l = 0; l < n2; ++l) Fhat[l] = Fhat[l] + tstep * Q[l];"),
    "DUFORT start-up t += dt": ("double tstep = scheme == 5 ? dt / 8 : dt;", "double tstep = dt;"),
    "DUFORT F_old tracks the start-up": ("entropy0 = post_step(&p, t, fixable && step > 0, entropy0);\n      step++;\n    } while (scheme == 5 ? step < 8 : t < TEnd);",
                                         "entropy0 = post_step(&p, t, fixable && step > 0, entropy0);\n      step++;\n      if (scheme == 5 && step < 8) memcpy(Fold, Fhat, cb);\n    } while (scheme == 5 ? step < 8 : t < TEnd);"),
    "DuFort-Frankel sign": ("beta_cfe * dt * (F1[l] - Fold[l]));", "beta_cfe * dt * (Fold[l] - F1[l]));"),
}


def _build_variant(tmp, name, old, new):
    src = open(ORACLE_SRC).read()
    assert src.count(old) == 1, (name, src.count(old))
    d = os.path.join(tmp, "v%08x" % (abs(hash(name)) & 0xFFFFFFFF))
    os.makedirs(d, exist_ok=True)
    shutil.copy(os.path.join(ROOT, "oracle", "edb_oracle.h"), d)
    with open(os.path.join(d, "edb_oracle.c"), "w") as fh:
        fh.write(src.replace(old, new))
    so = os.path.join(d, "liboracle_variant.so")
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-fopenmp", "-shared", "-o", so,
                    os.path.join(d, "edb_oracle.c"), "-lm"], check=True)
    return so


@pytest.mark.parametrize("name", sorted(PLANTED))
def test_planted_mistakes_are_caught(tmp_path, name):
    old, new = PLANTED[name]
    so = _build_variant(str(tmp_path), name, old, new)
    failed = []
    with oracle.using(so):
        for pin in PINS:
            try:
                pin()
            except AssertionError:
                failed.append(pin.__name__)
    assert failed, f"no pin catches: {name}"
