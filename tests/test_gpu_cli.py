"""The testBKW3D command-line driver (entropydecreaseboltzmann_amd/bin) against
the reference's output formats (CalQ/testBKW3D.cpp, BKW.h:163-248) and the
oracle's forward-Euler driver: header, per-step entropy table, F.dat and the
N<N>.dt<dt>.dat line cut."""
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "entropydecreaseboltzmann_amd", "bin", "testBKW3D")
R = 6.62132


def initial(N):
    """The reference's active initial data (BKW.h:96-112) on x_i = -R + i*2R/N."""
    h = 2 * R / N
    x = -R + h * np.arange(N)
    s = (x + R) / (2 * R) - 0.5
    g = sum(j / 55.0 * np.sin(j * s * np.pi) for j in range(11))
    return 3.2 + g[:, None, None] + g[None, :, None] + g[None, None, :]


def flogf_entropy(F, N):
    f = F.reshape(-1)
    v = np.where(np.abs(f) < 1e-6, 0.0, f * np.log(np.where(np.abs(f) < 1e-6, 1.0, f)) - f)
    return v.sum() * (2 * R) ** 3 / N ** 3


def test_testBKW3D_matches_reference_formats_and_oracle(tmp_path, edb):
    assert os.path.exists(EXE), "build() builds entropydecreaseboltzmann_amd/bin/testBKW3D"
    # the reference's initial data carries mass ~7e3 on [-R, R]^3, which makes
    # forward Euler stiff: dt = 1e-4 is stable where 1e-3 already diverges
    N, dt, TEnd = 9, 0.0001, 0.0008
    r = subprocess.run([EXE, str(N), str(dt), str(TEnd)], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.splitlines()
    # header of CalQ/testBKW3D.cpp:26-39 (default stream formatting)
    assert lines[:8] == [
        "##################################################",
        "# Entropy fixed method for BKW model",
        "# spatial domain  :\t[-6.62132, 6.62132]",
        "# number of grid N:\t9^3",
        "# time interval   :\t[0, 0.0008]",
        "# time step     dt:\t0.0001",
        "# Entropy fixing stratege is adopted.",
        "##################################################",
    ]
    assert "# Initialization..." in lines and "# QCalculator calculate B..." in lines
    assert "# Start the evolution using Forward Euler..." in lines
    i0 = lines.index("# Step\tt\tentropy_{n}\tentropy_{n-1}-entropy_{n}\tfixedEntropy\tfixedDiff")
    rows = [l.split("\t") for l in lines[i0 + 1:] if not l.startswith("#")]
    F0 = initial(N)
    B = oracle.genB(N, 0, R, True)  # QCalculator(N, 0): filter on
    fo, t, H, Hf = oracle.run(N, B, R, "FEuler", 0.0, TEnd, dt, np.fft.fftn(F0), fix_mode=1)
    n = len(t)
    assert len(rows) == n + 1
    # row 00: t0 and calEntropy of the initial F
    assert rows[0][0] == "00" and rows[0][1] == "0.00000000"
    assert float(rows[0][2]) == pytest.approx(oracle.cal_entropy(N, R, F0.astype(np.complex128)), abs=2e-8)
    for k in range(n):
        row = rows[k + 1]
        assert int(row[0]) == k + 1
        assert float(row[1]) == pytest.approx(t[k], abs=1e-8)
        assert float(row[2]) == pytest.approx(H[k], abs=2e-8)
        if k == 0:
            assert row[3:6] == ["-", "-", "-"]
        else:
            assert float(row[4]) == pytest.approx(Hf[k], abs=2e-8)
            assert float(row[3]) == pytest.approx(Hf[k - 1] - H[k], abs=2e-8)
            assert float(row[5]) == pytest.approx(Hf[k - 1] - Hf[k], abs=2e-8)
    # F.dat: t then the N^3 values of the driver's F after each step
    Fdat = [l.split("\t") for l in (tmp_path / "F.dat").read_text().splitlines()]
    assert len(Fdat) == n
    for k, row in enumerate(Fdat):
        vals = np.array([float(v) for v in row[1:] if v != ""])
        assert vals.size == N ** 3 and float(row[0]) == pytest.approx(t[k], rel=1e-5)
        fk, *_ = oracle.run(N, B, R, "FEuler", 0.0, (k + 0.5) * dt, dt, np.fft.fftn(F0), fix_mode=0)
        ref = np.fft.ifftn(fk).real.reshape(-1)
        if Hf[k] == H[k]:
            np.testing.assert_allclose(vals, ref, rtol=2e-5, atol=1e-6 * np.abs(ref).max())
        else:  # a fixed step: mass kept, entropy Hfix (6 printed digits)
            assert vals.mean() == pytest.approx(ref.mean(), rel=1e-5)
            assert flogf_entropy(vals, N) == pytest.approx(Hf[k], rel=1e-4)
    # outV1Direction: "# x\tf(t)\tf_0", x, F(i, N/2, N/2), f_0(x, h/2, h/2)
    cut = (tmp_path / "N9.dt0.000100.dat").read_text().splitlines()
    assert cut[0] == "# x\tf(t)\tf_0" and len(cut) == N + 1
    h = 2 * R / N
    Ffin = np.array([float(v) for v in Fdat[-1][1:] if v != ""]).reshape(N, N, N)
    for i, l in enumerate(cut[1:]):
        x, fx, f0 = (float(v) for v in l.split("\t"))
        assert l.split("\t")[0] == f"{-R + i * h:.8f}"
        assert fx == pytest.approx(Ffin[i, N // 2, N // 2], rel=1e-5)
        s = np.array([x, h / 2, h / 2])
        sd = (s + R) / (2 * R) - 0.5
        ref0 = 3.2 + sum(j / 55.0 * np.sin(j * sd * np.pi).sum() for j in range(11))
        assert f0 == pytest.approx(ref0, abs=1e-8)
