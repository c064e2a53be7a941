"""The C++ host side (include/edb.hpp) driven like the reference's own test
mains (tests/cpp/mirror_main.cpp), checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "mirror_main")


def test_cpp_mirror_matches_oracle(tmp_path, edb):
    assert os.path.exists(EXE), "build() builds tests/cpp/mirror_main"
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    # testEntropyFix.cpp prints for N = 1000 (2 decimals)
    assert "H(f^{n+1}) = 2700.92" in r.stdout and "H(hatf^{n+1}) = 2604.20" in r.stdout
    assert "rejected:" in r.stdout
    ld = lambda n, dt=np.float64: np.fromfile(tmp_path / n, dtype=dt)
    N = 9
    fin = ld("calq_fin.bin", np.complex128).reshape(N, N, N)
    q = ld("calq_q.bin", np.complex128).reshape(N, N, N)
    B = oracle.genB(N, 0, 6.62132, True)  # QCalculator(N, 0): filter on, R = 6.62132
    ref = oracle.calQ(N, B, fin)
    assert np.abs(q - ref).max() <= 1e-12 * np.abs(ref).max()
    qc = ld("calq_q_compressed.bin", np.complex128).reshape(N, N, N)  # calQCompressed, filter on
    assert np.abs(qc - ref).max() <= 1e-12 * np.abs(ref).max()
    F = ld("fft_F.bin", np.complex128).reshape(8, 8, 8)
    Fh = ld("fft_Fhat.bin", np.complex128).reshape(8, 8, 8)
    Ft = ld("fft_Ftest.bin", np.complex128).reshape(8, 8, 8)
    assert np.abs(Fh - np.fft.fftn(F)).max() <= 1e-12 * np.abs(Fh).max()
    assert np.abs(Ft - F).max() <= 1e-13 * np.abs(F).max()
    # default-constructed BKW3D RK4 (filter on) with the diagnostic fix vs the oracle driver
    Fb = ld("bkw_F.bin").reshape(N, N, N)
    h = 16.0 / N
    x = -8.0 + h * np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    s = 1 - 0.4
    v2 = X ** 2 + Y ** 2 + Z ** 2
    F0 = (2 * np.pi * s) ** -1.5 * np.exp(-v2 / (2 * s)) * ((5 * s - 3) / (2 * s) + (1 - s) / (2 * s * s) * v2)
    Bn = oracle.genB(N, 0, 8.0, True)  # BKW3D(N0, R) builds QCalculator(N0, 0): filter on
    fo, t, H, Hf = oracle.run(N, Bn, 8.0, "RK4", 0.0, 0.1, 0.02, np.fft.fftn(F0), fix_mode=1)
    ref_F = np.fft.ifftn(fo).real
    if not (np.isfinite(H[-1]) and Hf[-1] != H[-1]):  # no fix at the last step: F() is the transform of the state
        assert np.abs(Fb - ref_F).max() <= 1e-12 * np.abs(ref_F).max()
    else:  # F() holds the fixed values: same mass, entropy Hfix
        assert abs(Fb.mean() - ref_F.mean()) <= 1e-12 * abs(ref_F.mean())
        assert oracle.cal_entropy(N, 8.0, Fb.astype(np.complex128)) == pytest.approx(Hf[-1], rel=1e-9)
    np.testing.assert_allclose(ld("bkw_H.bin"), Hf, rtol=1e-9)


SCHEMES = ["FEuler", "RK2", "SSP3", "RK4", "CFE", "DUFORT"]


def _table_rows(path):
    """Data rows of one driver table (lines that are not '#' comments), split on tabs."""
    rows, sections = [], []
    for line in open(path).read().splitlines():
        if line.startswith("# Start the evolution using"):
            sections.append(line)
        elif not line.startswith("#"):
            rows.append(line.split("\t"))
    return rows, sections


def test_cpp_driver_tables_match_oracle(tmp_path, edb):
    """BKW3D::setOutput: each driver's progress lines and per-step table in the
    reference's format (BKW.h:212-633), every printed number against the
    oracle's driver (diagnostic fix, which fires on this data)."""
    r = subprocess.run([EXE, str(tmp_path), "tables"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    N, R, dt = 11, 6.0, 0.05
    T0 = 6 * np.log(2.0)
    F0 = oracle.grid_bkw(N, R, 0.8).real
    B = oracle.genB(N, 0, R, False)
    mean0 = F0.mean()
    fixes = 0
    for sc in SCHEMES:
        rows, sections = _table_rows(tmp_path / f"table_{sc}.txt")
        fo, t, H, Hf = oracle.run(N, B, R, sc, T0, T0 + 0.4, dt, np.fft.fftn(F0), beta=2.0, fix_mode=1)
        name = {"FEuler": "Forward Euler", "CFE": "C-Forward Euler"}.get(sc, sc)
        if sc == "DUFORT":
            assert sections == ["# Start the evolution using RK2...", "# Start the evolution using DUFORT..."]
        else:
            assert sections == [f"# Start the evolution using {name}..."]
        fixed = (Hf != H) & np.isfinite(H)
        fixes += int(fixed.sum())
        if sc == "FEuler":  # row 00: t0 and calEntropy of the initial F
            assert rows[0][0] == "00" and float(rows[0][2]) == pytest.approx(
                oracle.cal_entropy(N, R, F0.astype(np.complex128)), abs=2e-8)
            prev = float(rows[0][2])
            rows = rows[1:]
        else:
            prev = 1.0
        assert len(rows) == len(t), sc
        for k, row in enumerate(rows):
            kk = k - 8 if (sc == "DUFORT" and k >= 8) else k  # DUFORT restarts its count
            if sc == "FEuler":
                assert int(row[0]) == k + 1
                row = row[1:]
            assert float(row[0]) == pytest.approx(t[k], abs=1e-8)
            if sc in ("RK2", "DUFORT", "CFE") and not (sc == "CFE" and kk == 0):
                m = float(row[1])
                if sc != "DUFORT":
                    assert m == pytest.approx(mean0, abs=2e-8)  # Q_0 = 0 keeps the mean
                row = [row[0]] + row[2:]
            assert float(row[1]) == pytest.approx(H[k], abs=2e-8), (sc, k)
            first = (sc == "FEuler" and k == 0) or (sc != "FEuler" and kk == 0)
            if first:
                assert row[2] == "-"
                prev = Hf[k]  # entropy0 = entropy after every row
                continue
            assert float(row[2]) == pytest.approx(prev - H[k], abs=2e-8), (sc, k)
            if fixed[k]:
                assert float(row[3]) == pytest.approx(Hf[k], abs=2e-8)
                assert float(row[4]) == pytest.approx(prev - Hf[k], abs=2e-8)
            elif sc == "FEuler":
                assert float(row[3]) == pytest.approx(H[k], abs=2e-8)
            else:
                assert row[3:] in ([], [""]), (sc, k, row)
            prev = Hf[k]
        # the driver's F after the last step: the fixed values if the fix ran
        Fb = np.fromfile(tmp_path / f"table_F_{sc}.bin").reshape(N, N, N)
        ref = np.fft.ifftn(fo).real
        if fixed[-1]:
            assert oracle.cal_entropy(N, R, Fb.astype(np.complex128)) == pytest.approx(Hf[-1], rel=1e-9)
        else:
            assert np.abs(Fb - ref).max() <= 1e-12 * np.abs(ref).max(), sc
    assert fixes >= 3  # FEuler, SSP3, RK4 and CFE each fix once on this data
