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
    F = ld("fft_F.bin", np.complex128).reshape(8, 8, 8)
    Fh = ld("fft_Fhat.bin", np.complex128).reshape(8, 8, 8)
    Ft = ld("fft_Ftest.bin", np.complex128).reshape(8, 8, 8)
    assert np.abs(Fh - np.fft.fftn(F)).max() <= 1e-12 * np.abs(Fh).max()
    assert np.abs(Ft - F).max() <= 1e-13 * np.abs(F).max()
    # BKW3D RK4 with the diagnostic fix vs the oracle driver
    Fb = ld("bkw_F.bin").reshape(N, N, N)
    h = 16.0 / N
    x = -8.0 + h * np.arange(N)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    s = 1 - 0.4
    v2 = X ** 2 + Y ** 2 + Z ** 2
    F0 = (2 * np.pi * s) ** -1.5 * np.exp(-v2 / (2 * s)) * ((5 * s - 3) / (2 * s) + (1 - s) / (2 * s * s) * v2)
    Bn = oracle.genB(N, 0, 8.0, False)
    fo, t, H, Hf = oracle.run(N, Bn, 8.0, "RK4", 0.0, 0.1, 0.02, np.fft.fftn(F0), fix_mode=1)
    ref_F = np.fft.ifftn(fo).real
    if not (np.isfinite(H[-1]) and Hf[-1] != H[-1]):  # no fix at the last step: F() is the transform of the state
        assert np.abs(Fb - ref_F).max() <= 1e-12 * np.abs(ref_F).max()
    else:  # F() holds the fixed values: same mass, entropy Hfix
        assert abs(Fb.mean() - ref_F.mean()) <= 1e-12 * abs(ref_F.mean())
        assert oracle.cal_entropy(N, 8.0, Fb.astype(np.complex128)) == pytest.approx(Hf[-1], rel=1e-9)
    np.testing.assert_allclose(ld("bkw_H.bin"), Hf, rtol=1e-9)
