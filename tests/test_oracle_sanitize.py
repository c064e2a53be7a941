"""The CPU oracle (the checker every parity test relies on) built with
-fsanitize=address,undefined and driven through every entry point on odd and
even grids by tests/cpp/oracle_sanitize_main.c: it must finish with no
AddressSanitizer, LeakSanitizer or UndefinedBehaviorSanitizer report and with
all of the driver's own consistency checks holding."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_oracle_under_sanitizers(tmp_path):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc is not available")
    exe = tmp_path / "oracle_sanitize"
    cmd = [gcc, "-O1", "-g", "-std=c11", "-ffp-contract=off", "-fopenmp", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-I", os.path.join(ROOT, "oracle"),
           os.path.join(ROOT, "oracle", "edb_oracle.c"), os.path.join(HERE, "cpp", "oracle_sanitize_main.c"),
           "-lm", "-o", str(exe)]
    build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-4000:]
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=600)
    report = run.stderr[-4000:]
    assert run.returncode == 0, report
    for marker in ("AddressSanitizer", "LeakSanitizer", "runtime error"):
        assert marker not in run.stderr, report
    assert "all checks passed" in run.stdout
