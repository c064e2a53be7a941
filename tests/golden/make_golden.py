"""Write the golden vectors of the CalQ path from the reference's own code.

Run in a checkout where the reference repository is mounted (it is compiled
into oracle/_ref by ``make -C oracle ref``).  The fixtures it writes (data
only: inputs and the reference's outputs) are committed next to this script
and are what the CPU and GPU tests compare against; the reference itself is
not needed afterwards.

    python tests/golden/make_golden.py

Contents
  golden_genB.npz   QCalculator::genB (QCalculator.h:158-229), odd N, alpha
                    0/1, filter on/off: sha256 of the N^6 doubles, sums and
                    64 sampled entries per case.
  golden_calQ.npz   QCalculator::calQ (QCalculator.h:231-257) on seeded
                    complex inputs: the full outputs.
  golden_efix.npz   EntropyFix (EntropyFix.h:124-253): the testEntropyFix.cpp
                    sequence for N = 1000 and 13824, and FixSol on seeded
                    vectors.
"""
import ctypes
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

_dp = ctypes.POINTER(ctypes.c_double)


def P(a):
    return a.ctypes.data_as(_dp)


def main():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    r = oracle.ref()
    out = {}
    # ---- genB
    g = {}
    for N in (3, 5, 7, 9):
        for a in (0, 1):
            for fil in (0, 1):
                B = np.empty(N ** 6)
                r.edbr_genB(N, a, fil, P(B))
                key = f"N{N}_a{a}_f{fil}"
                rng = np.random.default_rng(1000 + N * 10 + a * 2 + fil)
                idx = rng.choice(N ** 6, size=64, replace=False)
                g[key + "_sha256"] = np.frombuffer(hashlib.sha256(B.astype("<f8").tobytes()).digest(),
                                                   dtype=np.uint8)
                g[key + "_sum"] = np.array([B.sum(), np.abs(B).sum()])
                g[key + "_idx"] = idx.astype(np.int64)
                g[key + "_val"] = B[idx]
    np.savez_compressed(os.path.join(HERE, "golden_genB.npz"), **g)
    # ---- calQ
    c = {}
    cases = [(N, a, fil) for N in (3, 5, 7, 9) for a in (0, 1) for fil in (0, 1)]
    cases += [(15, 0, 0), (15, 1, 1)]
    for (N, a, fil) in cases:
        rng = np.random.default_rng(2000 + N * 10 + a * 2 + fil)
        f = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
        f = np.ascontiguousarray(f)
        q = np.empty_like(f)
        r.edbr_calQ(N, a, fil, P(f.view(np.float64)), P(q.view(np.float64)))
        key = f"N{N}_a{a}_f{fil}"
        c[key + "_fin"] = f
        c[key + "_q"] = q
    # integer inputs as in CalQ/test.cpp:19-21 (values 0..9)
    rng = np.random.default_rng(7)
    f = (rng.integers(0, 10, (5, 5, 5)) + 1j * rng.integers(0, 10, (5, 5, 5))).astype(np.complex128)
    q = np.empty_like(f)
    r.edbr_calQ(5, 0, 1, P(f.view(np.float64)), P(q.view(np.float64)))
    c["int_N5_a0_f1_fin"] = f
    c["int_N5_a0_f1_q"] = q
    np.savez_compressed(os.path.join(HERE, "golden_calQ.npz"), **c)
    # ---- EntropyFix
    e = {}
    for N in (1000, 13824):  # testEntropyFix.cpp:65-81
        f1 = np.arange(N, dtype=np.float64)
        f2 = np.empty(N)
        f2[: N // 2] = np.floor((N - 1.0) / 2.0)
        f2[N // 2:] = np.ceil((N - 1.0) / 2.0)
        res = np.empty(3)
        r.edbr_entropy_sequence(N, P(f1), P(f2), P(res))
        e[f"seq{N}_H"] = res
        e[f"seq{N}_fixed_sha256"] = np.frombuffer(hashlib.sha256(f1.astype("<f8").tobytes()).digest(),
                                                  dtype=np.uint8)
        e[f"seq{N}_fixed_head"] = f1[:16].copy()
    for k, N in enumerate((64, 729, 4096)):
        rng = np.random.default_rng(3000 + k)
        f = rng.gamma(2.0, 0.05, N)
        f[rng.choice(N, N // 10, replace=False)] = 0.0  # zeros, below eps
        avg = f.mean()
        Hf = np.sum(np.where(f > 1e-10, f * np.log(np.where(f > 0, f, 1.0)), 0.0)) / N - avg
        Hc = avg * np.log(avg) - avg
        aim = Hc + 0.3 * (Hf - Hc)  # between the constant state and f
        out2 = np.empty(2)
        fx = f.copy()
        r.edbr_fixsol_ptr(N, P(fx), aim, P(out2))
        e[f"fix{N}_f"] = f
        e[f"fix{N}_aim"] = np.array([aim])
        e[f"fix{N}_out"] = fx
        e[f"fix{N}_avg_H"] = out2
    np.savez_compressed(os.path.join(HERE, "golden_efix.npz"), **e)
    meta = {"generator": "tests/golden/make_golden.py", "reference_lib": "oracle/_ref/libedbref.so",
            "compiler": subprocess.run(["g++", "--version"], capture_output=True, text=True).stdout.splitlines()[0],
            "flags": "-O3 -std=c++14 -ffp-contract=off"}
    with open(os.path.join(HERE, "golden_meta.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    for fn in ("golden_genB.npz", "golden_calQ.npz", "golden_efix.npz"):
        print(fn, os.path.getsize(os.path.join(HERE, fn)))


if __name__ == "__main__":
    main()
