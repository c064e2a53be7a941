"""Continuous batching (edb_server_*): results are the same alone and in the
mix, admission by HBM budget, mixed (N, alpha) groups overlapping."""
import numpy as np
import pytest

from conftest import maxwellian_mixture

pytestmark = pytest.mark.gpu

L, DT = 8.0, 0.01


def solve_alone(edb, N, alpha, t_end, F0):
    s = edb.Solver(N, alpha, L, False, 1, edb.PREC_FP32_BATCHED, "RK4", DT, 0.0, edb.FIX_APPLIED,
                   hermitian=True)
    s.set_initial_real(F0[None])
    n = s.run(t_end)
    t, steps = s.time()
    assert steps == n
    return n, t, s.get_real()[0]


def drain(srv, expect, max_polls=100000):
    out = {}
    for _ in range(max_polls):
        if srv.poll():
            while True:
                try:
                    rid, N, steps, t, F = srv.fetch()
                except Exception:
                    break
                out[rid] = (N, steps, t, F)
        if len(out) >= expect:
            return out
    raise AssertionError("server did not finish")


def test_server_results_equal_solo_solves(edb):
    rng = np.random.default_rng(7)
    srv = edb.Server(L=L, dt=DT, fix_mode=edb.FIX_APPLIED, max_slots=64, max_chunk=3)
    reqs = {}
    for rid in range(30):
        N = int(rng.choice([9, 16]))
        alpha = int(rng.integers(0, 2))
        t_end = float(rng.uniform(0.03, 0.12))
        F0 = maxwellian_mixture(N, L, 500 + rid)
        reqs[rid] = (N, alpha, t_end, F0)
        assert srv.submit(rid, N, alpha, t_end, F0) == 0
        if rid % 7 == 6:  # requests arriving while others run
            srv.poll()
    done = drain(srv, len(reqs))
    st = srv.stats()
    assert st["finished"] == 30 and st["active"] == 0 and st["waiting"] == 0
    for rid in (0, 5, 13, 22, 29):
        N, alpha, t_end, F0 = reqs[rid]
        n, ta, Fa = solve_alone(edb, N, alpha, t_end, F0)
        Ns, steps, t, Fs = done[rid]
        assert Ns == N and steps == n
        assert t == ta  # the accumulated t of the drivers' loop, not t_end
        assert np.abs(Fs - Fa).max() <= 1e-12 * np.abs(Fa).max(), rid
    # mass conservation of every served result
    for rid, (N, _a, _t, F0) in reqs.items():
        h3 = (2 * L / N) ** 3
        assert abs(done[rid][3].sum() * h3 - F0.sum() * h3) <= 1e-12


def test_server_hbm_budget(edb):
    F0 = maxwellian_mixture(9, L, 1)
    tiny = edb.Server(L=L, dt=DT, hbm_budget=1000)
    assert tiny.submit(1, 9, 0, 0.05, F0) == 2  # EDB_ENOMEM: can never fit
    one = edb.Server(L=L, dt=DT, max_slots=64, hbm_budget=0)
    assert one.submit(1, 9, 0, 0.05, F0) == 0
    used = one.stats()["hbm_used"]
    # a budget that holds one N=9 group: a second group must wait until the first is idle
    srv = edb.Server(L=L, dt=DT, max_slots=64, hbm_budget=int(used * 1.5))
    assert srv.submit(1, 9, 0, 0.05, F0) == 0
    F1 = maxwellian_mixture(9, L, 2)
    assert srv.submit(2, 9, 1, 0.05, F1) == 6  # EDB_EBUSY while group (9, 0) is active
    drain(srv, 1)
    assert srv.submit(2, 9, 1, 0.05, F1) == 0  # idle group evicted
    out = drain(srv, 1)
    assert 2 in out


def test_server_admission_limit(edb):
    """max_pending turns requests away (EDB_EBUSY) while that many wait, and
    admits again once the backlog drains."""
    srv = edb.Server(L=L, dt=DT, fix_mode=edb.FIX_APPLIED, max_slots=64, max_chunk=2, max_pending=3)
    F0 = maxwellian_mixture(9, L, 900)
    st = [srv.submit(i, 9, 0, 0.02, F0) for i in range(5)]
    assert st[:3] == [0, 0, 0] and st[3] != 0 and st[4] != 0
    done = drain(srv, 3)
    assert sorted(done) == [0, 1, 2]
    assert srv.submit(10, 9, 0, 0.02, F0) == 0
    assert 10 in drain(srv, 1)


def test_server_bench_sizes_equal_solo_solves(edb):
    """Groups at the serving workload's larger sizes, N = 24 and 32 with
    gamma 0 and 1, requests submitted while others are on the device: each
    served result equals its solo solve to 1e-12, with the same steps and t."""
    rng = np.random.default_rng(2432)
    srv = edb.Server(L=L, dt=DT, fix_mode=edb.FIX_APPLIED, max_chunk=2)
    reqs = {}
    for rid in range(12):
        N = (24, 32)[rid % 2]
        alpha = (rid // 2) % 2
        t_end = float(rng.uniform(0.015, 0.05))
        F0 = maxwellian_mixture(N, L, 2400 + rid)
        reqs[rid] = (N, alpha, t_end, F0)
        assert srv.submit(rid, N, alpha, t_end, F0) == 0
        srv.poll()  # arrivals interleaved with chunks of the groups already running
    done = drain(srv, len(reqs))
    assert srv.stats()["finished"] == 12
    for rid in (0, 1, 2, 3, 10, 11):  # every (N, alpha), first and last arrivals
        N, alpha, t_end, F0 = reqs[rid]
        n, ta, Fa = solve_alone(edb, N, alpha, t_end, F0)
        Ns, steps, t, Fs = done[rid]
        assert Ns == N and steps == n and t == ta
        assert np.abs(Fs - Fa).max() <= 1e-12 * np.abs(Fa).max(), rid
