"""Serving loop: admission by HBM budget, resident factors, continuous batching.

The CPU tests drive the scheduler with a stand-in factor (dense host solves
with a delay) so admission, eviction, merging and per-factor exclusivity are
checked without a device; the GPU tests run the real C-ABI path and check that
every result is the same alone and in the mix."""
import threading
import time
from types import SimpleNamespace

import numpy as np
import pytest

import taucs_amd as T
from taucs_amd import serving as S


class FakeFactor:
    """Stands in for a device factor: bytes proportional to n, dense solves
    with a delay, and a check that no two threads use one factor at once."""

    live = 0
    peak = 0
    lock = threading.Lock()

    def __init__(self, PA, delay):
        self.M = PA.to_scipy_full().toarray()
        self.n = PA.n
        self.delay = delay
        self.active = False
        self.freed = False
        self.bytes = None

    def device_bytes(self, precision, nrhs, refine=False):
        return 1000 * self.n

    def numeric(self, PA, precision):
        self.bytes = 1000 * self.n
        with FakeFactor.lock:
            FakeFactor.live += self.bytes
            FakeFactor.peak = max(FakeFactor.peak, FakeFactor.live)
        time.sleep(self.delay)
        return 0

    def solve_refined(self, PA, B, method=0, tol=1e-12, maxit=50, factored=False):
        assert not self.active, "two threads on one factor"
        assert not self.freed, "solve on an evicted factor"
        self.active = True
        try:
            time.sleep(self.delay)
            X = np.linalg.solve(self.M, B)
        finally:
            self.active = False
        return X, SimpleNamespace(iterations=0, converged=1, backward_error=0.0)

    def free(self):
        assert not self.active
        self.freed = True
        if self.bytes:
            with FakeFactor.lock:
                FakeFactor.live -= self.bytes
            self.bytes = None


def fake_backend(delay):
    FakeFactor.live = FakeFactor.peak = 0
    return lambda PA: FakeFactor(PA, delay)


def small_mix(count, seed, rate=2000.0, frac_solve=0.5):
    return S.make_mix(count, rate, seed=seed, frac_solve=frac_solve, side2=(6, 14), side3=(3, 5),
                      rhs=(1, 4))


def solution_ok(r, tol=1e-10):
    A = r.make()
    M = A.to_scipy_full()
    return np.abs(M @ r.X - r.B).max() <= tol * (1 + np.abs(r.B).max())


def test_budget_is_never_exceeded_and_idle_factors_are_evicted():
    mix = small_mix(60, seed=1)
    sizes = {r.key: r.B.shape[0] for _, r in mix}
    budget = 1000 * 3 * max(sizes.values())      # room for about three factors
    srv = S.Server(hbm_budget=budget, slots=4, symbolic=fake_backend(0.002))
    summ = S.run_mix(mix, srv)
    srv.close()
    assert summ["failed"] == 0 and summ["requests"] == len(mix)
    assert FakeFactor.peak <= budget
    assert srv.stats["peak_reserved"] <= budget
    assert summ["evictions"] > 0
    assert FakeFactor.live == 0
    for _, r in mix:
        assert solution_ok(r), r.rid


def test_queued_solves_merge_into_one_job():
    mix = small_mix(40, seed=2, frac_solve=0.9)
    srv = S.Server(hbm_budget=1 << 40, slots=2, symbolic=fake_backend(0.01))
    summ = S.run_mix(mix, srv)
    srv.close()
    assert summ["failed"] == 0
    assert summ["merged_solves"] > 0
    assert summ["solve_jobs"] < len(mix)
    for _, r in mix:
        assert solution_ok(r), r.rid


def test_request_larger_than_budget_fails_alone():
    mix = small_mix(20, seed=3)
    big = max((r for _, r in mix if not r.solve_only), key=lambda r: r.B.shape[0])
    budget = 1000 * big.B.shape[0] - 1           # every other system fits
    srv = S.Server(hbm_budget=budget, slots=3, symbolic=fake_backend(0.001))
    S.run_mix(mix, srv)
    srv.close()
    failed = [r for _, r in mix if r.error is not None]
    assert failed and all(r.key == big.key for r in failed)
    assert all("budget" in r.error for r in failed)
    for _, r in mix:
        if r.error is None:
            assert solution_ok(r)


def test_mix_statistics_follow_the_stated_distribution():
    mix = S.make_mix(2000, 100.0, seed=5, materialize=False)
    solve_only = np.mean([r.solve_only for _, r in mix])
    new = [r for _, r in mix if not r.solve_only]
    two_d = np.mean([r.pattern[0] == "2d" for r in new])
    k = np.array([r.nrhs for _, r in mix])
    gaps = np.diff([t for t, _ in mix])
    assert abs(solve_only - 0.4) < 0.04
    assert abs(two_d - 0.7) < 0.04
    assert k.min() == 1 and k.max() == 16
    assert abs(gaps.mean() - 0.01) < 0.001
    sides2 = [r.pattern[1] for r in new if r.pattern[0] == "2d"]
    sides3 = [r.pattern[1] for r in new if r.pattern[0] == "3d"]
    assert min(sides2) >= 32 and max(sides2) <= 256
    assert min(sides3) >= 8 and max(sides3) <= 40
    few = S.make_mix(20, 100.0, seed=5)
    assert all(r.B.shape == (r.make().n, r.nrhs) for _, r in few)
    assert all(np.abs(r.B).max() <= 1.0 for _, r in few)
    assert [r.nrhs for _, r in few] == [r.nrhs for _, r in mix[:20]]


def test_device_bytes_estimate_is_host_only():
    A = T.mesh3d(12)
    perm, inv = T.order(A, "genmmd")
    F = T.factor_symbolic(T.permute(A, perm, inv))
    fi = F.info()
    b64 = F.device_bytes(T.FP64, 4, refine=True)
    b32 = F.device_bytes(T.FP32, 4, refine=True)
    assert b64 > fi.relax_nnz * 8
    assert b32 < b64
    assert F.device_bytes(T.FP64, 4, refine=False) < b64


# ---- GPU: the real path --------------------------------------------------------

def alone(r, precision):
    srv = S.Server(hbm_budget=64 * S.GIB, slots=1, precision=precision)
    q = S.Request(r.rid, r.key, r.pattern, r.make, r.B, r.solve_only)
    srv.submit(q)
    q.done.wait()
    srv.close()
    assert q.error is None, q.error
    return q


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [T.FP64, T.FP32])
def test_results_are_the_same_alone_and_in_the_mix(precision):
    print(f"case: serving mix vs alone, precision {precision}", flush=True)
    mix = S.make_mix(40, 400.0, seed=11, side2=(16, 64), side3=(6, 14), rhs=(1, 8))
    srv = S.Server(hbm_budget=64 * S.GIB, slots=4, precision=precision)
    summ = S.run_mix(mix, srv)
    srv.close()
    print(summ, flush=True)
    assert summ["failed"] == 0
    assert summ["max_backward_error"] <= 1e-12
    for _, r in mix:
        q = alone(r, precision)
        assert np.array_equal(q.X, r.X), r.rid
        assert q.backward_error <= 1e-12


@pytest.mark.gpu
def test_budget_admission_on_the_device():
    print("case: serving under a tight HBM budget", flush=True)
    mix = S.make_mix(30, 400.0, seed=12, side2=(16, 48), side3=(6, 12), rhs=(1, 4))
    need = []
    for _, r in mix:
        if not r.solve_only:
            A = r.make()
            perm, inv = T.order(A, "genmmd")
            F = T.factor_symbolic(T.permute(A, perm, inv))
            need.append(F.device_bytes(T.FP64, 16, refine=True))
            F.free()
    budget = 2 * max(need)
    base = T.lib.taucs_amd_device_bytes_in_use()
    srv = S.Server(hbm_budget=budget, slots=3)
    summ = S.run_mix(mix, srv)
    live = T.lib.taucs_amd_device_bytes_in_use() - base
    assert live <= srv.reserved                      # estimates bound the real bytes
    srv.close()
    assert summ["failed"] == 0 and summ["evictions"] > 0
    assert summ["peak_reserved_gib"] * S.GIB <= budget
    assert T.lib.taucs_amd_device_bytes_in_use() == base
