"""Factor-and-solve serving loop over the C-ABI.

Requests ask for the solution of A X = B for a system A that is either new
(factor, then solve) or already factored (solve only).  The server

* admits work by an HBM budget: before a factor is built its device bytes are
  computed on the host (taucs_amd_factor_device_bytes); if they do not fit,
  idle resident factors are evicted least-recently-used first, and otherwise
  the request waits until running work frees its factors;
* keeps orderings (per sparsity pattern), symbolic structures and numeric
  factors resident on the device under the request's system key, so repeat
  right-hand sides pay only for the solve;
* batches continuously: `slots` worker threads each take the next admissible
  job as soon as they are free; all queued solves against one factor are
  merged into one multi-right-hand-side solve; jobs on distinct factors run on
  their factors' own HIP streams and overlap on the device (the library calls
  release the interpreter lock);
* reports per-request latency (arrival to result) and the backward error of
  every solve; `summary()` gives p50/p99 and solves/s.

Each right-hand side of a merged solve follows its own recurrence (the solve
and refinement treat columns independently), so a result is the same whether
its request ran alone or in the mix.

The reference has no serving layer; the factor/solve calls it makes are
taucs_ccs_factor_llt_mf (src/taucs_sn_llt.c:2979-3006) and
taucs_supernodal_solve_llt (:2844-2885), with the symbolic/numeric split of
:2128-2209 for refactorization.
"""
import collections
import threading
import time

import numpy as np

from . import FP32, FP64, factor_symbolic, mesh2d, mesh3d, order, permute, random_uniform

GIB = 1 << 30


class Request:
    """One factor-and-solve or solve-only request."""

    __slots__ = ("rid", "key", "pattern", "make", "B", "nrhs", "solve_only", "t_arrive", "t_start",
                 "t_done", "X", "backward_error", "iterations", "error", "done", "factored")

    def __init__(self, rid, key, pattern, make, B, solve_only=False, nrhs=None):
        self.rid = rid
        self.key = key            # identity of the system (its factor's key)
        self.pattern = pattern    # identity of the sparsity pattern (ordering cache key)
        self.make = make          # () -> unpermuted taucs matrix of the system
        self.B = None if B is None else \
            np.asfortranarray(np.asarray(B, dtype=np.float64).reshape(B.shape[0], -1))
        self.nrhs = self.B.shape[1] if self.B is not None else nrhs
        self.solve_only = solve_only
        self.t_arrive = None
        self.t_start = None
        self.t_done = None
        self.X = None
        self.backward_error = None
        self.iterations = None
        self.error = None
        self.factored = False     # this request paid for a factorization
        self.done = threading.Event()

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrive


class _Resident:
    __slots__ = ("key", "F", "PA", "perm", "invperm", "bytes", "busy", "last_use", "M", "anorm")

    def __init__(self, key, F, PA, perm, invperm, nbytes):
        self.key, self.F, self.PA, self.perm, self.invperm = key, F, PA, perm, invperm
        self.bytes = nbytes
        self.busy = True
        self.last_use = time.perf_counter()
        self.M = None             # host CSR of the full permuted matrix (error checks)
        self.anorm = None


class Server:
    """Continuous-batching request loop with admission by HBM budget."""

    def __init__(self, hbm_budget=32 * GIB, slots=4, precision=FP64, tol=1e-12, method=0,
                 maxit=50, max_rhs=64, symbolic=factor_symbolic):
        self.budget = int(hbm_budget)
        # symbolic(PA) -> factor handle with device_bytes/numeric/solve_refined/free
        self.symbolic = symbolic
        self.precision = precision
        self.tol, self.method, self.maxit, self.max_rhs = tol, method, maxit, max_rhs
        self.reserved = 0
        self.resident = collections.OrderedDict()   # key -> _Resident (LRU order)
        self.building = {}                           # key -> reserved bytes while factoring
        self.orderings = {}                          # pattern -> (perm, invperm)
        self.pending = collections.deque()
        self.finished = []
        self.stats = collections.Counter()
        self.cv = threading.Condition()
        self.closing = False
        self.workers = [threading.Thread(target=self._worker, daemon=True, name=f"slot{i}")
                        for i in range(slots)]
        for w in self.workers:
            w.start()

    # ---- client side -----------------------------------------------------
    def submit(self, req):
        with self.cv:
            req.t_arrive = time.perf_counter()
            self.pending.append(req)
            self.cv.notify_all()
        return req

    def drain(self):
        with self.cv:
            while self.pending or any(r.busy for r in self.resident.values()) or self.building:
                self.cv.wait(0.05)

    def close(self):
        self.drain()
        with self.cv:
            self.closing = True
            self.cv.notify_all()
        for w in self.workers:
            w.join()
        for r in self.resident.values():
            r.F.free()
        self.resident.clear()
        self.reserved = 0

    def summary(self):
        lat = np.array([r.latency for r in self.finished if r.error is None])
        ok = [r for r in self.finished if r.error is None]
        if not ok:
            return {}
        t0 = min(r.t_arrive for r in ok)
        t1 = max(r.t_done for r in ok)
        nrhs = sum(r.B.shape[1] for r in ok)
        return {
            "requests": len(ok),
            "failed": sum(1 for r in self.finished if r.error is not None),
            "rhs": nrhs,
            "wall_s": t1 - t0,
            "solves_per_s": nrhs / (t1 - t0) if t1 > t0 else None,
            "requests_per_s": len(ok) / (t1 - t0) if t1 > t0 else None,
            "latency_p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "latency_p99_ms": 1e3 * float(np.percentile(lat, 99)),
            "latency_max_ms": 1e3 * float(lat.max()),
            "max_backward_error": max(r.backward_error for r in ok),
            "max_refine_iterations": max(r.iterations for r in ok),
            "factorizations": self.stats["factor"],
            "evictions": self.stats["evict"],
            "merged_solves": self.stats["merged"],
            "solve_jobs": self.stats["solve_jobs"],
            "peak_reserved_gib": self.stats["peak_reserved"] / GIB,
            "budget_gib": self.budget / GIB,
        }

    # ---- scheduling (under self.cv) ----------------------------------------
    def _take_solves(self, key):
        """All queued requests for `key`, up to max_rhs columns (first always taken)."""
        taken, cols, keep = [], 0, collections.deque()
        for r in self.pending:
            if r.key == key and (not taken or cols + r.B.shape[1] <= self.max_rhs):
                taken.append(r)
                cols += r.B.shape[1]
            else:
                keep.append(r)
        self.pending = keep
        return taken

    def _next_job(self):
        for r in self.pending:
            ent = self.resident.get(r.key)
            if ent is not None:
                if ent.busy:
                    continue          # merged into the next job on this factor
                ent.busy = True
                self.resident.move_to_end(r.key)
                return ("solve", ent, self._take_solves(r.key))
            if r.key in self.building:
                continue              # its factor is being built by another slot
            self.building[r.key] = 0
            self.pending.remove(r)
            return ("factor", None, [r])
        return None

    def _reserve(self, key, need):
        """Wait until `need` bytes fit the budget, evicting idle factors (LRU)."""
        while True:
            if self.reserved + need <= self.budget:
                self.reserved += need
                self.building[key] = need
                self.stats["peak_reserved"] = max(self.stats["peak_reserved"], self.reserved)
                return True
            victim = next((e for e in self.resident.values() if not e.busy), None)
            if victim is not None:
                del self.resident[victim.key]
                self.reserved -= victim.bytes
                victim.F.free()
                self.stats["evict"] += 1
                continue
            others = any(e.busy for e in self.resident.values()) or \
                any(v > 0 for k, v in self.building.items() if k != key)
            if not others:
                return False          # larger than the whole budget
            self.cv.wait(0.05)

    # ---- worker ------------------------------------------------------------
    def _worker(self):
        while True:
            with self.cv:
                job = self._next_job()
                while job is None:
                    if self.closing:
                        return
                    self.cv.wait(0.05)
                    job = self._next_job()
            kind, ent, reqs = job
            now = time.perf_counter()
            for r in reqs:
                r.t_start = now
            try:
                if kind == "factor":
                    ent = self._factor(reqs[0])
                    if ent is not None:
                        with self.cv:
                            # requests for this key queued meanwhile join the first solve
                            reqs = reqs + self._take_solves(ent.key)
                if ent is not None:
                    self._solve(ent, reqs)
            except Exception as e:  # report, never lose a request
                for r in reqs:
                    r.error = repr(e)
                if kind == "factor" and ent is None:
                    self._fail(reqs[0], repr(e))
            finally:
                with self.cv:
                    if ent is not None:
                        ent.busy = False
                        ent.last_use = time.perf_counter()
                    now = time.perf_counter()
                    for r in reqs:
                        if r.t_done is None:
                            r.t_done = now
                        self.finished.append(r)
                        r.done.set()
                    self.cv.notify_all()

    def _factor(self, r):
        with self.cv:
            cached = self.orderings.get(r.pattern)
        A = r.make()
        if cached is None:
            cached = order(A, "genmmd")
            with self.cv:
                self.orderings[r.pattern] = cached
        perm, invperm = cached
        PA = permute(A, perm, invperm)
        A.free()
        F = self.symbolic(PA)
        if F is None:
            return self._fail(r, "symbolic analysis failed")
        need = F.device_bytes(self.precision, min(self.max_rhs, 16), refine=True)
        with self.cv:
            ok = self._reserve(r.key, need)
        if not ok:
            F.free()
            return self._fail(r, f"system needs {need} device bytes, budget is {self.budget}")
        if F.numeric(PA, self.precision) != 0:
            F.free()
            return self._fail(r, "numeric factorization failed")
        r.factored = True
        ent = _Resident(r.key, F, PA, perm, invperm, need)
        with self.cv:
            del self.building[r.key]
            self.resident[r.key] = ent
            self.stats["factor"] += 1
        return ent

    def _fail(self, r, msg):
        with self.cv:
            v = self.building.pop(r.key, 0)
            if v:
                self.reserved -= v
            r.error = msg
        return None

    def _solve(self, ent, reqs):
        cols = [r.B.shape[1] for r in reqs]
        Bp = np.asfortranarray(np.concatenate([r.B[ent.perm, :] for r in reqs], axis=1))
        X, info = ent.F.solve_refined(ent.PA, Bp, method=self.method, tol=self.tol, factored=True,
                                      maxit=self.maxit)
        X = X.reshape(Bp.shape)
        # per-request backward error in the original ordering (host, FP64)
        now = time.perf_counter()
        c0 = 0
        for r, k in zip(reqs, cols):
            r.X = X[:, c0:c0 + k][ent.invperm, :]
            r.iterations = info.iterations
            c0 += k
            r.t_done = now
        self._backward_errors(ent, reqs)
        with self.cv:
            self.stats["solve_jobs"] += 1
            if len(reqs) > 1:
                self.stats["merged"] += len(reqs) - 1

    @staticmethod
    def _backward_errors(ent, reqs):
        """Normwise backward error |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf) of
        every column, in FP64 on the host, independent of the device's check."""
        if ent.M is None:
            ent.M = ent.PA.to_scipy_full()
            ent.anorm = float(abs(ent.M).sum(axis=1).max())
        for r in reqs:
            Xp = r.X[ent.perm, :]
            Bp = r.B[ent.perm, :]
            R = Bp - ent.M @ Xp
            den = ent.anorm * np.abs(Xp).max(axis=0) + np.abs(Bp).max(axis=0)
            r.backward_error = float((np.abs(R).max(axis=0) / den).max())


# ---- the request mix ----------------------------------------------------------

def make_mix(count, rate, seed=0, frac_solve=0.4, frac_2d=0.7, side2=(32, 256), side3=(8, 40),
             rhs=(1, 16), materialize=True):
    """Poisson arrivals (rate per second): 70% 2D 5-point grids with side in
    [32, 256], 30% 3D 7-point grids with side in [8, 40], 1-16 uniform[-1, 1]
    right-hand sides, 40% of requests solve-only against a factor an earlier
    request built.  Returns [(arrival offset s, Request)]; with materialize
    False the right-hand sides are not generated (only their count)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out, systems = [], []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        if systems and rng.random() < frac_solve:
            key, pattern, make, n = systems[int(rng.integers(len(systems)))]
            solve_only = True
        else:
            if rng.random() < frac_2d:
                s = int(rng.integers(side2[0], side2[1] + 1))
                pattern, n = ("2d", s), s * s
                make = (lambda s=s: mesh2d(s))
            else:
                s = int(rng.integers(side3[0], side3[1] + 1))
                pattern, n = ("3d", s), s ** 3
                make = (lambda s=s: mesh3d(s))
            key = ("sys", i)
            systems.append((key, pattern, make, n))
            solve_only = False
        k = int(rng.integers(rhs[0], rhs[1] + 1))
        B = np.stack([random_uniform(n, 1_000_003 * (i + 1) + c) for c in range(k)], 1) \
            if materialize else None
        out.append((t, Request(i, key, pattern, make, B, solve_only, nrhs=k)))
    return out


def run_mix(mix, server, time_scale=1.0):
    """Replay arrivals in real time (open loop) and wait for every result."""
    t0 = time.perf_counter()
    for t, r in mix:
        delay = t0 + t * time_scale - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        server.submit(r)
    for _, r in mix:
        r.done.wait()
    server.drain()
    return server.summary()


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="serving loop over a Poisson request mix")
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--rate", type=float, default=50.0, help="arrivals per second")
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--budget-gib", type=float, default=16.0)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    prec = FP64 if a.precision == "fp64" else FP32
    srv = Server(hbm_budget=int(a.budget_gib * GIB), slots=a.slots, precision=prec)
    s = run_mix(make_mix(a.requests, a.rate, a.seed), srv)
    srv.close()
    s["precision"] = a.precision
    print(json.dumps(s))


if __name__ == "__main__":
    main()
