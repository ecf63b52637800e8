"""Benchmark of the MI355X supernodal multifrontal Cholesky (one GPU).

Workload (default): the plain FP64 GPU path on the 3D 7-point Laplacian of a
60x60x60 grid (n = 216,000), genmmd ordering, 8 right-hand sides uniform[-1,1]
(seeds 1..8).  One step = numeric factorization of P A P^T (symbolic analysis
done once, outside the timed region, as the reference's symbolic/numeric split
allows) + the forward/backward supernodal solve of the 8 right-hand sides, with
A's values and the right-hand sides resident in HBM when the step starts
(taucs_amd_factor_llt_numeric_device; the TAUCS entry point
taucs_ccs_factor_llt_numeric takes host values and adds their 7 MB copy).

Printed JSON line: value = factorization GFLOP/s (reference flop count: the
fundamental-supernode count of src/taucs_sn_llt.c:3911-3921, over the mean
numeric time); also solves/s, p50/p99 step latency, the mixed-precision
(FP32 fronts + FP64 refinement) figures on the same workload, the roofline of
the dominant kernel (k_gemm_tile128<double>, the FP64 MFMA trailing-update and
SYRK tile; the group of all tile-GEMM launches beside it) from device events, and the
CPU baseline (the oracle's C restatement of the reference's multifrontal factor and
supernodal solve, host BLAS per front, on the same workload, all threads and one).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12   # MI355X dense FP64 MFMA
PEAK_FP32 = 157.3e12  # MI355X dense FP32 MFMA


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(side=60, nrhs=8):
    """The reference's CPU numeric path restated in C (oracle.c: the
    multifrontal factor with LAPACK dpotrf / BLAS dtrsm, dsyrk per front, the
    extend-adds, and the supernodal solve, src/taucs_sn_llt.c:1086-1240,
    1831-1977, 2597-2885) on the same workload as the GPU line: mesh3d(side),
    genmmd, numeric factor + an nrhs-column solve, timed on the host with all
    BLAS threads and with one."""
    from oracle import oracle as O
    n, cp, ri, v = O.mesh3d(side, side, side)
    perm, inv = O.genmmd(n, cp, ri)
    pcp, pri, pv = O.permute(n, cp, ri, v, inv)
    B = np.stack([O.uniform(n, 1 + r) for r in range(nrhs)], 1)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    runs = {}
    for th in (threads, 1):
        O.blas_threads(th)
        M = O.MF(n, pcp, pri, pv)
        t0 = time.perf_counter()
        M.solve(B)
        ts = time.perf_counter() - t0
        runs[th] = (M.t_numeric, ts, M.fund_flops, M.parts)
        M.free()
        log(f"cpu baseline ({th} threads): factor {M.t_numeric:.2f} s, {nrhs}-rhs solve {ts:.2f} s, "
            f"parts {({k: round(x, 2) for k, x in M.parts.items()})}")
    tf, ts, fl, parts = runs[threads]
    tf1, ts1, _, parts1 = runs[1]
    return {"value": round(fl / tf / 1e9, 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
            "sample": f"the whole workload: mesh3d {side}^3 numeric factor (C restatement of the "
                      f"reference's multifrontal LL^T, host BLAS per front; fundamental flops "
                      f"{fl:.4e}) in {tf:.2f} s and the {nrhs}-rhs supernodal solve in {ts:.2f} s "
                      f"({nrhs / ts:.2f} solves/s)",
            "single_thread": {"factor_gflops": round(fl / tf1 / 1e9, 3), "factor_s": round(tf1, 3),
                              "solve_s": round(ts1, 3), "solves_per_s": round(nrhs / ts1, 2)},
            "factor_s": round(tf, 3), "solve_s": round(ts, 3),
            "solves_per_s": round(nrhs / ts, 2),
            "time_by_part_s": {k: round(x, 3) for k, x in parts.items()},
            "host": {"nproc": os.cpu_count(), "cpu_model": _cpu_model(), "blas": O.blas_config(),
                     "blas_threads": threads}}


def host_backward_error(M, X, B):
    """Max over columns of |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf), FP64 on the host."""
    X = X.reshape(B.shape)
    R = B - M @ X
    anorm = abs(M).sum(1).max()
    return float((abs(R).max(0) / (anorm * abs(X).max(0) + abs(B).max(0))).max())


def oracle_forward_error(T, X, gold):
    """Relative forward error of X against the oracle's FP64 factor-and-solve,
    from the committed fixture (the oracle's x through a 32-row random-sign
    sketch, and at 4096 sampled rows): |S(x - x_oracle)| / |S x_oracle| per
    column, which estimates |x - x_oracle| / |x_oracle| within a factor of
    about 1.5, and max_i |x_i - x_oracle,i| / |x_oracle| over the sampled rows.
    S is drawn from the same splitmix64 stream as the fixture's (seed
    777 + 7919 i), here with the product's generator."""
    X = np.asarray(X).reshape(X.shape[0], -1)
    n = X.shape[0]
    SX = np.empty((32, X.shape[1]))
    for i in range(32):
        SX[i] = np.where(T.random_uniform(n, 777 + 7919 * i, 0.0, 1.0) < 0.5, 1.0, -1.0) @ X
    sk = np.linalg.norm(SX - gold["sketch"], axis=0) / np.linalg.norm(gold["sketch"], axis=0)
    sa = np.abs(X[gold["sample_idx"]] - gold["sample_x"]).max(0) / gold["xnorm"]
    return {"sketch": float(sk.max()), "sampled_rows": float(sa.max())}


def bench_varcoef(T, side=80, nrhs=4, steps=2, seed=2024):
    """Variable-coefficient 3D operator (log-uniform conductivities over
    [1e-2, 1e2], shift 1e-6): plain FP64 factor + solve against FP32 fronts +
    FP64 refinement (classical and CG) and the BF16-update variant, each to
    backward error <= 1e-12, with the forward error against the FP64 result."""
    t0 = time.perf_counter()
    A = T.mesh3d_varcoef(side, side, side, seed)
    perm, inv = T.order(A, "genmmd")
    PA = T.permute(A, perm, inv)
    F = T.factor_symbolic(PA)
    fi = F.info()
    n = PA.n
    log(f"varcoef {side}^3: n {n}, symbolic {time.perf_counter() - t0:.1f} s, "
        f"fund flops {fi.fund_flops:.3e}")
    B = np.stack([T.random_uniform(n, 100 + r) for r in range(nrhs)], 1)
    M = PA.to_scipy_full()
    dV = T.DeviceArray(PA.arrays()[3])   # A's values resident in HBM
    out = {"workload": f"mesh3d_varcoef {side}^3 (n={n}), conductivities log-uniform [1e-2,1e2], "
                       f"shift 1e-6, genmmd, {nrhs} rhs", "fund_flops": fi.fund_flops}

    def run(prec, method, tol, maxit):
        tf, ts, its, bws, ok = [], [], [], [], True
        X = None
        for i in range(steps + 1):
            t_a = time.perf_counter()
            if F.numeric_device(dV.ptr, prec) != 0:
                return None
            t_b = time.perf_counter()
            if method is None:
                X = F.solve(B)
                info = None
            else:
                X, info = F.solve_refined(PA, B, method=method, tol=tol, maxit=maxit, factored=True)
            t_c = time.perf_counter()
            if i == 0:
                continue          # warm-up (graph capture, buffers)
            tf.append(t_b - t_a)
            ts.append(t_c - t_b)
            if info is not None:
                its.append(info.iterations)
                ok = ok and bool(info.converged)
        bwd = host_backward_error(M, X, B)
        return {"factor_ms": round(1e3 * float(np.mean(tf)), 2),
                "solve_ms": round(1e3 * float(np.mean(ts)), 2),
                "factor_gflops": round(fi.fund_flops / float(np.mean(tf)) / 1e9, 1),
                "iterations": max(its) if its else 0, "converged": ok,
                "backward_error": bwd, "X": X}

    # each precision runs at the schedule the autotuner picks for it
    sched = {}
    for name, prec in (("fp64", T.FP64), ("fp32", T.FP32), ("bf16upd", T.BF16UPD)):
        sched[prec] = F.autotune(PA, prec, reps=1)
        log(f"varcoef autotune {name}: {sched[prec]}")
    out["schedules"] = {name: {"panel_width": sched[p][0][0], "gemm_tile": sched[p][0][1],
                               "autotune_ms": {f"w{w}_t{t}": round(v, 2) for (w, t), v in sched[p][1].items()}}
                        for name, p in (("fp64", T.FP64), ("fp32", T.FP32), ("bf16upd", T.BF16UPD))}

    def use(prec):
        (w, t), _ = sched[prec]
        F.autotune(PA, prec, candidates=((w, t),), reps=1)

    gp = os.path.join(ROOT, "tests", "golden", "c3_varcoef_80.npz")
    gold = np.load(gp) if side == 80 and seed == 2024 and nrhs == 4 and os.path.exists(gp) else None
    use(T.FP64)
    r64 = run(T.FP64, None, 0, 0)
    out["fp64"] = r64
    x64 = r64["X"]
    if gold is not None:
        r64["forward_error_vs_oracle_fp64"] = oracle_forward_error(T, x64, gold)
    base = r64["factor_ms"] + r64["solve_ms"]
    for name, prec, method in (("fp32_ir", T.FP32, 0), ("fp32_pcg", T.FP32, 1),
                               ("bf16upd_ir", T.BF16UPD, 0), ("bf16upd_pcg", T.BF16UPD, 1)):
        use(prec)
        r = run(prec, method, 1e-12, 200)
        if r is None:
            out[name] = {"failed": "factorization broke down (pivot not positive)"}
            log(f"varcoef {name}: breakdown")
            continue
        r["forward_error_vs_fp64"] = float(np.linalg.norm(r["X"] - x64) / np.linalg.norm(x64))
        if gold is not None:
            r["forward_error_vs_oracle_fp64"] = oracle_forward_error(T, r["X"], gold)
        r["fp64_fallback_blocks"] = T.lib.taucs_amd_factor_fp64_fallbacks(F.h)
        r["speedup_factor_plus_solve_vs_fp64"] = round(base / (r["factor_ms"] + r["solve_ms"]), 3)
        out[name] = r
        log(f"varcoef {name}: {({k: v for k, v in r.items() if k != 'X'})}")
    for v in out.values():
        if isinstance(v, dict):
            v.pop("X", None)
    log(f"varcoef fp64: {out['fp64']}")
    F.free()
    dV.free()
    return out


def bench_batch(T, count=512, lo=16, hi=128, steps=3, seed=7):
    """512 independent shifted 2D Laplacians (side uniform on [16, 128],
    diagonal shift uniform[0,1]) in the same level launches: FP64 and FP32 +
    refinement, one uniform[-1,1] right-hand side per system."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    sides = rng.integers(lo, hi + 1, count)
    t0 = time.perf_counter()
    mats = []
    for q, s in enumerate(sides):
        A = T.mesh2d_shifted(int(s), seed=10_000 + q)
        perm, inv = T.order(A, "genmmd")
        mats.append(T.permute(A, perm, inv))
    t_order = time.perf_counter() - t0
    out = {"workload": f"{count} shifted 2D 5-point Laplacians, side U[{lo},{hi}], genmmd, 1 rhs each"}
    res = {}
    Mbd = None
    for name, prec in (("fp64", T.FP64), ("fp32_ir", T.FP32)):
        t0 = time.perf_counter()
        Bt = T.Batch(mats, precision=prec)
        t_setup = time.perf_counter() - t0
        tuned = Bt.autotune(reps=2)
        log(f"batch {name} autotune: {tuned}")
        n = Bt.n_total
        b = np.concatenate([T.random_uniform(m.n, 500 + q) for q, m in enumerate(mats)])
        if Mbd is None:
            Mbd = sp.block_diag([m.to_scipy_full() for m in mats], format="csr")
        dB = T.DeviceArray(b.reshape(n, 1))
        dX = T.DeviceArray(np.zeros((n, 1)))
        dV = T.DeviceArray(np.concatenate([m.arrays()[3] for m in mats]))
        t_a = time.perf_counter()
        Bt.numeric(None, prec)
        t_host = time.perf_counter() - t_a
        tf, ts = [], []
        for i in range(steps + 1):
            t_a = time.perf_counter()
            if Bt.numeric_device(dV.ptr, prec) != 0:
                raise RuntimeError("batched factorization failed")
            t_b = time.perf_counter()
            if prec == T.FP64:
                Bt.solve_device(dX.ptr, dB.ptr, 1)
                x = None
            else:
                x, info = Bt.solve_refined(b, method=0, tol=1e-13, maxit=50)
            t_c = time.perf_counter()
            if i:
                tf.append(t_b - t_a)
                ts.append(t_c - t_b)
        if x is None:
            x = dX.get()[:, 0]
        bw = []
        for q in range(count):
            sl = slice(int(Bt.xoff[q]), int(Bt.xoff[q + 1]))
            Mq = Mbd[sl, sl]
            bw.append(host_backward_error(Mq, x[sl].reshape(-1, 1), b[sl].reshape(-1, 1)))
        r = {"numeric_ms": round(1e3 * float(np.mean(tf)), 3),
             "solve_ms": round(1e3 * float(np.mean(ts)), 3),
             "factor_gflops": round(Bt.fund_flops / float(np.mean(tf)) / 1e9, 2),
             "systems_per_s": round(count / (float(np.mean(tf)) + float(np.mean(ts))), 1),
             "max_system_backward_error": float(max(bw)),
             "numeric_ms_values_from_host": round(1e3 * t_host, 3),
             "setup_s": round(t_setup, 2), "levels": Bt.n_levels,
             "schedule": {"panel_width": tuned[0][0], "gemm_tile": tuned[0][1]},
             "autotune_ms": {f"w{w}_t{t}": round(v, 3) for (w, t), v in tuned[1].items()}}
        if prec != T.FP64:
            r["refine_iterations"] = int(info.iterations)
        res[name] = r
        log(f"batch {name}: {r}")
        dB.free()
        dX.free()
        dV.free()
        Bt.free()
    out.update(res)
    out["n_total"] = int(sum(m.n for m in mats))
    out["ordering_s"] = round(t_order, 2)
    return out


def bench_serving(T, requests=300, slots=4, budget_gib=128.0, seed=3, loads=(0.5, 0.8, 0.95)):
    """Poisson request mix through the serving loop, FP64 and FP32 + refinement.
    Capacity first: the same requests all arriving at once (a saturated queue),
    requests/s over the drain.  Then the mix at 50/80/95% of that capacity,
    with p50/p99 latency and solves/s at each offered load."""
    from taucs_amd import serving as S
    out = {}
    for name, prec in (("fp64", T.FP64), ("fp32_ir", T.FP32)):
        def one(rate):
            mix = S.make_mix(requests, rate, seed=seed)
            srv = S.Server(hbm_budget=int(budget_gib * S.GIB), slots=slots, precision=prec, tol=1e-12)
            summ = S.run_mix(mix, srv)
            srv.close()
            return {k: (float(f"{v:.4g}") if isinstance(v, float) else v) for k, v in summ.items()}

        sat = one(1e9)
        cap = sat["requests_per_s"]
        res = {"capacity_requests_per_s": cap, "capacity_solves_per_s": sat["solves_per_s"],
               "saturated": sat, "loads": {}}
        log(f"serving {name}: capacity {cap:.1f} requests/s ({sat['solves_per_s']:.0f} solves/s)")
        for f in loads:
            r = one(f * cap)
            r["offered_requests_per_s"] = round(f * cap, 2)
            res["loads"][f"{int(round(100 * f))}%"] = r
            log(f"serving {name} at {100 * f:.0f}% of capacity: {r}")
        out[name] = res
    out["workload"] = (f"{requests} requests: 70% 2D side U[32,256], 30% 3D side U[8,40], 1-16 rhs, "
                       f"40% solve-only; Poisson arrivals at 50/80/95% of the measured capacity; "
                       f"{slots} slots, HBM budget {budget_gib} GiB")
    return out


def profile_workload(T, name, steps, panel_width, gemm_tile):
    """Only one workload's factor + solve steps, at a fixed schedule, for a
    profiler run (kernel trace or counters): varcoef-{fp64,fp32,bf16upd}
    (mesh3d_varcoef 80^3, seed 2024, 4 rhs; mixed precision refined to
    backward error 1e-12 by CG) or batch-{fp64,fp32} (the 512 shifted 2D
    systems, side U[16,128]; FP32 refined to 1e-13)."""
    kind, prec_name = name.split("-")
    prec = {"fp64": T.FP64, "fp32": T.FP32, "bf16upd": T.BF16UPD}[prec_name]
    if panel_width:
        T.lib.taucs_amd_set_panel_width(panel_width)
        T.lib.taucs_amd_set_gemm_tile(gemm_tile or 64)
    if kind == "varcoef":
        A = T.mesh3d_varcoef(80, 80, 80, 2024)
        perm, inv = T.order(A, "genmmd")
        PA = T.permute(A, perm, inv)
        F = T.factor_symbolic(PA)
        B = np.stack([T.random_uniform(PA.n, 100 + r) for r in range(4)], 1)
        dV = T.DeviceArray(PA.arrays()[3])
        dB, dX = T.DeviceArray(B), T.DeviceArray(np.zeros_like(B))
        for i in range(steps):
            t0 = time.perf_counter()
            if F.numeric_device(dV.ptr, prec) != 0:
                raise RuntimeError("factorization failed")
            t1 = time.perf_counter()
            if prec == T.FP64:
                F.solve_device(dX.ptr, dB.ptr, PA.n, 4)
                its = 0
            else:
                _, info = F.solve_refined(PA, B, method=1, tol=1e-12, maxit=100, factored=True)
                its = info.iterations
            t2 = time.perf_counter()
            log(f"{name} step {i}: factor {1e3 * (t1 - t0):.2f} ms, solve {1e3 * (t2 - t1):.2f} ms, "
                f"iterations {its}")
        F.free()
    elif kind == "batch":
        rng = np.random.default_rng(7)
        sides = rng.integers(16, 129, 512)
        mats = []
        for q, s in enumerate(sides):
            A = T.mesh2d_shifted(int(s), seed=10_000 + q)
            perm, inv = T.order(A, "genmmd")
            mats.append(T.permute(A, perm, inv))
        Bt = T.Batch(mats, precision=prec)
        b = np.concatenate([T.random_uniform(m.n, 500 + q) for q, m in enumerate(mats)])
        dV = T.DeviceArray(np.concatenate([m.arrays()[3] for m in mats]))
        dB, dX = T.DeviceArray(b.reshape(-1, 1)), T.DeviceArray(np.zeros((b.size, 1)))
        for i in range(steps):
            t0 = time.perf_counter()
            if Bt.numeric_device(dV.ptr, prec) != 0:
                raise RuntimeError("batched factorization failed")
            t1 = time.perf_counter()
            if prec == T.FP64:
                Bt.solve_device(dX.ptr, dB.ptr, 1)
            else:
                Bt.solve_refined(b, method=0, tol=1e-13, maxit=50)
            t2 = time.perf_counter()
            log(f"{name} step {i}: factor {1e3 * (t1 - t0):.2f} ms, solve {1e3 * (t2 - t1):.2f} ms")
        Bt.free()
    else:
        raise SystemExit(f"unknown workload {name}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--side", type=int, default=60)
    ap.add_argument("--nrhs", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-varcoef", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--panel-width", type=int, default=0,
                    help="use this outer panel width instead of running the autotuner")
    ap.add_argument("--gemm-tile", type=int, default=0,
                    help="with --panel-width: the tile edge (64 or 128) of the trailing updates")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="before the timed steps, copy this many MiB device to device once: "
                         "the reference point for hardware byte counters under a profiler")
    ap.add_argument("--dump-ops", default=None, help="write the per-op profile of the FP64 "
                    "factorization (kind, level, tasks, flops, bytes, ms) to this JSON file")
    ap.add_argument("--profile-workload", default=None,
                    help="run only this workload's steps for a profiler: varcoef-fp64, varcoef-fp32, "
                         "varcoef-bf16upd, batch-fp64 or batch-fp32 (with --steps, --panel-width, "
                         "--gemm-tile); prints no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only; --gpus {args.gpus} is not supported "
              "(sharding the factorization over GPUs is designed, not implemented)")
        sys.exit(2)

    import taucs_amd as T
    if args.profile_workload:
        profile_workload(T, args.profile_workload, args.steps, args.panel_width, args.gemm_tile)
        return
    side, k = args.side, args.nrhs
    log(f"bench: mesh3d {side}^3, {k} rhs, steps {args.steps} warmup {args.warmup}")
    A = T.mesh3d(side)
    n = A.n
    perm, inv = T.order(A, "genmmd")
    PA = T.permute(A, perm, inv)
    F = T.factor_symbolic(PA)
    fi = F.info()
    # the schedule autotuner picks the outer panel width (timed numeric runs)
    if args.panel_width:
        T.lib.taucs_amd_set_panel_width(args.panel_width)
        T.lib.taucs_amd_set_gemm_tile(args.gemm_tile or 64)
        tuned_w, tune_ms = (args.panel_width, args.gemm_tile or 64), {}
    else:
        tuned_w, tune_ms = F.autotune(PA, T.FP64, reps=2)
    log(f"autotune: panel width {tuned_w}, candidates (ms) {tune_ms}")
    B = np.stack([T.random_uniform(n, 1 + r) for r in range(k)], 1)
    dB = T.DeviceArray(B)
    dX = T.DeviceArray(np.zeros((n, k)))
    dV = T.DeviceArray(PA.arrays()[3])   # A's values resident in HBM

    def step(factor_obj, prec):
        t0 = time.perf_counter()
        if factor_obj.numeric_device(dV.ptr, prec) != 0:
            raise RuntimeError("numeric factorization failed")
        t1 = time.perf_counter()
        factor_obj.solve_device(dX.ptr, dB.ptr, n, k)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1

    if args.copy_calibration_mib:
        nb = args.copy_calibration_mib << 20
        src, dst = T.lib.taucs_amd_device_malloc(nb), T.lib.taucs_amd_device_malloc(nb)
        T.lib.taucs_amd_memcpy_d2d(dst, src, nb)
        T.lib.taucs_amd_device_synchronize()
        T.lib.taucs_amd_device_free(src)
        T.lib.taucs_amd_device_free(dst)
    for _ in range(args.warmup):
        step(F, T.FP64)
    T.lib.taucs_amd_device_synchronize()
    tf, ts = [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        a, b = step(F, T.FP64)
        tf.append(a)
        ts.append(b)
    T.lib.taucs_amd_device_synchronize()
    t_total = time.perf_counter() - t_start
    X = dX.get()
    M = PA.to_scipy_full()
    R = B - M @ X
    relres = float((np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)).max())
    anorm = abs(M).sum(1).max()
    bwd = float((abs(R).max(0) / (anorm * abs(X).max(0) + abs(B).max(0))).max())
    tf_mean, ts_mean = float(np.mean(tf)), float(np.mean(ts))
    step_ms = 1e3 * t_total / args.steps
    lat = 1e3 * (np.array(tf) + np.array(ts))
    log(f"fp64: factor {tf_mean*1e3:.2f} ms ({fi.fund_flops/tf_mean/1e9:.1f} GFLOP/s), solve {ts_mean*1e3:.2f} ms, "
        f"relres {relres:.2e} bwd {bwd:.2e}")

    # one refinement step brings the FP64 residual to working precision
    x_ref, info_ref = F.solve_refined(PA, B, method=0, tol=1e-15, maxit=1, factored=True)

    # roofline of the dominant kernel, from device events on the factor's stream
    T.lib.taucs_amd_set_profiling(1)
    step(F, T.FP64)
    T.lib.taucs_amd_set_profiling(0)
    tm = F.timing()
    prof = F.op_profile()
    kinds = {1: "assemble", 2: "small_fronts", 3: "potrf", 4: "trsm", 5: "gemm_syrk", 6: "extend_add",
             7: "zero_update_blocks"}
    breakdown = {}
    for kd, name in kinds.items():
        sel = prof["kind"] == kd
        if not sel.any():
            continue
        ms = float(prof["ms"][sel].sum())
        e = {"launches": int(sel.sum()), "ms": round(ms, 3)}
        if prof["flops"][sel].sum() > 0:
            e["flops"] = float(prof["flops"][sel].sum())
            e["tflops"] = round(e["flops"] / (ms * 1e-3) / 1e12, 3) if ms > 0 else None
        if prof["bytes"][sel].sum() > 0:
            e["bytes"] = float(prof["bytes"][sel].sum())
            e["gbps"] = round(e["bytes"] / (ms * 1e-3) / 1e9, 1) if ms > 0 else None
        breakdown[name] = e
    log(f"breakdown: {breakdown}")
    if args.dump_ops:
        with open(args.dump_ops, "w") as fh:
            json.dump({k: v.tolist() for k, v in prof.items()}, fh)
    gemm_avg_ms = tm.gemm_ms / max(tm.gemm_launches, 1)
    # the dominant kernel: k_gemm_tile128<double> (the trailing-update and SYRK
    # launches with at least 512 tiles of 128x128), from the same per-op events
    sel128 = (prof["kind"] == 5) & (prof["tile"] == 128)
    n128 = int(sel128.sum())
    f128, ms128 = float(prof["flops"][sel128].sum()), float(prof["ms"][sel128].sum())
    rate128 = f128 / (ms128 * 1e-3) if ms128 > 0 else 0.0
    # memory traffic per launch, from the committed counter passes (FETCH_SIZE /
    # WRITE_SIZE, calibrated by a 1 GiB copy) of this workload and schedule
    traffic, traffic_all = None, None
    sched = list(tuned_w) if isinstance(tuned_w, tuple) else [tuned_w, 64]
    tp = os.path.join(ROOT, "profiles", "c2_gemm_traffic.json")
    if side == 60 and os.path.exists(tp):
        with open(tp) as fh:
            tj = json.load(fh)
        ent = tj.get("schedules", {}).get(f"{sched[0]}x{sched[1]}")
        if ent:
            traffic_all = round(ent["bytes_per_launch"] / 1e6, 2)
            k128 = ent.get("kernels", {}).get("k_gemm_tile128<double>")
            if k128:
                traffic = round(k128["bytes_per_launch"] / 1e6, 2)
    gemm_rate = tm.gemm_flops / (tm.gemm_ms * 1e-3) if tm.gemm_ms > 0 else 0.0
    log(f"profile: total {tm.total_ms:.2f} ms, front {tm.front_ms:.2f} ms, gemm {tm.gemm_ms:.2f} ms "
        f"({tm.gemm_launches} launches, {gemm_rate/1e12:.2f} TFLOP/s), assembly {tm.assembly_ms:.2f} ms, "
        f"launches {tm.launches}")

    mixed = None
    if not args.no_mixed:
        Fm = T.factor_symbolic(PA)
        m_sched, m_tune = Fm.autotune(PA, T.FP32, reps=1)   # FP32 picks its own schedule
        tmf, tms, its, bw = [], [], [], []
        for i in range(args.warmup + args.steps):
            t0 = time.perf_counter()
            if Fm.numeric_device(dV.ptr, T.FP32) != 0:
                raise RuntimeError("fp32 factorization failed")
            t1 = time.perf_counter()
            xm, info = Fm.solve_refined(PA, B, method=1, tol=1e-12, maxit=50, factored=True)
            t2 = time.perf_counter()
            if i >= args.warmup:
                tmf.append(t1 - t0)
                tms.append(t2 - t1)
                its.append(info.iterations)
                bw.append(info.backward_error)
        mixed = {"precision": "fp32 fronts + fp64 preconditioned CG",
                 "schedule": {"panel_width": m_sched[0], "gemm_tile": m_sched[1]},
                 "factor_gflops": round(fi.fund_flops / np.mean(tmf) / 1e9, 2),
                 "factor_ms": round(1e3 * float(np.mean(tmf)), 3),
                 "refined_solve_ms": round(1e3 * float(np.mean(tms)), 3),
                 "refine_iterations": int(max(its)),
                 "backward_error": float(max(bw)),
                 "speedup_factor_plus_solve_vs_fp64": round((tf_mean + ts_mean) / (np.mean(tmf) + np.mean(tms)), 3)}
        log(f"mixed: {mixed}")
        Fm.free()

    varcoef = batch = serving = None
    if not args.no_varcoef:
        varcoef = bench_varcoef(T)
    if not args.no_batch:
        batch = bench_batch(T)
    if not args.no_serving:
        serving = bench_serving(T)

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(side, k)
        except Exception as e:  # the baseline must not hide the GPU result
            cpu = {"error": str(e)}

    out = {
        "metric": "factor GFLOP/s & solves/s, mixed vs FP64 at bkwd err <=1e-12; p50/p99 latency",
        "value": round(fi.fund_flops / tf_mean / 1e9, 3),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"mesh3d {side}x{side}x{side} 7-point Laplacian, genmmd ordering, "
                               f"numeric factor + {k}-rhs supernodal solve (plain FP64 path)",
                   "n": n, "nrhs": k, "fund_flops": fi.fund_flops, "relax_flops": fi.relax_flops,
                   "n_supernodes": fi.n_live, "levels": fi.n_levels, "max_front": fi.max_front,
                   "schedule": {"panel_width": tuned_w[0] if isinstance(tuned_w, tuple) else tuned_w,
                                "gemm_tile": tuned_w[1] if isinstance(tuned_w, tuple) else 64},
                   "autotune_ms": {f"w{w}_t{t}": round(v, 3) for (w, t), v in tune_ms.items()}},
        "factor_ms": round(1e3 * tf_mean, 3),
        "solve_ms": round(1e3 * ts_mean, 3),
        "solves_per_s": round(k / ts_mean, 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                       "p99": round(float(np.percentile(lat, 99)), 3)},
        "rel_residual": relres,
        "backward_error": bwd,
        "rel_residual_after_1_refinement": info_ref.rel_residual,
        "mixed": mixed,
        "roofline": {"bound": "mfma",
                     "kernel": ("k_gemm_tile128<double>: trailing updates and SYRKs of the launches with at "
                                "least 512 tiles of 128x128 (eight waves of 64x32, v_mfma_f64_16x16x4f64)"),
                     "achieved": round(rate128 / 1e12, 3), "peak": PEAK_FP64 / 1e12,
                     "unit": "TFLOP/s", "frac": round(rate128 / PEAK_FP64, 4),
                     "avg_launch_ms": round(ms128 / max(n128, 1), 4), "launches": n128,
                     "traffic": traffic, "traffic_unit": "MB per launch (profiles/c2_mesh3d60_fp64_gemm_traffic.md)",
                     "algorithmic_flops_per_launch": round(f128 / max(n128, 1), 1),
                     "all_gemm_launches": {
                         "kernels": "k_gemm_tile128<double>, k_gemm_tile<double>, k_gemm_short<double> (K = 64)",
                         "achieved": round(gemm_rate / 1e12, 3), "frac": round(gemm_rate / PEAK_FP64, 4),
                         "avg_launch_ms": round(gemm_avg_ms, 4), "launches": tm.gemm_launches,
                         "traffic": traffic_all,
                         "algorithmic_flops_per_launch": round(tm.gemm_flops / max(tm.gemm_launches, 1), 1)}},
        "cpu_baseline": cpu,
        "numeric_breakdown": breakdown,
        "mixed_precision_varcoef": varcoef,
        "batched": batch,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
