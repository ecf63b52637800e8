"""GPU parity at the full sizes of the benchmark workloads, against the
oracle's committed fixtures (tests/golden/make_golden.py big-solutions):

  mesh3d 60^3, FP64, 8 right-hand sides (seeds 1..8): the supernode partition
      equals the oracle's digest; each solution is within 1e-10 relative of
      the exact solution (computed here by refinement with an extended-
      precision residual) and agrees with the oracle's FP64 factor-and-solve
      within 1e-10 plus the oracle's own distance from the exact solution (in
      full for seed 1, through a dense random-sign sketch and 4096 sampled
      rows for all eight: the operator's condition number is ~4.4e6, so two
      backward-stable FP64 solutions may differ by ~5e-10, and the oracle's
      seed-1 solution is 2.3e-10 from the exact one, ours 2.5e-11); the
      relative residual is <= max(1e-12, 1.5 x the oracle's own on the same
      right-hand side); the lookahead schedule's factor is bitwise the serial
      schedule's at outer panel widths 256 and 512.
  mesh3d_varcoef 80^3 (seed 2024), 4 right-hand sides (seeds 100..103): the
      FP64 solution against the oracle's, the lookahead factor bitwise the
      serial one at the benchmark's schedule, FP32 fronts + FP64 CG to
      backward error <= 1e-12 with its forward error against the oracle.
  512 shifted mesh2d systems, side U[16,128] (numpy default_rng(7)): every
      system's backward error, eight systems against the oracle's solutions,
      bitwise equality with the solo factor-and-solve on a sample.
  The serving mix at its stated distribution (2D sides up to 256, 3D up to
      40, 1-16 right-hand sides, 40% solve-only): every result is the same
      alone and in the mix.

Tolerances are written in each test.
"""
import json
import os

import numpy as np
import pytest

import taucs_amd as T
from tests.golden.make_golden import KEYS, digest, sketch_matrix_apply

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def relres(M, X, B):
    return np.linalg.norm(B - M @ X, axis=0) / np.linalg.norm(B, axis=0)


def bwd(M, X, B):
    R = B - M @ X
    return abs(R).max(0) / (abs(M).sum(1).max() * abs(X).max(0) + abs(B).max(0))


def sketch_err(X, g):
    """|S (x - x_oracle)| / |S x_oracle| per column: S is a dense random-sign
    projection, so every entry of x moves it; it estimates the relative
    forward error within a factor of about 1.5 either way."""
    SX = sketch_matrix_apply(X)
    return np.linalg.norm(SX - g["sketch"], axis=0) / np.linalg.norm(g["sketch"], axis=0)


def sample_err(X, g):
    """max over the sampled rows of |x_i - x_oracle,i| / |x_oracle|_2."""
    idx = g["sample_idx"]
    return np.abs(X[idx] - g["sample_x"]).max(0) / g["xnorm"]


def exact_solution(M, B, solve, iters=3):
    """x* of M X = B to about 1e-12 relative: iterative refinement with the
    residual in x87 extended precision on the host (64-bit significand) and
    the FP64 factor's solve for the corrections.  Each step shrinks the error
    by about cond(M) u; the extended residual leaves cond(M) 5e-20 |M| |x|,
    about 3e-12 relative for cond ~ 4e6."""
    Mc = M.tocsr()
    ip, ix, v = Mc.indptr, Mc.indices, Mc.data.astype(np.longdouble)
    Bl = np.asarray(B).astype(np.longdouble)
    X = solve(B).astype(np.longdouble)
    for _ in range(iters):
        R = np.empty_like(Bl)
        for c in range(B.shape[1]):
            R[:, c] = Bl[:, c] - np.add.reduceat(v * X[ix, c], ip[:-1])
        X = X + solve(np.ascontiguousarray(R.astype(np.float64))).astype(np.longdouble)
    return X.astype(np.float64)


def ordered(A):
    perm, inv = T.order(A, "genmmd")
    return T.permute(A, perm, inv)


@pytest.fixture(scope="module")
def c2():
    print("case: mesh3d 60^3 setup (genmmd, symbolic)", flush=True)
    g = np.load(os.path.join(GOLD, "c2_mesh3d_60.npz"))
    PA = ordered(T.mesh3d(60))
    n = PA.n
    B = np.stack([T.random_uniform(n, int(s)) for s in g["seeds"]], 1)
    return g, PA, B, PA.to_scipy_full()


def test_c2_partition_and_solution_against_the_oracle(c2):
    g, PA, B, M = c2
    with open(os.path.join(GOLD, "big.json")) as f:
        big = json.load(f)["mesh3d_60"]
    for sched in ((256, 64), (512, 128)):
        print(f"case: mesh3d 60^3 fp64 8 rhs, schedule {sched}", flush=True)
        T.lib.taucs_amd_set_panel_width(sched[0])
        T.lib.taucs_amd_set_gemm_tile(sched[1])
        try:
            F = T.factor_symbolic(PA)
            S = F.supernodes()
            assert digest(*[S[k] for k in KEYS]) == big["sn_sha"]
            assert F.numeric(PA, T.FP64) == 0
            X = F.solve(B)
        finally:
            T.lib.taucs_amd_set_panel_width(256)
            T.lib.taucs_amd_set_gemm_tile(64)
        fe0 = np.linalg.norm(X[:, 0] - g["x_first"]) / np.linalg.norm(g["x_first"])
        se, sa = sketch_err(X, g), sample_err(X, g)
        rr = relres(M, X, B)
        # the exact solution, to tell whose error a disagreement is
        Xs = exact_solution(M, B, F.solve)
        fx = np.linalg.norm(X - Xs, axis=0) / np.linalg.norm(Xs, axis=0)
        fo0 = np.linalg.norm(g["x_first"] - Xs[:, 0]) / np.linalg.norm(Xs[:, 0])
        so = np.linalg.norm(sketch_matrix_apply(Xs) - g["sketch"], axis=0) / \
            np.linalg.norm(g["sketch"], axis=0)
        print(f"  vs oracle: seed 1 (full) {fe0:.2e}, sketch max {se.max():.2e}, sampled max "
              f"{sa.max():.2e}; vs exact: ours max {fx.max():.2e}, oracle's seed 1 {fo0:.2e}, "
              f"sketch max {so.max():.2e}; relres max {rr.max():.2e} (oracle "
              f"{g['relres'].max():.2e}); bwd {bwd(M, X, B).max():.2e}", flush=True)
        # every solution is within 1e-10 of the exact one
        assert fx.max() <= 1e-10
        # agreement with the oracle within 1e-10, beyond the oracle's own
        # distance from the exact solution (cond ~ 4.4e6: cond u ~ 5e-10)
        assert fe0 <= 1e-10 + fo0
        assert np.all(se <= 1.5e-10 + so)
        assert sa.max() <= 1e-10
        assert np.all(rr <= np.maximum(1e-12, 1.5 * g["relres"]))
        assert bwd(M, X, B).max() <= 1e-14
        F.free()


@pytest.mark.parametrize("sched", [(256, 64), (512, 128)])
def test_c2_lookahead_factor_is_bitwise_the_serial_one(c2, sched):
    """Lookahead on (twice) and off at the full 60^3 size: the factor digests
    (a position-mixed sum of every stored bit) must agree."""
    _, PA, _, _ = c2
    print(f"case: mesh3d 60^3 lookahead bitwise, schedule {sched}", flush=True)
    dig = []
    try:
        T.lib.taucs_amd_set_panel_width(sched[0])
        T.lib.taucs_amd_set_gemm_tile(sched[1])
        for la in (1, 1, 0):
            T.lib.taucs_amd_set_lookahead(la)
            F = T.factor_symbolic(PA)
            assert F.numeric(PA, T.FP64) == 0
            dig.append(F.digest())
            F.free()
    finally:
        T.lib.taucs_amd_set_lookahead(1)
        T.lib.taucs_amd_set_panel_width(256)
        T.lib.taucs_amd_set_gemm_tile(64)
    assert dig[0] == dig[1] == dig[2], dig


def test_c2_ccs_solve_at_full_size(c2):
    """taucs_ccs_solve_llt on the exported 60^3 factor (nnz(L) 1.5e8, the size at
    which an earlier per-value wait bound of the sync-free sweeps was reached):
    both sweeps finish, and the solution equals the supernodal solve's to
    rounding (the two sum each row in different orders)."""
    _, PA, B, M = c2
    print("case: mesh3d 60^3 ccs solve of the exported factor", flush=True)
    F = T.factor_symbolic(PA)
    assert F.numeric(PA, T.FP64) == 0
    b = np.ascontiguousarray(B[:, 0])
    x_sn = F.solve(b)
    Lc = F.to_ccs()
    F.free()
    x_ccs = Lc.solve_llt(b)
    d = np.linalg.norm(x_ccs - x_sn) / np.linalg.norm(x_sn)
    rr = np.linalg.norm(M @ x_ccs - b) / np.linalg.norm(b)
    print(f"  ccs vs supernodal {d:.2e}, relres {rr:.2e}", flush=True)
    # cond ~ 4.4e6: two backward-stable solves agree to ~ cond * u
    assert d <= 1e-9
    assert rr <= 1e-11


@pytest.fixture(scope="module")
def c3():
    print("case: mesh3d_varcoef 80^3 setup (genmmd, symbolic)", flush=True)
    g = np.load(os.path.join(GOLD, "c3_varcoef_80.npz"))
    PA = ordered(T.mesh3d_varcoef(80, 80, 80, 2024))
    B = np.stack([T.random_uniform(PA.n, int(s)) for s in g["seeds"]], 1)
    return g, PA, B, PA.to_scipy_full()


def test_c3_fp64_and_mixed_against_the_oracle(c3):
    g, PA, B, M = c3
    print("case: varcoef 80^3 fp64, 4 rhs, schedule (512, 128)", flush=True)
    T.lib.taucs_amd_set_panel_width(512)
    T.lib.taucs_amd_set_gemm_tile(128)
    try:
        F = T.factor_symbolic(PA)
        assert F.numeric(PA, T.FP64) == 0
        X = F.solve(B)
        d_look = F.digest()
        T.lib.taucs_amd_set_lookahead(0)
        F2 = T.factor_symbolic(PA)
        assert F2.numeric(PA, T.FP64) == 0
        d_serial = F2.digest()
        F2.free()
    finally:
        T.lib.taucs_amd_set_lookahead(1)
        T.lib.taucs_amd_set_panel_width(256)
        T.lib.taucs_amd_set_gemm_tile(64)
    # above 4096 trailing tiles the FP64 lookahead runs without its LDS
    # reservation: the factor must still be bitwise the serial one
    assert d_look == d_serial
    se, sa, b64 = sketch_err(X, g), sample_err(X, g), bwd(M, X, B)
    print(f"  fp64: forward error vs oracle (sketch) {se.max():.2e}, sampled {sa.max():.2e}, "
          f"bwd {b64.max():.2e} (oracle bwd {g['bwd'].max():.2e})", flush=True)
    # both are backward stable to ~1e-15 on an operator with kappa ~ 1e10:
    # they agree to kappa * u, measured ~1e-7 or better; 1e-6 leaves margin
    assert b64.max() <= 1e-14
    assert se.max() <= 1e-6 and sa.max() <= 1e-6
    F.free()
    print("case: varcoef 80^3 fp32 fronts + fp64 CG, tol 1e-12", flush=True)
    Fm = T.factor_symbolic(PA)
    assert Fm.numeric(PA, T.FP32) == 0
    Xm, info = Fm.solve_refined(PA, B, method=1, tol=1e-12, maxit=100, factored=True)
    sm, bm = sketch_err(Xm, g), bwd(M, Xm, B)
    print(f"  fp32+cg: {info.iterations} iterations, bwd {bm.max():.2e}, forward error vs "
          f"oracle (sketch) {sm.max():.2e}", flush=True)
    assert info.converged and bm.max() <= 1e-12
    assert sm.max() <= 1e-5
    Fm.free()


def test_c4_all_512_systems():
    print("case: 512 shifted mesh2d systems, side U[16,128], fp64", flush=True)
    g = np.load(os.path.join(GOLD, "c4_batch.npz"))
    sides = np.random.default_rng(7).integers(16, 129, 512)
    assert np.array_equal(sides[g["systems"]], g["sides"])
    mats = [ordered(T.mesh2d_shifted(int(s), seed=10_000 + q)) for q, s in enumerate(sides)]
    Bt = T.Batch(mats, precision=T.FP64)
    b = np.concatenate([T.random_uniform(m.n, 500 + q) for q, m in enumerate(mats)])
    x = Bt.solve(b)
    worst = 0.0
    for q, m in enumerate(mats):
        sl = slice(int(Bt.xoff[q]), int(Bt.xoff[q + 1]))
        worst = max(worst, float(bwd(m.to_scipy_full(), x[sl].reshape(-1, 1), b[sl].reshape(-1, 1))[0]))
    print(f"  max per-system backward error {worst:.2e}", flush=True)
    assert worst <= 1e-14
    off = np.concatenate([[0], np.cumsum(g["n"])])
    for i, q in enumerate(g["systems"]):
        xo = g["x"][off[i]:off[i + 1]]
        sl = slice(int(Bt.xoff[q]), int(Bt.xoff[q + 1]))
        assert np.linalg.norm(x[sl] - xo) / np.linalg.norm(xo) <= 1e-10, q
    for q in (0, int(np.argmax(sides)), 255, 511):
        F = T.factor_llt_mf(mats[q])
        sl = slice(int(Bt.xoff[q]), int(Bt.xoff[q + 1]))
        assert np.array_equal(F.solve(b[sl]), x[sl]), q
        F.free()
    Bt.free()


@pytest.mark.parametrize("precision", [T.FP64, T.FP32])
def test_c5_stated_mix_same_alone_and_in_the_mix(precision):
    from taucs_amd import serving as S
    from tests.test_serving import alone
    print(f"case: serving, stated distribution, precision {precision}", flush=True)
    mix = S.make_mix(24, 200.0, seed=21)
    srv = S.Server(hbm_budget=64 * S.GIB, slots=4, precision=precision)
    summ = S.run_mix(mix, srv)
    srv.close()
    print(f"  {summ}", flush=True)
    assert summ["failed"] == 0 and summ["max_backward_error"] <= 1e-12
    assert max(r.B.shape[0] for _, r in mix) > 20_000      # the large sizes are in it
    for _, r in mix[::4]:
        q = alone(r, precision)
        assert np.array_equal(q.X, r.X), r.rid
