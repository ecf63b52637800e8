"""Batched factorization (many independent systems in shared level launches)
and hipGraph replay, against single-system runs."""
import numpy as np
import pytest

import taucs_amd as T
from tests.golden.make_golden import KEYS

pytestmark = pytest.mark.gpu


def shifted_systems(count, lo, hi, seed):
    rng = np.random.default_rng(seed)
    sides = rng.integers(lo, hi + 1, count)
    mats = []
    for q, s in enumerate(sides):
        A = T.mesh2d_shifted(int(s), seed=1000 + q)
        perm, inv = T.order(A, "genmmd")
        mats.append(T.permute(A, perm, inv))
    return mats


def test_batch_equals_alone_bitwise():
    print("case: batch of 24 shifted 2D systems vs one at a time", flush=True)
    mats = shifted_systems(24, 16, 64, 7)
    B = T.Batch(mats)
    b = np.concatenate([T.random_uniform(m.n, 50 + q) for q, m in enumerate(mats)])
    x = B.solve(b)
    for q, m in enumerate(mats):
        F = T.factor_llt_mf(m)
        xq = F.solve(b[B.xoff[q]:B.xoff[q + 1]])
        assert np.array_equal(x[B.xoff[q]:B.xoff[q + 1]], xq), q
        S1, S2 = B.system_supernodes(q), F.supernodes()
        for k in KEYS:
            assert np.array_equal(S1[k], S2[k])


def test_batch_fp32_refined():
    print("case: batch fp32 + refinement", flush=True)
    mats = shifted_systems(16, 16, 48, 3)
    B = T.Batch(mats, precision=T.FP32)
    b = np.concatenate([T.random_uniform(m.n, q) for q, m in enumerate(mats)])
    x, info = B.solve_refined(b, method=0, tol=1e-14, maxit=50)
    assert info.converged and info.backward_error <= 1e-14
    for q, m in enumerate(mats):
        xs, bs = x[B.xoff[q]:B.xoff[q + 1]], b[B.xoff[q]:B.xoff[q + 1]]
        M = m.to_scipy_full()
        r = bs - M @ xs
        eta = abs(r).max() / (abs(M).sum(1).max() * abs(xs).max() + abs(bs).max())
        assert eta <= 1e-13, q


def test_batch_refactor_new_values():
    print("case: batch numeric refactorization", flush=True)
    mats = shifted_systems(8, 20, 40, 11)
    B = T.Batch(mats)
    b = np.concatenate([T.random_uniform(m.n, q) for q, m in enumerate(mats)])
    x1 = B.solve(b)
    scaled = []
    for m in mats:
        n, cp, ri, v = m.arrays()
        scaled.append(T.from_arrays(n, cp, ri, 4.0 * v))
    assert B.numeric(scaled) == 0
    x2 = B.solve(b)
    assert np.allclose(x2, 0.25 * x1, rtol=1e-12, atol=1e-15)


def test_graph_replays_match_eager():
    """Captured numeric and solve graphs, replayed several times with eager
    work in between, give exactly the eager results."""
    print("case: hipGraph replay vs eager", flush=True)
    A = T.mesh3d(18)
    perm, inv = T.order(A, "genmmd")
    PA = T.permute(A, perm, inv)
    b = np.stack([T.random_uniform(A.n, 9), T.random_uniform(A.n, 10)], 1)
    T.lib.taucs_amd_set_graphs(0)
    try:
        Fe = T.factor_llt_mf(PA)
        xe = Fe.solve(b)
        Le = Fe.diag()
    finally:
        T.lib.taucs_amd_set_graphs(1)
    F = T.factor_symbolic(PA)
    for rep in range(4):
        assert F.numeric(PA) == 0
        assert np.array_equal(F.diag(), Le)
        for _ in range(2):
            assert np.array_equal(F.solve(b), xe)
        # eager work in between the replays
        T.lib.taucs_amd_set_graphs(0)
        try:
            G = T.factor_llt_mf(PA)
            assert np.array_equal(G.solve(b), xe)
            G.free()
        finally:
            T.lib.taucs_amd_set_graphs(1)


def test_lookahead_gives_identical_results():
    """The panel lookahead only reorders independent launches across two
    streams; factor and solution are bitwise those of the one-stream order."""
    print("case: lookahead on vs off", flush=True)
    A = T.mesh3d(24)
    perm, inv = T.order(A, "genmmd")
    PA = T.permute(A, perm, inv)
    b = T.random_uniform(A.n, 77)
    T.lib.taucs_amd_set_lookahead(0)
    try:
        F0 = T.factor_llt_mf(PA)
        x0, d0 = F0.solve(b), F0.diag()
        L0 = F0.to_ccs().arrays()[3]
        F0.free()
    finally:
        T.lib.taucs_amd_set_lookahead(1)
    for graphs in (1, 0):
        T.lib.taucs_amd_set_graphs(graphs)
        try:
            F1 = T.factor_llt_mf(PA)
            assert np.array_equal(F1.diag(), d0)
            assert np.array_equal(F1.to_ccs().arrays()[3], L0)
            assert np.array_equal(F1.solve(b), x0)
            F1.free()
        finally:
            T.lib.taucs_amd_set_graphs(1)


def test_refinement_follows_refactorization():
    """The refinement context keeps its CSR structure across numeric
    refactorizations and takes the new values from the factor's device copy."""
    print("case: refine after refactorization with new values", flush=True)
    A = T.mesh3d(14)
    perm, inv = T.order(A, "genmmd")
    PA = T.permute(A, perm, inv)
    n, cp, ri, v = PA.arrays()
    b = np.stack([T.random_uniform(n, 5), T.random_uniform(n, 6)], 1)
    F = T.factor_symbolic(PA)
    assert F.numeric(PA, T.FP32) == 0
    x1, i1 = F.solve_refined(PA, b, method=0, tol=1e-14, maxit=50)
    assert i1.converged
    B2 = T.from_arrays(n, cp, ri, 3.0 * v)
    assert F.numeric(B2, T.FP32) == 0
    x2, i2 = F.solve_refined(B2, b, method=0, tol=1e-14, maxit=50)
    assert i2.converged
    # both at backward error <= 1e-14: they agree to about cond(A) * 1e-14
    assert np.allclose(x2, x1 / 3.0, rtol=1e-9, atol=1e-12)
    # refining against a matrix that is not the factored one uploads its values
    x3, i3 = F.solve_refined(PA, b, method=1, tol=1e-14, maxit=100)
    assert i3.converged
    assert np.allclose(x3, x1, rtol=1e-9, atol=1e-12)


def test_autotune_picks_a_candidate_and_keeps_a_correct_factor():
    """The schedule autotuner times each outer panel width and keeps the
    fastest; every width gives the same factorization up to rounding."""
    print("case: schedule autotuner", flush=True)
    A = T.mesh3d(22)
    perm, inv = T.order(A, "genmmd")
    PA = T.permute(A, perm, inv)
    b = T.random_uniform(A.n, 3)
    x0 = T.factor_llt_mf(PA).solve(b)
    F = T.factor_symbolic(PA)
    cands = ((64, 64), (256, 64), (256, 128))
    best, ms = F.autotune(PA, candidates=cands, reps=1)
    assert best in cands and all(v > 0 for v in ms.values())
    assert ms[best] == min(ms.values())
    x = F.solve(b)
    assert np.linalg.norm(x - x0) / np.linalg.norm(x0) <= 1e-10
    mats = shifted_systems(12, 16, 64, 5)
    B = T.Batch(mats)
    bb = np.concatenate([T.random_uniform(m.n, q) for q, m in enumerate(mats)])
    xb0 = B.solve(bb)
    best, ms = B.autotune(candidates=((64, 64), (192, 128)), reps=1)
    assert best in ((64, 64), (192, 128))
    assert np.allclose(B.solve(bb), xb0, rtol=1e-10, atol=1e-13)


def test_numeric_from_device_values_equals_host_values():
    print("case: numeric from device-resident values", flush=True)
    A = T.mesh3d(16)
    perm, inv = T.order(A, "genmmd")
    PA = T.permute(A, perm, inv)
    b = T.random_uniform(A.n, 8)
    F1 = T.factor_llt_mf(PA)
    x1 = F1.solve(b)
    F2 = T.factor_symbolic(PA)
    dV = T.DeviceArray(PA.arrays()[3])
    assert F2.numeric_device(dV.ptr) == 0
    assert np.array_equal(F2.solve(b), x1)
    mats = shifted_systems(6, 16, 40, 9)
    B1 = T.Batch(mats)
    bb = np.concatenate([T.random_uniform(m.n, q) for q, m in enumerate(mats)])
    xb = B1.solve(bb)
    dVb = T.DeviceArray(np.concatenate([m.arrays()[3] for m in mats]))
    assert B1.numeric_device(dVb.ptr) == 0
    assert np.array_equal(B1.solve(bb), xb)


@pytest.mark.parametrize("prec", [T.FP64, T.FP32])
def test_front_packing_is_bitwise_identical(prec):
    """Fronts of at most 16 (32) rows factored four (two) to a wavefront give
    the bits of one front per wavefront: factor digest and solution."""
    for name, A in (("mesh2d 60", T.mesh2d(60)), ("mesh3d 16", T.mesh3d(16)),
                    ("shifted mesh2d 90", T.mesh2d_shifted(90, seed=3))):
        print(f"case: front packing, {name}, precision {prec}", flush=True)
        perm, inv = T.order(A, "genmmd")
        PA = T.permute(A, perm, inv)
        b = T.random_uniform(PA.n, 17)
        res = []
        try:
            for on in (1, 0):
                T.lib.taucs_amd_set_front_packing(on)
                F = T.factor_symbolic(PA)
                assert F.numeric(PA, prec) == 0
                res.append((F.digest(), F.solve(b)))
                F.free()
        finally:
            T.lib.taucs_amd_set_front_packing(1)
        assert res[0][0] == res[1][0]
        assert np.array_equal(res[0][1], res[1][1])
