"""GPU parity of the multifrontal factor and solve (through the C-ABI).

Tolerances (FP64 path): solutions agree with the oracle's FP64 factor-and-
solve within 1e-10 relative; relative residuals <= 1e-12 at the C1 size.
Mixed precision: FP64 refinement to a normwise backward error <= 1e-14.
"""
import ctypes as C
import os

import numpy as np
import pytest

import taucs_amd as T
from oracle import oracle as O
from tests.golden.make_golden import KEYS

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def prepared(A):
    perm, inv = T.order(A, "genmmd")
    return perm, inv, T.permute(A, perm, inv)


def relres(PA, x, b):
    n, cp, ri, v = PA.arrays()
    r = O.residual(n, cp, ri, v, x, b)
    return np.linalg.norm(r, axis=0) / np.linalg.norm(b, axis=0)


def oracle_solution(PA, b):
    n, cp, ri, v = PA.arrays()
    S = O.symbolic(n, cp, ri)
    panels = O.factor(S, cp, ri, v)
    return S, panels, O.solve(S, panels, b)


def test_c1_against_fixture():
    print("case: C1 mesh2d(100) fp64, golden x", flush=True)
    g = np.load(os.path.join(ROOT, "tests", "golden", "c1_mesh2d_100.npz"))
    A = T.mesh2d(100)
    perm, inv, PA = prepared(A)
    F = T.factor_llt_mf(PA)
    assert F is not None
    S = F.supernodes()
    for k in KEYS:
        assert np.array_equal(S[k], g[k])
    x = F.solve(g["b"])
    fe = np.linalg.norm(x - g["x"]) / np.linalg.norm(g["x"])
    assert fe <= 1e-10
    assert relres(PA, x, g["b"]) <= 1e-12
    # the single-rhs reference entry point gives the same answer
    xs = np.zeros(A.n)
    b = g["b"].copy()
    assert T.lib.taucs_supernodal_solve_llt(F.h, xs.ctypes.data_as(C.c_void_p),
                                            b.ctypes.data_as(C.c_void_p)) == 0
    assert np.array_equal(xs, x)
    assert np.array_equal(b, g["b"])   # b is not modified


@pytest.mark.parametrize("kind,side,nrhs", [("2d", 37, 3), ("3d", 12, 2), ("3d", 20, 4), ("3d", 30, 1)])
def test_meshes_against_oracle(kind, side, nrhs):
    print(f"case: {kind} side {side} nrhs {nrhs}", flush=True)
    A = T.mesh2d(side) if kind == "2d" else T.mesh3d(side)
    _, _, PA = prepared(A)
    b = np.stack([T.random_uniform(A.n, 10 + r) for r in range(nrhs)], 1)
    F = T.factor_llt_mf(PA)
    x = F.solve(b)
    S, panels, xo = oracle_solution(PA, b)
    fe = np.linalg.norm(x - xo, axis=0) / np.linalg.norm(xo, axis=0)
    assert fe.max() <= 1e-10
    assert relres(PA, x, b).max() <= 1e-12


def test_factor_export_matches_oracle():
    print("case: to_ccs / get_diag on 3d 14", flush=True)
    A = T.mesh3d(14)
    _, _, PA = prepared(A)
    F = T.factor_llt_mf(PA)
    n, cp, ri, v = PA.arrays()
    S, panels, _ = oracle_solution(PA, np.ones(n))
    Lo = np.zeros(n)
    Ld = {}
    for sn, P in panels.items():
        rows = O.rows_of(S, sn)
        for jj in range(S["sn_size"][sn]):
            Lo[rows[jj]] = P[jj, jj]
            for ii in range(jj, len(rows)):
                if P[ii, jj] != 0.0:
                    Ld[(rows[ii], rows[jj])] = P[ii, jj]
    d = F.diag()
    assert np.allclose(d, Lo, rtol=1e-12, atol=0)
    L = F.to_ccs()
    ln, lcp, lri, lv = L.arrays()
    assert L.ptr.contents.flags & T.TAUCS_TRIANGULAR and L.ptr.contents.flags & T.TAUCS_LOWER
    got = {}
    for j in range(n):
        assert lri[lcp[j]] == j   # diagonal first
        for k in range(lcp[j], lcp[j + 1]):
            got[(lri[k], j)] = lv[k]
    keys = set(got) | set(Ld)
    err = max(abs(got.get(k, 0.0) - Ld.get(k, 0.0)) for k in keys)
    assert err <= 1e-12 * max(abs(x) for x in Ld.values())
    # L L^T reproduces P A P^T
    import scipy.sparse as sp
    Ls = sp.csc_matrix((lv, lri, lcp), shape=(n, n))
    R = (Ls @ Ls.T - PA.to_scipy_full())
    assert abs(R).max() <= 1e-12 * abs(PA.to_scipy_full()).max()


def test_unsorted_random_patterns():
    from tests.test_host import random_graph
    for seed in range(4):
        print(f"case: random graph seed {seed}", flush=True)
        n, cp, ri, v = random_graph(300 + 50 * seed, 2 + seed, seed)
        A = T.from_arrays(n, cp, ri, v)
        _, _, PA = prepared(A)
        b = T.random_uniform(n, seed)
        F = T.factor_llt_mf(PA)
        x = F.solve(b)
        _, _, xo = oracle_solution(PA, b)
        assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-10


def test_identity_ordering_and_dense_like_front():
    print("case: identity ordering, 2d 20 (one long chain)", flush=True)
    A = T.mesh2d(20)
    perm, inv = T.order(A, "identity")
    PA = T.permute(A, perm, inv)
    b = T.random_uniform(A.n, 5)
    x = T.factor_llt_mf(PA).solve(b)
    _, _, xo = oracle_solution(PA, b)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-10


def test_not_positive_definite_returns_null():
    print("case: indefinite", flush=True)
    A = T.mesh2d(30)
    n, cp, ri, v = A.arrays()
    v = v.copy()
    v[cp[400] - 1] = -10.0
    B = T.from_arrays(n, cp, ri, v)
    _, _, PB = prepared(B)
    assert not T.lib.taucs_ccs_factor_llt_mf(PB.ptr)
    # a large front must fail too (blocked path)
    A3 = T.mesh3d(16)
    n, cp, ri, v = A3.arrays()
    v = v.copy()
    v[cp[n] - 1] = -1.0  # last column: ends in the root front
    B3 = T.from_arrays(n, cp, ri, v)
    _, _, PB3 = prepared(B3)
    assert not T.lib.taucs_ccs_factor_llt_mf(PB3.ptr)


def test_symbolic_once_numeric_many():
    print("case: symbolic/numeric split", flush=True)
    A = T.mesh3d(15)
    _, _, PA = prepared(A)
    n, cp, ri, v = PA.arrays()
    F = T.factor_symbolic(PA)
    b = T.random_uniform(n, 3)
    assert F.numeric(PA) == 0
    x1 = F.solve(b)
    P2 = T.from_arrays(n, cp, ri, 2.0 * v)
    assert F.numeric(P2) == 0
    x2 = F.solve(b)
    assert np.allclose(x2, 0.5 * x1, rtol=1e-12, atol=1e-14)
    T.lib.taucs_supernodal_factor_free_numeric(F.h)
    assert F.numeric(PA) == 0
    assert np.array_equal(F.solve(b), x1)


def test_duplicate_entries_are_summed_in_stored_order():
    """A pattern with a duplicated (row, column) pair assembles column by
    column, adding the duplicates in stored order as the reference adds every
    stored entry into its front; a pattern without one assembles one entry
    per thread.  Splitting an off-diagonal entry v into v/2 + v/2 (exact in
    binary) must give bitwise the factor and solution of the original."""
    print("case: duplicated entry, 3d 8", flush=True)
    A = T.mesh3d(8)
    _, _, PA = prepared(A)
    n, cp, ri, v = PA.arrays()
    j = next(c for c in range(n) if any(ri[q] != c for q in range(cp[c], cp[c + 1])))
    k = next(q for q in range(cp[j], cp[j + 1]) if ri[q] != j)
    cp2 = cp.copy()
    cp2[j + 1:] += 1
    ri2 = np.insert(ri, k + 1, ri[k])
    v2 = np.insert(v, k + 1, v[k] / 2)
    v2[k] = v[k] / 2
    D = T.from_arrays(n, cp2, ri2, v2)
    F1 = T.factor_symbolic(PA)
    assert F1.numeric(PA) == 0
    F2 = T.factor_symbolic(D)
    assert F2.numeric(D) == 0
    assert F1.digest() == F2.digest()
    b = T.random_uniform(n, 3)
    assert np.array_equal(F1.solve(b), F2.solve(b))
    F1.free()
    F2.free()


def test_results_are_deterministic():
    print("case: determinism 3d 24", flush=True)
    A = T.mesh3d(24)
    _, _, PA = prepared(A)
    b = np.stack([T.random_uniform(A.n, 7), T.random_uniform(A.n, 8)], 1)
    x1 = T.factor_llt_mf(PA).solve(b)
    x2 = T.factor_llt_mf(PA).solve(b)
    assert np.array_equal(x1, x2)


def test_many_rhs_equal_single_rhs():
    print("case: 20 rhs vs one at a time", flush=True)
    A = T.mesh3d(10)
    _, _, PA = prepared(A)
    F = T.factor_llt_mf(PA)
    B = np.stack([T.random_uniform(A.n, 100 + r) for r in range(20)], 1)
    X = F.solve(B)
    for r in (0, 7, 19):
        assert np.allclose(X[:, r], F.solve(B[:, r]), rtol=1e-14, atol=1e-15)


def test_linsolve_protocol():
    """progs/test_linsolve.c: 20^3 mesh, X ~ U[0,1], B = A X, 4 rhs."""
    print("case: linsolve protocol", flush=True)
    A = T.mesh3d(20)
    n = A.n
    X = np.stack([T.random_uniform(n, 20 + r, 0.0, 1.0) for r in range(4)], 1)
    B = np.stack([A.times(np.ascontiguousarray(X[:, r])) for r in range(4)], 1)
    M = A.to_scipy_full()

    def rres(Y):
        return np.linalg.norm(M @ Y - B) / np.linalg.norm(B)

    for opts in (["taucs.factor.LLT=true", "taucs.factor.mf=true"],
                 ["taucs.factor.LLT=true", "taucs.factor.ll=true"],
                 ["taucs.factor.LLT=true", "taucs.factor.mf=true", "taucs.factor.ordering=genmmd"]):
        rc, Y, F = T.linsolve(A, B, opts)
        assert rc == T.TAUCS_SUCCESS and rres(Y) <= 1e-12
    # factor and keep, then solve with the kept factor, then free
    rc, _, F = T.linsolve(A, None, ["taucs.factor.LLT=true", "taucs.factor.mf=true"])
    assert rc == T.TAUCS_SUCCESS and F
    rc, Y, _ = T.linsolve(A, B, ["taucs.factor=false"], F=F)
    assert rc == T.TAUCS_SUCCESS and rres(Y) <= 1e-12
    rc, _, _ = T.linsolve(None, None, [], F=F)
    assert rc == T.TAUCS_SUCCESS
    # expected failures
    rc, _, _ = T.linsolve(A, B, ["taucs.factor.LLT=true", "taucs.factor.mf=true", "taucs.maxdepth=5"])
    assert rc != T.TAUCS_SUCCESS
    rc, _, _ = T.linsolve(A, B, ["taucs.factor.LLT=true", "taucs.factor.ordering=colamd"])
    assert rc != T.TAUCS_SUCCESS
    # refinement and the reduced-precision factor through the same entry point
    rc, Y, _ = T.linsolve(A, B, ["taucs.factor.LLT=true", "taucs.factor.mf=true",
                                 "taucs.factor.precision=fp32", "taucs.solve.refine=true"])
    assert rc == T.TAUCS_SUCCESS and rres(Y) <= 1e-13


def test_linsolve_cg_stops_on_the_reference_rule():
    """taucs.solve.cg=true runs CG preconditioned by the factor and stops on
    |r|_2 <= convergetol |b|_2 (src/taucs_iter.c:139,186), for the FP64 and a
    reduced-precision factor.  The operator (conductivities over four decades,
    shift 1e-6) has a relative-residual floor near 1e-10 in FP64 (the direct
    solve's, printed), so the tolerances stay above it: a tolerance below the
    floor cannot be met by any FP64 iteration, and CG then returns its last
    iterate with success, as the reference's does."""
    print("case: linsolve cg convergetol", flush=True)
    A = T.mesh3d_varcoef(16, 16, 16, seed=11)
    n = A.n
    B = np.stack([T.random_uniform(n, 40 + r) for r in range(2)], 1)
    M = A.to_scipy_full()

    def rres(Y):
        return np.linalg.norm(B - M @ Y, axis=0) / np.linalg.norm(B, axis=0)

    base = ["taucs.factor.LLT=true", "taucs.factor.mf=true"]
    rc, Y, _ = T.linsolve(A, B, base)
    assert rc == T.TAUCS_SUCCESS
    print(f"  direct fp64 solve relres {rres(Y).max():.2e}", flush=True)
    for prec, tol in (("fp64", 1e-8), ("fp32", 1e-5), ("fp32", 1e-8)):
        rc, Y, _ = T.linsolve(A, B, base + [f"taucs.factor.precision={prec}", "taucs.solve.cg=true",
                                            f"taucs.solve.convergetol={tol}"])
        rr = rres(Y)
        print(f"  cg {prec} tol {tol:.0e}: relres {rr.max():.2e}", flush=True)
        assert rc == T.TAUCS_SUCCESS and rr.max() <= tol, (prec, tol, rr)


def test_linsolve_cg_with_a_kept_factor_on_a_new_matrix():
    """The reference's flow of reusing a factor as a preconditioner:
    taucs.factor=false + taucs.solve.cg=true on a matrix whose values differ
    from the factored one.  The residuals must be those of the new matrix,
    even when the new permuted matrix lands at the address of the old one."""
    print("case: linsolve cg, kept factor, new values", flush=True)
    A = T.mesh3d_varcoef(14, 14, 14, seed=3)
    n, cp, ri, v = A.arrays()
    rc, _, F = T.linsolve(A, None, ["taucs.factor.LLT=true", "taucs.factor.mf=true"])
    assert rc == T.TAUCS_SUCCESS and F
    B = np.stack([T.random_uniform(n, 60 + r) for r in range(2)], 1)
    d = 1.0 + 0.2 * np.random.default_rng(0).random(n)
    col = np.repeat(np.arange(n), np.diff(cp))
    for scale in ("uniform", "diagonal"):
        # 2A, and D A D with a random positive diagonal D: both stay SPD
        v2 = 2.0 * v if scale == "uniform" else v * d[ri] * d[col]
        A2 = T.from_arrays(n, cp, ri, v2)
        M2 = A2.to_scipy_full()
        rc, Y, _ = T.linsolve(A2, B, ["taucs.factor=false", "taucs.solve.cg=true",
                                      "taucs.solve.convergetol=1e-8", "taucs.solve.maxits=200"], F=F)
        rr = np.linalg.norm(B - M2 @ Y, axis=0) / np.linalg.norm(B, axis=0)
        # against the old values the residual of the new system would stay
        # near 0.5 (2A) or O(0.1) (D A D); 1e-8 is above this operator's FP64
        # relative-residual floor (about 1e-10)
        print(f"  {scale}: relres {rr.max():.2e}", flush=True)
        assert rc == T.TAUCS_SUCCESS and rr.max() <= 1e-8, (scale, rr)
    T.linsolve(None, None, [], F=F)


def test_refinement_reads_values_changed_in_place():
    """solve_refined without the factored-values flag uploads A's values on
    every call: changing them in place (same pointer) changes the residuals."""
    print("case: refine after in-place value change", flush=True)
    A = T.mesh3d_varcoef(12, 12, 12, seed=8)
    _, _, PA = prepared(A)
    n = PA.n
    F = T.factor_llt_mf(PA)
    b = T.random_uniform(n, 9)
    x1, i1 = F.solve_refined(PA, b, method=0, tol=1e-14, maxit=20)
    assert i1.converged
    vals = np.ctypeslib.as_array(T.C.cast(PA.ptr.contents.values, T.C.POINTER(T.C.c_double)),
                                 shape=(PA.nnz,))
    vals *= 1.5          # in place: same matrix object, same value pointer
    x2, i2 = F.solve_refined(PA, b, method=1, tol=1e-14, maxit=100)
    assert i2.converged
    M = PA.to_scipy_full()
    r = b - M @ x2
    bwd = abs(r).max() / (abs(M).sum(1).max() * abs(x2).max() + abs(b).max())
    # with the old values the iterate would solve A, not 1.5 A (x2 = x1);
    # both solutions are backward stable, so they agree to about cond(A) u
    assert bwd <= 1e-14
    assert np.linalg.norm(x2 - x1 / 1.5) <= 1e-6 * np.linalg.norm(x1 / 1.5)


@pytest.mark.parametrize("prec", [T.FP32])
@pytest.mark.parametrize("method", [0, 1])
def test_mixed_precision_refinement(prec, method):
    print(f"case: precision {prec} method {method}", flush=True)
    A = T.mesh3d_varcoef(24, 24, 24, seed=5)
    _, _, PA = prepared(A)
    b = np.stack([T.random_uniform(A.n, 30 + r) for r in range(4)], 1)
    F = T.factor_llt_mf(PA, precision=prec)
    assert F is not None and F.info().precision == prec
    x, info = F.solve_refined(PA, b, method=method, tol=1e-14, maxit=200)
    assert info.converged and info.backward_error <= 1e-14
    F64 = T.factor_llt_mf(PA)
    x64, info64 = F64.solve_refined(PA, b, method=0, tol=1e-15, maxit=10)
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) <= 1e-6   # kappa ~ 1e9 here


def test_induced_device_allocation_failures_are_clean():
    print("case: induced allocation failures", flush=True)
    A = T.mesh3d(10)
    _, _, PA = prepared(A)
    fails = 0
    for i in range(0, 120, 7):
        T.lib.taucs_allocation_induce_failure(i)
        h = T.lib.taucs_ccs_factor_llt_mf(PA.ptr)
        if h:
            T.lib.taucs_supernodal_factor_free(h)
        else:
            fails += 1
    T.lib.taucs_allocation_induce_failure(-1)
    assert fails > 0
    F = T.factor_llt_mf(PA)
    b = T.random_uniform(A.n, 1)
    assert relres(PA, F.solve(b), b) <= 1e-12


def _dense_lower(M):
    """The lower triangle of a dense symmetric matrix as a TAUCS matrix with
    every entry stored (explicit zeros too): one supernode of n columns."""
    n = M.shape[0]
    cp, ri, v = [0], [], []
    for j in range(n):
        for i in range(j, n):
            ri.append(i)
            v.append(M[i, j])
        cp.append(len(ri))
    return T.from_arrays(n, np.array(cp), np.array(ri), np.array(v))


def _neumann_biharmonic(n, shift):
    """D^T D + shift I, D the (n-2) x n second difference: integer entries plus
    a power-of-two shift, so the matrix is exact in FP32; its last pivots are
    about 1e-5 of the diagonal (condition number ~2e7)."""
    D = np.zeros((n - 2, n))
    for i in range(n - 2):
        D[i, i:i + 3] = (1.0, -2.0, 1.0)
    return D.T @ D + shift * np.eye(n)


@pytest.mark.parametrize("n", [40, 128])
def test_fp64_fallback_of_flagged_fp32_blocks(n):
    """A diagonal block whose FP32 pivots lose more than 12 bits to
    cancellation is refactored in FP64 (k_potrf_wave for the blocked front of
    128 columns, k_front_wave for the 40-column front).  The matrix is exact
    in FP32 and its flagged block has no FP32-rounded input (the coupling to
    the first block is zero), so with the fallback the FP32 factor's diagonal
    is the FP64 factor's rounded to FP32 (relative 1e-6); without it, FP32
    Cholesky loses several digits in the last pivots (relative error > 1e-3;
    about 6% in a CPU simulation of the same elimination order)."""
    print(f"case: fp64 fallback, dense {n}", flush=True)
    M = np.zeros((n, n))
    if n > 64:
        M[:64, :64] = 4.0 * np.eye(64)
        M[64:, 64:] = _neumann_biharmonic(n - 64, 2.0 ** -20)
    else:
        M = _neumann_biharmonic(n, 2.0 ** -20)
    A = _dense_lower(M)
    perm, inv = T.order(A, "identity")
    PA = T.permute(A, perm, inv)
    d64 = T.factor_llt_mf(PA).diag()
    try:
        T.lib.taucs_amd_set_fp64_fallback(0)
        F0 = T.factor_llt_mf(PA, precision=T.FP32)
        assert T.lib.taucs_amd_factor_fp64_fallbacks(F0.h) == 0
        e_off = (np.abs(F0.diag() - d64) / np.abs(d64)).max()
    finally:
        T.lib.taucs_amd_set_fp64_fallback(1)
    F1 = T.factor_llt_mf(PA, precision=T.FP32)
    nfb = T.lib.taucs_amd_factor_fp64_fallbacks(F1.h)
    rel = np.abs(F1.diag() - d64) / np.abs(d64)
    print(f"  fallbacks {nfb}; diag rel err with fallback {rel.max():.2e}, without {e_off:.2e}",
          flush=True)
    assert nfb == 1
    assert rel.max() <= 1e-6
    assert e_off > 1e-3
    b = T.random_uniform(n, 77)
    x, info = F1.solve_refined(PA, b, method=0, tol=1e-14, maxit=30, factored=True)
    assert info.converged and info.backward_error <= 1e-14
