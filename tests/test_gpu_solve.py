"""GPU: the two sweep schedules of the supernodal solve, and refinement that
must not report a diverging iterate as converged.

The chained sweeps (one launch per tree level, block chains followed inside
the launch) and the per-block-step sweeps compute the same triangular solves
in different summation orders: their solutions agree to a few ulps of the
condition number (rtol 1e-12 in FP64 on these small meshes, 1e-5 in FP32),
and each against a dense solve.  Each schedule is deterministic: a repeated
solve is bitwise equal."""
import numpy as np
import pytest

import taucs_amd as T

pytestmark = pytest.mark.gpu


def factored(A, precision, chained):
    perm, inv = T.order(A, "genmmd")
    PA = T.permute(A, perm, inv)
    T.lib.taucs_amd_set_solve_kernel(1 if chained else 0)
    try:
        F = T.factor_symbolic(PA)
        assert F.numeric(PA, precision) == 0
    finally:
        T.lib.taucs_amd_set_solve_kernel(1)
    return F, PA


@pytest.mark.parametrize("precision,rtol", [(T.FP64, 1e-12), (T.FP32, 1e-5)])
@pytest.mark.parametrize("nrhs", [1, 3, 8, 16, 20])
def test_chained_and_step_sweeps_agree(precision, rtol, nrhs):
    print(f"case: chained vs step sweeps, mesh3d 14, precision {precision}, nrhs {nrhs}", flush=True)
    A = T.mesh3d(14)
    Fc, PA = factored(A, precision, True)
    Fs, _ = factored(A, precision, False)
    rng = np.random.default_rng(nrhs)
    B = rng.uniform(-1, 1, (PA.n, nrhs))
    Xc = Fc.solve(B)
    Xs = Fs.solve(B)
    M = PA.to_scipy_full().toarray()
    Xd = np.linalg.solve(M, B)
    scale = np.abs(Xd).max()
    assert np.abs(Xc - Xs).max() <= rtol * scale
    if precision == T.FP64:
        assert np.abs(Xc - Xd).max() <= 1e-10 * scale
    else:  # FP32 factor and solve: normwise backward error of single precision
        anorm = np.abs(M).sum(1).max()
        eta = np.abs(B - M @ Xc).max(0) / (anorm * np.abs(Xc).max(0) + np.abs(B).max(0))
        assert eta.max() <= 1e-5
    assert np.array_equal(Fc.solve(B), Xc)
    assert np.array_equal(Fs.solve(B), Xs)
    Fc.free()
    Fs.free()


def test_chained_sweeps_on_a_deep_front_chain():
    """Fronts wider than 64 columns at every top level: the root of mesh3d 24
    has a chain of tens of 64-column blocks in both sweeps."""
    print("case: chained sweeps, mesh3d 24, 4 rhs", flush=True)
    A = T.mesh3d(24)
    Fc, PA = factored(A, T.FP64, True)
    Fs, _ = factored(A, T.FP64, False)
    B = np.stack([T.random_uniform(PA.n, 40 + r) for r in range(4)], 1)
    Xc = Fc.solve(B)
    Xs = Fs.solve(B)
    assert np.abs(Xc - Xs).max() <= 1e-11 * np.abs(Xs).max()
    M = PA.to_scipy_full()
    R = B - M @ Xc
    assert (np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)).max() <= 1e-11
    Fc.free()
    Fs.free()


@pytest.mark.parametrize("method", [0, 1])
def test_diverging_refinement_is_not_reported_converged(method):
    """A preconditioner that is the factor of 0.3 A makes classical
    refinement's error grow by 7/3 per step: the iterate's backward error can
    still look small once |A||x| dwarfs the residual, so the growth of the
    residual must stop it as diverged.  CG with the same (SPD, merely
    mis-scaled) preconditioner converges."""
    print(f"case: refinement with a mis-scaled factor, method {method}", flush=True)
    A = T.mesh3d(10)
    perm, inv = T.order(A, "genmmd")
    PA = T.permute(A, perm, inv)
    n, cp, ri, v = PA.arrays()
    scaled = T.from_arrays(n, cp, ri, 0.3 * v)
    F = T.factor_symbolic(scaled)
    assert F.numeric(scaled, T.FP64) == 0
    B = np.stack([T.random_uniform(n, 7 + r) for r in range(2)], 1)
    X, info = F.solve_refined(PA, B, method=method, tol=1e-13, maxit=200)
    if method == 0:
        assert not info.converged
        assert info.iterations < 200
    else:
        assert info.converged
        assert info.backward_error <= 1e-13
        M = PA.to_scipy_full()
        assert np.abs(B - M @ X).max() <= 1e-10 * np.abs(B).max()
    F.free()


def test_device_solve_in_place():
    """x may alias b in the device-pointer solve: b is consumed before x is
    written (the chained schedule's readiness fill of x comes after)."""
    print("case: in-place device solve, mesh3d 16, 5 rhs", flush=True)
    A = T.mesh3d(16)
    F, PA = factored(A, T.FP64, True)
    n = PA.n
    B = np.stack([T.random_uniform(n, 90 + r) for r in range(5)], 1)
    X = F.solve(B)
    dB = T.DeviceArray(B)
    F.solve_device(dB.ptr, dB.ptr, n, 5)
    assert np.array_equal(dB.get(), X)
    dB.free()
    F.free()


@pytest.mark.parametrize("mesh", [1, 2, 9, 14, 30])
def test_ccs_solve_llt(mesh):
    """taucs_ccs_solve_llt on the CCS export of the factor: equals the oracle's
    column sweeps (src/taucs_ccs_solve_llt.c:19-130) and the supernodal solve
    to rounding (rtol 1e-12), and is deterministic."""
    from oracle import oracle as O
    print(f"case: ccs solve, mesh3d {mesh}", flush=True)
    A = T.mesh3d(mesh)
    F, PA = factored(A, T.FP64, True)
    L = F.to_ccs()
    n, cp, ri, v = L.arrays()
    assert np.array_equal(ri[cp[:-1]], np.arange(n))
    b = T.random_uniform(n, 3 + mesh)
    x = L.solve_llt(b)
    xo = O.ccs_solve_llt(n, cp, ri, v, b)
    assert np.abs(x - xo).max() <= 1e-12 * np.abs(xo).max()
    assert np.abs(x - F.solve(b)).max() <= 1e-12 * np.abs(xo).max()
    assert np.array_equal(L.solve_llt(b), x)
    L.free()
    F.free()



@pytest.mark.parametrize("width", [128, 512])
def test_lookahead_factor_is_bitwise_identical(width):
    """The lookahead runs each outer panel's trailing update beside the next
    panel's POTRF/TRSM/update chain (the update's workgroups reserve LDS so
    the chain fits beside them).  The two touch disjoint columns, so the
    factor must be bitwise the one of the serial schedule, and a repeat
    bitwise the same: a race would show up here first (mesh3d 44: fronts of
    thousands of rows, trailing updates of thousands of tiles)."""
    print(f"case: lookahead bitwise, mesh3d 44, width {width}", flush=True)
    A = T.mesh3d(44)
    perm, inv = T.order(A, "genmmd")
    PA = T.permute(A, perm, inv)
    vals = []
    try:
        T.lib.taucs_amd_set_panel_width(width)
        for la in (1, 1, 0):
            T.lib.taucs_amd_set_lookahead(la)
            F = T.factor_symbolic(PA)
            assert F.numeric(PA, T.FP64) == 0
            vals.append(F.to_ccs().arrays()[3])
            F.free()
    finally:
        T.lib.taucs_amd_set_lookahead(1)
        T.lib.taucs_amd_set_panel_width(256)
    assert np.array_equal(vals[0], vals[1])
    assert np.array_equal(vals[0], vals[2])


@pytest.mark.parametrize("prec", [T.FP64, T.FP32])
def test_potrf_kernels_are_bitwise_identical(prec):
    """The four-wave diagonal-block kernel (16-column panels, MFMA updates,
    the inverse in LDS) and the one-wave kernel (register broadcasts) give
    the same bits: the MFMA steps apply their products as fused
    multiply-adds in k order, as the one-wave kernel does.  mesh3d 36 has
    fronts of several 64-column blocks and partial last blocks; the dense
    FP32 case takes the FP64 fallback in one block."""
    cases = [("mesh3d 36", T.mesh3d(36), "genmmd")]
    if prec == T.FP32:
        M = np.zeros((128, 128))
        M[:64, :64] = 4.0 * np.eye(64)
        D = np.zeros((62, 64))
        for i in range(62):
            D[i, i:i + 3] = (1.0, -2.0, 1.0)
        M[64:, 64:] = D.T @ D + 2.0 ** -20 * np.eye(64)
        cp, ri, v = [0], [], []
        for j in range(128):
            for i in range(j, 128):
                ri.append(i)
                v.append(M[i, j])
            cp.append(len(ri))
        cases.append(("dense 128, flagged block",
                      T.from_arrays(128, np.array(cp), np.array(ri), np.array(v)), "identity"))
    for name, A, ordering in cases:
        print(f"case: potrf kernels bitwise, {name}, precision {prec}", flush=True)
        perm, inv = T.order(A, ordering)
        PA = T.permute(A, perm, inv)
        b = T.random_uniform(PA.n, 5)
        res = []
        try:
            for kind in (2, 1):
                T.lib.taucs_amd_set_potrf_kernel(kind)
                F = T.factor_symbolic(PA)
                assert F.numeric(PA, prec) == 0
                res.append((F.digest(), F.solve(b), T.lib.taucs_amd_factor_fp64_fallbacks(F.h)))
                F.free()
        finally:
            T.lib.taucs_amd_set_potrf_kernel(2)
        assert res[0][0] == res[1][0]
        assert np.array_equal(res[0][1], res[1][1])
        assert res[0][2] == res[1][2]


def test_matrix_file_through_the_drop_in(tmp_path):
    """The flow of the reference's direct driver (progs/direct.c:540-550) on a
    matrix read from a MatrixMarket file: order, factor on the GPU, solve with
    the supernodal factor and with its CCS export (taucs_ccs_solve_llt); both
    solutions against a dense solve of the matrix as read."""
    print("case: mtx file -> genmmd -> factor -> supernodal and CCS solves", flush=True)
    A = T.mesh3d(10)
    n, cp, ri, v = A.arrays()
    lines = [f"{n} {n} {len(v)}"]
    for j in range(n):
        for p in range(cp[j], cp[j + 1]):
            lines.append(f"{ri[p] + 1} {j + 1} {float(v[p]):.17g}")
    f = tmp_path / "mesh3d10.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real symmetric\n" + "\n".join(lines) + "\n")
    M = T.read_matrix(f, "mtx", T.TAUCS_SYMMETRIC | T.TAUCS_LOWER | T.TAUCS_DOUBLE)
    assert np.array_equal(M.to_scipy_full().toarray(), A.to_scipy_full().toarray())
    perm, inv = T.order(M, "genmmd")
    PM = T.permute(M, perm, inv)
    F = T.factor_symbolic(PM)
    assert F.numeric(PM, T.FP64) == 0
    b = T.random_uniform(n, 77)
    x1 = F.solve(b)
    L = F.to_ccs()
    x2 = L.solve_llt(b)
    xd = np.linalg.solve(PM.to_scipy_full().toarray(), b)
    for x in (x1, x2):
        assert np.abs(x - xd).max() <= 1e-10 * np.abs(xd).max()
    L.free()
    F.free()
