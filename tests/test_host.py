"""Host side of libtaucs_amd.so against the oracle and the committed fixtures
(CPU only: no test here launches a GPU kernel)."""
import ctypes as C
import json
import os
import re

import numpy as np
import pytest

import taucs_amd as T
from oracle import oracle as O
from tests.golden.make_golden import KEYS, digest, splitmix_uniform

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
needs_ref = pytest.mark.skipif(not O.have_ref(), reason="reference GENMMD not built")


def header_functions():
    names = set()
    for h in ("taucs.h", "taucs_amd.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(taucs_\w+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) > 50
    missing = [n for n in sorted(names) if not hasattr(T.lib, n)]
    assert not missing, missing


def test_struct_layout_matches_reference_abi():
    # src/taucs.h:537-554: n, m, flags (int), colptr, rowind, values (pointers)
    assert C.sizeof(T.CcsMatrix) == 40
    assert T.CcsMatrix.colptr.offset == 16 and T.CcsMatrix.values.offset == 32


@pytest.mark.parametrize("side", [1, 2, 7, 100])
def test_mesh2d_equals_oracle(side):
    n, cp, ri, v = T.mesh2d(side).arrays()
    on, ocp, ori, ov = O.mesh2d(side)
    assert n == on and np.array_equal(cp, ocp) and np.array_equal(ri, ori) and np.array_equal(v, ov)


@pytest.mark.parametrize("side", [1, 3, 20])
def test_mesh3d_equals_oracle(side):
    n, cp, ri, v = T.mesh3d(side).arrays()
    on, ocp, ori, ov = O.mesh3d(side, side, side)
    assert np.array_equal(cp, ocp) and np.array_equal(ri, ori) and np.array_equal(v, ov)


def test_mesh2d_variants():
    d = T.mesh2d(4, "dirichlet").arrays()
    assert np.all(d[3][d[1][1:] - 1] == 4.0)
    ax = T.mesh2d(4, "anisotropic_x").to_scipy_full()
    assert ax[0, 1] == -100.0 and ax[0, 4] == -1.0


def test_varcoef_closed_form():
    A = T.mesh3d_varcoef(6, 5, 4, seed=7, log10_lo=-2, log10_hi=2, shift=1e-6)
    n, cp, ri, v = A.arrays()
    B = T.mesh3d(6, 5, 4).arrays()
    assert np.array_equal(cp, B[1]) and np.array_equal(ri, B[2])   # mesh3d pattern and order
    M = A.to_scipy_full()
    assert np.allclose(np.asarray(M.sum(1)).ravel(), 1e-6, rtol=0, atol=1e-12)  # row sums = shift
    off = v[ri != np.repeat(np.arange(n), np.diff(cp))]
    assert off.max() <= -1e-2 and off.min() >= -1e2
    A0 = T.mesh3d_varcoef(5, 5, 5, seed=0).arrays()
    assert np.array_equal(A0[3], T.mesh3d(5).arrays()[3])


def test_shifted_mesh2d():
    A = T.mesh2d_shifted(9, seed=3)
    n, cp, ri, v = A.arrays()
    base = T.mesh2d(9).arrays()[3]
    diag = cp[1:] - 1
    shift = v[diag] - base[diag]
    assert np.array_equal(np.delete(v, diag), np.delete(base, diag))
    assert shift.min() >= 0 and shift.max() < 1
    assert np.allclose(shift, T.random_uniform(n, 3, 0.0, 1.0), rtol=0, atol=1e-14)


def test_rng_matches_restatement():
    assert np.array_equal(T.random_uniform(1000, 42), splitmix_uniform(1000, 42))


def test_c1_fixture():
    g = np.load(os.path.join(GOLD, "c1_mesh2d_100.npz"))
    A = T.mesh2d(100)
    perm, inv = T.order(A, "genmmd")
    assert np.array_equal(perm, g["perm"]) and np.array_equal(inv, g["invperm"])
    PA = T.permute(A, perm, inv)
    assert np.array_equal(T.etree(PA), g["parent"])
    F = T.factor_symbolic(PA)
    S = F.supernodes()
    assert S["n_sn"] == int(g["n_sn"])
    for k in KEYS:
        assert np.array_equal(S[k], g[k]), k
    fi = F.info()
    assert fi.fund_flops == float(g["fund_flops"]) and fi.relax_flops == float(g["relax_flops"])
    assert fi.fund_nnz == float(g["fund_nnz"]) and fi.relax_nnz == float(g["relax_nnz"])
    assert np.array_equal(g["b"], T.random_uniform(A.n, 1))


@pytest.mark.parametrize("kind", ["mesh2d", "mesh3d"])
def test_size_sweeps_match_fixtures(kind):
    g = np.load(os.path.join(GOLD, f"sweep_{kind}.npz"))
    for i, side in enumerate(g["side"]):
        side = int(side)
        if kind == "mesh3d" and side > 30:
            continue  # covered by test_large_meshes_match_fixtures
        A = T.mesh2d(side) if kind == "mesh2d" else T.mesh3d(side)
        perm, inv = T.order(A, "genmmd")
        assert digest(perm) == g["perm_sha"][i], side
        PA = T.permute(A, perm, inv)
        assert digest(T.etree(PA)) == g["parent_sha"][i], side
        F = T.factor_symbolic(PA)
        S = F.supernodes()
        assert digest(*[S[k] for k in KEYS]) == g["sn_sha"][i], side
        assert F.info().fund_flops == g["fund_flops"][i]


@pytest.mark.parametrize("side", [60, 80])
def test_large_meshes_match_fixtures(side):
    big = json.load(open(os.path.join(GOLD, "big.json")))[f"mesh3d_{side}"]
    A = T.mesh3d(side)
    perm, inv = T.order(A, "genmmd")
    assert digest(perm) == big["perm_sha"]
    PA = T.permute(A, perm, inv)
    assert digest(T.etree(PA)) == big["parent_sha"]
    F = T.factor_symbolic(PA)
    S = F.supernodes()
    assert digest(*[S[k] for k in KEYS]) == big["sn_sha"]
    fi = F.info()
    assert fi.n_sn == big["n_sn"] and fi.etree_depth == big["depth"]
    assert fi.fund_flops == big["fund_flops"] and fi.relax_flops == big["relax_flops"]


def random_graph(n, deg, seed):
    """Random symmetric pattern, lower CSC, rows unsorted."""
    rng = np.random.default_rng(seed)
    edges = set()
    for i in range(n):
        for j in rng.integers(0, n, deg):
            if i != j:
                edges.add((max(i, j), min(i, j)))
    cols = [[j] for j in range(n)]
    for (i, j) in edges:
        cols[j].append(i)
    cp, ri, v = [0], [], []
    for j in range(n):
        rows = cols[j]
        rng.shuffle(rows)
        for i in rows:
            ri.append(i)
            v.append(float(deg * 4) if i == j else -1.0)
        cp.append(len(ri))
    return n, np.array(cp, np.int32), np.array(ri, np.int32), np.array(v)


@needs_ref
@pytest.mark.parametrize("seed", range(12))
def test_mmd_and_symbolic_on_random_graphs(seed):
    n, cp, ri, v = random_graph(200 + 97 * seed, 1 + seed % 4, seed)
    A = T.from_arrays(n, cp, ri, v)
    perm, inv = T.order(A, "genmmd")
    operm, oinv = O.genmmd(n, cp, ri)
    assert np.array_equal(perm, operm)
    PA = T.permute(A, perm, inv)
    pcp, pri, pv = O.permute(n, cp, ri, v, oinv)
    n2, cp2, ri2, v2 = PA.arrays()
    assert np.array_equal(cp2, pcp) and np.array_equal(ri2, pri) and np.array_equal(v2, pv)
    assert np.array_equal(T.etree(PA), O.etree(n, pcp, pri))
    S = T.factor_symbolic(PA).supernodes()
    R = O.symbolic(n, pcp, pri)
    assert S["n_sn"] == R["n_sn"]
    for k in KEYS:
        assert np.array_equal(S[k], R[k]), k


@needs_ref
def test_disconnected_forest_amalgamates_last_tree_only():
    # block diagonal: two grids; only the tree of the last supernode is
    # amalgamated (src/taucs_sn_llt.c:3927)
    a = T.mesh2d(12).arrays()
    n1 = a[0]
    cp = np.concatenate([a[1], a[1][1:] + a[1][-1]])
    ri = np.concatenate([a[2], a[2] + n1])
    v = np.concatenate([a[3], a[3]])
    n = 2 * n1
    A = T.from_arrays(n, cp, ri, v)
    perm, inv = T.order(A, "genmmd")
    assert np.array_equal(perm, O.genmmd(n, cp, ri)[0])
    PA = T.permute(A, perm, inv)
    n_, pcp, pri, pv = PA.arrays()
    S = T.factor_symbolic(PA).supernodes()
    R = O.symbolic(n, pcp, pri)
    for k in KEYS:
        assert np.array_equal(S[k], R[k]), k


def test_times_vec_matches_scipy():
    A = T.mesh3d_varcoef(7, 6, 5, seed=11)
    x = np.random.default_rng(0).standard_normal(A.n)
    assert np.allclose(A.times(x), A.to_scipy_full() @ x, rtol=1e-13, atol=1e-13)


def test_vec_norm2_scaled():
    x = np.array([3e200, 4e200, 0.0])
    assert T.lib.taucs_vec_norm2(3, T.TAUCS_DOUBLE, x.ctypes.data_as(C.c_void_p)) == pytest.approx(5e200)


def test_getopt_semantics():
    x = C.c_int(0)
    args = (C.c_void_p * 2)(C.cast(C.pointer(C.c_int(7)), C.c_void_p), None)
    assert T.lib.taucs_getopt_boolean(b"taucs.factor.mf=true", None, b"taucs.factor.mf", C.byref(x)) == 1 and x.value == 1
    assert T.lib.taucs_getopt_boolean(b"taucs.factor.mf=false", None, b"taucs.factor.mf", C.byref(x)) == 1 and x.value == 0
    # a longer option with the same prefix is not this option
    assert T.lib.taucs_getopt_boolean(b"taucs.factor.mf.x=true", None, b"taucs.factor.mf", C.byref(x)) == 0
    assert T.lib.taucs_getopt_boolean(b"taucs.factor.mf=#0", args, b"taucs.factor.mf", C.byref(x)) == 1 and x.value == 7
    d = C.c_double(0)
    assert T.lib.taucs_getopt_double(b"taucs.maxdepth=5", None, b"taucs.maxdepth", C.byref(d)) == 1 and d.value == 5
    s = C.c_char_p()
    assert T.lib.taucs_getopt_string(b"taucs.factor.ordering=genmmd", None, b"taucs.factor.ordering", C.byref(s)) == 1
    assert s.value == b"genmmd"


def test_linsolve_host_error_paths():
    A = T.mesh3d(6)
    B = np.ones(A.n)
    # colamd on a symmetric matrix must fail (progs/test_linsolve.c:53-91)
    rc, X, F = T.linsolve(A, B, ["taucs.factor.LLT=true", "taucs.factor.ordering=colamd"])
    assert rc != T.TAUCS_SUCCESS
    # depth limit (progs/test_linsolve.c:116-121)
    rc, X, F = T.linsolve(A, B, ["taucs.factor.LLT=true", "taucs.factor.mf=true", "taucs.maxdepth=5"])
    assert rc == T.TAUCS_ERROR
    # solve without a factorization (progs/test_linsolve.c:140-148)
    rc, X, F = T.linsolve(A, B, ["taucs.factor=false"])
    assert rc == T.TAUCS_ERROR
    # LU is not part of this library
    rc, X, F = T.linsolve(A, B, ["taucs.factor.LU=true"])
    assert rc == T.TAUCS_ERROR_BADARGS
    # freeing through the linsolve convention is a no-op on NULL
    rc, _, _ = T.linsolve(None, None, [])
    assert rc == T.TAUCS_SUCCESS


def test_maxdepth_fails_in_symbolic():
    A = T.mesh3d(10)
    perm, inv = T.order(A, "genmmd")
    PA = T.permute(A, perm, inv)
    assert not T.lib.taucs_ccs_factor_llt_mf_maxdepth(PA.ptr, 5)


def test_unknown_ordering_returns_null():
    A = T.mesh2d(5)
    for name in ("metis", "nonsense"):
        p, i = T.order(A, name)
        assert p is None and i is None


def test_induced_host_allocation_failures_are_clean():
    A = T.mesh2d(20)
    perm, inv = T.order(A, "genmmd")
    PA = T.permute(A, perm, inv)
    failures = 0
    for i in range(0, 60, 3):
        T.lib.taucs_allocation_induce_failure(i)
        h = T.lib.taucs_ccs_factor_llt_symbolic(PA.ptr)
        if h:
            T.lib.taucs_supernodal_factor_free(h)
        else:
            failures += 1
    T.lib.taucs_allocation_induce_failure(-1)
    assert failures > 0
    F = T.factor_symbolic(PA)
    assert F is not None and F.info().n_live > 0


@pytest.mark.skipif(T.gpu_available(), reason="checks the behaviour without a GPU")
def test_numeric_fails_loudly_without_gpu():
    A = T.mesh2d(10)
    perm, inv = T.order(A, "genmmd")
    PA = T.permute(A, perm, inv)
    assert not T.lib.taucs_ccs_factor_llt_mf(PA.ptr)


@pytest.mark.parametrize("kind,side", [("2d", 60), ("3d", 20), ("3d", 34)])
def test_update_blocks_with_overlapping_lifetimes_do_not_share_memory(kind, side):
    """Children extend-add into a front's update block at the end of their own
    levels and the front's own extend-add reads it at the end of its level, so
    a block lives from its lowest child's level through its own.  Blocks alive
    at a common level must be disjoint, every block lies inside the arena, and
    panels tile the L arena exactly."""
    A = T.mesh2d(side) if kind == "2d" else T.mesh3d(side)
    perm, inv = T.order(A, "genmmd")
    F = T.factor_symbolic(T.permute(A, perm, inv))
    fi = F.info()
    S = F.supernodes()
    sch = F.schedule()
    ns = fi.n_sn
    live = np.where(sch["level"] >= 0)[0]
    assert len(live) == fi.n_live
    s = S["sn_size"][:ns]
    m = S["sn_up_size"][:ns]
    u = (m - s).astype(np.int64)
    lev, par, oU, oL = sch["level"], sch["parent"], sch["off_U"], sch["off_L"]
    start = lev.copy()
    for k in live:
        if par[k] < ns:
            start[par[k]] = min(start[par[k]], lev[k])
    blocks = [(int(start[k]), int(lev[k]), int(oU[k]), int(oU[k] + u[k] * u[k]))
              for k in live if u[k] > 0]
    U_elems = fi.U_bytes // 8
    assert all(0 <= b[2] and b[3] <= U_elems for b in blocks)
    assert max(b[3] for b in blocks) == U_elems
    for l in range(fi.n_levels):
        alive = sorted((b[2], b[3]) for b in blocks if b[0] <= l <= b[1])
        for (a0, a1), (b0, b1) in zip(alive, alive[1:]):
            assert a1 <= b0, (l, a0, a1, b0, b1)
    # panels: disjoint and contiguous in level order
    pl = sorted((int(oL[k]), int(oL[k] + m[k] * s[k])) for k in live)
    assert pl[0][0] == 0 and pl[-1][1] == fi.L_bytes // 8
    assert all(a[1] == b[0] for a, b in zip(pl, pl[1:]))
    # reuse pays: far below the sum of all update blocks
    assert U_elems < sum(int(x) * int(x) for x in u[live])


def test_ccs_solve_llt_rejects_non_factors():
    """taucs_ccs_solve_llt checks flags and structure on the host before it
    touches the device (src/taucs_ccs_solve_llt.c:30-37 checks the flags; the
    reference asserts the diagonal-first order, here it is an error): a
    symmetric matrix, an upper one, a lower one whose columns end with their
    diagonal (the generators' order) and one with an entry above the diagonal
    all return -1."""
    A = T.mesh2d(4)
    n, cp, ri, v = A.arrays()
    low = T.TAUCS_TRIANGULAR | T.TAUCS_LOWER | T.TAUCS_DOUBLE
    cases = [A, T.from_arrays(n, cp, ri, v, T.TAUCS_TRIANGULAR | T.TAUCS_UPPER | T.TAUCS_DOUBLE),
             T.from_arrays(n, cp, ri, v, low)]
    # diagonal first, but column 1 also holds row 0
    cp3 = np.array([0, 1, 3]); ri3 = np.array([0, 1, 0]); v3 = np.array([2.0, 2.0, 1.0])
    cases.append(T.from_arrays(2, cp3, ri3, v3, low))
    for M in cases:
        x = np.zeros(M.n)
        assert T.lib.taucs_ccs_solve_llt(M.ptr, T._dptr(x), T._dptr(np.ones(M.n))) == -1


@pytest.mark.parametrize("name", ["comb20", "two_children", "star5", "path5", "mesh2d_3"])
def test_product_supernodes_equal_the_hand_derived_ones(name):
    """The product's symbolic phase on the small patterns whose supernodes,
    row order and child lists are derived by hand from the reference's rules
    (tests/test_oracle_order.py holds the derivations)."""
    from tests.test_oracle_order import EXAMPLES, ccs, structure
    n, cols, _, final = EXAMPLES[name]()
    cp, ri, v = ccs(n, cols)
    S = T.factor_symbolic(T.from_arrays(n, cp, ri, v)).supernodes()
    n_sn, live, kids = final
    got_live, got_kids = structure(S)
    assert S["n_sn"] == n_sn and got_live == live and got_kids == kids
