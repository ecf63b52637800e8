"""Pins of the oracle itself (CPU only).

The reference's tests hold no golden vectors for this path (its only check is
relative residual <= 1e-8 in progs/test_linsolve.c:45), so each oracle
function is pinned by what the mathematics fixes, chosen so that a plausible
mistake fails:
  generators   -- closed forms: entry counts, row sums, the corner anchor and
                  the per-column entry order of src/taucs_ccs_generators.c
  GENMMD       -- it IS the reference routine (external/src/genmmd.c built
                  unmodified); pinned further by being a permutation whose
                  fill is far below the natural order's on meshes
  permutation  -- P A P^T equals the dense permutation of A
  etree        -- parent[j] = min{i > j : L(i,j) != 0} of a brute-force dense
                  symbolic Cholesky
  supernodes   -- fundamental partition: chains with nested structure;
                  row structures equal the brute-force column structures of L;
                  relaxed partition re-derived merge by merge from the
                  published thresholds
  numeric      -- factor equals numpy.linalg.cholesky of P A P^T; solve equals
                  numpy.linalg.solve
  CCS solve    -- equals scipy's dense triangular solves L^-T L^-1 b on a
                  random sparse factor (a transposed sweep or a dropped
                  diagonal fails it)
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.skipif(not O.have_ref(), reason="reference GENMMD not built (make -C oracle)")


def dense(n, cp, ri, v):
    A = np.zeros((n, n))
    for j in range(n):
        for k in range(cp[j], cp[j + 1]):
            A[ri[k], j] = v[k]
            A[j, ri[k]] = v[k]
    return A


def random_spd_pattern(n, density, seed):
    """Random symmetric pattern (lower CSC) with a dominant diagonal."""
    rng = np.random.default_rng(seed)
    M = np.triu(rng.random((n, n)) < density, 1)
    A = np.where(M | M.T, -rng.random((n, n)), 0.0)
    A = np.tril(A + A.T) / 2
    A = np.tril(A, -1)
    d = np.abs(A).sum(0) + np.abs(A).sum(1) + 1.0
    cp, ri, v = [0], [], []
    for j in range(n):
        rows = [j] + [i for i in range(j + 1, n) if A[i, j] != 0]
        rng.shuffle(rows)  # unsorted within the column, like the generators
        for i in rows:
            ri.append(i)
            v.append(d[j] if i == j else A[i, j])
        cp.append(len(ri))
    return n, np.array(cp, np.int32), np.array(ri, np.int32), np.array(v)


def test_mesh2d_closed_form():
    for s in (1, 2, 5, 17):
        n, cp, ri, v = O.mesh2d(s)
        assert n == s * s and cp[n] == 3 * s * s - 2 * s
        A = dense(n, cp, ri, v)
        rs = A.sum(1)
        assert rs[0] == 1.0 and np.all(rs[1:] == 0.0)          # Neumann + corner anchor
        for y in range(s):
            for x in range(s):
                j = x + y * s
                rows = list(ri[cp[j]:cp[j + 1]])
                exp = ([x + (y + 1) * s] if y < s - 1 else []) + ([x + 1 + y * s] if x < s - 1 else []) + [j]
                assert rows == exp


def test_mesh3d_closed_form():
    for s in (1, 3, 6):
        n, cp, ri, v = O.mesh3d(s, s, s)
        assert cp[n] == 4 * n - 3 * s * s
        A = dense(n, cp, ri, v)
        rs = A.sum(1)
        assert rs[0] == 1.0 and np.all(rs[1:] == 0.0)
        j = 0
        for z in range(s):
            for y in range(s):
                for x in range(s):
                    rows = list(ri[cp[j]:cp[j + 1]])
                    exp = ([j + 1] if x < s - 1 else []) + ([j + s] if y < s - 1 else []) + \
                          ([j + s * s] if z < s - 1 else []) + [j]
                    assert rows == exp
                    j += 1


def fill_count(n, cp, ri, perm):
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    pcp, pri, pv = O.permute(n, cp, ri, np.ones(cp[n]), inv)
    S = O.symbolic(n, pcp, pri)
    return S["fund_nnz"]


def test_genmmd_is_a_fill_reducing_permutation():
    for n, cp, ri, v in (O.mesh2d(30), O.mesh3d(8, 8, 8)):
        perm, inv = O.genmmd(n, cp, ri)
        assert sorted(perm) == list(range(n))
        assert np.all(inv[perm] == np.arange(n))
        assert fill_count(n, cp, ri, perm) < 0.5 * fill_count(n, cp, ri, np.arange(n))


def test_permute_matches_dense():
    n, cp, ri, v = random_spd_pattern(40, 0.1, 3)
    perm = np.random.default_rng(0).permutation(n).astype(np.int32)
    inv = np.empty(n, np.int32)
    inv[perm] = np.arange(n)
    pcp, pri, pv = O.permute(n, cp, ri, v, inv)
    A = dense(n, cp, ri, v)
    B = dense(n, pcp, pri, pv)
    assert np.array_equal(B, A[np.ix_(perm, perm)])
    assert all(pri[k] >= j for j in range(n) for k in range(pcp[j], pcp[j + 1]))


def brute_structure(A):
    """Column structures of the Cholesky factor by dense symbolic elimination."""
    n = A.shape[0]
    S = [set(np.nonzero(A[j:, j])[0] + j) for j in range(n)]
    for j in range(n):
        below = sorted(i for i in S[j] if i > j)
        if below:
            p = below[0]
            S[p] |= set(i for i in S[j] if i > j)
    return S


@pytest.mark.parametrize("seed", range(6))
def test_etree_and_structure_brute_force(seed):
    n, cp, ri, v = random_spd_pattern(30 + 5 * seed, 0.08, seed)
    A = dense(n, cp, ri, v)
    S = brute_structure(A)
    parent = O.etree(n, cp, ri)
    for j in range(n):
        below = sorted(i for i in S[j] if i > j)
        assert parent[j] == (below[0] if below else n)
    sym = O.symbolic(n, cp, ri)
    # every live supernode: its columns' structures are prefixes of its rows
    covered = []
    for sn in O.postorder(sym):
        rows = list(O.rows_of(sym, sn))
        s = sym["sn_size"][sn]
        cols = rows[:s]
        covered += cols
        for jj, c in enumerate(cols):
            assert S[c] <= set(rows[jj:])   # ancestors come after descendants
        # structure union equals the merged supernode's row set
        assert set().union(*[S[c] for c in cols]) | set(cols) == set(rows)
    assert sorted(covered) == list(range(n))


def test_fundamental_supernodes_definition():
    """The fundamental supernodes carry the exact column counts of L: the
    reference's statistics sum_j (1 + c_j^2) and sum_j c_j over them
    (src/taucs_sn_llt.c:3892-3921) equal the brute-force counts."""
    for n, cp, ri, v in (random_spd_pattern(50, 0.06, 11), O.mesh2d(12)):
        perm, inv = (np.arange(n), np.arange(n)) if n == 50 else O.genmmd(n, cp, ri)
        pcp, pri, pv = O.permute(n, cp, ri, v, inv)
        S = brute_structure(dense(n, pcp, pri, pv))
        cc = [len(s) for s in S]
        flops = sum(1.0 + c * c for c in cc)
        sym = O.symbolic(n, pcp, pri)
        assert sym["fund_flops"] == flops
        assert sym["fund_nnz"] == sum(cc)


def _children(sym, sn):
    c = sym["first_child"][sn]
    while c != -1:
        yield int(c)
        c = sym["next_child"][c]


@pytest.mark.parametrize("mesh", [("2d", 40), ("2d", 64), ("3d", 10), ("3d", 14)])
def test_amalgamation_thresholds_rederived(mesh):
    """Relaxed supernodes only add explicit zeros, and every parent that kept
    its children separate fails the merge rule of src/taucs_sn_llt.c:1716-1725
    when the rule is re-evaluated from first principles: zeros = padded
    entries of the would-be merged supernode minus the true nonzeros of its
    columns (brute-force structure), nonzeros = those true nonzeros."""
    kind, side = mesh
    n, cp, ri, v = O.mesh2d(side) if kind == "2d" else O.mesh3d(side, side, side)
    perm, inv = O.genmmd(n, cp, ri)
    pcp, pri, pv = O.permute(n, cp, ri, v, inv)
    sym = O.symbolic(n, pcp, pri)
    S = brute_structure(dense(n, pcp, pri, pv)) if n <= 4100 else None
    if S is None:
        return
    true_nnz = lambda cols: sum(len(S[c]) for c in cols)
    padded = lambda s, m: s * (s + 1) // 2 + (m - s) * s
    explicit = 0
    checked = 0
    for sn in O.postorder(sym):
        rows = list(O.rows_of(sym, sn))
        s = int(sym["sn_size"][sn])
        explicit += padded(s, len(rows)) - true_nnz(rows[:s])
        kids = list(_children(sym, sn))
        if not kids:
            continue
        kid_cols = [c for k in kids for c in O.rows_of(sym, k)[: sym["sn_size"][k]]]
        new_size = s + len(kid_cols)
        new_up = len(rows) + len(kid_cols)
        nonzeros = true_nnz(rows[:s]) + true_nnz(kid_cols)
        zeros = padded(new_size, new_up) - nonzeros
        merge = (new_size < 16 or (s < 50 and zeros < 0.5 * nonzeros)
                 or (s < 250 and zeros < 0.25 * nonzeros) or (s < 500 and zeros < 0.10 * nonzeros)
                 or zeros < 0.05 * nonzeros)
        assert not merge, f"supernode {sn} should have absorbed its children"
        checked += 1
    assert checked > 0
    assert explicit == sym["relax_nnz"] - sym["fund_nnz"]
    live = set(O.postorder(sym))
    for sn in range(sym["n_sn"]):
        assert (sym["sn_size"][sn] > 0) == (sn in live)


def test_numeric_factor_matches_dense_cholesky():
    for n, cp, ri, v in (O.mesh2d(9), O.mesh3d(4, 4, 4), random_spd_pattern(45, 0.1, 5)):
        inv = np.arange(n, dtype=np.int32) if n == 45 else O.genmmd(n, cp, ri)[1]
        pcp, pri, pv = O.permute(n, cp, ri, v, inv)
        sym = O.symbolic(n, pcp, pri)
        panels = O.factor(sym, pcp, pri, pv)
        Ld = np.zeros((n, n))
        for sn, P in panels.items():
            rows = O.rows_of(sym, sn)
            s = sym["sn_size"][sn]
            for jj in range(s):
                Ld[rows[jj:], rows[jj]] = P[jj:, jj]
        Ad = dense(n, pcp, pri, pv)
        Lref = np.linalg.cholesky(Ad)
        assert np.allclose(Ld, Lref, rtol=1e-12, atol=1e-12)
        b = np.random.default_rng(2).uniform(-1, 1, (n, 3))
        x = O.solve(sym, panels, b)
        assert np.allclose(x, np.linalg.solve(Ad, b), rtol=1e-10, atol=1e-12)


def test_numeric_factor_rejects_indefinite():
    n, cp, ri, v = O.mesh2d(6)
    v = v.copy()
    v[cp[10 + 1] - 1] = -5.0  # diagonal of column 10 made negative
    sym = O.symbolic(n, cp, ri)
    with pytest.raises(np.linalg.LinAlgError):
        O.factor(sym, cp, ri, v)


@pytest.mark.parametrize("seed", [1, 2])
def test_ccs_solve_llt_matches_dense_triangular_solves(seed):
    import scipy.linalg as sl
    rng = np.random.default_rng(seed)
    n = 60
    L = np.tril(rng.uniform(-1, 1, (n, n)) * (rng.uniform(0, 1, (n, n)) < 0.15), -1)
    L += np.diag(rng.uniform(1, 3, n))
    cp, ri, v = [0], [], []
    for j in range(n):                 # diagonal first, then the rows below
        rows = [j] + [i for i in range(j + 1, n) if L[i, j] != 0]
        ri += rows
        v += [L[i, j] for i in rows]
        cp.append(len(ri))
    b = rng.uniform(-1, 1, n)
    x = O.ccs_solve_llt(n, np.array(cp), np.array(ri), np.array(v), b)
    xd = sl.solve_triangular(L, sl.solve_triangular(L, b, lower=True), lower=True, trans="T")
    assert np.abs(x - xd).max() <= 1e-12 * np.abs(xd).max()


def test_c_numeric_oracle_matches_dense_cholesky_and_the_python_restatement():
    """oracle.c's multifrontal factor (host BLAS per front, F1/F2 kept apart as
    the reference keeps them) and its supernodal solve: every panel equals the
    dense Cholesky factor of P A P^T, and the solution equals a dense solve
    and the numpy restatement's."""
    for n, cp, ri, v in (O.mesh2d(9), O.mesh3d(5, 5, 5), random_spd_pattern(45, 0.1, 5),
                         O.mesh3d_varcoef(4, 5, 6, 9)):
        inv = O.genmmd(n, cp, ri)[1]
        pcp, pri, pv = O.permute(n, cp, ri, v, inv)
        sym = O.symbolic(n, pcp, pri)
        M = O.MF(n, pcp, pri, pv)
        Ad = dense(n, pcp, pri, pv)
        Lref = np.linalg.cholesky(Ad)
        for sn in O.postorder(sym):
            rows = O.rows_of(sym, sn)
            s = int(sym["sn_size"][sn])
            P = M.panel(sn, len(rows), s)
            for jj in range(s):
                # |l_ij| <= sqrt(a_ii): rounding of the two factorizations differs
                # by a few ulps of that scale (a tiny last pivot cancels)
                ref = Lref[rows[jj:], rows[jj]]
                assert np.abs(P[jj:, jj] - ref).max() <= 1e-12 * np.sqrt(np.diag(Ad).max())
        b = np.random.default_rng(3).uniform(-1, 1, (n, 4))
        x = M.solve(b)
        # agreement to about cond(A) * u (the variable-coefficient case has
        # cond ~ 1e7: solutions of size 3e4 differ in the 1e-10 digit)
        xd = np.linalg.solve(Ad, b)
        xp = O.solve(sym, O.factor(sym, pcp, pri, pv), b)
        tol = 1e-16 * np.linalg.cond(Ad) * np.abs(xd).max() + 1e-13
        assert np.abs(x - xd).max() <= tol and np.abs(x - xp).max() <= tol
        M.free()


def test_c_numeric_oracle_rejects_indefinite():
    n, cp, ri, v = O.mesh2d(6)
    v = v.copy()
    v[cp[10 + 1] - 1] = -5.0
    with pytest.raises(np.linalg.LinAlgError):
        O.MF(n, cp, ri, v)
