"""ORACLE -- test infrastructure only (imported by tests/, __graft_entry__.smoke()
and bench.py's CPU baseline; never by the product).

CPU restatement of the TAUCS 2.2.1 multifrontal LL^T path:
  * symbolic phase: liboracle.so (oracle.c), a C restatement of
    src/taucs_sn_llt.c:3663-4001 and GENMMD itself (the reference's
    external/src/genmmd.c compiled unmodified into _ref/);
  * numeric phase: `factor` below restates the recursive multifrontal driver
    src/taucs_sn_llt.c:1831-1977 front by front: zeroed F1/F2/U
    (:1013-1047), extend-add of each child in child-list order
    (:1200-1240, :1797-1827), assembly of A's columns (:1100-1119), then
    LAPACK dpotrf / dtrsm / dsyrk exactly as :1131-1186 call them;
  * `solve` restates taucs_dsupernodal_solve_llt (:2844-2885) with its
    forward (:2597-2720) and backward (:2723-2841) recursions.
The dense kernels are LAPACK/BLAS through scipy (the reference calls the
Fortran BLAS/LAPACK of the host), so the numeric oracle is pinned by the
uniqueness of the Cholesky factor (checked against numpy.linalg.cholesky on
small matrices in tests/test_oracle.py).
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise ImportError("oracle not built: make -C oracle")
        _lib = C.CDLL(path)
        _lib.oracle_symbolic.restype = C.c_void_p
        _lib.oracle_symbolic.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        _lib.oracle_symbolic_ex.restype = C.c_void_p
        _lib.oracle_symbolic_ex.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        _lib.oracle_sym_merge_log.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        _lib.oracle_sym_info.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        _lib.oracle_sym_export.argtypes = [C.c_void_p] + [C.c_void_p] * 6
        _lib.oracle_sym_free.argtypes = [C.c_void_p]
        _lib.oracle_genmmd.argtypes = [C.c_int] + [C.c_void_p] * 4
        _lib.oracle_load_ref.argtypes = [C.c_char_p]
        _lib.oracle_etree.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.oracle_permute.argtypes = [C.c_int] + [C.c_void_p] * 7
        for g in ("oracle_mesh2d",):
            getattr(_lib, g).argtypes = [C.c_int] + [C.c_void_p] * 3
        _lib.oracle_mesh3d.argtypes = [C.c_int] * 3 + [C.c_void_p] * 3
        _lib.oracle_free.argtypes = [C.c_void_p]
        _lib.oracle_load_blas.argtypes = [C.c_char_p, C.c_char_p]
        _lib.oracle_blas_threads.argtypes = [C.c_int]
        _lib.oracle_mf_factor.restype = C.c_void_p
        _lib.oracle_mf_factor.argtypes = [C.c_void_p] * 4
        _lib.oracle_mf_failed.argtypes = [C.c_void_p]
        _lib.oracle_mf_solve.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        _lib.oracle_mf_panel.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        _lib.oracle_mf_free.argtypes = [C.c_void_p]
        _lib.oracle_mf_times.argtypes = [C.c_void_p]
        _lib.oracle_blas_config.restype = C.c_char_p
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


REF_SO = os.path.join(HERE, "_ref", "libref_order.so")


def have_ref():
    return os.path.exists(REF_SO)


_ref_loaded = False


def _load_ref():
    global _ref_loaded
    if not _ref_loaded:
        if not have_ref() or lib().oracle_load_ref(REF_SO.encode()) != 0:
            raise RuntimeError("reference GENMMD not built (make -C oracle with /root/reference)")
        _ref_loaded = True


def _take(cp, ri, v, N, nnz):
    colptr = np.ctypeslib.as_array(C.cast(cp, C.POINTER(C.c_int)), shape=(N + 1,)).copy()
    rowind = np.ctypeslib.as_array(C.cast(ri, C.POINTER(C.c_int)), shape=(nnz,)).copy()
    vals = np.ctypeslib.as_array(C.cast(v, C.POINTER(C.c_double)), shape=(nnz,)).copy()
    for q in (cp, ri, v):
        lib().oracle_free(q)
    return colptr, rowind, vals


def mesh2d(n):
    cp, ri, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nnz = lib().oracle_mesh2d(n, C.byref(cp), C.byref(ri), C.byref(v))
    return (n * n,) + _take(cp, ri, v, n * n, nnz)


def mesh3d(X, Y, Z):
    cp, ri, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nnz = lib().oracle_mesh3d(X, Y, Z, C.byref(cp), C.byref(ri), C.byref(v))
    return (X * Y * Z,) + _take(cp, ri, v, X * Y * Z, nnz)


_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix_u01(count, seed):
    """The seeded uniform[0,1) stream of the workloads' generators: splitmix64
    (state += 0x9E3779B97F4A7C15, then the xor-shift-multiply finalizer), top
    53 bits scaled by 2^-53.  Vectorized over the stream position; numpy's
    uint64 arithmetic wraps modulo 2^64 like the C code's."""
    i = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform(count, seed, lo=-1.0, hi=1.0):
    return lo + (hi - lo) * splitmix_u01(count, seed)


def mesh3d_varcoef(X, Y, Z, seed, log10_lo=-2.0, log10_hi=2.0, shift=1e-6):
    """Variable-coefficient 7-point operator on the mesh3d pattern and entry
    order (the seeded workload generator, not a reference routine): every grid
    edge draws c = 10^(lo + (hi-lo) u) from one stream in the generator's visit
    order (z, y, x loops; x+1, y+1, z+1 edges of each point); off-diagonal -c;
    diagonal = incident c summed in the order +x, +y, +z, -x, -y, -z, then the
    shift."""
    N = X * Y * Z
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    has = np.stack([x < X - 1, y < Y - 1, z < Z - 1], 1)        # per point: x+1, y+1, z+1
    u = splitmix_u01(int(has.sum()), seed)
    w = np.zeros((N, 3))
    expo = log10_lo + (log10_hi - log10_lo) * u
    # C's pow (numpy's vectorized power may differ from it in the last bit)
    w[has] = np.fromiter((math.pow(10.0, e) for e in expo), np.float64, len(expo))
    wx, wy, wz = w[:, 0], w[:, 1], w[:, 2]
    d = np.zeros(N)
    for term in (wx, wy, wz):
        d = d + term
    j = np.arange(N)
    for sel, off, ww in ((x > 0, 1, wx), (y > 0, X, wy), (z > 0, X * Y, wz)):
        t = np.zeros(N)
        t[sel] = ww[j[sel] - off]
        d = d + t
    d = d + shift
    cnt = has.sum(1) + 1
    cp = np.zeros(N + 1, np.int64)
    cp[1:] = np.cumsum(cnt)
    ri = np.empty(cp[-1], np.int32)
    v = np.empty(cp[-1])
    pos = cp[:-1].copy()
    for k, off in enumerate((1, X, X * Y)):
        sel = has[:, k]
        ri[pos[sel]] = j[sel] + off
        v[pos[sel]] = -w[sel, k]
        pos[sel] += 1
    ri[pos] = j
    v[pos] = d
    return N, cp.astype(np.int32), ri, v


def mesh2d_shifted(s, seed):
    """mesh2d(s, "neumann") plus an i.i.d. uniform[0,1) shift on each diagonal
    entry (the last entry of each column), drawn in column order."""
    n, cp, ri, v = mesh2d(s)
    v = v.copy()
    v[cp[1:] - 1] += splitmix_u01(n, seed)
    return n, cp, ri, v


def genmmd(n, colptr, rowind):
    _load_ref()
    colptr = np.ascontiguousarray(colptr, np.int32)
    rowind = np.ascontiguousarray(rowind, np.int32)
    perm = np.empty(n, np.int32)
    inv = np.empty(n, np.int32)
    if lib().oracle_genmmd(n, _p(colptr), _p(rowind), _p(perm), _p(inv)) != 0:
        raise RuntimeError("genmmd failed")
    return perm, inv


def permute(n, colptr, rowind, vals, invperm):
    nnz = int(colptr[n])
    oc = np.empty(n + 1, np.int32)
    orr = np.empty(nnz, np.int32)
    ov = np.empty(nnz, np.float64)
    lib().oracle_permute(n, _p(np.ascontiguousarray(colptr, np.int32)),
                         _p(np.ascontiguousarray(rowind, np.int32)),
                         _p(np.ascontiguousarray(vals, np.float64)),
                         _p(np.ascontiguousarray(invperm, np.int32)), _p(oc), _p(orr), _p(ov))
    return oc, orr, ov


def etree(n, colptr, rowind):
    parent = np.empty(n, np.int32)
    lib().oracle_etree(n, _p(np.ascontiguousarray(colptr, np.int32)),
                       _p(np.ascontiguousarray(rowind, np.int32)), _p(parent))
    return parent


def symbolic(n, colptr, rowind, max_depth=0, amalgamate=True, merge_log=False):
    """Supernodal structure exactly as taucs_ccs_symbolic_elimination leaves it.
    amalgamate=False stops after the fundamental supernodes
    (recursive_symbolic_elimination alone); merge_log=True also returns, per
    supernode with children that recursive_amalgamate_supernodes visits, the
    decision it took: (sn, own size, candidate size, candidate update size,
    merged, zeros, nonzeros, candidate columns)."""
    colptr = np.ascontiguousarray(colptr, np.int32)
    rowind = np.ascontiguousarray(rowind, np.int32)
    flags = (0 if amalgamate else 1) | (2 if merge_log else 0)
    h = lib().oracle_symbolic_ex(n, _p(colptr), _p(rowind), max_depth, flags)
    if not h:
        raise RuntimeError("oracle symbolic failed")
    ns, dp, fl = C.c_int(), C.c_int(), C.c_int()
    st = np.zeros(4)
    lib().oracle_sym_info(h, C.byref(ns), C.byref(dp), C.byref(fl), _p(st))
    ns = ns.value
    out = dict(n=n, n_sn=ns, depth=dp.value, failed=bool(fl.value), fund_nnz=st[0],
               fund_flops=st[1], relax_nnz=st[2], relax_flops=st[3])
    if not fl.value:
        sz, up, fc, nc, sp_ = (np.zeros(ns + 1, np.int32) for _ in range(5))
        tot = lib().oracle_sym_export(h, _p(sz), _p(up), _p(fc), _p(nc), _p(sp_), None)
        idx = np.zeros(max(tot, 1), np.int32)
        lib().oracle_sym_export(h, None, None, None, None, None, _p(idx))
        out.update(sn_size=sz[:ns], sn_up_size=up[:ns], first_child=fc, next_child=nc,
                   struct_ptr=sp_, struct_idx=idx[:tot])
    if merge_log:
        nc_ = C.c_int()
        cnt = lib().oracle_sym_merge_log(h, None, None, None, C.byref(nc_))
        ri_ = np.zeros((max(cnt, 1), 6), np.int32)
        rd_ = np.zeros((max(cnt, 1), 2))
        cols = np.zeros(max(nc_.value, 1), np.int32)
        lib().oracle_sym_merge_log(h, _p(ri_), _p(rd_), _p(cols), None)
        out["merge_log"] = [dict(sn=int(r[0]), own_size=int(r[1]), new_size=int(r[2]),
                                 new_up=int(r[3]), merged=bool(r[4]), zeros=float(d[0]),
                                 nonzeros=float(d[1]), cols=cols[r[5]:r[5] + r[2]].copy())
                            for r, d in zip(ri_[:cnt], rd_[:cnt])]
    lib().oracle_sym_free(h)
    return out


def _children(sym, sn):
    c = sym["first_child"][sn]
    while c != -1:
        yield int(c)
        c = sym["next_child"][c]


def postorder(sym):
    """Live supernodes, children before parents, in the reference's
    recursion order (children in child-list order)."""
    root = sym["n_sn"]
    out, stack = [], [(root, iter(list(_children(sym, root))))]
    while stack:
        sn, it = stack[-1]
        c = next(it, None)
        if c is None:
            stack.pop()
            if sn != root:
                out.append(sn)
        else:
            stack.append((c, iter(list(_children(sym, c)))))
    return out


def rows_of(sym, sn):
    a = sym["struct_ptr"][sn]
    return sym["struct_idx"][a: a + sym["sn_up_size"][sn]]


def factor(sym, colptr, rowind, vals, dtype=np.float64):
    """Multifrontal LL^T of PAPT (lower CSC) on the given structure.
    Returns {sn: panel (m x s) = [L11; L21]}; raises on a nonpositive pivot."""
    import scipy.linalg as sl
    fronts, panels = {}, {}
    n = sym["n"]
    pos = np.empty(n, np.int64)
    parent = {}
    for sn in range(sym["n_sn"] + 1):
        for c in _children(sym, sn):
            parent[c] = sn
    for sn in postorder(sym):
        s = int(sym["sn_size"][sn])
        rows = rows_of(sym, sn)
        m = len(rows)
        F = fronts.pop(sn, None)
        if F is None:
            F = (np.zeros((m, s), dtype), np.zeros((m - s, m - s), dtype))
        P, U = F
        pos[rows] = np.arange(m)
        for jj in range(s):                        # A assembly
            col = rows[jj]
            a, b = colptr[col], colptr[col + 1]
            P[pos[rowind[a:b]], jj] += vals[a:b]
        if s:
            try:
                L11 = sl.cholesky(P[:s, :s], lower=True, check_finite=False)
            except np.linalg.LinAlgError:
                raise np.linalg.LinAlgError(f"not positive definite in supernode {sn}")
            P[:s, :s] = L11
            if m > s:
                P[s:, :] = sl.solve_triangular(L11, P[s:, :].T, lower=True,
                                               check_finite=False).T
                U -= P[s:, :] @ P[s:, :].T
        panels[sn] = P
        p = parent.get(sn, sym["n_sn"])
        if p != sym["n_sn"] and m > s:             # extend-add into the parent
            prow = rows_of(sym, p)
            ps, pm = int(sym["sn_size"][p]), len(prow)
            if p not in fronts:
                fronts[p] = (np.zeros((pm, ps), dtype), np.zeros((pm - ps, pm - ps), dtype))
            PP, PU = fronts[p]
            pos[prow] = np.arange(pm)
            rel = pos[rows[s:]]
            I, J = np.tril_indices(m - s)
            pi, pj = rel[I], rel[J]
            r_, c_ = np.maximum(pi, pj), np.minimum(pi, pj)
            v = U[I, J]
            inF = c_ < ps
            PP[r_[inF], c_[inF]] += v[inF]
            PU[r_[~inF] - ps, c_[~inF] - ps] += v[~inF]
    return panels


def solve(sym, panels, b):
    """x = (L L^T)^-1 b in the permuted space; b of shape (n,) or (n, k)."""
    import scipy.linalg as sl
    b = np.asarray(b, dtype=np.float64)
    x = b.copy().reshape(b.shape[0], -1)
    y = np.zeros_like(x)
    order = postorder(sym)
    for sn in order:                               # forward, children first
        s = int(sym["sn_size"][sn])
        rows = rows_of(sym, sn)
        P = panels[sn]
        xd = sl.solve_triangular(P[:s, :s], x[rows[:s]], lower=True, check_finite=False)
        y[rows[:s]] = xd
        if len(rows) > s:
            x[rows[s:]] -= P[s:, :] @ xd
    out = np.zeros_like(x)
    for sn in reversed(order):                     # backward, parents first
        s = int(sym["sn_size"][sn])
        rows = rows_of(sym, sn)
        P = panels[sn]
        bd = y[rows[:s]].copy()
        if len(rows) > s:
            bd -= P[s:, :].T @ out[rows[s:]]
        out[rows[:s]] = sl.solve_triangular(P[:s, :s], bd, lower=True, trans="T",
                                            check_finite=False)
    return out.reshape(b.shape)


def ccs_solve_llt(n, colptr, rowind, vals, b):
    """x = (L L^T)^-1 b for a lower CSC factor whose columns start with their
    diagonal (src/taucs_ccs_solve_llt.c:19-130): forward by columns, each
    column scattering y_j L_ij into the rows below it; backward by columns
    from the last, each gathering L_ji x_j before dividing by L_ii."""
    x = np.array(b, dtype=np.float64)
    y = np.zeros(n)
    for j in range(n):
        p0, p1 = colptr[j], colptr[j + 1]
        assert rowind[p0] == j
        y[j] = x[j] / vals[p0]
        x[rowind[p0 + 1:p1]] -= y[j] * vals[p0 + 1:p1]
    for i in range(n - 1, -1, -1):
        p0, p1 = colptr[i], colptr[i + 1]
        y[i] -= np.dot(x[rowind[p0 + 1:p1]], vals[p0 + 1:p1])
        x[i] = y[i] / vals[p0]
    return x


_blas_loaded = False


def _load_blas():
    """The host BLAS/LAPACK of the C numeric oracle: scipy's bundled OpenBLAS
    (exported as scipy_dpotrf_, ...)."""
    global _blas_loaded
    if _blas_loaded:
        return
    import glob
    import scipy
    libs = sorted(glob.glob(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs",
                                         "libscipy_openblas*.so")))
    if not libs or lib().oracle_load_blas(os.path.abspath(libs[0]).encode(), b"scipy_") != 0:
        raise RuntimeError("no host BLAS/LAPACK for the C numeric oracle")
    _blas_loaded = True


def blas_config():
    _load_blas()
    return lib().oracle_blas_config().decode(errors="replace")


def blas_threads(t):
    """Threads of the host BLAS used by the C numeric oracle."""
    _load_blas()
    lib().oracle_blas_threads(int(t))


class MF:
    """The C restatement of the numeric path (oracle.c: oracle_mf_factor /
    oracle_mf_solve): the reference's multifrontal factor with dpotrf / dtrsm /
    dsyrk per front and its supernodal solve, on the host BLAS.  Construction
    runs the symbolic phase and the numeric factorization; `t_symbolic` and
    `t_numeric` are their wall times (s)."""

    def __init__(self, n, colptr, rowind, vals):
        import time
        _load_blas()
        L = lib()
        self.colptr = np.ascontiguousarray(colptr, np.int32)
        self.rowind = np.ascontiguousarray(rowind, np.int32)
        self.vals = np.ascontiguousarray(vals, np.float64)
        self.n = n
        t0 = time.perf_counter()
        self.S = L.oracle_symbolic(n, _p(self.colptr), _p(self.rowind), 0)
        t1 = time.perf_counter()
        if not self.S:
            raise RuntimeError("oracle symbolic failed")
        self.N = L.oracle_mf_factor(self.S, _p(self.colptr), _p(self.rowind), _p(self.vals))
        t2 = time.perf_counter()
        self.t_symbolic, self.t_numeric = t1 - t0, t2 - t1
        if not self.N or L.oracle_mf_failed(self.N) != -1:
            self.free()
            raise np.linalg.LinAlgError("not positive definite")
        ns, dp, fl = C.c_int(), C.c_int(), C.c_int()
        st = np.zeros(4)
        L.oracle_sym_info(self.S, C.byref(ns), C.byref(dp), C.byref(fl), _p(st))
        self.fund_flops, self.relax_flops = st[1], st[3]
        tp = np.zeros(6)
        L.oracle_mf_times(_p(tp))
        self.parts = dict(zip(("assemble", "potrf", "trsm", "syrk", "extend_add", "alloc"), tp))

    def solve(self, b):
        b = np.asarray(b, dtype=np.float64)
        B = np.asfortranarray(b.reshape(b.shape[0], -1))
        X = np.zeros_like(B, order="F")
        if lib().oracle_mf_solve(self.N, B.shape[1], _p(B), _p(X)) != 0:
            raise RuntimeError("oracle solve failed")
        return X.reshape(b.shape)

    def panel(self, sn, m, s):
        out = np.zeros((m, s), order="F")
        if lib().oracle_mf_panel(self.N, sn, _p(out)) != 0:
            raise KeyError(sn)
        return out

    def free(self):
        if getattr(self, "N", None):
            lib().oracle_mf_free(self.N)
            self.N = None
        if getattr(self, "S", None):
            lib().oracle_sym_free(self.S)
            self.S = None

    def __del__(self):
        self.free()


def residual(n, colptr, rowind, vals, x, b):
    """b - A x for symmetric lower CSC (src/taucs_ccs_ops.c:299-344)."""
    import scipy.sparse as sp
    Lw = sp.csc_matrix((vals, rowind, colptr), shape=(n, n))
    A = Lw + Lw.T - sp.diags(Lw.diagonal())
    return b - A @ x
