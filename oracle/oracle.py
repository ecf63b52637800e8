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


def symbolic(n, colptr, rowind, max_depth=0):
    """Supernodal structure exactly as taucs_ccs_symbolic_elimination leaves it."""
    colptr = np.ascontiguousarray(colptr, np.int32)
    rowind = np.ascontiguousarray(rowind, np.int32)
    h = lib().oracle_symbolic(n, _p(colptr), _p(rowind), max_depth)
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


def residual(n, colptr, rowind, vals, x, b):
    """b - A x for symmetric lower CSC (src/taucs_ccs_ops.c:299-344)."""
    import scipy.sparse as sp
    Lw = sp.csc_matrix((vals, rowind, colptr), shape=(n, n))
    A = Lw + Lw.T - sp.diags(Lw.diagonal())
    return b - A @ x
