"""Python view of libtaucs_amd.so, the MI355X multifrontal Cholesky behind the
TAUCS C-ABI.

Everything here is a thin ctypes binding over the C entry points declared in
include/taucs.h and include/taucs_amd.h; the numeric work runs in the HIP
kernels of the library.  Importing fails loudly when the library has not
been built (python taucs_amd/build.py).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtaucs_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `python taucs_amd/build.py`")
lib = C.CDLL(LIB_PATH)

TAUCS_SUCCESS, TAUCS_ERROR, TAUCS_ERROR_NOMEM, TAUCS_ERROR_BADARGS = 0, -1, -2, -3
TAUCS_INT, TAUCS_DOUBLE, TAUCS_SINGLE = 1024, 2048, 4096
TAUCS_LOWER, TAUCS_UPPER, TAUCS_TRIANGULAR, TAUCS_SYMMETRIC = 1, 2, 4, 8
TAUCS_HERMITIAN, TAUCS_PATTERN = 16, 32
FP64, FP32, BF16UPD = 0, 1, 2


class CcsMatrix(C.Structure):
    """taucs_ccs_matrix (src/taucs.h:537-554)."""
    _fields_ = [("n", C.c_int), ("m", C.c_int), ("flags", C.c_int),
                ("colptr", C.POINTER(C.c_int)), ("rowind", C.POINTER(C.c_int)),
                ("values", C.c_void_p)]


class FactorInfo(C.Structure):
    _fields_ = [("n", C.c_int), ("n_sn", C.c_int), ("n_live", C.c_int), ("n_levels", C.c_int),
                ("max_front", C.c_int), ("etree_depth", C.c_int), ("precision", C.c_int),
                ("has_numeric", C.c_int), ("fund_nnz", C.c_double), ("fund_flops", C.c_double),
                ("relax_nnz", C.c_double), ("relax_flops", C.c_double),
                ("t_symbolic", C.c_double), ("t_numeric", C.c_double),
                ("L_bytes", C.c_int64), ("U_bytes", C.c_int64)]


class RefineInfo(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int),
                ("backward_error", C.c_double), ("rel_residual", C.c_double)]


class NumericTiming(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("front_ms", C.c_double), ("assembly_ms", C.c_double),
                ("gemm_ms", C.c_double), ("gemm_launches", C.c_int), ("gemm_flops", C.c_double),
                ("assembly_bytes", C.c_double), ("launches", C.c_int)]


P = C.POINTER
_ccs_p = P(CcsMatrix)
_sigs = {
    "taucs_ccs_create": (_ccs_p, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "taucs_ccs_free": (None, [_ccs_p]),
    "taucs_ccs_generate_mesh2d": (_ccs_p, [C.c_int, C.c_char_p]),
    "taucs_ccs_generate_mesh3d": (_ccs_p, [C.c_int, C.c_int, C.c_int]),
    "taucs_amd_generate_mesh3d_varcoef": (_ccs_p, [C.c_int, C.c_int, C.c_int, C.c_uint64,
                                                   C.c_double, C.c_double, C.c_double]),
    "taucs_amd_generate_mesh2d_shifted": (_ccs_p, [C.c_int, C.c_uint64]),
    "taucs_amd_random_uniform": (None, [C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_void_p]),
    "taucs_ccs_order": (None, [_ccs_p, P(P(C.c_int)), P(P(C.c_int)), C.c_char_p]),
    "taucs_ccs_permute_symmetrically": (_ccs_p, [_ccs_p, P(C.c_int), P(C.c_int)]),
    "taucs_ccs_times_vec": (None, [_ccs_p, C.c_void_p, C.c_void_p]),
    "taucs_ccs_etree": (C.c_int, [_ccs_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "taucs_vec_norm2": (C.c_double, [C.c_int, C.c_int, C.c_void_p]),
    "taucs_ccs_factor_llt_mf": (C.c_void_p, [_ccs_p]),
    "taucs_ccs_factor_llt_mf_maxdepth": (C.c_void_p, [_ccs_p, C.c_int]),
    "taucs_ccs_factor_llt_symbolic": (C.c_void_p, [_ccs_p]),
    "taucs_ccs_factor_llt_numeric": (C.c_int, [_ccs_p, C.c_void_p]),
    "taucs_supernodal_solve_llt": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "taucs_supernodal_factor_free": (None, [C.c_void_p]),
    "taucs_supernodal_factor_free_numeric": (None, [C.c_void_p]),
    "taucs_supernodal_factor_to_ccs": (_ccs_p, [C.c_void_p]),
    "taucs_supernodal_factor_get_diag": (C.c_void_p, [C.c_void_p]),
    "taucs_ccs_solve_llt": (C.c_int, [_ccs_p, C.c_void_p, C.c_void_p]),
    "taucs_ccs_read_binary": (_ccs_p, [C.c_char_p]),
    "taucs_ccs_write_binary": (C.c_int, [_ccs_p, C.c_char_p]),
    "taucs_ccs_write_ijv": (C.c_int, [_ccs_p, C.c_char_p]),
    "taucs_ccs_read_ijv": (_ccs_p, [C.c_char_p, C.c_int]),
    "taucs_ccs_read_mtx": (_ccs_p, [C.c_char_p, C.c_int]),
    "taucs_ccs_read_ccs": (_ccs_p, [C.c_char_p, C.c_int]),
    "taucs_ccs_read_hb": (_ccs_p, [C.c_char_p, C.c_int]),
    "taucs_vec_read_binary": (C.c_void_p, [C.c_int, C.c_int, C.c_char_p]),
    "taucs_vec_write_binary": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_char_p]),
    "taucs_linsolve": (C.c_int, [_ccs_p, P(C.c_void_p), C.c_int, C.c_void_p, C.c_void_p,
                                 P(C.c_char_p), P(C.c_void_p)]),
    "taucs_logfile": (None, [C.c_char_p]),
    "taucs_free_stub": (None, [C.c_void_p]),
    "taucs_allocation_induce_failure": (None, [C.c_int]),
    "taucs_allocation_attempts": (C.c_int, []),
    "taucs_amd_factor_get_info": (C.c_int, [C.c_void_p, P(FactorInfo)]),
    "taucs_amd_factor_get_supernodes": (C.c_int, [C.c_void_p] + [C.c_void_p] * 6),
    "taucs_amd_factor_get_schedule": (C.c_int, [C.c_void_p] + [C.c_void_p] * 4),
    "taucs_amd_supernodal_solve_many": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                                  C.c_void_p, C.c_int]),
    "taucs_amd_supernodal_solve_many_device": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                                         C.c_void_p, C.c_int]),
    "taucs_amd_supernodal_solve_many_device_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p,
                                                                 C.c_int, C.c_void_p, C.c_int]),
    "taucs_amd_noalloc_scratch_bytes": (C.c_longlong, []),
    "taucs_amd_set_fp64_fallback": (None, [C.c_int]),
    "taucs_amd_factor_fp64_fallbacks": (C.c_int, [C.c_void_p]),
    "taucs_amd_factor_llt_mf_precision": (C.c_void_p, [_ccs_p, C.c_int]),
    "taucs_amd_factor_llt_numeric_precision": (C.c_int, [_ccs_p, C.c_void_p, C.c_int]),
    "taucs_amd_solve_refined": (C.c_int, [C.c_void_p, _ccs_p, C.c_int, C.c_void_p, C.c_int,
                                          C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int,
                                          P(RefineInfo)]),
    "taucs_amd_factor_digest": (C.c_int, [C.c_void_p, P(C.c_uint64)]),
    "taucs_amd_solve_refined_device": (C.c_int, [C.c_void_p, _ccs_p, C.c_int, C.c_void_p, C.c_int,
                                                 C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int,
                                                 P(RefineInfo)]),
    "taucs_amd_set_profiling": (None, [C.c_int]),
    "taucs_amd_set_graphs": (None, [C.c_int]),
    "taucs_amd_set_lookahead": (None, [C.c_int]),
    "taucs_amd_set_potrf_kernel": (None, [C.c_int]),
    "taucs_amd_set_front_packing": (None, [C.c_int]),
    "taucs_amd_set_solve_kernel": (None, [C.c_int]),
    "taucs_amd_set_panel_width": (None, [C.c_int]),
    "taucs_amd_memcpy_d2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "taucs_amd_factor_llt_numeric_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "taucs_amd_batch_numeric_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "taucs_amd_set_gemm_tile": (None, [C.c_int]),
    "taucs_amd_autotune": (C.c_int, [C.c_void_p, _ccs_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                     C.c_int, C.c_void_p, C.c_void_p]),
    "taucs_amd_batch_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int,
                                           C.c_void_p, C.c_void_p]),
    "taucs_amd_get_op_profile": (C.c_int, [C.c_void_p, C.c_int] + [C.c_void_p] * 6),
    "taucs_amd_get_op_tiles": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "taucs_amd_batch_factor_llt_mf": (C.c_void_p, [C.c_void_p, C.c_void_p, C.c_int]),
    "taucs_amd_batch_numeric": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "taucs_amd_batch_solve": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "taucs_amd_batch_solve_device": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                               C.c_int]),
    "taucs_amd_batch_solve_refined": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                                C.c_int, C.c_int, C.c_double, C.c_int, P(RefineInfo)]),
    "taucs_amd_batch_info": (C.c_int, [C.c_void_p, P(C.c_int), P(C.c_int), C.c_void_p, P(C.c_double),
                                       P(C.c_double), P(C.c_int), P(C.c_double)]),
    "taucs_amd_batch_system_factor": (C.c_void_p, [C.c_void_p, C.c_int]),
    "taucs_amd_batch_free": (None, [C.c_void_p]),
    "taucs_amd_batch_get_timing": (C.c_int, [C.c_void_p, P(NumericTiming)]),
    "taucs_amd_device_malloc": (C.c_void_p, [C.c_size_t]),
    "taucs_amd_device_free": (None, [C.c_void_p]),
    "taucs_amd_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "taucs_amd_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "taucs_amd_device_synchronize": (C.c_int, []),
    "taucs_amd_factor_device_bytes": (C.c_longlong, [C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "taucs_amd_device_bytes_in_use": (C.c_longlong, []),
    "taucs_amd_get_numeric_timing": (C.c_int, [C.c_void_p, P(NumericTiming)]),
}
for _name, (_res, _args) in _sigs.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def _dptr(a):
    return a.ctypes.data_as(C.c_void_p)


def ccs_to_numpy(Ap):
    """(n, colptr, rowind, values) copies of a taucs_ccs_matrix."""
    A = Ap.contents
    n = A.n
    colptr = np.ctypeslib.as_array(A.colptr, shape=(n + 1,)).copy()
    nnz = int(colptr[n])
    rowind = np.ctypeslib.as_array(A.rowind, shape=(max(nnz, 1),))[:nnz].copy()
    vals = np.ctypeslib.as_array(C.cast(A.values, P(C.c_double)), shape=(max(nnz, 1),))[:nnz].copy()
    return n, colptr, rowind, vals


class Matrix:
    """Owner of a library-allocated taucs_ccs_matrix."""

    def __init__(self, ptr):
        if not ptr:
            raise RuntimeError("matrix creation failed")
        self.ptr = ptr

    @property
    def n(self):
        return self.ptr.contents.n

    @property
    def nnz(self):
        return self.ptr.contents.colptr[self.n]

    def arrays(self):
        return ccs_to_numpy(self.ptr)

    def to_scipy_full(self):
        import scipy.sparse as sp
        n, cp, ri, v = self.arrays()
        Lw = sp.csc_matrix((v, ri, cp), shape=(n, n))
        return (Lw + Lw.T - sp.diags(Lw.diagonal())).tocsr()

    def times(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        lib.taucs_ccs_times_vec(self.ptr, _dptr(x), _dptr(y))
        return y

    def solve_llt(self, b):
        """x with L L^T x = b for a lower-triangular factor L (this matrix, as
        taucs_supernodal_factor_to_ccs exports it), both sweeps on the GPU:
        taucs_ccs_solve_llt."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        rc = lib.taucs_ccs_solve_llt(self.ptr, _dptr(x), _dptr(b))
        if rc != 0:
            raise RuntimeError(f"taucs_ccs_solve_llt failed ({rc})")
        return x

    def write(self, path, fmt="binary"):
        """Write in the binary or ijv format (taucs_ccs_write_binary/_ijv)."""
        f = {"binary": lib.taucs_ccs_write_binary, "ijv": lib.taucs_ccs_write_ijv}[fmt]
        if f(self.ptr, os.fsencode(path)) != 0:
            raise RuntimeError(f"writing {path} failed")

    def free(self):
        if self.ptr:
            lib.taucs_ccs_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def from_arrays(n, colptr, rowind, vals, flags=TAUCS_SYMMETRIC | TAUCS_LOWER | TAUCS_DOUBLE):
    nnz = int(colptr[n])
    p = lib.taucs_ccs_create(n, n, nnz, flags)
    if not p:
        raise MemoryError
    A = p.contents
    # the converted arrays are held in names until the copies are done (a
    # temporary's buffer is freed as soon as its address has been taken)
    cp = np.ascontiguousarray(colptr, dtype=np.int32)
    ri = np.ascontiguousarray(rowind, dtype=np.int32)
    v = np.ascontiguousarray(vals, dtype=np.float64)
    C.memmove(A.colptr, cp.ctypes.data, 4 * (n + 1))
    C.memmove(A.rowind, ri.ctypes.data, 4 * nnz)
    C.memmove(A.values, v.ctypes.data, 8 * nnz)
    return Matrix(p)


def read_matrix(path, fmt="binary", flags=TAUCS_SYMMETRIC | TAUCS_LOWER | TAUCS_DOUBLE):
    """Read a matrix file: fmt is "binary" (the flags are in the file),
    "ijv", "mtx", "ccs" or "hb" (taucs_ccs_read_*)."""
    name = os.fsencode(path)
    if fmt == "binary":
        p = lib.taucs_ccs_read_binary(name)
    else:
        p = getattr(lib, "taucs_ccs_read_" + fmt)(name, flags)
    if not p:
        raise RuntimeError(f"reading {path} as {fmt} failed")
    return Matrix(p)


def write_vector(path, v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    if lib.taucs_vec_write_binary(v.size, TAUCS_DOUBLE, _dptr(v), os.fsencode(path)) != 0:
        raise RuntimeError(f"writing {path} failed")


def read_vector(path, n):
    p = lib.taucs_vec_read_binary(n, TAUCS_DOUBLE, os.fsencode(path))
    if not p:
        raise RuntimeError(f"reading {path} failed")
    v = np.ctypeslib.as_array(C.cast(p, P(C.c_double)), shape=(n,)).copy()
    lib.taucs_free_stub(p)
    return v


def mesh2d(n, which="neumann"):
    return Matrix(lib.taucs_ccs_generate_mesh2d(n, which.encode()))


def mesh3d(X, Y=None, Z=None):
    return Matrix(lib.taucs_ccs_generate_mesh3d(X, Y or X, Z or X))


def mesh3d_varcoef(X, Y, Z, seed, log10_lo=-2.0, log10_hi=2.0, shift=1e-6):
    return Matrix(lib.taucs_amd_generate_mesh3d_varcoef(X, Y, Z, seed, log10_lo, log10_hi, shift))


def mesh2d_shifted(n, seed):
    return Matrix(lib.taucs_amd_generate_mesh2d_shifted(n, seed))


def random_uniform(n, seed, lo=-1.0, hi=1.0):
    out = np.empty(n, dtype=np.float64)
    lib.taucs_amd_random_uniform(n, seed, lo, hi, _dptr(out))
    return out


def order(A, which="genmmd"):
    """(perm, invperm) as numpy arrays, or (None, None) on failure."""
    pp, ip = P(C.c_int)(), P(C.c_int)()
    lib.taucs_ccs_order(A.ptr, C.byref(pp), C.byref(ip), which.encode())
    if not pp:
        return None, None
    n = A.n
    perm = np.ctypeslib.as_array(pp, shape=(n,)).copy()
    inv = np.ctypeslib.as_array(ip, shape=(n,)).copy()
    lib.taucs_free_stub(C.cast(pp, C.c_void_p))
    lib.taucs_free_stub(C.cast(ip, C.c_void_p))
    return perm, inv


def permute(A, perm, invperm):
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    invperm = np.ascontiguousarray(invperm, dtype=np.int32)
    return Matrix(lib.taucs_ccs_permute_symmetrically(
        A.ptr, perm.ctypes.data_as(P(C.c_int)), invperm.ctypes.data_as(P(C.c_int))))


def etree(A):
    parent = np.empty(A.n, dtype=np.int32)
    if lib.taucs_ccs_etree(A.ptr, _dptr(parent), None, None, None) != 0:
        raise RuntimeError("etree failed")
    return parent


REFINE_FACTORED_VALUES = 0x100
REFINE_STOP_RELRES = 0x200


def _refine_method(method, factored, stop_relres):
    return int(method) | (REFINE_FACTORED_VALUES if factored else 0) | \
        (REFINE_STOP_RELRES if stop_relres else 0)


class Factor:
    """Owner of an opaque device-resident supernodal factor."""

    def __init__(self, handle):
        if not handle:
            raise RuntimeError("factorization failed")
        self.h = C.c_void_p(handle)

    def info(self):
        fi = FactorInfo()
        lib.taucs_amd_factor_get_info(self.h, C.byref(fi))
        return fi

    def supernodes(self):
        fi = self.info()
        ns = fi.n_sn
        sz = np.zeros(ns + 1, np.int32)
        up = np.zeros(ns + 1, np.int32)
        fc = np.zeros(ns + 1, np.int32)
        nc = np.zeros(ns + 1, np.int32)
        sp_ = np.zeros(ns + 1, np.int32)
        lib.taucs_amd_factor_get_supernodes(self.h, _dptr(sz), _dptr(up), _dptr(fc), _dptr(nc),
                                            _dptr(sp_), None)
        idx = np.zeros(max(int(sp_[ns]), 1), np.int32)
        lib.taucs_amd_factor_get_supernodes(self.h, None, None, None, None, None, _dptr(idx))
        return dict(n_sn=ns, sn_size=sz[:ns], sn_up_size=up[:ns], first_child=fc, next_child=nc,
                    struct_ptr=sp_, struct_idx=idx[: int(sp_[ns])])

    def schedule(self):
        """Per supernode: level, parent, L offset, U offset (live supernodes only)."""
        ns = self.info().n_sn
        lev = np.zeros(ns, dtype=np.int32)
        par = np.zeros(ns, dtype=np.int32)
        oL = np.zeros(ns, dtype=np.int64)
        oU = np.zeros(ns, dtype=np.int64)
        lib.taucs_amd_factor_get_schedule(self.h, _dptr(lev), _dptr(par), _dptr(oL), _dptr(oU))
        return {"level": lev, "parent": par, "off_L": oL, "off_U": oU}

    def numeric(self, A, precision=None):
        if precision is None:
            return lib.taucs_ccs_factor_llt_numeric(A.ptr, self.h)
        return lib.taucs_amd_factor_llt_numeric_precision(A.ptr, self.h, precision)

    def solve(self, b):
        """x = (L L^T)^-1 b in the permuted space; b of shape (n,) or (n, k)."""
        b = np.asarray(b, dtype=np.float64)
        one = b.ndim == 1
        B = np.asfortranarray(b.reshape(b.shape[0], -1))
        X = np.empty_like(B, order="F")
        n, k = B.shape
        rc = lib.taucs_amd_supernodal_solve_many(self.h, k, _dptr(X), n, _dptr(B), n)
        if rc != 0:
            raise RuntimeError("solve failed")
        return X[:, 0].copy() if one else X

    def solve_refined(self, A, b, method=0, tol=1e-14, maxit=100, factored=False,
                      stop_relres=False):
        """FP64 refinement (method 0) or preconditioned CG (method 1) on A.
        factored=True: A's values are those this factor was last factored from
        (the factor's device copy is used instead of uploading A's values).
        stop_relres=True: stop on |r|_2 <= tol |b|_2 instead of the backward error."""
        b = np.asarray(b, dtype=np.float64)
        one = b.ndim == 1
        B = np.asfortranarray(b.reshape(b.shape[0], -1))
        X = np.empty_like(B, order="F")
        n, k = B.shape
        info = RefineInfo()
        m = _refine_method(method, factored, stop_relres)
        rc = lib.taucs_amd_solve_refined(self.h, A.ptr, k, _dptr(X), n, _dptr(B), n, m, tol,
                                         maxit, C.byref(info))
        if rc != 0:
            raise RuntimeError("refined solve failed")
        return (X[:, 0].copy() if one else X), info

    def solve_refined_device(self, A, dX, dB, n, nrhs, method=0, tol=1e-14, maxit=100,
                             factored=False, stop_relres=False):
        """solve_refined with right-hand sides and solutions in HBM."""
        info = RefineInfo()
        m = _refine_method(method, factored, stop_relres)
        rc = lib.taucs_amd_solve_refined_device(self.h, A.ptr, nrhs, dX, n, dB, n, m, tol, maxit,
                                                C.byref(info))
        if rc != 0:
            raise RuntimeError("refined solve failed")
        return info

    def digest(self):
        """Digest of the numeric factor's stored bits (equal iff bitwise equal)."""
        d = C.c_uint64()
        if lib.taucs_amd_factor_digest(self.h, C.byref(d)) != 0:
            raise RuntimeError("factor digest failed")
        return d.value

    def device_bytes(self, precision=FP64, nrhs=1, refine=False):
        """Device bytes this factor holds once factored and solved (host estimate)."""
        return lib.taucs_amd_factor_device_bytes(self.h, precision, nrhs, int(bool(refine)))

    def solve_device(self, dX, dB, n, nrhs):
        """Solve with right-hand sides and solutions already in HBM."""
        rc = lib.taucs_amd_supernodal_solve_many_device(self.h, nrhs, dX, n, dB, n)
        if rc != 0:
            raise RuntimeError("device solve failed")

    def to_ccs(self):
        return Matrix(lib.taucs_supernodal_factor_to_ccs(self.h))

    def diag(self):
        p = lib.taucs_supernodal_factor_get_diag(self.h)
        if not p:
            raise RuntimeError("get_diag failed")
        n = self.info().n
        d = np.ctypeslib.as_array(C.cast(p, P(C.c_double)), shape=(n,)).copy()
        lib.taucs_free_stub(p)
        return d

    def timing(self):
        t = NumericTiming()
        lib.taucs_amd_get_numeric_timing(self.h, C.byref(t))
        return t

    def numeric_device(self, d_vals, precision=FP64):
        """Numeric factorization from values already in device memory."""
        return lib.taucs_amd_factor_llt_numeric_device(self.h, d_vals, precision)

    def autotune(self, A, precision=FP64, candidates=((64, 64), (128, 64), (256, 64), (256, 128),
                                                       (512, 128)), reps=2):
        """Pick the schedule shape (panel width, GEMM tile edge) by timing the
        numeric phase; L ends factored.  Returns (best pair, {pair: ms})."""
        cand = [(c, 0) if np.isscalar(c) else tuple(c) for c in candidates]
        w = np.asarray([c[0] for c in cand], dtype=np.int32)
        tl = np.asarray([c[1] for c in cand], dtype=np.int32)
        ms = np.zeros(len(cand))
        best = C.c_int()
        if lib.taucs_amd_autotune(self.h, A.ptr, precision, _dptr(w), _dptr(tl), len(cand), reps,
                                  _dptr(ms), C.byref(best)):
            raise RuntimeError("autotune failed")
        return cand[best.value], dict(zip(cand, ms.tolist()))

    def op_profile(self):
        """Per-op kind, level, tasks, flops, bytes and ms of the last profiled numeric run, and
        the tile edge and K <= 64 flag of the GEMM ops."""
        n = lib.taucs_amd_get_op_profile(self.h, 0, None, None, None, None, None, None)
        if n < 0:
            raise RuntimeError("no device schedule")
        k, lv, c = (np.zeros(n, dtype=np.int32) for _ in range(3))
        f, b, ms = (np.zeros(n) for _ in range(3))
        lib.taucs_amd_get_op_profile(self.h, n, _dptr(k), _dptr(lv), _dptr(c), _dptr(f), _dptr(b),
                                     _dptr(ms))
        tile, shortk = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
        lib.taucs_amd_get_op_tiles(self.h, n, _dptr(tile), _dptr(shortk))
        return {"kind": k, "level": lv, "count": c, "flops": f, "bytes": b, "ms": ms, "tile": tile,
                "shortk": shortk}

    def free(self):
        if self.h:
            lib.taucs_supernodal_factor_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def factor_llt_mf(A, precision=FP64, max_depth=0):
    if precision == FP64 and max_depth:
        h = lib.taucs_ccs_factor_llt_mf_maxdepth(A.ptr, max_depth)
    elif precision == FP64:
        h = lib.taucs_ccs_factor_llt_mf(A.ptr)
    else:
        h = lib.taucs_amd_factor_llt_mf_precision(A.ptr, precision)
    return Factor(h) if h else None


def factor_symbolic(A):
    h = lib.taucs_ccs_factor_llt_symbolic(A.ptr)
    return Factor(h) if h else None


def linsolve(A, B, options, F=None, opt_arg=None):
    """taucs_linsolve(A, &F, nrhs, X, B, options, opt_arg).  Returns (rc, X, F)."""
    opts = (C.c_char_p * (len(options) + 1))(*[o.encode() for o in options], None)
    args = (C.c_void_p * (len(opt_arg or []) + 1))(*(opt_arg or []), None)
    Fh = C.c_void_p(F) if F is not None else C.c_void_p()
    if A is None:
        rc = lib.taucs_linsolve(None, C.byref(Fh), 0, None, None, opts, args)
        return rc, None, None
    if B is None:
        rc = lib.taucs_linsolve(A.ptr, C.byref(Fh), 0, None, None, opts, args)
        return rc, None, Fh.value
    B = np.asfortranarray(np.asarray(B, dtype=np.float64).reshape(A.n, -1))
    X = np.zeros_like(B, order="F")
    rc = lib.taucs_linsolve(A.ptr, C.byref(Fh), B.shape[1], _dptr(X), _dptr(B), opts, args)
    return rc, X, Fh.value


def gpu_available():
    try:
        import subprocess
        r = subprocess.run(["rocminfo"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                           text=True, timeout=60)
        return "gfx950" in r.stdout
    except Exception:
        return False


class DeviceArray:
    """A column-major float64 array in HBM (n x k)."""

    def __init__(self, host):
        host = np.asfortranarray(np.asarray(host, dtype=np.float64))
        self.shape = host.shape
        self.nbytes = host.nbytes
        self.ptr = lib.taucs_amd_device_malloc(max(self.nbytes, 8))
        if not self.ptr:
            raise MemoryError("device allocation failed")
        lib.taucs_amd_memcpy_h2d(self.ptr, host.ctypes.data, self.nbytes)

    def get(self):
        out = np.empty(self.shape, dtype=np.float64, order="F")
        lib.taucs_amd_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes)
        return out

    def free(self):
        if self.ptr:
            lib.taucs_amd_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """Many independent systems (permuted matrices) factored in shared launches."""

    def __init__(self, mats, precision=FP64):
        self.mats = list(mats)
        arr = (C.c_void_p * len(self.mats))(*[C.cast(m.ptr, C.c_void_p) for m in self.mats])
        self._arr = arr
        h = lib.taucs_amd_batch_factor_llt_mf(len(self.mats), arr, precision)
        if not h:
            raise RuntimeError("batched factorization failed")
        self.h = C.c_void_p(h)
        cnt, nt, lv = C.c_int(), C.c_int(), C.c_int()
        ff, rf, tn = C.c_double(), C.c_double(), C.c_double()
        self.xoff = np.zeros(len(self.mats) + 1, dtype=np.int64)
        lib.taucs_amd_batch_info(self.h, C.byref(cnt), C.byref(nt), _dptr(self.xoff), C.byref(ff),
                                 C.byref(rf), C.byref(lv), C.byref(tn))
        self.n_total = nt.value
        self.fund_flops = ff.value
        self.relax_flops = rf.value
        self.n_levels = lv.value

    def t_numeric(self):
        tn = C.c_double()
        lib.taucs_amd_batch_info(self.h, None, None, None, None, None, None, C.byref(tn))
        return tn.value

    def numeric(self, mats=None, precision=FP64):
        mats = list(mats) if mats is not None else self.mats
        arr = (C.c_void_p * len(mats))(*[C.cast(m.ptr, C.c_void_p) for m in mats])
        return lib.taucs_amd_batch_numeric(self.h, arr, precision)

    def solve(self, b):
        b = np.asarray(b, dtype=np.float64)
        one = b.ndim == 1
        B = np.asfortranarray(b.reshape(b.shape[0], -1))
        X = np.empty_like(B, order="F")
        if lib.taucs_amd_batch_solve(self.h, B.shape[1], _dptr(X), B.shape[0], _dptr(B), B.shape[0]):
            raise RuntimeError("batched solve failed")
        return X[:, 0].copy() if one else X

    def solve_device(self, dX, dB, nrhs):
        if lib.taucs_amd_batch_solve_device(self.h, nrhs, dX, self.n_total, dB, self.n_total):
            raise RuntimeError("batched device solve failed")

    def solve_refined(self, b, method=0, tol=1e-14, maxit=100):
        b = np.asarray(b, dtype=np.float64)
        one = b.ndim == 1
        B = np.asfortranarray(b.reshape(b.shape[0], -1))
        X = np.empty_like(B, order="F")
        info = RefineInfo()
        if lib.taucs_amd_batch_solve_refined(self.h, B.shape[1], _dptr(X), B.shape[0], _dptr(B),
                                             B.shape[0], method, tol, maxit, C.byref(info)):
            raise RuntimeError("batched refined solve failed")
        return (X[:, 0].copy() if one else X), info

    def numeric_device(self, d_vals, precision=FP64):
        return lib.taucs_amd_batch_numeric_device(self.h, d_vals, precision)

    def autotune(self, candidates=((64, 64), (128, 64), (256, 64), (256, 128)), reps=2):
        """Pick the batch schedule shape (panel width, GEMM tile edge) by timing;
        the batch ends factored.  Returns (best pair, {pair: ms})."""
        cand = [(c, 0) if np.isscalar(c) else tuple(c) for c in candidates]
        w = np.asarray([c[0] for c in cand], dtype=np.int32)
        tl = np.asarray([c[1] for c in cand], dtype=np.int32)
        ms = np.zeros(len(cand))
        best = C.c_int()
        if lib.taucs_amd_batch_autotune(self.h, _dptr(w), _dptr(tl), len(cand), reps, _dptr(ms),
                                        C.byref(best)):
            raise RuntimeError("batch autotune failed")
        return cand[best.value], dict(zip(cand, ms.tolist()))

    def system_supernodes(self, q):
        f = Factor(lib.taucs_amd_batch_system_factor(self.h, q))
        try:
            return f.supernodes()
        finally:
            f.h = None  # owned by the batch

    def timing(self):
        t = NumericTiming()
        lib.taucs_amd_batch_get_timing(self.h, C.byref(t))
        return t

    def free(self):
        if self.h:
            lib.taucs_amd_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
