/*
 * taucs.h -- C-ABI of the MI355X supernodal multifrontal Cholesky path.
 *
 * This header is the drop-in boundary: every declaration below has the
 * prototype of the TAUCS 2.2.1 function it replaces, so code written against
 * TAUCS (src/taucs.h + src/taucs_private.h) compiles and links unchanged
 * against libtaucs_amd.so.  Only double precision (TAUCS_DOUBLE) matrices are
 * accepted on this path; other datatypes return the reference's failure value.
 *
 * Struct layout and constants are binary-identical to the reference:
 *   error codes         src/taucs.h:90-95
 *   datatype/structure  src/taucs.h:97-108
 *   taucs_ccs_matrix    src/taucs.h:537-554
 */
#ifndef TAUCS_AMD_TAUCS_H
#define TAUCS_AMD_TAUCS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TAUCS_SUCCESS           0
#define TAUCS_ERROR            -1
#define TAUCS_ERROR_NOMEM      -2
#define TAUCS_ERROR_BADARGS    -3
#define TAUCS_ERROR_INDEFINITE -4
#define TAUCS_ERROR_MAXDEPTH   -5

#define TAUCS_INT       1024
#define TAUCS_DOUBLE    2048
#define TAUCS_SINGLE    4096
#define TAUCS_DCOMPLEX  8192
#define TAUCS_SCOMPLEX 16384

#define TAUCS_LOWER      1
#define TAUCS_UPPER      2
#define TAUCS_TRIANGULAR 4
#define TAUCS_SYMMETRIC  8
#define TAUCS_HERMITIAN  16
#define TAUCS_PATTERN    32

typedef double taucs_double;
typedef float  taucs_single;

/* src/taucs.h:537-554.  0-based CSC; symmetric matrices store the lower
   triangle.  The values union keeps the reference's member names. */
typedef struct {
  int   n;
  int   m;
  int   flags;
  int*  colptr;
  int*  rowind;
  union {
    void*         v;
    taucs_double* d;
    taucs_single* s;
    void*         z;
    void*         c;
  } values;
} taucs_ccs_matrix;

/* ---- containers, generators, vectors (host) ---------------------------- */

/* src/taucs_ccs_base.c:21-126 */
taucs_ccs_matrix* taucs_ccs_create(int m, int n, int nnz, int flags);
taucs_ccs_matrix* taucs_dccs_create(int m, int n, int nnz);
void              taucs_ccs_free(taucs_ccs_matrix* matrix);

/* src/taucs_ccs_generators.c:89-195 (which = "neumann", "dirichlet",
   "anisotropic_x", "anisotropic_y") and :197-261 */
taucs_ccs_matrix* taucs_ccs_generate_mesh2d(int n, char* which);
taucs_ccs_matrix* taucs_ccs_generate_mesh3d(int X, int Y, int Z);

/* src/taucs_ccs_ops.c:143-257 */
taucs_ccs_matrix* taucs_ccs_permute_symmetrically(taucs_ccs_matrix* A,
                                                  int* perm, int* invperm);
taucs_ccs_matrix* taucs_dccs_permute_symmetrically(taucs_ccs_matrix* A,
                                                   int* perm, int* invperm);
/* src/taucs_ccs_ops.c:267-344 (B = A*X, symmetric lower storage) */
void taucs_ccs_times_vec(taucs_ccs_matrix* m, void* X, void* B);
void taucs_dccs_times_vec(taucs_ccs_matrix* m, taucs_double* X, taucs_double* B);

/* src/taucs_vec_base.c:28-206 */
taucs_double taucs_vec_norm2(int n, int flags, void* x);
void  taucs_vec_axpby(int n, int flags, taucs_double a, void* x,
                      taucs_double b, void* y, void* axpby);
void* taucs_vec_create(int n, int flags);
void  taucs_vec_permute(int n, int flags, void* v, void* pv, int p[]);
void  taucs_vec_ipermute(int n, int flags, void* pv, void* v, int invp[]);

/* ---- ordering ----------------------------------------------------------- */

/* src/taucs_ccs_order.c:843-880.  Supported: "genmmd", "identity".
   Others print a message and return NULL permutations, as the reference
   does for orderings that were not linked. */
void taucs_ccs_order(taucs_ccs_matrix* m, int** perm, int** invperm,
                     char* which);

/* ---- supernodal multifrontal LL^T (the hot path) ------------------------ */

/* src/taucs_sn_llt.c:3243-3520 (parent only; counts optional) */
int taucs_ccs_etree(taucs_ccs_matrix* A, int* parent, int* l_colcount,
                    int* l_rowcount, int* l_nnz);

/* src/taucs_sn_llt.c:2979-3006 / 2894-2922 and the d-variants :1979-2122.
   Returns an opaque factor whose numeric blocks live in GPU memory, or NULL
   on any failure (not SPD, out of memory, depth limit, no GPU). */
void* taucs_ccs_factor_llt_mf(taucs_ccs_matrix* A);
void* taucs_ccs_factor_llt_mf_maxdepth(taucs_ccs_matrix* A, int max_depth);
void* taucs_dccs_factor_llt_mf(taucs_ccs_matrix* A);
void* taucs_dccs_factor_llt_mf_maxdepth(taucs_ccs_matrix* A, int max_depth);

/* src/taucs_sn_llt.c:2128-2209: symbolic analysis once, numeric many */
void* taucs_ccs_factor_llt_symbolic(taucs_ccs_matrix* A);
void* taucs_ccs_factor_llt_symbolic_maxdepth(taucs_ccs_matrix* A, int max_depth);
int   taucs_ccs_factor_llt_numeric(taucs_ccs_matrix* A, void* L);
void* taucs_dccs_factor_llt_symbolic(taucs_ccs_matrix* A);
void* taucs_dccs_factor_llt_symbolic_maxdepth(taucs_ccs_matrix* A, int max_depth);
int   taucs_dccs_factor_llt_numeric(taucs_ccs_matrix* A, void* L);

/* src/taucs_sn_llt.c:2844-2885 and :3091.  x, b: host arrays of length n. */
int   taucs_supernodal_solve_llt(void* L, void* x, void* b);
int   taucs_dsupernodal_solve_llt(void* L, void* x, void* b);

/* src/taucs_sn_llt.c:803-1006: every consumer of the opaque factor */
void              taucs_supernodal_factor_free(void* L);
void              taucs_supernodal_factor_free_numeric(void* L);
taucs_ccs_matrix* taucs_supernodal_factor_to_ccs(void* L);
void*             taucs_supernodal_factor_get_diag(void* L);
void              taucs_dsupernodal_factor_free(void* L);
void              taucs_dsupernodal_factor_free_numeric(void* L);
taucs_ccs_matrix* taucs_dsupernodal_factor_to_ccs(void* L);
taucs_double*     taucs_dsupernodal_factor_get_diag(void* L);

/* src/taucs_ccs_solve_llt.c:19-130, :420-447.  Solves L L^T x = b with a
   lower-triangular CCS factor whose columns start with their diagonal entry
   (the export of taucs_supernodal_factor_to_ccs).  Both sweeps run on the
   GPU; x, b are host arrays of length n.  Returns 0, or -1 with a message
   for a factor that is not lower triangular, not double, or malformed. */
int taucs_ccs_solve_llt(void* L, void* x, void* b);

/* ---- on-disk formats (host) --------------------------------------------- */

/* src/taucs_ccs_io.c:34-121 (binary: nrows, ncols, flags, colptr, rowind,
   values); the writer is this library's own, in the same format */
taucs_ccs_matrix* taucs_ccs_read_binary(char* filename);
int               taucs_ccs_write_binary(taucs_ccs_matrix* A, char* filename);
/* src/taucs_ccs_io.c:280-575 ("i j v" text, 1-based) */
int               taucs_ccs_write_ijv(taucs_ccs_matrix* A, char* filename);
taucs_ccs_matrix* taucs_ccs_read_ijv(char* filename, int flags);
/* src/taucs_ccs_io.c:580-820 (MatrixMarket coordinate entries) */
taucs_ccs_matrix* taucs_ccs_read_mtx(char* filename, int flags);
/* src/taucs_ccs_io.c:826-960 (N, colptr, 0-based rowind, values) */
taucs_ccs_matrix* taucs_ccs_read_ccs(char* filename, int flags);
/* src/taucs_ccs_io.c:128-275; the Harwell-Boeing reader is not linked, so
   this returns NULL, as the reference does without it */
taucs_ccs_matrix* taucs_ccs_read_hb(char* filename, int flags);
/* src/taucs_ccs_io.c:967-1047 (n raw values) */
void* taucs_vec_read_binary(int n, int flags, char* filename);
int   taucs_vec_write_binary(int n, int flags, void* v, char* filename);

/* ---- unified solver ----------------------------------------------------- */

/* src/taucs_linsolve.c:41-151 */
int taucs_getopt_boolean(char* cmd, void* args[], char* name, int* x);
int taucs_getopt_double (char* cmd, void* args[], char* name, double* x);
int taucs_getopt_pointer(char* cmd, void* args[], char* name, void** x);
int taucs_getopt_string (char* cmd, void* args[], char* name, char** x);

/* src/taucs_linsolve.c:189-622 (LL^T, in-core, mf branch) */
int taucs_linsolve(taucs_ccs_matrix* A, void** F, int nrhs, void* X, void* B,
                   char* options[], void* opt_arg[]);

/* ---- utilities ---------------------------------------------------------- */

/* src/taucs_logging.c:32-79 */
void   taucs_logfile(char* file_prefix);
int    taucs_printf(char* fmt, ...);
/* src/taucs_timer.c:250-297 */
double taucs_wtime(void);
double taucs_ctime(void);
/* src/taucs_malloc.c (allocation stubs and failure injection) */
void*  taucs_malloc_stub(size_t size);
void*  taucs_calloc_stub(size_t nmemb, size_t size);
void*  taucs_realloc_stub(void* ptr, size_t size);
void   taucs_free_stub(void* ptr);
void   taucs_allocation_induce_failure(int i);
int    taucs_allocation_attempts(void);

#ifdef __cplusplus
}
#endif
#endif
