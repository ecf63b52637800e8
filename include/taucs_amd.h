/*
 * taucs_amd.h -- additive entry points of the MI355X multifrontal Cholesky.
 *
 * None of these replace a reference function; they extend the TAUCS ABI
 * (include/taucs.h) without changing any existing signature:
 *   - seeded generators for the variable-coefficient and shifted workloads,
 *   - factor statistics and a structure export for parity tests,
 *   - multi-right-hand-side and device-pointer solves,
 *   - reduced-precision factors with FP64 iterative refinement,
 *   - many independent systems factored in the same level launches,
 *   - a serving loop with HBM-budget admission and resident factors.
 */
#ifndef TAUCS_AMD_H
#define TAUCS_AMD_H

#include <stdint.h>
#include "taucs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded inputs ------------------------------------------------------ */
uint64_t taucs_amd_splitmix64(uint64_t* state);
double   taucs_amd_uniform01(uint64_t* state);
void     taucs_amd_random_uniform(int n, uint64_t seed, double lo, double hi, double* out);
/* 7-point operator on the mesh3d pattern and entry order; seed 0 gives
   exactly taucs_ccs_generate_mesh3d. */
taucs_ccs_matrix* taucs_amd_generate_mesh3d_varcoef(int X, int Y, int Z, uint64_t seed,
                                                    double log10_lo, double log10_hi,
                                                    double shift);
/* mesh2d(n,"neumann") + i.i.d. uniform[0,1) on each diagonal entry */
taucs_ccs_matrix* taucs_amd_generate_mesh2d_shifted(int n, uint64_t seed);

/* ---- factor inspection -------------------------------------------------- */
typedef struct {
  int     n;
  int     n_sn;          /* supernode slots, dead merged slots included */
  int     n_live;        /* supernodes reachable from the root */
  int     n_levels;      /* height of the supernodal tree + 1 */
  int     max_front;     /* max (sn_size + update size) */
  int     etree_depth;
  int     precision;     /* 0 fp64, 1 fp32, 2 bf16-update */
  int     has_numeric;
  double  fund_nnz, fund_flops;
  double  relax_nnz, relax_flops;
  double  t_symbolic, t_numeric;   /* seconds, host wall clock */
  int64_t L_bytes, U_bytes;        /* device storage */
} taucs_amd_factor_info;

int taucs_amd_factor_get_info(void* L, taucs_amd_factor_info* info);

/* Copy the supernodal structure out.  Arrays sized n_sn+1 (sizes, tree) and
   struct_ptr n_sn+1, struct_idx sum of sn_up_size. Any pointer may be NULL. */
int taucs_amd_factor_get_supernodes(void* L, int* sn_size, int* sn_up_size,
                                    int* first_child, int* next_child,
                                    int* struct_ptr, int* struct_idx);
/* Level schedule of the factor, per supernode (-1 for supernodes merged away):
   its level (children before parents), its parent (n_sn = the virtual root),
   and the element offsets of its panel in the L arena and of its update block
   in the U arena (update blocks of disjoint lifetimes share space). */
int taucs_amd_factor_get_schedule(void* L, int* sn_level, int* sn_parent, long long* off_L,
                                  long long* off_U);

/* Digest of the numeric factor's stored bits (a position-mixed sum modulo
   2^64 over the device L arena): two factors of one symbolic structure are
   bitwise equal iff their digests are (up to 2^-64). */
int taucs_amd_factor_digest(void* L, unsigned long long* digest);

/* ---- solves ------------------------------------------------------------- */
/* X, B column-major host arrays (n x nrhs); permuted space like
   taucs_supernodal_solve_llt. */
int taucs_amd_supernodal_solve_many(void* L, int nrhs, double* X, int ldx,
                                    const double* B, int ldb);
/* Same, device pointers; runs on the factor's stream and returns when done. */
int taucs_amd_supernodal_solve_many_device(void* L, int nrhs, double* dX, int ldx,
                                           const double* dB, int ldb);

/* Allocation-free variant of taucs_amd_supernodal_solve_many_device.  No
   counterpart in the reference: it is for callers that cannot allocate device
   memory (code under graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  The sweeps run eagerly on the factor's stream with their
   workspace in one fixed device array per GPU of
   taucs_amd_noalloc_scratch_bytes() bytes (64 MiB), held until the process
   ends; right-hand sides go through it as many at a time as fit (at most 16).
   Returns -1 with a message when one right-hand side's workspace exceeds the
   array.  All calls on one GPU share the array: calls that overlap on one GPU
   race and must be serialized by the caller. */
int       taucs_amd_supernodal_solve_many_device_noalloc(void* L, int nrhs, double* dX, int ldx,
                                                         const double* dB, int ldb);
long long taucs_amd_noalloc_scratch_bytes(void);

/* ---- precision and refinement ------------------------------------------- */
#define TAUCS_AMD_FP64    0
#define TAUCS_AMD_FP32    1
#define TAUCS_AMD_BF16UPD 2
void* taucs_amd_factor_llt_mf_precision(taucs_ccs_matrix* A, int precision);
int   taucs_amd_factor_llt_numeric_precision(taucs_ccs_matrix* A, void* L, int precision);
/* Numeric factorization from A's values already in device memory (nnz
   doubles in the stored order of the permuted matrix of the symbolic factor). */
int   taucs_amd_factor_llt_numeric_device(void* L, const double* d_vals, int precision);
/* FP64 fallback of reduced-precision factors (default on): a 64 x 64 diagonal
   block (or a whole front of at most 64 rows) whose FP32 Cholesky meets a
   nonpositive pivot, or a pivot below 2^-12 of the block's input diagonal
   entry, is refactored in FP64 from its input and rounded back to FP32; the
   factorization fails only if the FP64 Cholesky fails too.  Applies to
   schedules launched afterwards.  The count is that of the factor's last
   numeric factorization (batch handles included), -1 without one. */
void  taucs_amd_set_fp64_fallback(int on);
int   taucs_amd_factor_fp64_fallbacks(void* L);

typedef struct {
  int    iterations;     /* outer iterations used (max over right-hand sides) */
  int    converged;      /* all right-hand sides reached the tolerance */
  double backward_error; /* max over rhs of |b-Ax|_inf / (|A|_inf |x|_inf + |b|_inf) */
  double rel_residual;   /* max over rhs of |b-Ax|_2 / |b|_2 */
} taucs_amd_refine_info;

/* Solve A X = B (A in the permuted space of the factor; its pattern must be
   the factored one, its values may differ) with FP64 residuals on the device
   and the factor as preconditioner.
   method: 0 = classical refinement, 1 = preconditioned conjugate gradients,
   ORed with the flags below.  By default each right-hand side stops once its
   normwise backward error |b-Ax|_inf / (|A|_inf |x|_inf + |b|_inf) <= tol.
   A's values are uploaded on every call (A may have changed in place since
   the factorization) unless FACTORED_VALUES is given. */
#define TAUCS_AMD_REFINE_FACTORED_VALUES 0x100 /* A's values are exactly those of the
                                                  factor's last numeric factorization:
                                                  use the factor's device copy */
#define TAUCS_AMD_REFINE_STOP_RELRES     0x200 /* stop on |b-Ax|_2 <= tol |b|_2, the rule
                                                  of the reference's conjugate gradients
                                                  (src/taucs_iter.c:139,186) */
int taucs_amd_solve_refined(void* L, taucs_ccs_matrix* A, int nrhs, double* X, int ldx,
                            const double* B, int ldb, int method, double tol, int maxit,
                            taucs_amd_refine_info* info);
/* Same, with X and B in device memory. */
int taucs_amd_solve_refined_device(void* L, taucs_ccs_matrix* A, int nrhs, double* dX, int ldx,
                                   const double* dB, int ldb, int method, double tol, int maxit,
                                   taucs_amd_refine_info* info);

/* ---- many independent systems in the same launches ---------------------- */
/* A[q] are the permuted matrices P A P^T of `count` independent systems.  Each
   system keeps the reference's own symbolic analysis; the numeric phase runs
   the systems' fronts level by level in shared launches (hipGraph-replayed).
   Vectors are concatenated: system q owns rows xoff[q] .. xoff[q+1]-1. */
void* taucs_amd_batch_factor_llt_mf(int count, taucs_ccs_matrix** A, int precision);
int   taucs_amd_batch_numeric(void* B, taucs_ccs_matrix** A, int precision);
int   taucs_amd_batch_numeric_device(void* B, const double* d_vals, int precision);
int   taucs_amd_batch_solve(void* B, int nrhs, double* X, int ldx, const double* Bv, int ldb);
int   taucs_amd_batch_solve_device(void* B, int nrhs, double* dX, int ldx, const double* dB,
                                   int ldb);
int   taucs_amd_batch_solve_refined(void* B, int nrhs, double* X, int ldx, const double* Bv,
                                    int ldb, int method, double tol, int maxit,
                                    taucs_amd_refine_info* info);
int   taucs_amd_batch_info(void* B, int* count, int* n_total, long long* xoff,
                           double* fund_flops, double* relax_flops, int* n_levels,
                           double* t_numeric);
void* taucs_amd_batch_system_factor(void* B, int q);
void  taucs_amd_batch_free(void* B);

/* ---- schedule shape and its autotuner ------------------------------------ */
/* Large fronts are factored in outer panels of w columns (64-column POTRF /
   TRSM steps inside, one trailing update of depth w per panel), and the
   trailing updates and SYRKs use GEMM tiles of edge 64 or 128.  Defaults
   256 and 64; they apply to schedules built afterwards. */
void  taucs_amd_set_panel_width(int w);
void  taucs_amd_set_gemm_tile(int edge);
/* Time the numeric factorization of A with each candidate schedule (panel
   width cand_w[i], multiple of 64; GEMM tile edge cand_tile[i], or NULL for
   the default) and keep the fastest schedule in L, factored.  ms[i]: best
   device time of candidate i over `reps` runs; *best: its index. */
int   taucs_amd_autotune(void* L, taucs_ccs_matrix* A, int precision, const int* cand_w,
                         const int* cand_tile, int ncand, int reps, double* ms, int* best);
int   taucs_amd_batch_autotune(void* B, const int* cand_w, const int* cand_tile, int ncand,
                               int reps, double* ms, int* best);

/* Replay the numeric schedule and the solves as hipGraphs built node by node (default on) */
void  taucs_amd_set_graphs(int on);
/* Panel lookahead: the trailing update of a panel step runs on a second stream
   while the next panel is factored (default on; results are identical). */
void  taucs_amd_set_lookahead(int on);
/* Diagonal-block POTRF kernel for the block that starts each outer panel:
   2 = four waves, 16-column panels with MFMA updates (default); 1 = one
   wave, register broadcasts; 0 = 256 threads through LDS.  2 and 1 give
   bitwise equal factors.  The other diagonal blocks are factored by the
   four-wave code at the end of the previous panel step's launch.  Applies
   to numeric factorizations started afterwards. */
void  taucs_amd_set_potrf_kernel(int kind);
/* Fronts with at most 16 (32) rows are factored four (two) to a wavefront,
   and fronts with at most 48 rows with 48-wide staging (default on); the
   results are bitwise those of the 64-row one-front-per-wavefront kernel.
   Applies to schedules built afterwards. */
void  taucs_amd_set_front_packing(int on);
/* Triangular sweeps of the solve: 1 = one launch per tree level, the block
   chain of each front followed inside the launch through per-block flags
   (default); 0 = one launch per 64-column block step.  Applies to schedules
   built afterwards. */
void  taucs_amd_set_solve_kernel(int chained);

/* ---- device buffers (for right-hand sides kept in HBM) ------------------- */
void* taucs_amd_device_malloc(size_t bytes);
void  taucs_amd_device_free(void* p);
int   taucs_amd_memcpy_h2d(void* dst, const void* src, size_t bytes);
int   taucs_amd_memcpy_d2h(void* dst, const void* src, size_t bytes);
int   taucs_amd_memcpy_d2d(void* dst, const void* src, size_t bytes);
int   taucs_amd_device_synchronize(void);

/* ---- device memory accounting (admission by HBM budget) ----------------- */
/* Bytes of device memory the factor L (symbolic or numeric) holds once it is
   factored in `precision` and has solved up to `nrhs` right-hand sides from
   host buffers; `refine` adds the FP64 refinement context.  Computed on the
   host from the schedule, without a device. */
long long taucs_amd_factor_device_bytes(void* L, int precision, int nrhs, int refine);
/* Live bytes of all device allocations this library holds right now. */
long long taucs_amd_device_bytes_in_use(void);
/* The library is thread-safe across distinct factors: each factor runs on its
   own HIP stream, so calls on different factors from different threads
   overlap on the device.  Calls on one factor must not overlap. */

/* ---- timing of the last numeric factorization (device events) ----------- */
typedef struct {
  double total_ms;       /* numeric factorization, stream time */
  double front_ms;       /* dense front kernels (POTRF/TRSM/GEMM/SYRK) */
  double assembly_ms;    /* A scatter + extend-add */
  double gemm_ms;        /* tile GEMM/SYRK kernels only */
  int    gemm_launches;
  double gemm_flops;     /* algorithmic tile-GEMM/SYRK flops (lower part of diagonal tiles) */
  double assembly_bytes; /* algorithmic bytes of A scatter + extend-add */
  int    launches;
} taucs_amd_numeric_timing;
void taucs_amd_set_profiling(int on);
int  taucs_amd_get_numeric_timing(void* L, taucs_amd_numeric_timing* t);
int  taucs_amd_batch_get_timing(void* B, taucs_amd_numeric_timing* t);
/* Per-op record of the numeric schedule (ops in launch order): kind
   (1 assemble, 2 small fronts, 3 POTRF, 4 TRSM, 5 GEMM/SYRK, 6 extend-add,
   7 zero), tree level, task count, algorithmic flops and bytes, and the
   device time (ms) of the last factorization run with profiling on.
   Returns the number of ops (arrays hold up to cap entries). */
int  taucs_amd_get_op_profile(void* L, int cap, int* kind, int* level, int* count,
                              double* flops, double* bytes, double* ms);
/* The tile shape of the same ops: the tile edge of a GEMM op (64 or 128; 0
   for other kinds) and whether its tasks all have K <= 64 (the panel-internal
   updates, k_gemm_short).  Returns the op count, or -1 without a device
   schedule.  No counterpart in the reference (inspection). */
int  taucs_amd_get_op_tiles(void* L, int cap, int* tile, int* shortk);

#ifdef __cplusplus
}
#endif
#endif
