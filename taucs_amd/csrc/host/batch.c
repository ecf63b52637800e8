/*
 * Many independent SPD systems factored and solved together.
 *
 * Each system gets the reference's own symbolic analysis
 * (src/taucs_sn_llt.c:3663-4001) exactly as if it were factored alone, so
 * every per-system structure is bit-identical to taucs_ccs_factor_llt_mf's;
 * concatenating the systems into one block-diagonal matrix instead would let
 * the reference's last-tree-only amalgamation (:3927) change all but one of
 * them.  The device schedule then merges the systems level by level: all
 * fronts of level l of every system run in the same launches, and the whole
 * numeric phase and every solve replay as captured hipGraphs.
 *
 * Vectors of the batch are concatenated: system q owns rows
 * [xoff[q], xoff[q] + n_q).  Every matrix passed in is the permuted P A P^T
 * of its system (the caller orders, as with taucs_ccs_factor_llt_mf).
 */
#include <stdio.h>
#include <string.h>
#include "internal.h"
#include "../../../include/taucs_amd.h"

int tamd_dev_batch_setup(tamd_factor* FB, int count, tamd_factor* const* F, const long long* xoff);
int tamd_dev_numeric_values(tamd_factor* F, const double* vals);
int tamd_dev_numeric_device(tamd_factor* F, const double* d_vals);
void* taucs_dccs_factor_llt_symbolic(taucs_ccs_matrix* A);

typedef struct {
  int           count;
  int           n_total;
  long long     nnz_total;
  tamd_factor** sys;      /* per-system symbolic factors (no device state) */
  long long*    xoff;     /* [count+1] */
  tamd_factor   all;      /* the merged device schedule; n = n_total */
  double*       vals;     /* concatenated A values (the block-diagonal matrix's) */
  taucs_ccs_matrix* blockdiag; /* concatenated pattern for refinement */
} tamd_batch;

static void batch_free(tamd_batch* B)
{
  int q;
  if (!B) return;
  tamd_dev_free(&B->all);
  for (q = 0; q < B->count; q++)
    if (B->sys && B->sys[q]) taucs_supernodal_factor_free(B->sys[q]);
  tamd_free(B->sys);
  tamd_free(B->xoff);
  taucs_ccs_free(B->blockdiag);   /* owns vals */
  tamd_free(B);
}

void taucs_amd_batch_free(void* vB) { batch_free((tamd_batch*)vB); }

static int gather_values(tamd_batch* B, taucs_ccs_matrix** A)
{
  int q;
  long long off = 0;
  for (q = 0; q < B->count; q++) {
    int nnz = A[q]->colptr[A[q]->n];
    if (A[q]->n != B->sys[q]->n || nnz != B->sys[q]->plan.a_nnz
        || tamd_pattern_hash(A[q]) != B->sys[q]->plan.a_pattern_hash) {
      taucs_printf("taucs_amd_batch: system %d does not match its symbolic factor\n", q);
      return -1;
    }
    memcpy(B->vals + off, A[q]->values.d, sizeof(double) * (size_t)nnz);
    off += nnz;
  }
  return 0;
}

void* taucs_amd_batch_factor_llt_mf(int count, taucs_ccs_matrix** A, int precision)
{
  tamd_batch* B;
  int q;
  long long nnz = 0, n = 0;
  if (count <= 0 || !A) return NULL;
  B = (tamd_batch*)tamd_calloc(1, sizeof(tamd_batch));
  if (!B) return NULL;
  B->count = count;
  B->sys = (tamd_factor**)tamd_calloc((size_t)count, sizeof(tamd_factor*));
  B->xoff = (long long*)tamd_malloc(sizeof(long long) * (size_t)(count + 1));
  if (!B->sys || !B->xoff) { batch_free(B); return NULL; }
  for (q = 0; q < count; q++) {
    B->sys[q] = (tamd_factor*)taucs_dccs_factor_llt_symbolic(A[q]);
    if (!B->sys[q]) { batch_free(B); return NULL; }
    B->xoff[q] = n;
    n += A[q]->n;
    nnz += A[q]->colptr[A[q]->n];
  }
  B->xoff[count] = n;
  if (n > 0x7fffffffLL || nnz > 0x7fffffffLL) { batch_free(B); return NULL; }
  B->n_total = (int)n;
  B->nnz_total = nnz;

  /* block-diagonal matrix; its value array is the concatenated values the
     device factors from, so refinement against it uses the device copy */
  B->blockdiag = taucs_ccs_create((int)n, (int)n, (int)nnz, A[0]->flags);
  if (!B->blockdiag) { batch_free(B); return NULL; }
  B->vals = B->blockdiag->values.d;
  {
    long long off = 0;
    int j;
    for (q = 0; q < count; q++) {
      int nq = A[q]->n;
      for (j = 0; j < nq; j++) B->blockdiag->colptr[B->xoff[q] + j] = (int)(off + A[q]->colptr[j]);
      for (j = 0; j < A[q]->colptr[nq]; j++)
        B->blockdiag->rowind[off + j] = (int)(B->xoff[q] + A[q]->rowind[j]);
      off += A[q]->colptr[nq];
    }
    B->blockdiag->colptr[n] = (int)off;
  }

  B->all.flags = TAUCS_DOUBLE;
  B->all.uplo = 'l';
  B->all.n = (int)n;
  B->all.precision = precision;
  if (tamd_dev_batch_setup(&B->all, count, B->sys, B->xoff) != 0) { batch_free(B); return NULL; }
  if (taucs_amd_batch_numeric(B, A, precision) != 0) { batch_free(B); return NULL; }
  return B;
}

int taucs_amd_batch_numeric(void* vB, taucs_ccs_matrix** A, int precision)
{
  tamd_batch* B = (tamd_batch*)vB;
  double t0 = taucs_wtime();
  if (!B || !A) return -1;
  if (gather_values(B, A) != 0) return -1;
  B->all.precision = precision;
  if (tamd_dev_numeric_values(&B->all, B->vals) != 0) {
    B->all.has_numeric = 0;
    return -1;
  }
  B->all.has_numeric = 1;
  B->all.t_numeric = taucs_wtime() - t0;
  return 0;
}

int taucs_amd_batch_solve(void* vB, int nrhs, double* X, int ldx, const double* Bv, int ldb)
{
  tamd_batch* B = (tamd_batch*)vB;
  if (!B || !B->all.has_numeric) return -1;
  return tamd_dev_solve(&B->all, nrhs, X, ldx, Bv, ldb);
}

int taucs_amd_batch_solve_device(void* vB, int nrhs, double* dX, int ldx, const double* dB, int ldb)
{
  tamd_batch* B = (tamd_batch*)vB;
  if (!B || !B->all.has_numeric) return -1;
  return taucs_amd_supernodal_solve_many_device(&B->all, nrhs, dX, ldx, dB, ldb);
}

int taucs_amd_batch_solve_refined(void* vB, int nrhs, double* X, int ldx, const double* Bv, int ldb,
                                  int method, double tol, int maxit, taucs_amd_refine_info* info)
{
  tamd_batch* B = (tamd_batch*)vB;
  if (!B || !B->all.has_numeric) return -1;
  /* the residuals are those of the systems as last factored (from host or
     device values), read from the factor's own device copy */
  return taucs_amd_solve_refined(&B->all, B->blockdiag, nrhs, X, ldx, Bv, ldb,
                                 method | TAUCS_AMD_REFINE_FACTORED_VALUES, tol, maxit, info);
}

int taucs_amd_batch_info(void* vB, int* count, int* n_total, long long* xoff, double* fund_flops,
                         double* relax_flops, int* n_levels, double* t_numeric)
{
  tamd_batch* B = (tamd_batch*)vB;
  int q;
  double ff = 0, rf = 0;
  int lv = 0;
  if (!B) return -1;
  for (q = 0; q < B->count; q++) {
    ff += B->sys[q]->sym.fund_flops;
    rf += B->sys[q]->sym.relax_flops;
    if (B->sys[q]->plan.n_levels > lv) lv = B->sys[q]->plan.n_levels;
  }
  if (count) *count = B->count;
  if (n_total) *n_total = B->n_total;
  if (xoff) memcpy(xoff, B->xoff, sizeof(long long) * (size_t)(B->count + 1));
  if (fund_flops) *fund_flops = ff;
  if (relax_flops) *relax_flops = rf;
  if (n_levels) *n_levels = lv;
  if (t_numeric) *t_numeric = B->all.t_numeric;
  return 0;
}

void* taucs_amd_batch_system_factor(void* vB, int q)
{
  tamd_batch* B = (tamd_batch*)vB;
  if (!B || q < 0 || q >= B->count) return NULL;
  return B->sys[q];
}

int taucs_amd_batch_get_timing(void* vB, taucs_amd_numeric_timing* t)
{
  tamd_batch* B = (tamd_batch*)vB;
  if (!B) return -1;
  return taucs_amd_get_numeric_timing(&B->all, t);
}

/* The schedule autotuner (taucs_amd_autotune) for a batch: all systems share
   one schedule, so one schedule shape is chosen for the whole batch. */
int taucs_amd_batch_autotune(void* vB, const int* cand_w, const int* cand_tile, int ncand, int reps,
                             double* ms, int* best)
{
  tamd_batch* B = (tamd_batch*)vB;
  int c, bi = -1;
  double best_ms = 1e300;
  if (!B || !cand_w || ncand <= 0 || !B->all.has_numeric) return -1;
  for (c = 0; c < ncand; c++) {
    double t;
    if (cand_w[c] < 64 || cand_w[c] % 64) return -1;
    tamd_dev_free(&B->all);
    B->all.has_numeric = 0;
    B->all.panel_w = cand_w[c];
    B->all.gemm_tile = cand_tile ? cand_tile[c] : 0;
    if (tamd_dev_batch_setup(&B->all, B->count, B->sys, B->xoff) != 0
        || tamd_dev_time_numeric(&B->all, B->vals, reps, &t) != 0)
      return -1;
    if (ms) ms[c] = t;
    if (t < best_ms) { best_ms = t; bi = c; }
  }
  tamd_dev_free(&B->all);
  B->all.panel_w = cand_w[bi];
  B->all.gemm_tile = cand_tile ? cand_tile[bi] : 0;
  if (tamd_dev_batch_setup(&B->all, B->count, B->sys, B->xoff) != 0
      || tamd_dev_numeric_values(&B->all, B->vals) != 0)
    return -1;
  B->all.has_numeric = 1;
  if (best) *best = bi;
  return 0;
}

/* Numeric phase of a batch from the concatenated values already in device
   memory (system q's stored entries at the offset of its nnz prefix sum). */
int taucs_amd_batch_numeric_device(void* vB, const double* d_vals, int precision)
{
  tamd_batch* B = (tamd_batch*)vB;
  double t0 = taucs_wtime();
  if (!B || !d_vals) return -1;
  B->all.precision = precision;
  if (tamd_dev_numeric_device(&B->all, d_vals) != 0) {
    B->all.has_numeric = 0;
    return -1;
  }
  B->all.has_numeric = 1;
  B->all.t_numeric = taucs_wtime() - t0;
  return 0;
}
