/*
 * The TAUCS supernodal-factor entry points over the device-resident factor.
 *
 *   taucs_ccs_factor_llt_mf[_maxdepth]      src/taucs_sn_llt.c:2894-3006, 2051-2122
 *   taucs_ccs_factor_llt_symbolic/_numeric  src/taucs_sn_llt.c:2128-2209
 *   taucs_supernodal_solve_llt              src/taucs_sn_llt.c:2844-2885, 3091
 *   taucs_supernodal_factor_free[_numeric]  src/taucs_sn_llt.c:803-849
 *   taucs_supernodal_factor_to_ccs          src/taucs_sn_llt.c:851-978
 *   taucs_supernodal_factor_get_diag        src/taucs_sn_llt.c:982-1006
 *
 * Error conventions follow the reference: factor entry points return NULL on
 * any failure (not SPD, allocation, depth limit) and print through
 * taucs_printf; numeric returns 0 / -1 and frees the numeric data on failure;
 * solve returns 0 / -1.  The symbolic phase runs on the host; everything
 * numeric runs on the GPU and fails loudly (NULL / -1 with a message) when no
 * GPU is present -- there is no CPU fallback.
 */
#include <stdio.h>
#include <string.h>
#include "internal.h"
#include "../../../include/taucs_amd.h"

static int check_input(const taucs_ccs_matrix* A, const char* who)
{
  if (!A) return -1;
  if (!(A->flags & TAUCS_DOUBLE)) {
    taucs_printf("%s: only TAUCS_DOUBLE matrices are supported on this path\n", who);
    return -1;
  }
  if (!(A->flags & (TAUCS_SYMMETRIC | TAUCS_HERMITIAN)) || !(A->flags & TAUCS_LOWER)) {
    taucs_printf("%s: matrix must be symmetric with its lower part stored\n", who);
    return -1;
  }
  return 0;
}

static tamd_factor* factor_symbolic(taucs_ccs_matrix* A, int max_depth)
{
  tamd_factor* F;
  double t0 = taucs_wtime();
  if (check_input(A, "taucs_ccs_factor_llt") != 0) return NULL;
  F = (tamd_factor*)tamd_calloc(1, sizeof(tamd_factor));
  if (!F) return NULL;
  F->flags = TAUCS_DOUBLE;
  F->uplo = 'l';
  F->n = A->n;
  if (tamd_symbolic_analyze(A, max_depth, &F->sym) != 0) {
    tamd_free(F);
    return NULL;
  }
  if (tamd_plan_build(&F->sym, A, &F->plan) != 0) {
    taucs_printf("taucs_ccs_factor_llt: building the level schedule failed\n");
    tamd_symbolic_free(&F->sym);
    tamd_free(F);
    return NULL;
  }
  F->t_symbolic = taucs_wtime() - t0;
  taucs_printf("\t\tSymbolic Analysis            = % 10.3f seconds\n", F->t_symbolic);
  return F;
}

static int factor_numeric(tamd_factor* F, taucs_ccs_matrix* A, int precision)
{
  double t0 = taucs_wtime();
  if (check_input(A, "taucs_ccs_factor_llt_numeric") != 0) return -1;
  if (A->n != F->n) {
    taucs_printf("taucs_ccs_factor_llt_numeric: dimension mismatch\n");
    return -1;
  }
  if (tamd_pattern_hash(A) != F->plan.a_pattern_hash) {
    /* same ordering, new pattern: recompute the assembly map (it fails if a
       new entry falls outside the symbolic structure) */
    if (tamd_plan_amap(&F->sym, A, &F->plan) != 0) {
      taucs_printf("taucs_ccs_factor_llt_numeric: pattern does not fit the symbolic factor\n");
      return -1;
    }
    tamd_dev_free(F);
  }
  F->precision = precision;
  if (!F->dev && tamd_dev_setup(F) != 0) return -1;
  if (tamd_dev_numeric(F, A) != 0) {
    tamd_dev_free_numeric(F);
    F->has_numeric = 0;
    return -1;
  }
  F->has_numeric = 1;
  F->t_numeric = taucs_wtime() - t0;
  taucs_printf("\t\tSupernodal Multifrontal LL^T = % 10.3f seconds\n", F->t_numeric);
  return 0;
}

static void factor_destroy(tamd_factor* F)
{
  if (!F) return;
  tamd_dev_free(F);
  tamd_plan_free(&F->plan);
  tamd_symbolic_free(&F->sym);
  tamd_free(F);
}

void* taucs_dccs_factor_llt_symbolic_maxdepth(taucs_ccs_matrix* A, int max_depth)
{
  return factor_symbolic(A, max_depth);
}
void* taucs_dccs_factor_llt_symbolic(taucs_ccs_matrix* A)
{
  return factor_symbolic(A, 0);
}
void* taucs_ccs_factor_llt_symbolic_maxdepth(taucs_ccs_matrix* A, int max_depth)
{
  return factor_symbolic(A, max_depth);
}
void* taucs_ccs_factor_llt_symbolic(taucs_ccs_matrix* A)
{
  return factor_symbolic(A, 0);
}

int taucs_dccs_factor_llt_numeric(taucs_ccs_matrix* A, void* vL)
{
  if (!vL) return -1;
  return factor_numeric((tamd_factor*)vL, A, ((tamd_factor*)vL)->precision);
}
int taucs_ccs_factor_llt_numeric(taucs_ccs_matrix* A, void* vL)
{
  return taucs_dccs_factor_llt_numeric(A, vL);
}
int taucs_amd_factor_llt_numeric_precision(taucs_ccs_matrix* A, void* vL, int precision)
{
  if (!vL) return -1;
  return factor_numeric((tamd_factor*)vL, A, precision);
}

void* taucs_amd_factor_llt_mf_precision(taucs_ccs_matrix* A, int precision)
{
  tamd_factor* F = factor_symbolic(A, 0);
  if (!F) return NULL;
  if (factor_numeric(F, A, precision) != 0) {
    factor_destroy(F);
    return NULL;
  }
  return F;
}

void* taucs_dccs_factor_llt_mf_maxdepth(taucs_ccs_matrix* A, int max_depth)
{
  tamd_factor* F = factor_symbolic(A, max_depth);
  if (!F) return NULL;
  if (factor_numeric(F, A, TAMD_PREC_FP64) != 0) {
    factor_destroy(F);
    return NULL;
  }
  return F;
}
void* taucs_dccs_factor_llt_mf(taucs_ccs_matrix* A) { return taucs_dccs_factor_llt_mf_maxdepth(A, 0); }
void* taucs_ccs_factor_llt_mf_maxdepth(taucs_ccs_matrix* A, int max_depth)
{
  if (A && !(A->flags & TAUCS_DOUBLE)) {
    taucs_printf("taucs_ccs_factor_llt_mf: only TAUCS_DOUBLE is supported\n");
    return NULL;
  }
  return taucs_dccs_factor_llt_mf_maxdepth(A, max_depth);
}
void* taucs_ccs_factor_llt_mf(taucs_ccs_matrix* A) { return taucs_ccs_factor_llt_mf_maxdepth(A, 0); }

int taucs_amd_supernodal_solve_many(void* vL, int nrhs, double* X, int ldx, const double* B, int ldb)
{
  tamd_factor* F = (tamd_factor*)vL;
  if (!F || !F->has_numeric) {
    taucs_printf("taucs_supernodal_solve_llt: no numeric factor\n");
    return -1;
  }
  if (nrhs <= 0) return 0;
  return tamd_dev_solve(F, nrhs, X, ldx, B, ldb);
}

int taucs_dsupernodal_solve_llt(void* vL, void* vx, void* vb)
{
  tamd_factor* F = (tamd_factor*)vL;
  if (!F) return -1;
  return taucs_amd_supernodal_solve_many(vL, 1, (double*)vx, F->n, (const double*)vb, F->n);
}
int taucs_supernodal_solve_llt(void* vL, void* vx, void* vb)
{
  tamd_factor* F = (tamd_factor*)vL;
  if (!F || !(F->flags & TAUCS_DOUBLE)) return -1;
  return taucs_dsupernodal_solve_llt(vL, vx, vb);
}

void taucs_dsupernodal_factor_free(void* vL) { factor_destroy((tamd_factor*)vL); }
void taucs_supernodal_factor_free(void* vL) { factor_destroy((tamd_factor*)vL); }

void taucs_dsupernodal_factor_free_numeric(void* vL)
{
  tamd_factor* F = (tamd_factor*)vL;
  if (!F) return;
  tamd_dev_free_numeric(F);
  F->has_numeric = 0;
}
void taucs_supernodal_factor_free_numeric(void* vL) { taucs_dsupernodal_factor_free_numeric(vL); }

/* Host copy of all panels (FP64, whatever the device precision). */
static double* download_panels(tamd_factor* F)
{
  double* h;
  if (!F || !F->has_numeric) return NULL;
  h = (double*)tamd_malloc(sizeof(double) * (size_t)(F->plan.L_elems > 0 ? F->plan.L_elems : 1));
  if (!h) return NULL;
  if (tamd_dev_download_blocks(F, h) != 0) { tamd_free(h); return NULL; }
  return h;
}

/* Columns in supernode order, each: its rows from the diagonal down in the
   supernode's row order; exact zeros dropped (src/taucs_sn_llt.c:885-975). */
taucs_ccs_matrix* taucs_dsupernodal_factor_to_ccs(void* vL)
{
  tamd_factor* F = (tamd_factor*)vL;
  tamd_symbolic* S;
  taucs_ccs_matrix* C;
  double* h;
  int* len;
  int n, sn, jp, ip, j;
  int64_t nnz = 0;
  if (!F) return NULL;
  S = &F->sym;
  n = F->n;
  int* dpos;  /* global row -> position in the current front's device row order */
  h = download_panels(F);
  if (!h) return NULL;
  len = (int*)tamd_calloc((size_t)n + 1, sizeof(int));
  dpos = (int*)tamd_malloc(sizeof(int) * ((size_t)n + 1));
  if (!len || !dpos) { tamd_free(len); tamd_free(dpos); tamd_free(h); return NULL; }
  /* entries go out in the reference's row order (sn_struct), read from the
     device panel's row order */
  for (sn = 0; sn < S->n_sn; sn++) {
    int s = S->sn_size[sn], m = S->sn_up_size[sn];
    const double* P;
    if (s == 0 || !F->plan.drow[sn]) continue;
    P = h + F->plan.off_L[sn];
    for (ip = 0; ip < m; ip++) dpos[F->plan.drow[sn][ip]] = ip;
    for (jp = 0; jp < s; jp++) {
      j = S->sn_struct[sn][jp];
      len[j] = 0;
      for (ip = jp; ip < m; ip++)
        if (P[(int64_t)jp * m + dpos[S->sn_struct[sn][ip]]] != 0.0) { len[j]++; nnz++; }
    }
  }
  if (nnz > 0x7fffffff) { tamd_free(len); tamd_free(dpos); tamd_free(h); return NULL; }
  C = taucs_ccs_create(n, n, (int)nnz, TAUCS_DOUBLE);
  if (!C) { tamd_free(len); tamd_free(dpos); tamd_free(h); return NULL; }
  C->flags = TAUCS_DOUBLE | TAUCS_TRIANGULAR | TAUCS_LOWER;
  C->colptr[0] = 0;
  for (j = 0; j < n; j++) C->colptr[j + 1] = C->colptr[j] + len[j];
  for (sn = 0; sn < S->n_sn; sn++) {
    int s = S->sn_size[sn], m = S->sn_up_size[sn];
    const double* P;
    if (s == 0) continue;
    P = h + F->plan.off_L[sn];
    if (!F->plan.drow[sn]) continue;
    for (ip = 0; ip < m; ip++) dpos[F->plan.drow[sn][ip]] = ip;
    for (jp = 0; jp < s; jp++) {
      int next;
      j = S->sn_struct[sn][jp];
      next = C->colptr[j];
      for (ip = jp; ip < m; ip++) {
        double v = P[(int64_t)jp * m + dpos[S->sn_struct[sn][ip]]];
        if (v == 0.0) continue;
        C->rowind[next] = S->sn_struct[sn][ip];
        C->values.d[next] = v;
        next++;
      }
    }
  }
  tamd_free(len);
  tamd_free(dpos);
  tamd_free(h);
  return C;
}
taucs_ccs_matrix* taucs_supernodal_factor_to_ccs(void* vL) { return taucs_dsupernodal_factor_to_ccs(vL); }

taucs_double* taucs_dsupernodal_factor_get_diag(void* vL)
{
  tamd_factor* F = (tamd_factor*)vL;
  double *h, *d;
  int sn, jp;
  if (!F) return NULL;
  h = download_panels(F);
  if (!h) return NULL;
  d = (double*)tamd_malloc(sizeof(double) * (size_t)(F->n > 0 ? F->n : 1));
  if (!d) { tamd_free(h); return NULL; }
  for (sn = 0; sn < F->sym.n_sn; sn++) {
    int s = F->sym.sn_size[sn], m = F->sym.sn_up_size[sn];
    for (jp = 0; jp < s; jp++)
      d[F->sym.sn_struct[sn][jp]] = h[F->plan.off_L[sn] + (int64_t)jp * m + jp];
  }
  tamd_free(h);
  return d;
}
void* taucs_supernodal_factor_get_diag(void* vL) { return taucs_dsupernodal_factor_get_diag(vL); }

int taucs_amd_factor_get_info(void* vL, taucs_amd_factor_info* info)
{
  tamd_factor* F = (tamd_factor*)vL;
  if (!F || !info) return -1;
  memset(info, 0, sizeof(*info));
  info->n = F->n;
  info->n_sn = F->sym.n_sn;
  info->n_live = F->plan.n_live;
  info->n_levels = F->plan.n_levels;
  info->max_front = F->plan.max_front;
  info->etree_depth = F->sym.etree_depth;
  info->precision = F->precision;
  info->has_numeric = F->has_numeric;
  info->fund_nnz = F->sym.fund_nnz;
  info->fund_flops = F->sym.fund_flops;
  info->relax_nnz = F->sym.relax_nnz;
  info->relax_flops = F->sym.relax_flops;
  info->t_symbolic = F->t_symbolic;
  info->t_numeric = F->t_numeric;
  info->L_bytes = F->plan.L_elems * (F->precision == TAMD_PREC_FP64 ? 8 : 4);
  info->U_bytes = F->plan.U_elems * (F->precision == TAMD_PREC_FP64 ? 8 : 4);
  return 0;
}

int taucs_amd_factor_get_supernodes(void* vL, int* sn_size, int* sn_up_size, int* first_child,
                                    int* next_child, int* struct_ptr, int* struct_idx)
{
  tamd_factor* F = (tamd_factor*)vL;
  const tamd_symbolic* S;
  int sn, p = 0;
  if (!F) return -1;
  S = &F->sym;
  for (sn = 0; sn <= S->n_sn; sn++) {
    if (sn_size) sn_size[sn] = sn < S->n_sn ? S->sn_size[sn] : 0;
    if (sn_up_size) sn_up_size[sn] = sn < S->n_sn ? S->sn_up_size[sn] : 0;
    if (first_child) first_child[sn] = S->first_child[sn];
    if (next_child) next_child[sn] = S->next_child[sn];
  }
  for (sn = 0; sn < S->n_sn; sn++) {
    if (struct_ptr) struct_ptr[sn] = p;
    if (struct_idx && S->sn_up_size[sn] > 0)
      memcpy(struct_idx + p, S->sn_struct[sn], sizeof(int) * (size_t)S->sn_up_size[sn]);
    p += S->sn_up_size[sn];
  }
  if (struct_ptr) struct_ptr[S->n_sn] = p;
  return 0;
}

int taucs_amd_factor_get_schedule(void* vL, int* sn_level, int* sn_parent, long long* off_L,
                                  long long* off_U)
{
  tamd_factor* F = (tamd_factor*)vL;
  const tamd_plan* P;
  int sn;
  if (!F) return -1;
  P = &F->plan;
  for (sn = 0; sn < P->n_sn; sn++) {
    if (sn_level) sn_level[sn] = P->sn_level[sn];
    if (sn_parent) sn_parent[sn] = P->sn_parent[sn];
    if (off_L) off_L[sn] = (long long)P->off_L[sn];
    if (off_U) off_U[sn] = (long long)P->off_U[sn];
  }
  return 0;
}

/* Autotuner of the schedule shape: the numeric phase is timed with each
   candidate schedule (outer panel width cand_w[i], multiple of 64, and tile
   edge cand_tile[i] of the large GEMMs, 64 or 128; cand_tile may be NULL for
   the default), and the fastest schedule is kept in L, factored from A.
   ms[i] receives candidate i's best time over `reps` timed runs; *best its
   index. */
int taucs_amd_autotune(void* vL, taucs_ccs_matrix* A, int precision, const int* cand_w,
                       const int* cand_tile, int ncand, int reps, double* ms, int* best)
{
  tamd_factor* F = (tamd_factor*)vL;
  int c, bi = -1;
  double best_ms = 1e300;
  if (!F || !A || !cand_w || ncand <= 0 || check_input(A, "taucs_amd_autotune") != 0) return -1;
  if (A->n != F->n || tamd_pattern_hash(A) != F->plan.a_pattern_hash) return -1;
  F->precision = precision;
  for (c = 0; c < ncand; c++) {
    double t;
    if (cand_w[c] < 64 || cand_w[c] % 64) return -1;
    tamd_dev_free(F);
    F->has_numeric = 0;
    F->panel_w = cand_w[c];
    F->gemm_tile = cand_tile ? cand_tile[c] : 0;
    if (tamd_dev_setup(F) != 0 || tamd_dev_time_numeric(F, A->values.d, reps, &t) != 0) return -1;
    if (ms) ms[c] = t;
    if (t < best_ms) { best_ms = t; bi = c; }
  }
  tamd_dev_free(F);
  F->panel_w = cand_w[bi];
  F->gemm_tile = cand_tile ? cand_tile[bi] : 0;
  if (tamd_dev_setup(F) != 0 || tamd_dev_numeric(F, A) != 0) return -1;
  F->has_numeric = 1;
  if (best) *best = bi;
  return 0;
}

/* Numeric factorization from A's values already in device memory (d_vals:
   nnz doubles in the stored order of the permuted matrix the symbolic
   factor was built from), for callers whose inputs live in HBM. */
int taucs_amd_factor_llt_numeric_device(void* vL, const double* d_vals, int precision)
{
  tamd_factor* F = (tamd_factor*)vL;
  double t0 = taucs_wtime();
  if (!F || !d_vals) return -1;
  F->precision = precision;
  if (!F->dev && tamd_dev_setup(F) != 0) return -1;
  if (tamd_dev_numeric_device(F, d_vals) != 0) {
    tamd_dev_free_numeric(F);
    F->has_numeric = 0;
    return -1;
  }
  F->has_numeric = 1;
  F->t_numeric = taucs_wtime() - t0;
  return 0;
}
