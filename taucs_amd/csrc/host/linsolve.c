/*
 * taucs_linsolve and its option parser (src/taucs_linsolve.c:41-151,
 * 165-187, 189-622), restricted to in-core LL^T on the GPU.
 *
 * Option strings keep the reference's grammar: "name=value", "name=#k"
 * fetches the value from opt_arg[k], a name followed by '.' belongs to a
 * longer option.  Understood:
 *   taucs.factor, taucs.factor.symbolic, taucs.factor.numeric,
 *   taucs.factor.LLT, taucs.factor.mf, taucs.factor.ll, taucs.factor.ordering,
 *   taucs.maxdepth, taucs.solve.cg, taucs.solve.maxits, taucs.solve.convergetol
 * and, additively,
 *   taucs.factor.precision=fp64|fp32|bf16upd  device factor precision
 *   taucs.solve.refine=true                   FP64 refinement to convergetol
 * Both LL^T variants (mf and the left-looking default) run the GPU
 * multifrontal factor; their results differ only by rounding.  LU, out-of-core
 * and the support-graph preconditioners are not part of this library and
 * return TAUCS_ERROR_BADARGS.
 */
#include <stdio.h>
#include <string.h>
#include "internal.h"
#include "../../../include/taucs_amd.h"

static int getopt_ptr_index(const char* cmd, int ln, void* args[], void** out)
{
  unsigned int p, i;
  if (sscanf(cmd + ln + 2, "%u", &p) == 1) {
    for (i = 0; args && args[i]; i++)
      if (i == p) { *out = args[i]; return 1; }
    taucs_printf("taucs: WARNING, pointer argument out of range in [%s]\n", cmd);
  }
  taucs_printf("taucs: WARNING, illegal pointer argument in [%s]\n", cmd);
  return 0;
}

/* 1: the option is `name` and was parsed; 0 otherwise.  *rest points at the
   value text when the option is `name=value` (not #k). */
static int getopt_match(char* cmd, void* args[], char* name, const char** rest, void** ptr)
{
  int lc = (int)strlen(cmd), ln = (int)strlen(name);
  *rest = NULL;
  *ptr = NULL;
  if (strncmp(cmd, name, (size_t)ln)) return 0;
  if (lc > ln && cmd[ln] == '.') return 0;
  if (lc > ln && cmd[ln] == '=') {
    if (cmd[ln + 1] == '#') return getopt_ptr_index(cmd, ln, args, ptr) ? 2 : -1;
    *rest = cmd + ln + 1;
    return 1;
  }
  taucs_printf("taucs: WARNING, illegal argument in [%s]\n", cmd);
  return -1;
}

int taucs_getopt_boolean(char* cmd, void* args[], char* name, int* x)
{
  const char* v;
  void* p;
  int r = getopt_match(cmd, args, name, &v, &p);
  if (r == 2) { *x = *(int*)p; return 1; }
  if (r != 1) return 0;
  if (!strcmp(v, "true")) { *x = 1; return 1; }
  if (!strcmp(v, "false")) { *x = 0; return 1; }
  taucs_printf("taucs: WARNING, illegal argument in [%s]\n", cmd);
  return 0;
}

int taucs_getopt_double(char* cmd, void* args[], char* name, double* x)
{
  const char* v;
  void* p;
  int r = getopt_match(cmd, args, name, &v, &p);
  if (r == 2) { *x = *(double*)p; return 1; }
  if (r != 1) return 0;
  if (sscanf(v, "%le", x) == 1) return 1;
  taucs_printf("taucs: WARNING, illegal argument in [%s]\n", cmd);
  return 0;
}

int taucs_getopt_pointer(char* cmd, void* args[], char* name, void** x)
{
  const char* v;
  void* p;
  int r = getopt_match(cmd, args, name, &v, &p);
  if (r == 2) { *x = *(void**)p; return 1; }
  if (r == 1) taucs_printf("taucs: WARNING, illegal argument in [%s]\n", cmd);
  return 0;
}

int taucs_getopt_string(char* cmd, void* args[], char* name, char** x)
{
  const char* v;
  void* p;
  int r = getopt_match(cmd, args, name, &v, &p);
  if (r == 2) { *x = *(char**)p; return 1; }
  if (r != 1) return 0;
  *x = (char*)v;
  return 1;
}

/* the taucs_factorization handle of src/taucs_linsolve.c:165-172 */
typedef struct {
  int   n;
  int   flags;
  int   type;     /* 1: supernodal LL^T */
  int*  rowperm;
  int*  colperm;
  void* L;
} linsolve_factorization;

static void linsolve_free(linsolve_factorization* f)
{
  if (!f) return;
  if (f->type == 1) taucs_supernodal_factor_free(f->L);
  tamd_free(f->rowperm);
  tamd_free(f->colperm);
  tamd_free(f);
}

int taucs_linsolve(taucs_ccs_matrix* A, void** F, int nrhs, void* X, void* B, char* options[],
                   void* opt_arg[])
{
  int rc = TAUCS_SUCCESS, i, j;
  int opt_factor = 1, opt_symbolic = 1, opt_numeric = 1;
  int opt_llt = 0, opt_lu = 0, opt_mf = 0, opt_ll = 0, opt_ooc = 0;
  int opt_cg = 0, opt_minres = 0, opt_refine = 0;
  double opt_maxdepth = 0.0, opt_maxits = 300.0, opt_tol = 1e-6;
  char* opt_ordering = NULL;
  char* opt_precision = NULL;
  int precision = TAMD_PREC_FP64;
  linsolve_factorization* f = NULL;
  taucs_ccs_matrix* PAPT = NULL;
  int *rowperm = NULL, *colperm = NULL;
  double *PX = NULL, *PB = NULL;

  if (!A && nrhs == 0) {
    if (F) linsolve_free((linsolve_factorization*)*F);
    if (F) *F = NULL;
    return TAUCS_SUCCESS;
  }

  if (options) {
    for (i = 0; options[i]; i++) {
      int u = 0;
      u |= taucs_getopt_boolean(options[i], opt_arg, "taucs.factor", &opt_factor);
      u |= taucs_getopt_boolean(options[i], opt_arg, "taucs.factor.symbolic", &opt_symbolic);
      u |= taucs_getopt_boolean(options[i], opt_arg, "taucs.factor.numeric", &opt_numeric);
      u |= taucs_getopt_boolean(options[i], opt_arg, "taucs.factor.LLT", &opt_llt);
      u |= taucs_getopt_boolean(options[i], opt_arg, "taucs.factor.LU", &opt_lu);
      u |= taucs_getopt_boolean(options[i], opt_arg, "taucs.factor.mf", &opt_mf);
      u |= taucs_getopt_boolean(options[i], opt_arg, "taucs.factor.ll", &opt_ll);
      u |= taucs_getopt_string(options[i], opt_arg, "taucs.factor.ordering", &opt_ordering);
      u |= taucs_getopt_string(options[i], opt_arg, "taucs.factor.precision", &opt_precision);
      u |= taucs_getopt_double(options[i], opt_arg, "taucs.maxdepth", &opt_maxdepth);
      u |= taucs_getopt_boolean(options[i], opt_arg, "taucs.ooc", &opt_ooc);
      u |= taucs_getopt_boolean(options[i], opt_arg, "taucs.solve.cg", &opt_cg);
      u |= taucs_getopt_boolean(options[i], opt_arg, "taucs.solve.minres", &opt_minres);
      u |= taucs_getopt_boolean(options[i], opt_arg, "taucs.solve.refine", &opt_refine);
      u |= taucs_getopt_double(options[i], opt_arg, "taucs.solve.maxits", &opt_maxits);
      u |= taucs_getopt_double(options[i], opt_arg, "taucs.solve.convergetol", &opt_tol);
      if (!u) taucs_printf("taucs_linsolve: illegal option [[%s]]\n", options[i]);
    }
  }
  (void)opt_mf;
  (void)opt_ll;

  if (opt_precision) {
    if (!strcmp(opt_precision, "fp64")) precision = TAMD_PREC_FP64;
    else if (!strcmp(opt_precision, "fp32")) precision = TAMD_PREC_FP32;
    else if (!strcmp(opt_precision, "bf16upd")) precision = TAMD_PREC_BF16UPD;
    else {
      taucs_printf("taucs_linsolve: unknown precision [%s]\n", opt_precision);
      return TAUCS_ERROR_BADARGS;
    }
  }

  if (opt_factor) {
    if (opt_lu || opt_ooc || !opt_llt) {
      taucs_printf("taucs_linsolve: only in-core LL^T factorization is available\n");
      return TAUCS_ERROR_BADARGS;
    }
    f = (linsolve_factorization*)tamd_calloc(1, sizeof(linsolve_factorization));
    if (!f) return TAUCS_ERROR_NOMEM;
    f->n = A->n;
    f->flags = A->flags;
    if (!opt_numeric && nrhs > 0) {
      taucs_printf("taucs_linsolve: WARNING, you can't solve without a numeric factorization\n");
      opt_numeric = 1;
    }
    if (!opt_ordering) opt_ordering = "genmmd";
    taucs_ccs_order(A, &rowperm, &colperm, opt_ordering);
    if (!rowperm) {
      taucs_printf("taucs_factor: ordering failed\n");
      rc = TAUCS_ERROR_NOMEM;
      goto fail;
    }
    f->rowperm = rowperm;
    f->colperm = colperm;
    PAPT = taucs_ccs_permute_symmetrically(A, rowperm, colperm);
    if (!PAPT) { rc = TAUCS_ERROR_NOMEM; goto fail; }

    if (!opt_numeric && opt_symbolic) {
      f->L = taucs_ccs_factor_llt_symbolic_maxdepth(PAPT, (int)opt_maxdepth);
    } else if (opt_numeric && !opt_symbolic) {
      linsolve_factorization* g = F ? (linsolve_factorization*)*F : NULL;
      if (!g || g->type != 1) {
        taucs_printf("taucs_linsolve: ERROR, you need to provide a symbolic factorization for a numeric factorization\n");
        rc = TAUCS_ERROR_BADARGS;
        goto fail;
      }
      /* the new factorization takes over the symbolic object */
      f->L = g->L;
      g->L = NULL;
      g->type = 0;
      taucs_supernodal_factor_free_numeric(f->L);
      if (taucs_amd_factor_llt_numeric_precision(PAPT, f->L, precision) != 0) {
        taucs_supernodal_factor_free(f->L);
        f->L = NULL;
      }
    } else {
      f->L = taucs_ccs_factor_llt_symbolic_maxdepth(PAPT, (int)opt_maxdepth);
      if (f->L && taucs_amd_factor_llt_numeric_precision(PAPT, f->L, precision) != 0) {
        taucs_supernodal_factor_free(f->L);
        f->L = NULL;
      }
    }
    if (!f->L) {
      taucs_printf("taucs_factor: factorization failed\n");
      rc = TAUCS_ERROR;
      goto fail;
    }
    f->type = 1;
  }

  if (nrhs > 0) {
    int n = A->n;
    if (!f) {
      if (!F || !*F) {
        taucs_printf("taucs_linsolve: can't solve, no factorization\n");
        rc = TAUCS_ERROR;
        goto fail;
      }
      f = (linsolve_factorization*)*F;
    }
    if (f->type != 1 || opt_minres) {
      taucs_printf("taucs_linsolve: I don't know how to solve!\n");
      rc = TAUCS_ERROR;
      goto fail_solve;
    }
    PX = (double*)tamd_malloc(sizeof(double) * (size_t)n * (size_t)nrhs);
    PB = (double*)tamd_malloc(sizeof(double) * (size_t)n * (size_t)nrhs);
    if (!PX || !PB) { rc = TAUCS_ERROR_NOMEM; goto fail_solve; }
    for (j = 0; j < nrhs; j++)
      taucs_vec_permute(n, TAUCS_DOUBLE, (double*)B + (size_t)j * n, PB + (size_t)j * n,
                        f->rowperm);
    if (opt_cg || opt_refine || precision != TAMD_PREC_FP64) {
      taucs_amd_refine_info info;
      int method;
      double tol;
      if (!PAPT) PAPT = taucs_ccs_permute_symmetrically(A, f->rowperm, f->colperm);
      if (!PAPT) { rc = TAUCS_ERROR_NOMEM; goto fail_solve; }
      if (opt_cg) {
        /* conjugate gradients preconditioned by the factor, stopping like the
           reference's on |r|_2 <= convergetol |b|_2 (src/taucs_iter.c:139,186;
           x0 = 0, so r0 = b) */
        method = 1 | TAUCS_AMD_REFINE_STOP_RELRES;
        tol = opt_tol;
      } else {
        /* refinement to working accuracy in the normwise backward error */
        method = 0;
        tol = 1e-14;
      }
      /* the factor holds this PAPT's values only when this call factored it */
      if (opt_factor) method |= TAUCS_AMD_REFINE_FACTORED_VALUES;
      if (taucs_amd_solve_refined(f->L, PAPT, nrhs, PX, n, PB, n, method, tol, (int)opt_maxits,
                                  &info) != 0)
        rc = TAUCS_ERROR;
    } else {
      if (taucs_amd_supernodal_solve_many(f->L, nrhs, PX, n, PB, n) != 0) rc = TAUCS_ERROR;
    }
    for (j = 0; j < nrhs; j++)
      taucs_vec_ipermute(n, TAUCS_DOUBLE, PX + (size_t)j * n, (double*)X + (size_t)j * n,
                         f->rowperm);
  fail_solve:
    tamd_free(PX);
    tamd_free(PB);
    PX = PB = NULL;
    if (rc != TAUCS_SUCCESS && opt_factor) goto fail;
  }

  if (opt_factor) {
    if (F) *F = f;
    else linsolve_free(f);
  }
  taucs_ccs_free(PAPT);
  return rc;

fail:
  taucs_printf("taucs_linsolve: an error occured, releasing resources and bailing out\n");
  taucs_ccs_free(PAPT);
  if (opt_factor) linsolve_free(f);
  else if (!f) { tamd_free(rowperm); tamd_free(colperm); }
  return rc;
}
