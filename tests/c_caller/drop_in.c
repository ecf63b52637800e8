/*
 * A C caller of the drop-in, built with gcc against include/taucs.h and
 * linked with -ltaucs_amd, exactly as existing TAUCS code links the
 * reference library.  It runs the protocol of the reference's own test
 * program (progs/test_linsolve.c:128-186 and its main: a 20^3 mesh, four
 * right-hand sides B = A X with X uniform on [0,1], relative residual
 * <= 1e-8 after every solve, the expected failures), the flow of the direct
 * driver (progs/direct.c:540-550: order, permute, taucs_ccs_factor_llt_mf,
 * taucs_supernodal_solve_llt, then the CCS export solved with
 * taucs_ccs_solve_llt), the symbolic/numeric split, and the conjugate
 * gradients of taucs_linsolve (taucs.solve.cg) with the factor as the
 * preconditioner.
 *
 * Orderings and factorizations this build leaves out (METIS, AMD, the
 * out-of-core factor) are reported as not available and must fail cleanly.
 * Prints "test succeeded" and exits 0, or names the failing step and exits 1.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "taucs.h"

static int failures = 0;

static void expect(int ok, const char* what)
{
  printf("%-64s %s\n", what, ok ? "ok" : "FAILED");
  if (!ok) failures++;
}

/* |A y - b|_2 / |b|_2 with the library's own vector kernels */
static double relres(taucs_ccs_matrix* A, double* y, double* b, double* tmp)
{
  taucs_ccs_times_vec(A, y, tmp);
  taucs_vec_axpby(A->n, A->flags, 1.0, tmp, -1.0, b, tmp);
  return taucs_vec_norm2(A->n, A->flags, tmp) / taucs_vec_norm2(A->n, A->flags, b);
}

static int good(taucs_ccs_matrix* A, double* y, double* b, double* tmp, int nrhs, double tol)
{
  int r;
  for (r = 0; r < nrhs; r++) {
    const size_t o = (size_t)r * A->n;
    if (!(relres(A, y + o, b + o, tmp) <= tol)) return 0;
  }
  return 1;
}

int main(int argc, char** argv)
{
  const int side = argc > 1 ? atoi(argv[1]) : 20;
  taucs_ccs_matrix* A;
  double *X, *B, *Y, *Z;
  int n, i, rc;
  void* F = NULL;
  void* opt_arg[] = {NULL};
  char* factor[] = {"taucs.factor.LLT=true", NULL};
  char* solve[] = {"taucs.factor=false", NULL};
  char* mf[] = {"taucs.factor.LLT=true", "taucs.factor.mf=true", NULL};
  char* ll[] = {"taucs.factor.LLT=true", "taucs.factor.ll=true", NULL};
  char* mf_depth[] = {"taucs.factor.LLT=true", "taucs.factor.mf=true", "taucs.maxdepth=5", NULL};
  char* ll_depth[] = {"taucs.factor.LLT=true", "taucs.factor.ll=true", "taucs.maxdepth=5", NULL};
  char* genmmd[] = {"taucs.factor.LLT=true", "taucs.factor.ordering=genmmd", NULL};
  char* amd[] = {"taucs.factor.LLT=true", "taucs.factor.ordering=amd", NULL};
  char* metis[] = {"taucs.factor.LLT=true", "taucs.factor.ordering=metis", NULL};
  char* colamd[] = {"taucs.factor.LLT=true", "taucs.factor.ordering=colamd", NULL};
  char* ooc[] = {"taucs.factor.LLT=true", "taucs.ooc=true", "taucs.ooc.basename=taucs-test", NULL};
  char* cg[] = {"taucs.factor.LLT=true", "taucs.solve.cg=true", "taucs.solve.convergetol=1e-12",
                NULL};

  taucs_logfile("none");
  A = taucs_ccs_generate_mesh3d(side, side, side);
  if (!A) { printf("matrix generation failed\n"); return 1; }
  n = A->n;
  X = malloc(sizeof(double) * 4 * n);
  B = malloc(sizeof(double) * 4 * n);
  Y = malloc(sizeof(double) * 4 * n);
  Z = malloc(sizeof(double) * 4 * n);
  if (!X || !B || !Y || !Z) { printf("vector allocation failed\n"); return 1; }
  srand(12345);
  for (i = 0; i < 4 * n; i++) X[i] = (double)rand() / RAND_MAX;
  for (i = 0; i < 4; i++) taucs_ccs_times_vec(A, X + (size_t)i * n, B + (size_t)i * n);

  /* factor-solve protocol */
  expect(taucs_linsolve(A, NULL, 1, Y, B, solve, opt_arg) != TAUCS_SUCCESS,
         "solve without a factorization fails (no handle)");
  expect(taucs_linsolve(A, &F, 1, Y, B, solve, opt_arg) != TAUCS_SUCCESS,
         "solve without a factorization fails (empty handle)");
  rc = taucs_linsolve(A, &F, 1, Y, B, factor, opt_arg);
  expect(rc == TAUCS_SUCCESS && good(A, Y, B, Z, 1, 1e-8), "factor + solve, 1 rhs");
  expect(taucs_linsolve(NULL, &F, 0, NULL, NULL, factor, opt_arg) == TAUCS_SUCCESS, "free the factor");
  expect(taucs_linsolve(A, &F, 0, NULL, NULL, factor, opt_arg) == TAUCS_SUCCESS, "factor only");
  rc = taucs_linsolve(A, &F, 1, Y, B, solve, opt_arg);
  expect(rc == TAUCS_SUCCESS && good(A, Y, B, Z, 1, 1e-8), "solve with the kept factor");
  expect(taucs_linsolve(NULL, &F, 0, NULL, NULL, factor, opt_arg) == TAUCS_SUCCESS, "free the factor");
  rc = taucs_linsolve(A, &F, 4, Y, B, factor, opt_arg);
  expect(rc == TAUCS_SUCCESS && good(A, Y, B, Z, 4, 1e-8), "factor + solve, 4 rhs");
  expect(taucs_linsolve(NULL, &F, 0, NULL, NULL, factor, opt_arg) == TAUCS_SUCCESS, "free the factor");

  /* orderings */
  rc = taucs_linsolve(A, NULL, 1, Y, B, genmmd, opt_arg);
  expect(rc == TAUCS_SUCCESS && good(A, Y, B, Z, 1, 1e-8), "ordering genmmd");
  expect(taucs_linsolve(A, NULL, 1, Y, B, amd, opt_arg) != TAUCS_SUCCESS,
         "ordering amd: not in this build, fails cleanly");
  expect(taucs_linsolve(A, NULL, 1, Y, B, metis, opt_arg) != TAUCS_SUCCESS,
         "ordering metis: not in this build, fails cleanly");
  expect(taucs_linsolve(A, NULL, 1, Y, B, colamd, opt_arg) != TAUCS_SUCCESS,
         "ordering colamd on a symmetric matrix fails");

  /* factorizations */
  rc = taucs_linsolve(A, NULL, 1, Y, B, mf, opt_arg);
  expect(rc == TAUCS_SUCCESS && good(A, Y, B, Z, 1, 1e-8), "multifrontal LL^T");
  rc = taucs_linsolve(A, NULL, 1, Y, B, ll, opt_arg);
  expect(rc == TAUCS_SUCCESS && good(A, Y, B, Z, 1, 1e-8), "left-looking LL^T (served by the GPU MF)");
  expect(taucs_linsolve(A, NULL, 1, Y, B, mf_depth, opt_arg) != TAUCS_SUCCESS, "mf, maxdepth 5 fails");
  expect(taucs_linsolve(A, NULL, 1, Y, B, ll_depth, opt_arg) != TAUCS_SUCCESS, "ll, maxdepth 5 fails");
  expect(taucs_linsolve(A, NULL, 1, Y, B, ooc, opt_arg) != TAUCS_SUCCESS,
         "out-of-core: not in this build, fails cleanly");

  /* conjugate gradients preconditioned by the factor */
  rc = taucs_linsolve(A, NULL, 2, Y, B, cg, opt_arg);
  expect(rc == TAUCS_SUCCESS && good(A, Y, B, Z, 2, 1e-12), "cg, convergetol 1e-12");

  /* the direct driver's flow */
  {
    int *perm = NULL, *invperm = NULL;
    taucs_ccs_matrix *PAPT, *Lc;
    double *pb = malloc(sizeof(double) * n), *px = malloc(sizeof(double) * n),
           *x = malloc(sizeof(double) * n), *px2 = malloc(sizeof(double) * n);
    void* L;
    double dmax = 0.0;
    taucs_ccs_order(A, &perm, &invperm, "genmmd");
    expect(perm && invperm, "taucs_ccs_order genmmd");
    PAPT = taucs_ccs_permute_symmetrically(A, perm, invperm);
    expect(PAPT != NULL, "taucs_ccs_permute_symmetrically");
    L = taucs_ccs_factor_llt_mf(PAPT);
    expect(L != NULL, "taucs_ccs_factor_llt_mf");
    taucs_vec_permute(n, TAUCS_DOUBLE, B, pb, perm);
    expect(taucs_supernodal_solve_llt(L, px, pb) == 0, "taucs_supernodal_solve_llt");
    taucs_vec_ipermute(n, TAUCS_DOUBLE, px, x, perm);
    expect(good(A, x, B, Z, 1, 1e-12), "supernodal solve residual <= 1e-12");
    Lc = taucs_supernodal_factor_to_ccs(L);
    expect(Lc != NULL, "taucs_supernodal_factor_to_ccs");
    expect(taucs_ccs_solve_llt(Lc, px2, pb) == 0, "taucs_ccs_solve_llt on the export");
    for (i = 0; i < n; i++) {
      double d = px2[i] - px[i];
      if (d < 0) d = -d;
      if (d > dmax) dmax = d;
    }
    expect(dmax <= 1e-10, "CCS solve equals the supernodal solve");
    taucs_ccs_free(Lc);
    taucs_supernodal_factor_free(L);
    /* the symbolic / numeric split on the same structure */
    L = taucs_ccs_factor_llt_symbolic(PAPT);
    expect(L != NULL && taucs_ccs_factor_llt_numeric(PAPT, L) == 0, "symbolic, then numeric");
    expect(L && taucs_supernodal_solve_llt(L, px2, pb) == 0, "solve after the split");
    dmax = 0.0;
    for (i = 0; i < n; i++) {
      double d = px2[i] - px[i];
      if (d < 0) d = -d;
      if (d > dmax) dmax = d;
    }
    expect(dmax == 0.0, "split factor solves bitwise like the one-call factor");
    taucs_supernodal_factor_free(L);
    taucs_ccs_free(PAPT);
    free(perm); free(invperm); free(pb); free(px); free(x); free(px2);
  }

  taucs_ccs_free(A);
  free(X); free(B); free(Y); free(Z);
  if (failures) {
    printf("%d checks failed\n", failures);
    return 1;
  }
  printf("test succeeded\n");
  return 0;
}
