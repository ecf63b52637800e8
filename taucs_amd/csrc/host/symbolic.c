/*
 * Host symbolic analysis: elimination tree, fundamental supernodes,
 * relaxed amalgamation, and the level schedule / HBM layout the device
 * numeric phase runs on.
 *
 * The supernode partition, the supernodal tree (including child-list order)
 * and every supernode's row order must equal TAUCS's
 * (src/taucs_sn_llt.c:3663-4001), because the TAUCS double-precision
 * multifrontal path never sorts row indices (:2074-2076) and the tests pin
 * the structure bit-exactly.  The decision rules are therefore the
 * reference's:
 *   fundamental supernodes  -- :1413-1451 (single child whose structure
 *                              covers the column), row list = {j}, then new
 *                              rows from the children in child-list order,
 *                              then A's column :1455-1480
 *   child-list order        -- first etree child kept last, later children
 *                              pushed in front :1483-1492
 *   amalgamation            -- :1527-1786, thresholds :1716-1725, started at
 *                              the last supernode only :3927
 * The traversals are iterative (explicit stacks) so deep elimination trees
 * of 3D meshes need no large thread stack.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "internal.h"

/* ---- elimination tree ---------------------------------------------------
 * Liu's algorithm with path-halving ancestors.  The tree is unique for a
 * given pattern, so any correct algorithm reproduces src/taucs_sn_llt.c:
 * 3243-3340; roots get parent n. */
int tamd_etree(const taucs_ccs_matrix* A, int* parent)
{
  int n = A->n, i, j, ip;
  int *anc, *rowcnt, *rowptr, *colind;
  int nnz = A->colptr[n];
  anc = (int*)tamd_malloc(sizeof(int) * (size_t)(3 * (n + 1) + nnz + 1));
  if (!anc) return -1;
  rowcnt = anc + (n + 1);
  rowptr = rowcnt + (n + 1);
  colind = rowptr + (n + 1);
  /* row structure of the strict lower triangle: row i lists columns j < i */
  for (i = 0; i <= n; i++) rowcnt[i] = 0;
  for (j = 0; j < n; j++)
    for (ip = A->colptr[j]; ip < A->colptr[j + 1]; ip++)
      if (A->rowind[ip] > j) rowcnt[A->rowind[ip]]++;
  rowptr[0] = 0;
  for (i = 0; i < n; i++) rowptr[i + 1] = rowptr[i] + rowcnt[i];
  for (i = 0; i < n; i++) rowcnt[i] = rowptr[i];
  for (j = 0; j < n; j++)
    for (ip = A->colptr[j]; ip < A->colptr[j + 1]; ip++) {
      int r = A->rowind[ip];
      if (r > j) colind[rowcnt[r]++] = j;
    }
  for (i = 0; i < n; i++) {
    parent[i] = n;
    anc[i] = -1;
    for (ip = rowptr[i]; ip < rowptr[i + 1]; ip++) {
      int k = colind[ip];
      while (k != -1 && k < i) {
        int next = anc[k];
        anc[k] = i;
        if (next == -1) { parent[k] = i; break; }
        k = next;
      }
    }
  }
  tamd_free(anc);
  return 0;
}

int taucs_ccs_etree(taucs_ccs_matrix* A, int* parent, int* l_colcount, int* l_rowcount,
                    int* l_nnz)
{
  (void)l_colcount; (void)l_rowcount; (void)l_nnz;
  return tamd_etree(A, parent);
}

void tamd_symbolic_free(tamd_symbolic* S)
{
  int s;
  if (!S) return;
  if (S->sn_struct) {
    for (s = 0; s < S->n_sn; s++) tamd_free(S->sn_struct[s]);
  }
  tamd_free(S->sn_struct);
  tamd_free(S->sn_size);
  tamd_free(S->sn_up_size);
  tamd_free(S->first_child);
  tamd_free(S->next_child);
  memset(S, 0, sizeof(*S));
}

static void symbolic_stats(const tamd_symbolic* S, double* nnz, double* flops)
{
  int sn, i;
  *nnz = 0.0;
  *flops = 0.0;
  for (sn = 0; sn < S->n_sn; sn++) {
    int colnnz = S->sn_up_size[sn];
    for (i = 0; i < S->sn_size[sn]; i++, colnnz--) {
      *flops += 1.0 + (double)colnnz * (double)colnnz;
      *nnz += (double)colnnz;
    }
  }
}

/* ---- fundamental supernodes (postorder over the column etree) ---------- */

static int fundamental_supernodes(const taucs_ccs_matrix* A, const int* first_child,
                                  const int* next_child, tamd_symbolic* S, int* col2sn,
                                  int* map, int* rowbuf, int* stack)
{
  int n = A->n;
  int sp = 0;
  int* iter = stack + (n + 1); /* per-node child iterator */
  int j;
  for (j = 0; j <= n; j++) iter[j] = -2;
  stack[sp++] = n;
  while (sp > 0) {
    int v = stack[sp - 1];
    int c;
    if (iter[v] == -2) iter[v] = first_child[v];
    else iter[v] = next_child[iter[v]];
    c = iter[v];
    if (c != -1) { stack[sp++] = c; continue; }
    sp--;

    /* all children of v are done: decide v's supernode */
    {
      int inprev = 1, ip, nnz = 0;
      if (v == n) inprev = 0;
      else if (first_child[v] == -1) inprev = 0;
      else if (next_child[first_child[v]] != -1) inprev = 0;
      else {
        int fc = first_child[v];
        for (ip = A->colptr[v]; ip < A->colptr[v + 1]; ip++)
          if (map[A->rowind[ip]] != fc) { inprev = 0; break; }
      }

      if (inprev) {
        int csn = col2sn[first_child[v]];
        int* r = S->sn_struct[csn];
        int sz = S->sn_size[csn], up = S->sn_up_size[csn];
        col2sn[v] = csn;
        for (ip = sz; ip < up; ip++)
          if (r[ip] == v) break;
        if (ip >= up) return -1; /* structure not nested: cannot happen */
        r[ip] = r[sz];
        r[sz] = v;
        for (ip = sz; ip < up; ip++) map[r[ip]] = v;
        S->sn_size[csn]++;
        continue;
      }

      if (v < n) {
        rowbuf[nnz++] = v;
        map[v] = v;
        for (c = first_child[v]; c != -1; c = next_child[c]) {
          int csn = col2sn[c];
          for (ip = S->sn_size[csn]; ip < S->sn_up_size[csn]; ip++) {
            int i = S->sn_struct[csn][ip];
            if (i > v && map[i] != v) { map[i] = v; rowbuf[nnz++] = i; }
          }
        }
        for (ip = A->colptr[v]; ip < A->colptr[v + 1]; ip++) {
          int i = A->rowind[ip];
          if (map[i] != v) { map[i] = v; rowbuf[nnz++] = i; }
        }
      }

      for (c = first_child[v]; c != -1; c = next_child[c]) {
        int csn = col2sn[c];
        if (c == first_child[v]) {
          S->first_child[S->n_sn] = csn;
        } else {
          S->next_child[csn] = S->first_child[S->n_sn];
          S->first_child[S->n_sn] = csn;
        }
      }

      if (v < n) {
        int sn = S->n_sn;
        col2sn[v] = sn;
        S->sn_size[sn] = 1;
        S->sn_up_size[sn] = nnz;
        S->sn_struct[sn] = (int*)tamd_malloc(sizeof(int) * (size_t)nnz);
        if (!S->sn_struct[sn]) return -1;
        memcpy(S->sn_struct[sn], rowbuf, sizeof(int) * (size_t)nnz);
        S->n_sn++;
      }
    }
  }
  return 0;
}

/* ---- relaxed amalgamation ----------------------------------------------
 * Post-order over the supernodal tree rooted at `root`.  A parent absorbs
 * ALL its children when the merged supernode is small or the explicit zeros
 * it introduces are a small fraction of its nonzeros.  The running
 * (zeros, nonzeros) a subtree reports upward follows the reference exactly,
 * including that an unmerged supernode reports zero zeros. */
typedef struct { double zeros, nonzeros; } znz_t;

static int amalgamate(tamd_symbolic* S, int root, int* col2sn, int* map, int* rowbuf,
                      int* stack, znz_t* zret, double* zcount)
{
  int sp = 0;
  int* iter = stack + (S->n_sn + 1);
  int k;
  for (k = 0; k <= S->n_sn; k++) iter[k] = -2;
  stack[sp++] = root;
  while (sp > 0) {
    int sn = stack[sp - 1], c;
    if (iter[sn] == -2) iter[sn] = S->first_child[sn];
    else iter[sn] = S->next_child[iter[sn]];
    c = iter[sn];
    if (c != -1) { stack[sp++] = c; continue; }
    sp--;

    {
      int s_sz = S->sn_size[sn], s_up = S->sn_up_size[sn];
      znz_t own, merged;
      int nnz = 0, new_size, new_up, ip, i, nch = 0;
      own.zeros = 0.0;
      own.nonzeros = (double)((int64_t)(s_up - s_sz) * s_sz + ((int64_t)s_sz * (s_sz + 1)) / 2);
      zret[sn] = own;
      if (S->first_child[sn] == -1) continue;

      merged = own;
      for (c = S->first_child[sn]; c != -1; c = S->next_child[c]) {
        merged.nonzeros += zret[c].nonzeros;
        merged.zeros += zret[c].zeros;
      }

      /* merged row list: children's columns, own columns, then new rows */
      for (c = S->first_child[sn]; c != -1; c = S->next_child[c])
        for (ip = 0; ip < S->sn_size[c]; ip++) {
          i = S->sn_struct[c][ip];
          map[i] = sn;
          rowbuf[nnz++] = i;
        }
      for (ip = 0; ip < s_sz; ip++) {
        i = S->sn_struct[sn][ip];
        map[i] = sn;
        rowbuf[nnz++] = i;
      }
      new_size = nnz;
      for (c = S->first_child[sn]; c != -1; c = S->next_child[c])
        for (ip = S->sn_size[c]; ip < S->sn_up_size[c]; ip++) {
          i = S->sn_struct[c][ip];
          if (map[i] != sn) { map[i] = sn; rowbuf[nnz++] = i; }
        }
      for (ip = s_sz; ip < s_up; ip++) {
        i = S->sn_struct[sn][ip];
        if (map[i] != sn) { map[i] = sn; rowbuf[nnz++] = i; }
      }
      new_up = nnz;

      /* explicit zeros the merge would introduce */
      for (ip = 0; ip < new_size; ip++) zcount[rowbuf[ip]] = (double)(ip + 1);
      for (ip = new_size; ip < new_up; ip++) zcount[rowbuf[ip]] = (double)new_size;
      for (c = S->first_child[sn]; c != -1; c = S->next_child[c]) {
        for (ip = 0; ip < S->sn_size[c]; ip++) zcount[S->sn_struct[c][ip]] -= (double)(ip + 1);
        for (ip = S->sn_size[c]; ip < S->sn_up_size[c]; ip++)
          zcount[S->sn_struct[c][ip]] -= (double)S->sn_size[c];
      }
      for (ip = 0; ip < s_sz; ip++) zcount[S->sn_struct[sn][ip]] -= (double)(ip + 1);
      for (ip = s_sz; ip < s_up; ip++) zcount[S->sn_struct[sn][ip]] -= (double)s_sz;
      for (ip = 0; ip < new_up; ip++) merged.zeros += zcount[rowbuf[ip]];

      if (!((new_size < 16)
            || ((s_sz < 50) && (merged.zeros < 0.5 * merged.nonzeros))
            || ((s_sz < 250) && (merged.zeros < 0.25 * merged.nonzeros))
            || ((s_sz < 500) && (merged.zeros < 0.10 * merged.nonzeros))
            || (merged.zeros < 0.05 * merged.nonzeros)))
        continue; /* keep separate; report own counts */

      /* merge: sn takes the union, children become dead slots */
      {
        int* nr = (int*)tamd_realloc(S->sn_struct[sn], sizeof(int) * (size_t)new_up);
        if (!nr) return -1;
        S->sn_struct[sn] = nr;
        memcpy(nr, rowbuf, sizeof(int) * (size_t)new_up);
        S->sn_size[sn] = new_size;
        S->sn_up_size[sn] = new_up;
      }
      for (c = S->first_child[sn]; c != -1; c = S->next_child[c]) {
        int g;
        for (ip = 0; ip < S->sn_size[c]; ip++) col2sn[S->sn_struct[c][ip]] = sn;
        for (g = S->first_child[c]; g != -1; g = S->next_child[g]) rowbuf[nch++] = g;
      }
      for (c = S->first_child[sn]; c != -1; c = S->next_child[c]) {
        tamd_free(S->sn_struct[c]);
        S->sn_struct[c] = NULL;
        S->sn_size[c] = 0;
        S->sn_up_size[c] = 0;
      }
      S->first_child[sn] = -1;
      for (i = 0; i < nch; i++) {
        S->next_child[rowbuf[i]] = S->first_child[sn];
        S->first_child[sn] = rowbuf[i];
      }
      zret[sn] = merged;
    }
  }
  return 0;
}

int tamd_symbolic_analyze(const taucs_ccs_matrix* A, int max_depth, tamd_symbolic* S)
{
  int n = A->n, j, rc = -1;
  int *parent = NULL, *first_child = NULL, *next_child = NULL, *col2sn = NULL, *map = NULL;
  int *rowbuf = NULL, *stack = NULL, *depth = NULL;
  znz_t* zret = NULL;
  double* zcount = NULL;

  memset(S, 0, sizeof(*S));
  S->n = n;
  S->sn_struct   = (int**)tamd_calloc((size_t)n + 1, sizeof(int*));
  S->sn_size     = (int*)tamd_malloc(sizeof(int) * (size_t)(n + 1));
  S->sn_up_size  = (int*)tamd_malloc(sizeof(int) * (size_t)(n + 1));
  S->first_child = (int*)tamd_malloc(sizeof(int) * (size_t)(n + 1));
  S->next_child  = (int*)tamd_malloc(sizeof(int) * (size_t)(n + 1));
  parent      = (int*)tamd_malloc(sizeof(int) * (size_t)(n + 1));
  first_child = (int*)tamd_malloc(sizeof(int) * (size_t)(n + 1));
  next_child  = (int*)tamd_malloc(sizeof(int) * (size_t)(n + 1));
  col2sn      = (int*)tamd_malloc(sizeof(int) * (size_t)(n + 1));
  map         = (int*)tamd_malloc(sizeof(int) * (size_t)(n + 1));
  rowbuf      = (int*)tamd_malloc(sizeof(int) * (size_t)(n + 1));
  stack       = (int*)tamd_malloc(sizeof(int) * (size_t)(2 * (n + 2)));
  if (!S->sn_struct || !S->sn_size || !S->sn_up_size || !S->first_child || !S->next_child
      || !parent || !first_child || !next_child || !col2sn || !map || !rowbuf || !stack)
    goto done;

  if (tamd_etree(A, parent) != 0) goto done;

  for (j = 0; j <= n; j++) first_child[j] = -1;
  for (j = n - 1; j >= 0; j--) {
    next_child[j] = first_child[parent[j]];
    first_child[parent[j]] = j;
  }

  /* etree depth: the virtual root is depth 0 (src/taucs_sn_llt.c:3781-3815) */
  depth = col2sn; /* reuse as scratch */
  S->etree_depth = n > 0 ? 1 : 0;
  for (j = n - 1; j >= 0; j--) {
    depth[j] = parent[j] == n ? 1 : depth[parent[j]] + 1;
    if (depth[j] > S->etree_depth) S->etree_depth = depth[j];
  }
  if (max_depth && S->etree_depth > max_depth) {
    taucs_printf("taucs_ccs_symbolic_elimination: etree depth %d, maximum allowed is %d\n",
                 S->etree_depth, max_depth);
    goto done;
  }

  S->n_sn = 0;
  for (j = 0; j < n; j++) map[j] = -1;
  for (j = 0; j <= n; j++) S->first_child[j] = S->next_child[j] = -1;
  if (fundamental_supernodes(A, first_child, next_child, S, col2sn, map, rowbuf, stack) != 0)
    goto done;
  symbolic_stats(S, &S->fund_nnz, &S->fund_flops);

  for (j = 0; j < n; j++) map[j] = -1;
  zret = (znz_t*)tamd_malloc(sizeof(znz_t) * (size_t)(S->n_sn + 1));
  zcount = (double*)tamd_malloc(sizeof(double) * (size_t)(n + 1));
  if (!zret || !zcount) goto done;
  if (S->n_sn > 0
      && amalgamate(S, S->n_sn - 1, col2sn, map, rowbuf, stack, zret, zcount) != 0)
    goto done;
  symbolic_stats(S, &S->relax_nnz, &S->relax_flops);
  taucs_printf("\t\tSymbolic Analysis of LL^T: %.2e nonzeros, %.2e flops\n", S->fund_nnz,
               S->fund_flops);
  taucs_printf("\t\tRelaxed  Analysis of LL^T: %.2e nonzeros, %.2e flops\n", S->relax_nnz,
               S->relax_flops);
  rc = 0;

done:
  tamd_free(parent);
  tamd_free(first_child);
  tamd_free(next_child);
  tamd_free(col2sn);
  tamd_free(map);
  tamd_free(rowbuf);
  tamd_free(stack);
  tamd_free(zret);
  tamd_free(zcount);
  if (rc != 0) tamd_symbolic_free(S);
  return rc;
}

/* ---- level schedule and HBM layout ---------------------------------------
 * Levels are heights: leaves are level 0, a parent is one above its highest
 * child.  All fronts of one level are independent and run in the same
 * launches.  Two children of one parent at the same level would race in the
 * extend-add, so each child gets a rank among its same-level siblings and
 * ranks run in separate launches (a fixed order: results are deterministic,
 * with no floating-point atomics). */
uint64_t tamd_pattern_hash(const taucs_ccs_matrix* A)
{
  uint64_t h = 1469598103934665603ull;
  int j, ip;
  h = (h ^ (uint64_t)A->n) * 1099511628211ull;
  for (j = 0; j <= A->n; j++) h = (h ^ (uint64_t)(uint32_t)A->colptr[j]) * 1099511628211ull;
  for (ip = 0; ip < A->colptr[A->n]; ip++)
    h = (h ^ (uint64_t)(uint32_t)A->rowind[ip]) * 1099511628211ull;
  return h;
}

void tamd_plan_free(tamd_plan* P)
{
  if (!P) return;
  tamd_free(P->sn_parent);
  tamd_free(P->sn_level);
  tamd_free(P->order);
  tamd_free(P->level_ptr);
  tamd_free(P->off_L);
  tamd_free(P->off_U);
  tamd_free(P->u_start);
  if (P->drow) {
    int sn;
    for (sn = 0; sn < P->n_sn; sn++) tamd_free(P->drow[sn]);
  }
  tamd_free(P->drow);
  tamd_free(P->rel_ptr);
  tamd_free(P->rel);
  tamd_free(P->sib_rank);
  tamd_free(P->a_dest);
  tamd_free(P->a_colptr);
  memset(P, 0, sizeof(*P));
}

int tamd_plan_amap(const tamd_symbolic* S, const taucs_ccs_matrix* A, tamd_plan* P)
{
  int n = S->n, k, sn, jj, ip;
  int* pos = (int*)tamd_malloc(sizeof(int) * (size_t)(n + 1));
  tamd_free(P->a_dest);
  tamd_free(P->a_colptr);
  P->a_nnz = A->colptr[n];
  P->a_dest = (int64_t*)tamd_malloc(sizeof(int64_t) * (size_t)(P->a_nnz + 1));
  P->a_colptr = (int*)tamd_malloc(sizeof(int) * (size_t)(n + 1));
  if (!pos || !P->a_dest || !P->a_colptr) { tamd_free(pos); return -1; }
  memcpy(P->a_colptr, A->colptr, sizeof(int) * (size_t)(n + 1));
  for (ip = 0; ip < P->a_nnz; ip++) P->a_dest[ip] = -1;
  for (k = 0; k < P->n_live; k++) {
    int m;
    sn = P->order[k];
    m = S->sn_up_size[sn];
    for (ip = 0; ip < m; ip++) pos[P->drow[sn][ip]] = ip;
    for (jj = 0; jj < S->sn_size[sn]; jj++) {
      int col = S->sn_struct[sn][jj];
      for (ip = A->colptr[col]; ip < A->colptr[col + 1]; ip++) {
        int p = pos[A->rowind[ip]];
        if (p < jj) { tamd_free(pos); return -1; } /* would land above the diagonal */
        P->a_dest[ip] = P->off_L[sn] + (int64_t)jj * m + p;
      }
    }
  }
  for (ip = 0; ip < P->a_nnz; ip++)
    if (P->a_dest[ip] < 0) { tamd_free(pos); return -1; }
  P->a_pattern_hash = tamd_pattern_hash(A);
  tamd_free(pos);
  return 0;
}

/* Sort v[0..len) by key[v[i]] (keys distinct): insertion sort for short
   lists, else a heap sort; deterministic. */
static void sift_down(int* v, int start, int end, const int* key)
{
  int root = start;
  while (2 * root + 1 < end) {
    int child = 2 * root + 1, sw = root;
    if (key[v[sw]] < key[v[child]]) sw = child;
    if (child + 1 < end && key[v[sw]] < key[v[child + 1]]) sw = child + 1;
    if (sw == root) return;
    { int t = v[root]; v[root] = v[sw]; v[sw] = t; }
    root = sw;
  }
}

static void sort_by_key(int* v, int len, const int* key)
{
  int i, j;
  if (len < 32) {
    for (i = 1; i < len; i++) {
      int x = v[i];
      for (j = i; j > 0 && key[v[j - 1]] > key[x]; j--) v[j] = v[j - 1];
      v[j] = x;
    }
    return;
  }
  for (i = len / 2 - 1; i >= 0; i--) sift_down(v, i, len, key);
  for (i = len - 1; i > 0; i--) {
    int t = v[0];
    v[0] = v[i];
    v[i] = t;
    sift_down(v, 0, i, key);
  }
}

/* Offsets of the update blocks U_sn (u x u, u = m - s) in one arena whose
   space is reused once a block is dead.  Children extend-add into U_sn at the
   end of their own levels, the front's SYRK updates it during level(sn), and
   its own extend-add into the parent reads it at the end of level(sn); so U_sn
   lives from u_start(sn) (the lowest child level, or level(sn) without
   children) through level(sn), and is zeroed when its life starts.  Blocks
   dead after level l-1 are handed to blocks starting at level l (first fit
   over a coalescing free list; deterministic).  The reference keeps each
   update matrix on the heap from its front's return until the parent has
   assembled it (src/taucs_sn_llt.c:1925-1940); here the parent's block exists
   already while its children assemble into it. */
typedef struct { int64_t off, len; } uspan;

static int uspan_free(uspan** fl, int* nfree, int* cap, uspan b)
{
  int i, j;
  if (*nfree == *cap) {
    uspan* nf = (uspan*)tamd_realloc(*fl, sizeof(uspan) * (size_t)(*cap) * 2);
    if (!nf) return -1;
    *fl = nf;
    *cap *= 2;
  }
  for (i = *nfree; i > 0 && (*fl)[i - 1].off > b.off; i--) (*fl)[i] = (*fl)[i - 1];
  (*fl)[i] = b;
  (*nfree)++;
  if (i + 1 < *nfree && (*fl)[i].off + (*fl)[i].len == (*fl)[i + 1].off) {
    (*fl)[i].len += (*fl)[i + 1].len;
    for (j = i + 1; j + 1 < *nfree; j++) (*fl)[j] = (*fl)[j + 1];
    (*nfree)--;
  }
  if (i > 0 && (*fl)[i - 1].off + (*fl)[i - 1].len == (*fl)[i].off) {
    (*fl)[i - 1].len += (*fl)[i].len;
    for (j = i; j + 1 < *nfree; j++) (*fl)[j] = (*fl)[j + 1];
    (*nfree)--;
  }
  return 0;
}

static int plan_update_blocks(const tamd_symbolic* S, tamd_plan* P, int64_t* total)
{
  int nsn = S->n_sn, nl = P->n_levels, k, lev, e;
  int64_t top = 0, peak = 0;   /* current end of the used space, and its maximum */
  int nfree = 0, cap = 64;
  uspan* fl = (uspan*)tamd_malloc(sizeof(uspan) * (size_t)cap);
  int* byst = (int*)tamd_malloc(sizeof(int) * (size_t)(P->n_live + 1));  /* by start level */
  int* sptr = (int*)tamd_calloc((size_t)nl + 2, sizeof(int));
  int* fill = (int*)tamd_malloc(sizeof(int) * (size_t)(nl + 1));
  P->u_start = (int*)tamd_malloc(sizeof(int) * (size_t)(nsn + 1));
  if (!fl || !byst || !sptr || !fill || !P->u_start) goto fail;
  for (k = 0; k <= nsn; k++) P->u_start[k] = -1;
  for (k = 0; k < P->n_live; k++) {
    int sn = P->order[k], c, st = P->sn_level[sn];
    for (c = S->first_child[sn]; c != -1; c = S->next_child[c])
      if (P->sn_level[c] < st) st = P->sn_level[c];
    P->u_start[sn] = st;
    sptr[st + 1]++;
  }
  for (lev = 0; lev <= nl; lev++) sptr[lev + 1] += sptr[lev];
  memcpy(fill, sptr, sizeof(int) * (size_t)(nl + 1));
  for (k = 0; k < P->n_live; k++) {   /* level order inside each start level */
    int sn = P->order[k];
    byst[fill[P->u_start[sn]]++] = sn;
  }
  for (lev = 0; lev < nl; lev++) {
    /* blocks whose own level was lev-1 are dead now */
    if (lev > 0)
      for (k = P->level_ptr[lev - 1]; k < P->level_ptr[lev]; k++) {
        int sn = P->order[k];
        int64_t u = (int64_t)(S->sn_up_size[sn] - S->sn_size[sn]);
        uspan b;
        if (u == 0) continue;
        b.off = P->off_U[sn];
        b.len = u * u;
        if (uspan_free(&fl, &nfree, &cap, b) != 0) goto fail;
      }
    /* a free block that ends at the top shrinks the used space instead */
    if (nfree > 0 && fl[nfree - 1].off + fl[nfree - 1].len == top) {
      top = fl[nfree - 1].off;
      nfree--;
    }
    for (e = sptr[lev]; e < sptr[lev + 1]; e++) {
      int sn = byst[e], i, j;
      int64_t u = (int64_t)(S->sn_up_size[sn] - S->sn_size[sn]), need = u * u;
      if (need == 0) { P->off_U[sn] = 0; continue; }
      for (i = 0; i < nfree && fl[i].len < need; i++) {}
      if (i < nfree) {
        P->off_U[sn] = fl[i].off;
        fl[i].off += need;
        fl[i].len -= need;
        if (fl[i].len == 0) {
          for (j = i; j + 1 < nfree; j++) fl[j] = fl[j + 1];
          nfree--;
        }
      } else {
        P->off_U[sn] = top;
        top += need;
        if (top > peak) peak = top;
      }
    }
  }
  tamd_free(fl);
  tamd_free(byst);
  tamd_free(sptr);
  tamd_free(fill);
  *total = peak;
  return 0;
fail:
  tamd_free(fl);
  tamd_free(byst);
  tamd_free(sptr);
  tamd_free(fill);
  return -1;
}

int tamd_plan_build(const tamd_symbolic* S, const taucs_ccs_matrix* A, tamd_plan* P)
{
  int nsn = S->n_sn, k, sn, c, nlive = 0, lev;
  int *stack = NULL, *cnt = NULL, *pos = NULL;
  int64_t offl = 0, offu = 0, nrel = 0;

  memset(P, 0, sizeof(*P));
  P->n_sn = nsn;
  P->sn_parent = (int*)tamd_malloc(sizeof(int) * (size_t)(nsn + 1));
  P->sn_level  = (int*)tamd_malloc(sizeof(int) * (size_t)(nsn + 1));
  P->order     = (int*)tamd_malloc(sizeof(int) * (size_t)(nsn + 1));
  P->off_L     = (int64_t*)tamd_malloc(sizeof(int64_t) * (size_t)(nsn + 1));
  P->off_U     = (int64_t*)tamd_malloc(sizeof(int64_t) * (size_t)(nsn + 1));
  P->rel_ptr   = (int*)tamd_malloc(sizeof(int) * (size_t)(nsn + 2));
  P->sib_rank  = (int*)tamd_malloc(sizeof(int) * (size_t)(nsn + 1));
  stack        = (int*)tamd_malloc(sizeof(int) * (size_t)(nsn + 2));
  pos          = (int*)tamd_malloc(sizeof(int) * (size_t)(S->n + 1));
  if (!P->sn_parent || !P->sn_level || !P->order || !P->off_L || !P->off_U || !P->rel_ptr
      || !P->sib_rank || !stack || !pos)
    goto fail;

  for (sn = 0; sn <= nsn; sn++) { P->sn_parent[sn] = -1; P->sn_level[sn] = -1; P->sib_rank[sn] = 0; }

  /* live slots in preorder from the virtual root; reverse gives children
     before parents */
  {
    int sp = 0;
    for (c = S->first_child[nsn]; c != -1; c = S->next_child[c]) { P->sn_parent[c] = nsn; stack[sp++] = c; }
    while (sp > 0) {
      sn = stack[--sp];
      P->order[nlive++] = sn;
      for (c = S->first_child[sn]; c != -1; c = S->next_child[c]) {
        P->sn_parent[c] = sn;
        stack[sp++] = c;
      }
    }
  }
  P->n_live = nlive;
  P->n_levels = 0;
  for (k = nlive - 1; k >= 0; k--) {
    int l = 0;
    sn = P->order[k];
    for (c = S->first_child[sn]; c != -1; c = S->next_child[c])
      if (P->sn_level[c] + 1 > l) l = P->sn_level[c] + 1;
    P->sn_level[sn] = l;
    if (l + 1 > P->n_levels) P->n_levels = l + 1;
  }

  /* group by level (stable: keeps the reverse-preorder within a level) */
  P->level_ptr = (int*)tamd_calloc((size_t)P->n_levels + 2, sizeof(int));
  cnt = (int*)tamd_calloc((size_t)P->n_levels + 2, sizeof(int));
  if (!P->level_ptr || !cnt) goto fail;
  for (k = 0; k < nlive; k++) P->level_ptr[P->sn_level[P->order[k]] + 1]++;
  for (lev = 0; lev < P->n_levels; lev++) P->level_ptr[lev + 1] += P->level_ptr[lev];
  for (lev = 0; lev < P->n_levels; lev++) cnt[lev] = P->level_ptr[lev];
  for (k = nlive - 1; k >= 0; k--) {
    sn = P->order[k];
    stack[cnt[P->sn_level[sn]]++] = sn;
  }
  memcpy(P->order, stack, sizeof(int) * (size_t)nlive);

  /* device row order: parents before children (decreasing level), each
     front's update rows sorted by their position in the parent's order */
  P->drow = (int**)tamd_calloc((size_t)nsn + 1, sizeof(int*));
  if (!P->drow) goto fail;
  for (k = nlive - 1; k >= 0; k--) {
    int m, s, ip, p;
    sn = P->order[k];
    s = S->sn_size[sn];
    m = S->sn_up_size[sn];
    p = P->sn_parent[sn];
    P->drow[sn] = (int*)tamd_malloc(sizeof(int) * (size_t)(m > 0 ? m : 1));
    if (!P->drow[sn]) goto fail;
    memcpy(P->drow[sn], S->sn_struct[sn], sizeof(int) * (size_t)m);
    if (p < nsn && m > s) {
      for (ip = 0; ip < S->sn_up_size[p]; ip++) pos[P->drow[p][ip]] = ip;
      sort_by_key(P->drow[sn] + s, m - s, pos);
    }
  }

  /* panels, update blocks, relative indices */
  P->max_front = 0;
  for (sn = 0; sn <= nsn; sn++) { P->off_L[sn] = -1; P->off_U[sn] = -1; }
  for (k = 0; k < nlive; k++) {
    int s, m, u;
    sn = P->order[k];
    s = S->sn_size[sn];
    m = S->sn_up_size[sn];
    u = m - s;
    P->off_L[sn] = offl;
    offl += (int64_t)m * s;
    if (m > P->max_front) P->max_front = m;
    if (P->sn_parent[sn] < nsn) nrel += u;
  }
  P->L_elems = offl;
  if (plan_update_blocks(S, P, &offu) != 0) goto fail;
  P->U_elems = offu;
  if (nrel > 0x7fffffff) goto fail;
  P->rel = (int*)tamd_malloc(sizeof(int) * (size_t)(nrel + 1));
  if (!P->rel) goto fail;
  {
    int r = 0;
    for (sn = 0; sn <= nsn; sn++) P->rel_ptr[sn] = -1;
    for (k = 0; k < nlive; k++) {
      int p, ip;
      sn = P->order[k];
      p = P->sn_parent[sn];
      P->rel_ptr[sn] = r;
      if (p >= nsn) continue;
      for (ip = 0; ip < S->sn_up_size[p]; ip++) pos[P->drow[p][ip]] = ip;
      for (ip = S->sn_size[sn]; ip < S->sn_up_size[sn]; ip++)
        P->rel[r++] = pos[P->drow[sn][ip]];
    }
    P->rel_ptr[nsn] = r;
  }

  /* sibling ranks among children of one parent that share a level */
  for (k = 0; k < nlive; k++) {
    int p = P->order[k];
    for (c = S->first_child[p]; c != -1; c = S->next_child[c]) {
      int c2, rk = 0;
      for (c2 = S->first_child[p]; c2 != c; c2 = S->next_child[c2])
        if (P->sn_level[c2] == P->sn_level[c]) rk++;
      P->sib_rank[c] = rk;
    }
  }

  if (tamd_plan_amap(S, A, P) != 0) goto fail;
  tamd_free(stack);
  tamd_free(cnt);
  tamd_free(pos);
  return 0;

fail:
  tamd_free(stack);
  tamd_free(cnt);
  tamd_free(pos);
  tamd_plan_free(P);
  return -1;
}
