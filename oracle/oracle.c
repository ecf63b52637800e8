/*
 * ORACLE -- test infrastructure only.  A plain-C restatement of the TAUCS
 * 2.2.1 symbolic path, used by tests/, __graft_entry__.smoke() and bench.py's
 * CPU baseline as the checker of libtaucs_amd.so.  Nothing in the product
 * links or calls this file.
 *
 * It follows the reference's own (recursive) formulation function by
 * function, so that the product's iterative implementation is checked against
 * an independent restatement:
 *
 *   oracle_mesh2d / oracle_mesh3d     src/taucs_ccs_generators.c:89-195, 197-261
 *   oracle_genmmd                     src/taucs_ccs_order.c:262-442 (adjacency
 *                                     build, delta=1, maxint=2^31-1) calling
 *                                     GENMMD itself: external/src/genmmd.c,
 *                                     compiled unmodified into _ref/ by the
 *                                     Makefile and loaded with dlopen
 *   oracle_permute                    src/taucs_ccs_ops.c:174-257
 *   oracle_etree                      src/taucs_sn_llt.c:3243-3340 (union-find)
 *   oracle_symbolic                   src/taucs_sn_llt.c:3663-4001 with
 *                                     recursive_postorder :1269-1287,
 *                                     recursive_symbolic_elimination :1377-1517,
 *                                     recursive_amalgamate_supernodes :1527-1786
 *
 * Parity pins (see tests/test_oracle.py): GENMMD is the reference binary
 * itself; the generators are pinned by closed forms (row sums, counts, entry
 * order); the etree and column counts by brute-force symbolic Cholesky on
 * small matrices; the fundamental supernodes by their definition (chains with
 * nested structure); amalgamation by re-deriving every merge decision from
 * the published thresholds on the final partition.  The reference's own
 * tests hold no golden vectors for this path.
 *
 * The recursion depth equals the elimination-tree height, so the symbolic
 * entry point runs on a thread with a 1 GiB stack.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---- generators ----------------------------------------------------------- */

/* 5-point, lower CSC; per column (x,y+1), (x+1,y), diag; Neumann diagonal =
   degree, +1 at (0,0).  Arrays are malloc'ed; caller frees with free(). */
int oracle_mesh2d(int n, int** colptr, int** rowind, double** vals)
{
  int N = n * n, x, y, ip = 0;
  int* cp = malloc(sizeof(int) * (N + 1));
  int* ri = malloc(sizeof(int) * 3 * N);
  double* v = malloc(sizeof(double) * 3 * N);
  for (y = 0; y < n; y++)
    for (x = 0; x < n; x++) {
      int j = x + y * n;
      double d = 0;
      cp[j] = ip;
      if (y < n - 1) { ri[ip] = x + (y + 1) * n; v[ip++] = -1.0; }
      if (x < n - 1) { ri[ip] = (x + 1) + y * n; v[ip++] = -1.0; }
      if (x > 0) d += 1.0;
      if (y > 0) d += 1.0;
      if (x < n - 1) d += 1.0;
      if (y < n - 1) d += 1.0;
      if (x == 0 && y == 0) d += 1.0;
      ri[ip] = j;
      v[ip++] = d;
    }
  cp[N] = ip;
  *colptr = cp; *rowind = ri; *vals = v;
  return ip;
}

/* 7-point; per column x+1, y+1, z+1, diag; diagonal = degree, +1 at origin */
int oracle_mesh3d(int X, int Y, int Z, int** colptr, int** rowind, double** vals)
{
  int N = X * Y * Z, x, y, z, ip = 0;
  int* cp = malloc(sizeof(int) * (N + 1));
  int* ri = malloc(sizeof(int) * 4 * N);
  double* v = malloc(sizeof(double) * 4 * N);
  for (z = 0; z < Z; z++)
    for (y = 0; y < Y; y++)
      for (x = 0; x < X; x++) {
        int j = z * X * Y + y * X + x;
        double d = 0;
        cp[j] = ip;
        if (x < X - 1) { ri[ip] = j + 1; v[ip++] = -1.0; }
        if (y < Y - 1) { ri[ip] = j + X; v[ip++] = -1.0; }
        if (z < Z - 1) { ri[ip] = j + X * Y; v[ip++] = -1.0; }
        if (x < X - 1) d += 1.0;
        if (y < Y - 1) d += 1.0;
        if (z < Z - 1) d += 1.0;
        if (x > 0) d += 1.0;
        if (y > 0) d += 1.0;
        if (z > 0) d += 1.0;
        if (x == 0 && y == 0 && z == 0) d += 1.0;
        ri[ip] = j;
        v[ip++] = d;
      }
  cp[N] = ip;
  *colptr = cp; *rowind = ri; *vals = v;
  return ip;
}

/* ---- GENMMD through the reference's own compiled routine ------------------ */

/* genmmd.c is the f2c translation with `integer` = long; the routine is
   called with long arrays so no header of the reference needs changing. */
typedef int (*genmmd_fn)(long*, long*, long*, long*, long*, long*, long*, long*, long*, long*,
                         long*, long*);
static genmmd_fn ref_genmmd = NULL;

int oracle_load_ref(const char* so_path)
{
  void* h = dlopen(so_path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return -1;
  ref_genmmd = (genmmd_fn)dlsym(h, "genmmd_");
  return ref_genmmd ? 0 : -1;
}

/* src/taucs_ccs_order.c:262-442: 1-based adjacency, both directions of each
   off-diagonal entry appended in column-scan order. */
int oracle_genmmd(int n, const int* colptr, const int* rowind, int* perm, int* invperm)
{
  long nn = n, delta = 1, maxint = 2147483647, nofsub = 0;
  long nnz = colptr[n];
  long *xadj, *adj, *invp, *prm, *dhead, *qsize, *llist, *marker, *len, *next;
  long i, j, ip;
  if (!ref_genmmd) return -1;
  xadj = calloc(n + 1, sizeof(long));
  adj = calloc(2 * nnz - n + 1, sizeof(long));
  invp = calloc(n + 1, sizeof(long));
  prm = calloc(n + 1, sizeof(long));
  dhead = calloc(n + 1, sizeof(long));
  qsize = calloc(n + 1, sizeof(long));
  llist = calloc(n + 1, sizeof(long));
  marker = calloc(n + 1, sizeof(long));
  len = dhead;
  next = qsize;
  for (j = 0; j < n; j++)
    for (ip = colptr[j]; ip < colptr[j + 1]; ip++) {
      i = rowind[ip];
      if (i != j) { len[i]++; len[j]++; }
    }
  xadj[0] = 1;
  for (i = 1; i <= n; i++) xadj[i] = xadj[i - 1] + len[i - 1];
  for (i = 0; i < n; i++) next[i] = xadj[i] - 1;
  for (j = 0; j < n; j++)
    for (ip = colptr[j]; ip < colptr[j + 1]; ip++) {
      i = rowind[ip];
      if (i != j) {
        adj[next[i]++] = j + 1;
        adj[next[j]++] = i + 1;
      }
    }
  ref_genmmd(&nn, xadj, adj, invp, prm, &delta, dhead, qsize, llist, marker, &maxint, &nofsub);
  for (i = 0; i < n; i++) perm[i] = (int)prm[i] - 1;
  for (i = 0; i < n; i++) invperm[perm[i]] = (int)i;
  free(xadj); free(adj); free(invp); free(prm); free(dhead); free(qsize); free(llist); free(marker);
  return 0;
}

/* ---- symmetric permutation ------------------------------------------------ */

void oracle_permute(int n, const int* colptr, const int* rowind, const double* vals,
                    const int* invperm, int* ocolptr, int* orowind, double* ovals)
{
  int j, ip;
  int* len = calloc(n + 1, sizeof(int));
  for (j = 0; j < n; j++)
    for (ip = colptr[j]; ip < colptr[j + 1]; ip++) {
      int I = invperm[rowind[ip]], J = invperm[j];
      if (I < J) { int t = I; I = J; J = t; }
      len[J]++;
    }
  ocolptr[0] = 0;
  for (j = 1; j <= n; j++) ocolptr[j] = ocolptr[j - 1] + len[j - 1];
  for (j = 0; j < n; j++) len[j] = ocolptr[j];
  for (j = 0; j < n; j++)
    for (ip = colptr[j]; ip < colptr[j + 1]; ip++) {
      int I = invperm[rowind[ip]], J = invperm[j];
      if (I < J) { int t = I; I = J; J = t; }
      orowind[len[J]] = I;
      ovals[len[J]] = vals[ip];
      len[J]++;
    }
  free(len);
}

/* ---- elimination tree (union-find over the row structure) ----------------- */

static int uf_find(int* uf, int i)
{
  if (uf[i] != i) uf[i] = uf_find(uf, uf[i]);
  return uf[i];
}
static int uf_union(int* uf, int s, int t)
{
  uf[uf_find(uf, s)] = uf_find(uf, t);
  return uf_find(uf, t);
}

void oracle_etree(int n, const int* colptr, const int* rowind, int* parent)
{
  int nnz = colptr[n], i, j, ip, kp;
  int* uf = malloc(sizeof(int) * (n + 1));
  int* rowcount = calloc(n + 1, sizeof(int));
  int* rowptr = malloc(sizeof(int) * (n + 1));
  int* colind = malloc(sizeof(int) * (nnz + 1));
  int* realroot;
  for (j = 0; j < n; j++)
    for (ip = colptr[j]; ip < colptr[j + 1]; ip++)
      if (j < rowind[ip]) rowcount[rowind[ip]]++;
  ip = 0;
  for (i = 0; i <= n; i++) {
    int nx = ip + rowcount[i];
    rowcount[i] = ip;
    rowptr[i] = ip;
    ip = nx;
  }
  for (j = 0; j < n; j++)
    for (ip = colptr[j]; ip < colptr[j + 1]; ip++) {
      i = rowind[ip];
      if (i == j) continue;
      colind[rowcount[i]++] = j;
    }
  realroot = rowcount;
  for (i = 0; i < n; i++) {
    int vroot = i;
    uf[i] = i;
    realroot[i] = i;
    parent[i] = n;
    for (kp = rowptr[i]; kp < rowptr[i + 1]; kp++) {
      int u = uf_find(uf, colind[kp]);
      int t = realroot[u];
      if (parent[t] == n && t != i) {
        parent[t] = i;
        vroot = uf_union(uf, vroot, u);
        realroot[vroot] = i;
      }
    }
  }
  free(uf); free(rowcount); free(rowptr); free(colind);
}

/* ---- symbolic elimination ------------------------------------------------- */

typedef struct { double zeros, nonzeros; } znz;

typedef struct {
  int n, n_sn, depth;
  int *sn_size, *sn_up_size, *first_child, *next_child;
  int** sn_struct;
  double fund_nnz, fund_flops, relax_nnz, relax_flops;
  int failed;
  /* merge log: one record per supernode with children that amalgamation
     visits, in visit order: sn, its own (fundamental) size, the candidate
     merged size and update size, the decision, and the candidate columns */
  int nlog, caplog, ncols, capcols;
  int* log_i;     /* 6 per record: sn, own_size, new_size, new_up, merged, cols_off */
  double* log_d;  /* 2 per record: zeros, nonzeros */
  int* log_cols;
} oracle_sym;

static void log_decision(oracle_sym* L, int sn, int new_size, int new_up, znz m, int merged,
                         const int* cols);

typedef struct {
  const int *colptr, *rowind;
  int n;
  int *first_child, *next_child, *col2sn, *map, *rowind_buf;
  oracle_sym* L;
} sym_ctx;

static int rec_symbolic(sym_ctx* c, int j)
{
  oracle_sym* L = c->L;
  int ch, ip, i, csn, in_prev, nnz = 0;
  for (ch = c->first_child[j]; ch != -1; ch = c->next_child[ch])
    if (rec_symbolic(c, ch) == -1) return -1;

  in_prev = 1;
  if (j == c->n) in_prev = 0;
  else if (c->first_child[j] == -1) in_prev = 0;
  else if (c->next_child[c->first_child[j]] != -1) in_prev = 0;
  else {
    ch = c->first_child[j];
    for (ip = c->colptr[j]; ip < c->colptr[j + 1]; ip++)
      in_prev = in_prev && (c->map[c->rowind[ip]] == ch);
  }

  if (in_prev) {
    ch = c->first_child[j];
    csn = c->col2sn[ch];
    c->col2sn[j] = csn;
    for (ip = L->sn_size[csn]; ip < L->sn_up_size[csn]; ip++)
      if (L->sn_struct[csn][ip] == j) break;
    L->sn_struct[csn][ip] = L->sn_struct[csn][L->sn_size[csn]];
    L->sn_struct[csn][L->sn_size[csn]] = j;
    for (ip = L->sn_size[csn]; ip < L->sn_up_size[csn]; ip++) c->map[L->sn_struct[csn][ip]] = j;
    L->sn_size[csn]++;
    return 0;
  }

  if (j < c->n) {
    nnz = 1;
    c->rowind_buf[0] = j;
    c->map[j] = j;
    for (ch = c->first_child[j]; ch != -1; ch = c->next_child[ch]) {
      csn = c->col2sn[ch];
      for (ip = L->sn_size[csn]; ip < L->sn_up_size[csn]; ip++) {
        i = L->sn_struct[csn][ip];
        if (i > j && c->map[i] != j) { c->map[i] = j; c->rowind_buf[nnz++] = i; }
      }
    }
    for (ip = c->colptr[j]; ip < c->colptr[j + 1]; ip++) {
      i = c->rowind[ip];
      if (c->map[i] != j) { c->map[i] = j; c->rowind_buf[nnz++] = i; }
    }
  }

  for (ch = c->first_child[j]; ch != -1; ch = c->next_child[ch]) {
    csn = c->col2sn[ch];
    if (ch == c->first_child[j]) L->first_child[L->n_sn] = csn;
    else {
      L->next_child[csn] = L->first_child[L->n_sn];
      L->first_child[L->n_sn] = csn;
    }
  }

  if (j < c->n) {
    c->col2sn[j] = L->n_sn;
    L->sn_size[L->n_sn] = 1;
    L->sn_up_size[L->n_sn] = nnz;
    L->sn_struct[L->n_sn] = malloc(sizeof(int) * nnz);
    memcpy(L->sn_struct[L->n_sn], c->rowind_buf, sizeof(int) * nnz);
    L->n_sn++;
  }
  return 0;
}

static znz rec_amalgamate(sym_ctx* c, int sn)
{
  oracle_sym* L = c->L;
  int i, ip, csn, gc, nnz, nch = 0, new_size, new_up, n = 0;
  znz own, merged, *cz;
  double* zcount;
  own.zeros = 0.0;
  own.nonzeros = (double)(((L->sn_up_size[sn] - L->sn_size[sn]) * L->sn_size[sn])
                          + (L->sn_size[sn] * (L->sn_size[sn] + 1)) / 2);
  if (L->first_child[sn] == -1) return own;
  for (csn = L->first_child[sn]; csn != -1; csn = L->next_child[csn]) nch++;
  cz = malloc(sizeof(znz) * nch);
  i = 0;
  for (csn = L->first_child[sn]; csn != -1; csn = L->next_child[csn]) cz[i++] = rec_amalgamate(c, csn);
  merged = own;
  for (i = 0; i < nch; i++) { merged.nonzeros += cz[i].nonzeros; merged.zeros += cz[i].zeros; }
  free(cz);

  nnz = 0;
  for (csn = L->first_child[sn]; csn != -1; csn = L->next_child[csn])
    for (ip = 0; ip < L->sn_size[csn]; ip++) {
      i = L->sn_struct[csn][ip];
      c->map[i] = sn;
      c->rowind_buf[nnz++] = i;
    }
  for (ip = 0; ip < L->sn_size[sn]; ip++) {
    i = L->sn_struct[sn][ip];
    c->map[i] = sn;
    c->rowind_buf[nnz++] = i;
  }
  new_size = nnz;
  for (csn = L->first_child[sn]; csn != -1; csn = L->next_child[csn])
    for (ip = L->sn_size[csn]; ip < L->sn_up_size[csn]; ip++) {
      i = L->sn_struct[csn][ip];
      if (c->map[i] != sn) { c->map[i] = sn; c->rowind_buf[nnz++] = i; }
    }
  for (ip = L->sn_size[sn]; ip < L->sn_up_size[sn]; ip++) {
    i = L->sn_struct[sn][ip];
    if (c->map[i] != sn) { c->map[i] = sn; c->rowind_buf[nnz++] = i; }
  }
  new_up = nnz;

  for (ip = 0; ip < nnz; ip++)
    if (c->rowind_buf[ip] >= n) n = c->rowind_buf[ip] + 1;
  zcount = malloc(sizeof(double) * (n + 1));
  for (ip = 0; ip < new_size; ip++) zcount[c->rowind_buf[ip]] = (double)(ip + 1);
  for (ip = new_size; ip < new_up; ip++) zcount[c->rowind_buf[ip]] = (double)new_size;
  for (csn = L->first_child[sn]; csn != -1; csn = L->next_child[csn]) {
    for (ip = 0; ip < L->sn_size[csn]; ip++) zcount[L->sn_struct[csn][ip]] -= (double)(ip + 1);
    for (ip = L->sn_size[csn]; ip < L->sn_up_size[csn]; ip++)
      zcount[L->sn_struct[csn][ip]] -= (double)L->sn_size[csn];
  }
  for (ip = 0; ip < L->sn_size[sn]; ip++) zcount[L->sn_struct[sn][ip]] -= (double)(ip + 1);
  for (ip = L->sn_size[sn]; ip < L->sn_up_size[sn]; ip++)
    zcount[L->sn_struct[sn][ip]] -= (double)L->sn_size[sn];
  for (ip = 0; ip < new_up; ip++) merged.zeros += zcount[c->rowind_buf[ip]];
  free(zcount);

  {
    int take = (new_size < 16)
               || ((L->sn_size[sn] < 50) && (merged.zeros < 0.5 * merged.nonzeros))
               || ((L->sn_size[sn] < 250) && (merged.zeros < 0.25 * merged.nonzeros))
               || ((L->sn_size[sn] < 500) && (merged.zeros < 0.10 * merged.nonzeros))
               || (merged.zeros < 0.05 * merged.nonzeros);
    log_decision(L, sn, new_size, new_up, merged, take, c->rowind_buf);
    if (!take) return own;
  }

  L->sn_size[sn] = new_size;
  L->sn_up_size[sn] = new_up;
  L->sn_struct[sn] = realloc(L->sn_struct[sn], sizeof(int) * new_up);
  memcpy(L->sn_struct[sn], c->rowind_buf, sizeof(int) * new_up);

  nch = 0;
  for (csn = L->first_child[sn]; csn != -1; csn = L->next_child[csn]) {
    for (ip = 0; ip < L->sn_size[csn]; ip++) c->col2sn[L->sn_struct[csn][ip]] = sn;
    for (gc = L->first_child[csn]; gc != -1; gc = L->next_child[gc]) c->rowind_buf[nch++] = gc;
  }
  for (csn = L->first_child[sn]; csn != -1; csn = L->next_child[csn]) {
    free(L->sn_struct[csn]);
    L->sn_struct[csn] = NULL;
    L->sn_size[csn] = 0;
    L->sn_up_size[csn] = 0;
  }
  L->first_child[sn] = -1;
  for (i = 0; i < nch; i++) {
    L->next_child[c->rowind_buf[i]] = L->first_child[sn];
    L->first_child[sn] = c->rowind_buf[i];
  }
  return merged;
}

static void log_decision(oracle_sym* L, int sn, int new_size, int new_up, znz m, int merged,
                         const int* cols)
{
  if (L->caplog < 0) return;  /* logging off */
  if (L->nlog == L->caplog) {
    L->caplog = L->caplog ? 2 * L->caplog : 64;
    L->log_i = realloc(L->log_i, sizeof(int) * 6 * L->caplog);
    L->log_d = realloc(L->log_d, sizeof(double) * 2 * L->caplog);
  }
  if (L->ncols + new_size > L->capcols) {
    while (L->ncols + new_size > L->capcols) L->capcols = L->capcols ? 2 * L->capcols : 1024;
    L->log_cols = realloc(L->log_cols, sizeof(int) * L->capcols);
  }
  {
    int* r = L->log_i + 6 * L->nlog;
    r[0] = sn; r[1] = L->sn_size[sn]; r[2] = new_size; r[3] = new_up; r[4] = merged;
    r[5] = L->ncols;
  }
  L->log_d[2 * L->nlog] = m.zeros;
  L->log_d[2 * L->nlog + 1] = m.nonzeros;
  memcpy(L->log_cols + L->ncols, cols, sizeof(int) * new_size);
  L->ncols += new_size;
  L->nlog++;
}

static void stats(oracle_sym* L, double* nnz, double* flops)
{
  int sn, i, colnnz;
  *nnz = 0;
  *flops = 0;
  for (sn = 0; sn < L->n_sn; sn++)
    for (i = 0, colnnz = L->sn_up_size[sn]; i < L->sn_size[sn]; i++, colnnz--) {
      *flops += 1.0 + (double)colnnz * (double)colnnz;
      *nnz += (double)colnnz;
    }
}

typedef struct {
  int n;
  const int *colptr, *rowind;
  int max_depth;
  int amalgamate;  /* 0: stop after the fundamental supernodes */
  int log_merges;
  oracle_sym* out;
} sym_job;

static void* sym_thread(void* arg)
{
  sym_job* job = (sym_job*)arg;
  int n = job->n, j;
  oracle_sym* L = calloc(1, sizeof(oracle_sym));
  sym_ctx c;
  int* parent = malloc(sizeof(int) * (n + 1));
  job->out = L;
  L->n = n;
  L->caplog = job->log_merges ? 0 : -1;
  L->sn_struct = calloc(n + 1, sizeof(int*));
  L->sn_size = malloc(sizeof(int) * (n + 1));
  L->sn_up_size = malloc(sizeof(int) * (n + 1));
  L->first_child = malloc(sizeof(int) * (n + 1));
  L->next_child = malloc(sizeof(int) * (n + 1));
  c.colptr = job->colptr;
  c.rowind = job->rowind;
  c.n = n;
  c.L = L;
  c.first_child = malloc(sizeof(int) * (n + 1));
  c.next_child = malloc(sizeof(int) * (n + 1));
  c.col2sn = malloc(sizeof(int) * (n + 1));
  c.map = malloc(sizeof(int) * (n + 1));
  c.rowind_buf = malloc(sizeof(int) * (n + 1));

  oracle_etree(n, job->colptr, job->rowind, parent);
  for (j = 0; j <= n; j++) c.first_child[j] = -1;
  for (j = n - 1; j >= 0; j--) {
    int p = parent[j];
    c.next_child[j] = c.first_child[p];
    c.first_child[p] = j;
  }
  /* depth by breadth-first levels from the virtual root */
  {
    int *cur = malloc(sizeof(int) * (n + 1)), *nxt = malloc(sizeof(int) * (n + 1)), *t;
    int nc = 1, nn, i;
    cur[0] = n;
    L->depth = -1;
    while (nc) {
      nn = 0;
      for (i = 0; i < nc; i++)
        for (j = c.first_child[cur[i]]; j != -1; j = c.next_child[j]) nxt[nn++] = j;
      t = cur; cur = nxt; nxt = t;
      nc = nn;
      L->depth++;
    }
    free(cur);
    free(nxt);
  }
  if (job->max_depth && L->depth > job->max_depth) {
    L->failed = 1;
  } else {
    L->n_sn = 0;
    for (j = 0; j < n; j++) c.map[j] = -1;
    for (j = 0; j <= n; j++) L->first_child[j] = L->next_child[j] = -1;
    rec_symbolic(&c, n);
    stats(L, &L->fund_nnz, &L->fund_flops);
    for (j = 0; j < n; j++) c.map[j] = -1;
    if (L->n_sn > 0 && job->amalgamate) rec_amalgamate(&c, L->n_sn - 1);
    stats(L, &L->relax_nnz, &L->relax_flops);
  }
  free(parent); free(c.first_child); free(c.next_child); free(c.col2sn); free(c.map);
  free(c.rowind_buf);
  return NULL;
}

/* flags: 1 = skip amalgamation (fundamental supernodes only), 2 = log the
   amalgamation decisions */
oracle_sym* oracle_symbolic_ex(int n, const int* colptr, const int* rowind, int max_depth, int flags)
{
  pthread_t th;
  pthread_attr_t at;
  sym_job job = {n, colptr, rowind, max_depth, !(flags & 1), (flags & 2) != 0, NULL};
  pthread_attr_init(&at);
  pthread_attr_setstacksize(&at, (size_t)1 << 30);
  if (pthread_create(&th, &at, sym_thread, &job) != 0) return NULL;
  pthread_join(th, NULL);
  pthread_attr_destroy(&at);
  return job.out;
}

oracle_sym* oracle_symbolic(int n, const int* colptr, const int* rowind, int max_depth)
{
  return oracle_symbolic_ex(n, colptr, rowind, max_depth, 0);
}

/* The merge log (see oracle_sym).  Returns the record count; with non-NULL
   arrays copies 6 ints and 2 doubles per record and the candidate columns. */
int oracle_sym_merge_log(const oracle_sym* L, int* rec_i, double* rec_d, int* cols, int* ncols)
{
  if (rec_i && L->nlog) memcpy(rec_i, L->log_i, sizeof(int) * 6 * L->nlog);
  if (rec_d && L->nlog) memcpy(rec_d, L->log_d, sizeof(double) * 2 * L->nlog);
  if (cols && L->ncols) memcpy(cols, L->log_cols, sizeof(int) * L->ncols);
  if (ncols) *ncols = L->ncols;
  return L->nlog;
}

int oracle_sym_info(const oracle_sym* L, int* n_sn, int* depth, int* failed, double* st4)
{
  *n_sn = L->n_sn;
  *depth = L->depth;
  *failed = L->failed;
  st4[0] = L->fund_nnz; st4[1] = L->fund_flops; st4[2] = L->relax_nnz; st4[3] = L->relax_flops;
  return 0;
}

int oracle_sym_export(const oracle_sym* L, int* sn_size, int* sn_up_size, int* first_child,
                      int* next_child, int* struct_ptr, int* struct_idx)
{
  int sn, p = 0;
  for (sn = 0; sn <= L->n_sn; sn++) {
    if (sn_size) sn_size[sn] = sn < L->n_sn ? L->sn_size[sn] : 0;
    if (sn_up_size) sn_up_size[sn] = sn < L->n_sn ? L->sn_up_size[sn] : 0;
    if (first_child) first_child[sn] = L->first_child[sn];
    if (next_child) next_child[sn] = L->next_child[sn];
  }
  for (sn = 0; sn < L->n_sn; sn++) {
    if (struct_ptr) struct_ptr[sn] = p;
    if (struct_idx && L->sn_up_size[sn] > 0)
      memcpy(struct_idx + p, L->sn_struct[sn], sizeof(int) * L->sn_up_size[sn]);
    p += L->sn_up_size[sn];
  }
  if (struct_ptr) struct_ptr[L->n_sn] = p;
  return p;
}

void oracle_sym_free(oracle_sym* L)
{
  int sn;
  if (!L) return;
  for (sn = 0; sn < L->n; sn++) free(L->sn_struct[sn]);
  free(L->sn_struct); free(L->sn_size); free(L->sn_up_size); free(L->first_child);
  free(L->next_child); free(L->log_i); free(L->log_d); free(L->log_cols); free(L);
}

void oracle_free(void* p) { free(p); }

/* ---- numeric phase: the multifrontal factor and the supernodal solve ------
 *
 * A C restatement of the reference's numeric path, for the CPU baseline and
 * fast oracle solutions of large systems:
 *   oracle_mf_factor   recursive_multifrontal_supernodal_factor_llt
 *                      (src/taucs_sn_llt.c:1831-1977): children in child-list
 *                      order, each extend-added into the parent front as it
 *                      returns (extend_add_wrapper :1797-1827, the front
 *                      created zeroed at the first child, :1013-1047); then
 *                      multifrontal_supernodal_front_factor (:1086-1194):
 *                      A's columns added through the row bitmap, dpotrf on F1
 *                      (s x s), dtrsm R/L/T/N on F2 (u x s), dsyrk L/N into U
 *   oracle_mf_solve    taucs_dsupernodal_solve_llt (:2844-2885): forward
 *                      recursion :2597-2720 (gather, trsm L/L/N/N, gemm N/N,
 *                      scatter), backward :2723-2841 (gather, gemm T/N,
 *                      trsm L/L/T/N, scatter), here with k right-hand sides as
 *                      the columns of the dense blocks
 * The dense kernels are the host's BLAS/LAPACK (the reference links the
 * Fortran BLAS), resolved at run time from a shared library given by path
 * (scipy's bundled OpenBLAS exports them as scipy_dpotrf_ etc.). */

typedef void (*potrf_fn)(const char*, const int*, double*, const int*, int*);
typedef void (*trsm_fn)(const char*, const char*, const char*, const char*, const int*, const int*,
                        const double*, const double*, const int*, double*, const int*);
typedef void (*syrk_fn)(const char*, const char*, const int*, const int*, const double*,
                        const double*, const int*, const double*, double*, const int*);
typedef void (*gemm_fn)(const char*, const char*, const int*, const int*, const int*, const double*,
                        const double*, const int*, const double*, const int*, const double*, double*,
                        const int*);
static potrf_fn b_potrf;
static trsm_fn b_trsm;
static syrk_fn b_syrk;
static gemm_fn b_gemm;
static void (*b_threads)(int);
static const char* (*b_config)(void);

int oracle_load_blas(const char* so_path, const char* prefix)
{
  char nm[128];
  void* h = dlopen(so_path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return -1;
#define SYM(var, type, base)                                  \
  snprintf(nm, sizeof nm, "%s%s", prefix, base);              \
  var = (type)dlsym(h, nm);                                   \
  if (!var) return -1;
  SYM(b_potrf, potrf_fn, "dpotrf_")
  SYM(b_trsm, trsm_fn, "dtrsm_")
  SYM(b_syrk, syrk_fn, "dsyrk_")
  SYM(b_gemm, gemm_fn, "dgemm_")
#undef SYM
  snprintf(nm, sizeof nm, "%sopenblas_set_num_threads", prefix);
  b_threads = (void (*)(int))dlsym(h, nm);
  snprintf(nm, sizeof nm, "%sopenblas_get_config", prefix);
  b_config = (const char* (*)(void))dlsym(h, nm);
  return 0;
}

/* the BLAS library's own description of itself (OpenBLAS: version and build) */
const char* oracle_blas_config(void) { return b_config ? b_config() : "unknown"; }

int oracle_blas_threads(int t)
{
  if (!b_threads) return -1;
  b_threads(t);
  return 0;
}

typedef struct {
  const oracle_sym* S;
  int n;
  double** F1;   /* [n_sn] s x s, ld s (lower: L11) */
  double** F2;   /* [n_sn] u x s, ld u (L21) */
  int failed_sn; /* -1, or the supernode whose pivot was not positive */
} oracle_num;

typedef struct { double *F1, *F2, *U; } front_t;

/* wall time of the last factorization by part (s): assembly, dpotrf, dtrsm,
   dsyrk, extend-add, front allocation */
static double t_part[6];
static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}
void oracle_mf_times(double* out) { memcpy(out, t_part, sizeof t_part); }

static int front_create(const oracle_sym* S, int sn, front_t* f)
{
  size_t s = (size_t)S->sn_size[sn], u = (size_t)(S->sn_up_size[sn] - S->sn_size[sn]);
  f->F1 = calloc(s * s + 1, sizeof(double));
  f->F2 = calloc(u * s + 1, sizeof(double));
  f->U = calloc(u * u + 1, sizeof(double));
  return (f->F1 && f->F2 && f->U) ? 0 : -1;
}

/* child's U (lower) into the parent front, entries swapped to stay lower */
static void extend_add(const oracle_sym* S, int parent, front_t* pf, int child, const double* U,
                       int* bitmap)
{
  const int ps = S->sn_size[parent], pu = S->sn_up_size[parent] - ps;
  const int cs = S->sn_size[child], cu = S->sn_up_size[child] - cs;
  const int* prow = S->sn_struct[parent];
  const int* crow = S->sn_struct[child] + cs;
  int i, j;
  for (i = 0; i < ps + pu; i++) bitmap[prow[i]] = i;
  for (j = 0; j < cu; j++)
    for (i = j; i < cu; i++) {
      int pj = bitmap[crow[j]], pi = bitmap[crow[i]];
      double v = U[(size_t)cu * j + i];
      if (pj > pi) { int t = pj; pj = pi; pi = t; }
      if (pj < ps) {
        if (pi < ps) pf->F1[(size_t)ps * pj + pi] += v;
        else pf->F2[(size_t)pu * pj + (pi - ps)] += v;
      } else {
        pf->U[(size_t)pu * (pj - ps) + (pi - ps)] += v;
      }
    }
}

static int front_factor(const oracle_sym* S, int sn, front_t* f, const int* colptr,
                        const int* rowind, const double* vals, int* bitmap)
{
  const int s = S->sn_size[sn], u = S->sn_up_size[sn] - s;
  const int* rows = S->sn_struct[sn];
  int i, j, info = 0;
  const double one = 1.0, mone = -1.0;
  double t0 = now_s(), t1, t2, t3;
  for (i = 0; i < s + u; i++) bitmap[rows[i]] = i;
  for (j = 0; j < s; j++) {
    const int c = rows[j];
    for (i = colptr[c]; i < colptr[c + 1]; i++) {
      const int r = bitmap[rowind[i]];
      if (r < s) f->F1[(size_t)s * j + r] += vals[i];
      else f->F2[(size_t)u * j + (r - s)] += vals[i];
    }
  }
  t1 = now_s();
  if (s) b_potrf("L", &s, f->F1, &s, &info);
  t2 = now_s();
  t_part[0] += t1 - t0;
  t_part[1] += t2 - t1;
  if (info) return -1;
  if (u && s) {
    b_trsm("R", "L", "T", "N", &u, &s, &one, f->F1, &s, f->F2, &u);
    t3 = now_s();
    b_syrk("L", "N", &u, &s, &mone, f->F2, &u, &one, f->U, &u);
    t_part[2] += t3 - t2;
    t_part[3] += now_s() - t3;
  }
  return 0;
}

/* live supernodes, children before parents, children in child-list order */
static int* postorder_sn(const oracle_sym* S, int* count)
{
  int* out = malloc(sizeof(int) * (S->n_sn + 1));
  int* st = malloc(sizeof(int) * (S->n_sn + 2));
  int* it = malloc(sizeof(int) * (S->n_sn + 2));
  int top = 0, k = 0;
  st[0] = S->n_sn;
  it[0] = S->first_child[S->n_sn];
  while (top >= 0) {
    int c = it[top];
    if (c == -1) {
      if (st[top] != S->n_sn) out[k++] = st[top];
      top--;
    } else {
      it[top] = S->next_child[c];
      st[++top] = c;
      it[top] = S->first_child[c];
    }
  }
  free(st);
  free(it);
  *count = k;
  return out;
}

oracle_num* oracle_mf_factor(const oracle_sym* S, const int* colptr, const int* rowind,
                             const double* vals)
{
  int cnt, k, *order, *bitmap, *parent;
  front_t* fr;
  oracle_num* N;
  if (!b_potrf || !S || S->failed) return NULL;
  N = calloc(1, sizeof(oracle_num));
  N->S = S;
  N->n = S->n;
  N->failed_sn = -1;
  N->F1 = calloc(S->n_sn + 1, sizeof(double*));
  N->F2 = calloc(S->n_sn + 1, sizeof(double*));
  fr = calloc(S->n_sn + 1, sizeof(front_t));
  bitmap = malloc(sizeof(int) * (S->n + 1));
  parent = malloc(sizeof(int) * (S->n_sn + 1));
  for (k = 0; k <= S->n_sn; k++) parent[k] = -1;
  for (k = 0; k <= S->n_sn; k++) {
    int c;
    for (c = S->first_child[k]; c != -1; c = S->next_child[c]) parent[c] = k;
  }
  order = postorder_sn(S, &cnt);
  memset(t_part, 0, sizeof t_part);
  for (k = 0; k < cnt; k++) {
    const int sn = order[k], p = parent[sn];
    double ta = now_s();
    if (!fr[sn].F1 && front_create(S, sn, &fr[sn]) != 0) { N->failed_sn = sn; break; }
    t_part[5] += now_s() - ta;
    if (front_factor(S, sn, &fr[sn], colptr, rowind, vals, bitmap) != 0) {
      N->failed_sn = sn;
      break;
    }
    N->F1[sn] = fr[sn].F1;
    N->F2[sn] = fr[sn].F2;
    if (p != S->n_sn && p >= 0) {   /* the virtual root takes no extend-add */
      double tb = now_s(), tc;
      if (!fr[p].F1 && front_create(S, p, &fr[p]) != 0) { N->failed_sn = p; break; }
      tc = now_s();
      extend_add(S, p, &fr[p], sn, fr[sn].U, bitmap);
      t_part[5] += tc - tb;
      t_part[4] += now_s() - tc;
    }
    free(fr[sn].U);
    fr[sn].U = NULL;
  }
  for (k = 0; k <= S->n_sn; k++) {
    free(fr[k].U);
    if (fr[k].F1 && N->F1[k] != fr[k].F1) { free(fr[k].F1); free(fr[k].F2); }
  }
  free(fr); free(bitmap); free(parent); free(order);
  return N;
}

int oracle_mf_failed(const oracle_num* N) { return N ? N->failed_sn : -2; }

/* X (n x k, ld n) = (L L^T)^{-1} B in the permuted space */
int oracle_mf_solve(const oracle_num* N, int k, const double* B, double* X)
{
  const oracle_sym* S = N->S;
  const int n = N->n;
  int cnt, q, i, r, *order = postorder_sn(S, &cnt);
  double* b = malloc(sizeof(double) * (size_t)n * k);   /* forward: x <- b, updated */
  double* y = malloc(sizeof(double) * (size_t)n * k);   /* forward solution */
  double* t = malloc(sizeof(double) * (size_t)(n + 1) * k);
  const double one = 1.0, mone = -1.0, zero = 0.0;
  if (N->failed_sn != -1) { free(order); free(b); free(y); free(t); return -1; }
  memcpy(b, B, sizeof(double) * (size_t)n * k);
  for (q = 0; q < cnt; q++) {
    const int sn = order[q], s = S->sn_size[sn], u = S->sn_up_size[sn] - s;
    const int* rows = S->sn_struct[sn];
    double *xd = t, *bd = t + (size_t)s * k;
    for (r = 0; r < k; r++)
      for (i = 0; i < s; i++) xd[i + (size_t)s * r] = b[rows[i] + (size_t)n * r];
    b_trsm("L", "L", "N", "N", &s, &k, &one, N->F1[sn], &s, xd, &s);
    if (u > 0 && s > 0)
      b_gemm("N", "N", &u, &k, &s, &one, N->F2[sn], &u, xd, &s, &zero, bd, &u);
    for (r = 0; r < k; r++) {
      for (i = 0; i < s; i++) y[rows[i] + (size_t)n * r] = xd[i + (size_t)s * r];
      for (i = 0; i < u; i++) b[rows[s + i] + (size_t)n * r] -= bd[i + (size_t)u * r];
    }
  }
  for (q = cnt - 1; q >= 0; q--) {
    const int sn = order[q], s = S->sn_size[sn], u = S->sn_up_size[sn] - s;
    const int* rows = S->sn_struct[sn];
    double *bd = t, *xd = t + (size_t)s * k;
    for (r = 0; r < k; r++) {
      for (i = 0; i < s; i++) bd[i + (size_t)s * r] = y[rows[i] + (size_t)n * r];
      for (i = 0; i < u; i++) xd[i + (size_t)u * r] = X[rows[s + i] + (size_t)n * r];
    }
    if (u > 0 && s > 0)
      b_gemm("T", "N", &s, &k, &u, &mone, N->F2[sn], &u, xd, &u, &one, bd, &s);
    b_trsm("L", "L", "T", "N", &s, &k, &one, N->F1[sn], &s, bd, &s);
    for (r = 0; r < k; r++)
      for (i = 0; i < s; i++) X[rows[i] + (size_t)n * r] = bd[i + (size_t)s * r];
  }
  free(order); free(b); free(y); free(t);
  return 0;
}

/* panel of supernode sn as [L11; L21], (s+u) x s column-major, ld s+u */
int oracle_mf_panel(const oracle_num* N, int sn, double* out)
{
  const oracle_sym* S = N->S;
  const int s = S->sn_size[sn], u = S->sn_up_size[sn] - s, m = s + u;
  int i, j;
  if (!N->F1[sn]) return -1;
  for (j = 0; j < s; j++) {
    for (i = 0; i < s; i++) out[i + (size_t)m * j] = i >= j ? N->F1[sn][i + (size_t)s * j] : 0.0;
    for (i = 0; i < u; i++) out[s + i + (size_t)m * j] = N->F2[sn][i + (size_t)u * j];
  }
  return 0;
}

void oracle_mf_free(oracle_num* N)
{
  int k;
  if (!N) return;
  for (k = 0; k <= N->S->n_sn; k++) { free(N->F1[k]); free(N->F2[k]); }
  free(N->F1); free(N->F2); free(N);
}
