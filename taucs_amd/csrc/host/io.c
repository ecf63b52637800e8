/*
 * io.c -- on-disk matrix and vector formats of the TAUCS C-ABI (host code).
 *
 *   taucs_ccs_read_binary / taucs_ccs_write_binary   src/taucs_ccs_io.c:34-121
 *       int nrows, ncols, flags; int colptr[ncols+1]; int rowind[nnz];
 *       then the nnz values (the reference writes this format from MATLAB,
 *       matlab/taucs_ccs_write_binary.m; the C writer here is an addition)
 *   taucs_ccs_write_ijv / taucs_ccs_read_ijv         :280-575
 *       text lines "i j v", 1-based; a symmetric matrix is written with both
 *       triangles and read back keeping the lower one
 *   taucs_ccs_read_mtx                                :580-820
 *       "nrows ncols nnz" then "i j v" lines (1-based); with TAUCS_SYMMETRIC
 *       an upper entry is mirrored to the lower triangle; with TAUCS_PATTERN
 *       the values become nrows+1 on the diagonal and -1 elsewhere, nrows
 *       being the largest row index read before the entry.  Lines starting
 *       with '%' before the size line (the MatrixMarket banner and comments)
 *       are skipped, which the reference does not do; files it reads parse
 *       the same.
 *   taucs_ccs_read_ccs                                :826-960
 *       N; colptr[N+1]; rowind[nnz] (0-based); values[nnz]
 *   taucs_ccs_read_hb                                 :128-275
 *       needs the Harwell-Boeing Fortran reader, which is not linked here:
 *       returns NULL, as the reference does without it
 *   taucs_vec_read_binary / taucs_vec_write_binary    :967-1047
 *       n raw values
 *
 * Matrices read are real double (the datatype this library computes in);
 * the dimensions of the text formats are the largest indices read.  Entries
 * keep their file order within each column.  Any failure returns NULL (or
 * -1) after a taucs_printf message; nothing is half-built.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "taucs.h"

/* ---- binary ------------------------------------------------------------- */

static int read_exact(FILE* f, void* p, size_t size, size_t count)
{
  return count == 0 || fread(p, size, count, f) == count;
}

taucs_ccs_matrix* taucs_ccs_read_binary(char* filename)
{
  FILE* f = fopen(filename, "rb");
  int hdr[3];
  taucs_ccs_matrix* A = NULL;
  int* colptr = NULL;
  if (!f) {
    taucs_printf("taucs_ccs_read_binary: could not open %s\n", filename);
    return NULL;
  }
  if (!read_exact(f, hdr, sizeof(int), 3) || hdr[1] < 0 || hdr[0] < 0) goto bad;
  if (!(hdr[2] & TAUCS_DOUBLE)) {
    taucs_printf("taucs_ccs_read_binary: only double matrices are supported\n");
    goto fail;
  }
  colptr = (int*)malloc(sizeof(int) * ((size_t)hdr[1] + 1));
  if (!colptr || !read_exact(f, colptr, sizeof(int), (size_t)hdr[1] + 1)) goto bad;
  if (colptr[0] != 0 || colptr[hdr[1]] < 0) goto bad;
  A = taucs_ccs_create(hdr[0], hdr[1], colptr[hdr[1]], TAUCS_DOUBLE);
  if (!A) goto fail;
  memcpy(A->colptr, colptr, sizeof(int) * ((size_t)hdr[1] + 1));
  if (!read_exact(f, A->rowind, sizeof(int), (size_t)colptr[hdr[1]])
      || !read_exact(f, A->values.d, sizeof(double), (size_t)colptr[hdr[1]]))
    goto bad;
  {  /* pointers nondecreasing, row indices in range */
    int j, p;
    for (j = 0; j < hdr[1]; j++)
      if (colptr[j + 1] < colptr[j]) goto bad;
    for (p = 0; p < colptr[hdr[1]]; p++)
      if (A->rowind[p] < 0 || A->rowind[p] >= hdr[0]) goto bad;
  }
  A->flags = hdr[2];
  free(colptr);
  fclose(f);
  return A;
bad:
  taucs_printf("taucs_ccs_read_binary: %s is truncated or malformed\n", filename);
fail:
  free(colptr);
  if (A) taucs_ccs_free(A);
  fclose(f);
  return NULL;
}

int taucs_ccs_write_binary(taucs_ccs_matrix* A, char* filename)
{
  FILE* f;
  int hdr[3];
  int nnz;
  if (!A || !(A->flags & TAUCS_DOUBLE)) {
    taucs_printf("taucs_ccs_write_binary: only double matrices are supported\n");
    return -1;
  }
  f = fopen(filename, "wb");
  if (!f) {
    taucs_printf("taucs_ccs_write_binary: could not open %s\n", filename);
    return -1;
  }
  hdr[0] = A->m;  /* rows */
  hdr[1] = A->n;  /* columns */
  hdr[2] = A->flags;
  nnz = A->colptr[A->n];
  if (fwrite(hdr, sizeof(int), 3, f) != 3
      || fwrite(A->colptr, sizeof(int), (size_t)A->n + 1, f) != (size_t)A->n + 1
      || fwrite(A->rowind, sizeof(int), (size_t)nnz, f) != (size_t)nnz
      || fwrite(A->values.d, sizeof(double), (size_t)nnz, f) != (size_t)nnz) {
    fclose(f);
    taucs_printf("taucs_ccs_write_binary: write to %s failed\n", filename);
    return -1;
  }
  return fclose(f) == 0 ? 0 : -1;
}

/* ---- coordinate lists (ijv, mtx) ---------------------------------------- */

typedef struct {
  int *i, *j;
  double* v;
  int n, cap;
} triplets;

static int trip_push(triplets* t, int i, int j, double v)
{
  if (t->n == t->cap) {
    int cap = t->cap ? t->cap + t->cap / 2 : 4096;
    int* ni = (int*)realloc(t->i, sizeof(int) * (size_t)cap);
    int* nj = ni ? (int*)realloc(t->j, sizeof(int) * (size_t)cap) : NULL;
    double* nv = nj ? (double*)realloc(t->v, sizeof(double) * (size_t)cap) : NULL;
    if (ni) t->i = ni;
    if (nj) t->j = nj;
    if (!nv) return -1;
    t->v = nv;
    t->cap = cap;
  }
  t->i[t->n] = i;
  t->j[t->n] = j;
  t->v[t->n] = v;
  t->n++;
  return 0;
}

static void trip_free(triplets* t)
{
  free(t->i);
  free(t->j);
  free(t->v);
}

/* 1-based triplets -> CSC, file order kept within each column (stable
   counting sort by column) */
static taucs_ccs_matrix* trip_to_ccs(const triplets* t, int nrows, int ncols, int flags)
{
  taucs_ccs_matrix* A = taucs_ccs_create(nrows, ncols, t->n, TAUCS_DOUBLE);
  int* next;
  int k, j;
  if (!A) return NULL;
  next = (int*)calloc((size_t)ncols + 1, sizeof(int));
  if (!next) {
    taucs_ccs_free(A);
    return NULL;
  }
  for (k = 0; k < t->n; k++) next[t->j[k]]++;  /* column j-1 counted at j */
  A->colptr[0] = 0;
  for (j = 0; j < ncols; j++) A->colptr[j + 1] = A->colptr[j] + next[j + 1];
  for (j = 0; j < ncols; j++) next[j] = A->colptr[j];
  for (k = 0; k < t->n; k++) {
    const int c = t->j[k] - 1, p = next[c]++;
    A->rowind[p] = t->i[k] - 1;
    A->values.d[p] = t->v[k];
  }
  free(next);
  A->flags = flags;
  return A;
}

int taucs_ccs_write_ijv(taucs_ccs_matrix* A, char* filename)
{
  FILE* f;
  int j, p;
  if (!A || !(A->flags & TAUCS_DOUBLE)) {
    taucs_printf("taucs_ccs_write_ijv: only double matrices are supported\n");
    return -1;
  }
  f = fopen(filename, "w");
  if (!f) {
    taucs_printf("taucs_ccs_write_ijv: could not open ijv file %s\n", filename);
    return -1;
  }
  for (j = 0; j < A->n; j++)
    for (p = A->colptr[j]; p < A->colptr[j + 1]; p++) {
      const int i = A->rowind[p];
      const double v = A->values.d[p];
      fprintf(f, "%d %d %0.17e\n", i + 1, j + 1, v);
      if (i != j && (A->flags & TAUCS_SYMMETRIC)) fprintf(f, "%d %d %0.17e\n", j + 1, i + 1, v);
    }
  return fclose(f) == 0 ? 0 : -1;
}

static int bad_index(double di, double dj)
{
  return !(di >= 1 && dj >= 1 && di <= 2147483647.0 && dj <= 2147483647.0);
}

taucs_ccs_matrix* taucs_ccs_read_ijv(char* filename, int flags)
{
  FILE* f;
  triplets t = {0};
  double di, dj, v;
  int nrows = 0, ncols = 0, out_flags;
  taucs_ccs_matrix* A;
  if (!(flags & TAUCS_DOUBLE)) {
    taucs_printf("taucs_ccs_read_ijv: only double matrices are supported\n");
    return NULL;
  }
  f = fopen(filename, "r");
  if (!f) {
    taucs_printf("taucs_ccs_read_ijv: could not open ijv file %s\n", filename);
    return NULL;
  }
  while (fscanf(f, "%lg %lg %lg", &di, &dj, &v) == 3) {
    if (bad_index(di, dj)) {
      taucs_printf("taucs_ccs_read_ijv: bad index in %s\n", filename);
      fclose(f);
      trip_free(&t);
      return NULL;
    }
    /* a symmetric (or Hermitian) matrix keeps its lower triangle */
    if ((flags & (TAUCS_SYMMETRIC | TAUCS_HERMITIAN)) && di < dj) continue;
    if (trip_push(&t, (int)di, (int)dj, v) != 0) {
      taucs_printf("taucs_ccs_read_ijv: out of memory\n");
      fclose(f);
      trip_free(&t);
      return NULL;
    }
    if ((int)di > nrows) nrows = (int)di;
    if ((int)dj > ncols) ncols = (int)dj;
  }
  fclose(f);
  out_flags = TAUCS_DOUBLE;
  if (flags & TAUCS_SYMMETRIC) out_flags |= TAUCS_SYMMETRIC | TAUCS_LOWER;
  else if (flags & TAUCS_HERMITIAN) out_flags |= TAUCS_HERMITIAN | TAUCS_LOWER;
  A = trip_to_ccs(&t, nrows, ncols, out_flags);
  trip_free(&t);
  if (!A) taucs_printf("taucs_ccs_read_ijv: out of memory\n");
  return A;
}

taucs_ccs_matrix* taucs_ccs_read_mtx(char* filename, int flags)
{
  FILE* f;
  triplets t = {0};
  double di, dj, v;
  int hr, hc, hn, nrows = 0, ncols = 0, c;
  taucs_ccs_matrix* A;
  if (!(flags & TAUCS_DOUBLE)) {
    taucs_printf("taucs_ccs_read_mtx: only double matrices are supported\n");
    return NULL;
  }
  f = fopen(filename, "r");
  if (!f) {
    taucs_printf("taucs_ccs_read_mtx: could not open mtx file %s\n", filename);
    return NULL;
  }
  /* comment lines ('%') before the size line */
  while ((c = fgetc(f)) == '%')
    while ((c = fgetc(f)) != EOF && c != '\n') {
    }
  if (c != EOF) ungetc(c, f);
  if (fscanf(f, "%d %d %d", &hr, &hc, &hn) != 3) {
    taucs_printf("taucs_ccs_read_mtx: %s has no size line\n", filename);
    fclose(f);
    return NULL;
  }
  while (fscanf(f, "%lg %lg %lg", &di, &dj, &v) == 3) {
    int i, j;
    if (bad_index(di, dj)) {
      taucs_printf("taucs_ccs_read_mtx: bad index in %s\n", filename);
      fclose(f);
      trip_free(&t);
      return NULL;
    }
    i = (int)di;
    j = (int)dj;
    if ((flags & TAUCS_SYMMETRIC) && i < j) {  /* mirrored to the lower triangle */
      const int s = i;
      i = j;
      j = s;
    }
    if (flags & TAUCS_PATTERN) v = (i == j) ? (double)(nrows + 1) : -1.0;
    if (trip_push(&t, i, j, v) != 0) {
      taucs_printf("taucs_ccs_read_mtx: out of memory\n");
      fclose(f);
      trip_free(&t);
      return NULL;
    }
    if (i > nrows) nrows = i;
    if (j > ncols) ncols = j;
  }
  fclose(f);
  A = trip_to_ccs(&t, nrows, ncols,
                  TAUCS_DOUBLE | ((flags & TAUCS_SYMMETRIC) ? TAUCS_SYMMETRIC | TAUCS_LOWER : 0));
  trip_free(&t);
  if (!A) taucs_printf("taucs_ccs_read_mtx: out of memory\n");
  return A;
}

/* ---- ccs text ------------------------------------------------------------ */

taucs_ccs_matrix* taucs_ccs_read_ccs(char* filename, int flags)
{
  FILE* f;
  int N, i, nnz;
  taucs_ccs_matrix* A = NULL;
  if (!(flags & TAUCS_DOUBLE)) {
    taucs_printf("taucs_ccs_read_ccs: only double matrices are supported\n");
    return NULL;
  }
  f = fopen(filename, "r");
  if (!f) {
    taucs_printf("taucs_ccs_read_ccs: could not open ccs file %s\n", filename);
    return NULL;
  }
  if (fscanf(f, "%d", &N) != 1 || N < 0) goto bad;
  {
    int* ptr = (int*)malloc(sizeof(int) * ((size_t)N + 1));
    if (!ptr) goto bad;
    for (i = 0; i <= N; i++)
      if (fscanf(f, "%d", &ptr[i]) != 1 || (i > 0 && ptr[i] < ptr[i - 1])) {
        free(ptr);
        goto bad;
      }
    nnz = ptr[N];
    A = taucs_ccs_create(N, N, nnz, TAUCS_DOUBLE);
    if (!A) {
      free(ptr);
      goto bad;
    }
    memcpy(A->colptr, ptr, sizeof(int) * ((size_t)N + 1));
    free(ptr);
  }
  for (i = 0; i < nnz; i++)
    if (fscanf(f, "%d", &A->rowind[i]) != 1 || A->rowind[i] < 0 || A->rowind[i] >= N) goto bad;
  for (i = 0; i < nnz; i++)
    if (fscanf(f, "%lg", &A->values.d[i]) != 1) goto bad;
  fclose(f);
  A->flags = TAUCS_DOUBLE;
  if (flags & TAUCS_SYMMETRIC) {
    int j, p;
    for (j = 0; j < N; j++)
      for (p = A->colptr[j]; p < A->colptr[j + 1]; p++)
        if (A->rowind[p] < j) {
          taucs_printf("taucs_ccs_read_ccs: symmetric matrix with an upper entry\n");
          taucs_ccs_free(A);
          return NULL;
        }
    A->flags |= TAUCS_SYMMETRIC | TAUCS_LOWER;
  }
  return A;
bad:
  taucs_printf("taucs_ccs_read_ccs: %s is truncated or malformed\n", filename);
  if (A) taucs_ccs_free(A);
  fclose(f);
  return NULL;
}

taucs_ccs_matrix* taucs_ccs_read_hb(char* filename, int flags)
{
  (void)flags;
  taucs_printf("taucs_ccs_read_hb: the Harwell-Boeing reader is not linked (%s)\n", filename);
  return NULL;
}

/* ---- vectors -------------------------------------------------------------- */

static size_t elem_size(int flags)
{
  if (flags & TAUCS_DOUBLE) return sizeof(double);
  if (flags & TAUCS_SINGLE) return sizeof(float);
  if (flags & TAUCS_DCOMPLEX) return 2 * sizeof(double);
  if (flags & TAUCS_SCOMPLEX) return 2 * sizeof(float);
  return 0;
}

void* taucs_vec_read_binary(int n, int flags, char* filename)
{
  const size_t es = elem_size(flags);
  FILE* f;
  void* v;
  if (!es || n < 0) {
    taucs_printf("taucs_vec_read_binary: unknown datatype\n");
    return NULL;
  }
  f = fopen(filename, "rb");
  if (!f) {
    taucs_printf("taucs_vec_read_binary: could not open %s\n", filename);
    return NULL;
  }
  v = taucs_malloc_stub(es * (size_t)(n > 0 ? n : 1));
  if (!v || !read_exact(f, v, es, (size_t)n)) {
    taucs_printf("taucs_vec_read_binary: %s is shorter than %d values\n", filename, n);
    taucs_free_stub(v);
    fclose(f);
    return NULL;
  }
  fclose(f);
  return v;
}

int taucs_vec_write_binary(int n, int flags, void* v, char* filename)
{
  const size_t es = elem_size(flags);
  FILE* f;
  if (!es || n < 0) {
    taucs_printf("taucs_vec_write_binary: unknown datatype\n");
    return -1;
  }
  f = fopen(filename, "wb");
  if (!f) {
    taucs_printf("taucs_vec_write_binary: could not open %s\n", filename);
    return -1;
  }
  if (fwrite(v, es, (size_t)n, f) != (size_t)n) {
    fclose(f);
    return -1;
  }
  return fclose(f) == 0 ? 0 : -1;
}
