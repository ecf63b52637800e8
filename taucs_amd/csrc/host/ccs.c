/*
 * CSC containers, mesh generators, symmetric permutation, SpMV and vector
 * helpers (host side of the ABI).
 *
 *   taucs_ccs_create/free                  src/taucs_ccs_base.c:21-126
 *   taucs_ccs_generate_mesh2d / mesh3d     src/taucs_ccs_generators.c:89-261
 *   taucs_dccs_permute_symmetrically       src/taucs_ccs_ops.c:174-257
 *   taucs_dccs_times_vec                   src/taucs_ccs_ops.c:299-344
 *   taucs_vec_*                            src/taucs_vec_base.c:28-206
 *
 * The generators reproduce the reference's per-column entry order, because
 * that order propagates through the ordering and the permutation into the
 * supernodal row structure.  The variable-coefficient 3D operator and the
 * diagonally shifted 2D family are additions (taucs_amd_* below); their
 * random values come from one splitmix64 stream so any host reproduces them.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "internal.h"
#include "../../../include/taucs_amd.h"

taucs_ccs_matrix* taucs_ccs_create(int m, int n, int nnz, int flags)
{
  taucs_ccs_matrix* A;
  size_t esz;
  if (flags & TAUCS_DOUBLE) esz = sizeof(double);
  else if (flags & TAUCS_SINGLE) esz = sizeof(float);
  else if (flags & TAUCS_INT) esz = sizeof(int);
  else {
    taucs_printf("taucs_ccs_create: only real datatypes are supported\n");
    return NULL;
  }
  A = (taucs_ccs_matrix*)tamd_malloc(sizeof(taucs_ccs_matrix));
  if (!A) return NULL;
  A->n = n;
  A->m = m;
  A->flags = flags;
  A->colptr = (int*)tamd_malloc((size_t)(n + 1) * sizeof(int));
  A->rowind = (int*)tamd_malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
  A->values.v = tamd_malloc((size_t)(nnz > 0 ? nnz : 1) * esz);
  if (!A->colptr || !A->rowind || !A->values.v) {
    tamd_free(A->colptr);
    tamd_free(A->rowind);
    tamd_free(A->values.v);
    tamd_free(A);
    return NULL;
  }
  return A;
}

taucs_ccs_matrix* taucs_dccs_create(int m, int n, int nnz)
{
  return taucs_ccs_create(m, n, nnz, TAUCS_DOUBLE);
}

void taucs_ccs_free(taucs_ccs_matrix* A)
{
  if (!A) return;
  tamd_free(A->colptr);
  tamd_free(A->rowind);
  tamd_free(A->values.v);
  tamd_free(A);
}

/* ---- generators ---------------------------------------------------------- */

static taucs_ccs_matrix* sym_lower_create(int N, int nnz)
{
  return taucs_ccs_create(N, N, nnz, TAUCS_SYMMETRIC | TAUCS_LOWER | TAUCS_DOUBLE);
}

/* 5-point grid.  Column (x,y) holds (x,y+1), then (x+1,y), then the diagonal
   (src/taucs_ccs_generators.c:130-177). */
taucs_ccs_matrix* taucs_ccs_generate_mesh2d(int n, char* which)
{
  const double jump = 100.0;
  int aniso_x = which && !strcmp(which, "anisotropic_x");
  int aniso_y = which && !strcmp(which, "anisotropic_y");
  int dirichlet = which && !strcmp(which, "dirichlet");
  int N = n * n, x, y, ip = 0;
  taucs_ccs_matrix* A = sym_lower_create(N, 3 * N);
  if (!A) return NULL;
  for (y = 0; y < n; y++)
    for (x = 0; x < n; x++) {
      int j = x + y * n;
      double wy = aniso_y ? jump : 1.0, wx = aniso_x ? jump : 1.0, d = 0.0;
      A->colptr[j] = ip;
      if (y < n - 1) { A->rowind[ip] = x + (y + 1) * n; A->values.d[ip] = -wy; ip++; }
      if (x < n - 1) { A->rowind[ip] = (x + 1) + y * n; A->values.d[ip] = -wx; ip++; }
      if (dirichlet) {
        d = 4.0;
      } else {
        if (x > 0) d += wx;
        if (y > 0) d += wy;
        if (x < n - 1) d += wx;
        if (y < n - 1) d += wy;
        if (x == 0 && y == 0) d += 1.0;
      }
      A->rowind[ip] = j;
      A->values.d[ip] = d;
      ip++;
    }
  A->colptr[N] = ip;
  return A;
}

/* 7-point grid.  Column (x,y,z) holds x+1, y+1, z+1, then the diagonal
   (src/taucs_ccs_generators.c:229-257). */
taucs_ccs_matrix* taucs_ccs_generate_mesh3d(int X, int Y, int Z)
{
  return taucs_amd_generate_mesh3d_varcoef(X, Y, Z, 0, 0.0, 0.0, 0.0);
}

uint64_t taucs_amd_splitmix64(uint64_t* state)
{
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double taucs_amd_uniform01(uint64_t* state)
{
  return (double)(taucs_amd_splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
}

void taucs_amd_random_uniform(int n, uint64_t seed, double lo, double hi, double* out)
{
  uint64_t st = seed;
  int i;
  for (i = 0; i < n; i++) out[i] = lo + (hi - lo) * taucs_amd_uniform01(&st);
}

/* seed == 0: unit conductivities and the reference's corner anchor (this is
   exactly mesh3d).  Otherwise each edge draws c = 10^(lo + (hi-lo) u) in the
   generator's visit order (z, y, x loops; x+1, y+1, z+1 edges), off-diagonal
   -c, diagonal = sum of incident c + shift. */
taucs_ccs_matrix* taucs_amd_generate_mesh3d_varcoef(int X, int Y, int Z, uint64_t seed,
                                                    double log10_lo, double log10_hi,
                                                    double shift)
{
  int N = X * Y * Z, x, y, z, ip = 0;
  double *wx, *wy, *wz;
  uint64_t st = seed;
  taucs_ccs_matrix* A = sym_lower_create(N, 4 * N);
  if (!A) return NULL;
  wx = (double*)tamd_malloc(sizeof(double) * (size_t)N * 3);
  if (!wx) { taucs_ccs_free(A); return NULL; }
  wy = wx + N;
  wz = wy + N;
  for (z = 0; z < Z; z++)
    for (y = 0; y < Y; y++)
      for (x = 0; x < X; x++) {
        int j = z * X * Y + y * X + x;
        wx[j] = wy[j] = wz[j] = 0.0;
        if (x < X - 1) wx[j] = seed ? pow(10.0, log10_lo + (log10_hi - log10_lo) * taucs_amd_uniform01(&st)) : 1.0;
        if (y < Y - 1) wy[j] = seed ? pow(10.0, log10_lo + (log10_hi - log10_lo) * taucs_amd_uniform01(&st)) : 1.0;
        if (z < Z - 1) wz[j] = seed ? pow(10.0, log10_lo + (log10_hi - log10_lo) * taucs_amd_uniform01(&st)) : 1.0;
      }
  for (z = 0; z < Z; z++)
    for (y = 0; y < Y; y++)
      for (x = 0; x < X; x++) {
        int j = z * X * Y + y * X + x;
        double d = 0.0;
        A->colptr[j] = ip;
        if (x < X - 1) { A->rowind[ip] = j + 1;     A->values.d[ip] = -wx[j]; ip++; }
        if (y < Y - 1) { A->rowind[ip] = j + X;     A->values.d[ip] = -wy[j]; ip++; }
        if (z < Z - 1) { A->rowind[ip] = j + X * Y; A->values.d[ip] = -wz[j]; ip++; }
        /* same summation order as the reference: +x, +y, +z, -x, -y, -z */
        if (x < X - 1) d += wx[j];
        if (y < Y - 1) d += wy[j];
        if (z < Z - 1) d += wz[j];
        if (x > 0) d += wx[j - 1];
        if (y > 0) d += wy[j - X];
        if (z > 0) d += wz[j - X * Y];
        if (seed) d += shift;
        else if (x == 0 && y == 0 && z == 0) d += 1.0;
        A->rowind[ip] = j;
        A->values.d[ip] = d;
        ip++;
      }
  A->colptr[N] = ip;
  tamd_free(wx);
  return A;
}

/* mesh2d(n, "neumann") plus an i.i.d. uniform[0,1) shift on every diagonal
   entry, drawn in column order. */
taucs_ccs_matrix* taucs_amd_generate_mesh2d_shifted(int n, uint64_t seed)
{
  taucs_ccs_matrix* A = taucs_ccs_generate_mesh2d(n, "neumann");
  uint64_t st = seed;
  int j;
  if (!A) return NULL;
  for (j = 0; j < A->n; j++) {
    int last = A->colptr[j + 1] - 1; /* diagonal is stored last */
    A->values.d[last] += taucs_amd_uniform01(&st);
  }
  return A;
}

/* ---- symmetric permutation ---------------------------------------------- */

/* Entry (i,j) of A goes to (invperm[i], invperm[j]) swapped into the lower
   triangle; the new column receives its entries in the order the source is
   traversed (src/taucs_ccs_ops.c:231-253). */
taucs_ccs_matrix* taucs_dccs_permute_symmetrically(taucs_ccs_matrix* A, int* perm, int* invperm)
{
  int n = A->n, nnz = A->colptr[n], j, ip;
  int* fill;
  taucs_ccs_matrix* B;
  (void)perm;
  if (!(A->flags & (TAUCS_SYMMETRIC | TAUCS_HERMITIAN)) || !(A->flags & TAUCS_LOWER)) {
    taucs_printf("taucs_ccs_permute_symmetrically: matrix must be symmetric, lower\n");
    return NULL;
  }
  B = taucs_ccs_create(n, n, nnz, A->flags);
  if (!B) return NULL;
  fill = (int*)tamd_calloc((size_t)n + 1, sizeof(int));
  if (!fill) { taucs_ccs_free(B); return NULL; }
  for (j = 0; j < n; j++)
    for (ip = A->colptr[j]; ip < A->colptr[j + 1]; ip++) {
      int I = invperm[A->rowind[ip]], J = invperm[j];
      fill[(I < J ? I : J) + 1]++;
    }
  B->colptr[0] = 0;
  for (j = 0; j < n; j++) B->colptr[j + 1] = B->colptr[j] + fill[j + 1];
  for (j = 0; j < n; j++) fill[j] = B->colptr[j];
  for (j = 0; j < n; j++)
    for (ip = A->colptr[j]; ip < A->colptr[j + 1]; ip++) {
      int I = invperm[A->rowind[ip]], J = invperm[j];
      int lo = I < J ? I : J, hi = I < J ? J : I;
      B->rowind[fill[lo]] = hi;
      B->values.d[fill[lo]] = A->values.d[ip];
      fill[lo]++;
    }
  tamd_free(fill);
  return B;
}

taucs_ccs_matrix* taucs_ccs_permute_symmetrically(taucs_ccs_matrix* A, int* perm, int* invperm)
{
  if (A && (A->flags & TAUCS_DOUBLE)) return taucs_dccs_permute_symmetrically(A, perm, invperm);
  taucs_printf("taucs_ccs_permute_symmetrically: only TAUCS_DOUBLE is supported\n");
  return NULL;
}

/* ---- SpMV --------------------------------------------------------------- */

void taucs_dccs_times_vec(taucs_ccs_matrix* m, taucs_double* X, taucs_double* B)
{
  int n = m->n, i, j, ip;
  for (i = 0; i < n; i++) B[i] = 0.0;
  if (m->flags & (TAUCS_SYMMETRIC | TAUCS_HERMITIAN)) {
    for (j = 0; j < n; j++)
      for (ip = m->colptr[j]; ip < m->colptr[j + 1]; ip++) {
        double a = m->values.d[ip];
        i = m->rowind[ip];
        B[i] += X[j] * a;
        if (i != j) B[j] += X[i] * a;
      }
  } else {
    for (j = 0; j < n; j++)
      for (ip = m->colptr[j]; ip < m->colptr[j + 1]; ip++)
        B[m->rowind[ip]] += X[j] * m->values.d[ip];
  }
}

void taucs_ccs_times_vec(taucs_ccs_matrix* m, void* X, void* B)
{
  if (m->flags & TAUCS_DOUBLE) taucs_dccs_times_vec(m, (double*)X, (double*)B);
}

/* ---- vectors ------------------------------------------------------------ */

/* Scaled two-norm (the LAPACK dnrm2 recurrence, so no overflow for large
   entries); the reference calls dnrm2_ (src/taucs_vec_base.c:28-53). */
taucs_double taucs_vec_norm2(int n, int flags, void* vx)
{
  double scale = 0.0, ssq = 1.0;
  int i;
  if (flags & TAUCS_DOUBLE) {
    const double* x = (const double*)vx;
    for (i = 0; i < n; i++) {
      if (x[i] != 0.0) {
        double a = fabs(x[i]);
        if (scale < a) { ssq = 1.0 + ssq * (scale / a) * (scale / a); scale = a; }
        else ssq += (a / scale) * (a / scale);
      }
    }
    return scale * sqrt(ssq);
  }
  if (flags & TAUCS_SINGLE) {
    const float* x = (const float*)vx;
    for (i = 0; i < n; i++) {
      if (x[i] != 0.0f) {
        double a = fabs((double)x[i]);
        if (scale < a) { ssq = 1.0 + ssq * (scale / a) * (scale / a); scale = a; }
        else ssq += (a / scale) * (a / scale);
      }
    }
    return scale * sqrt(ssq);
  }
  return NAN;
}

void taucs_vec_axpby(int n, int flags, taucs_double a, void* x, taucs_double b, void* y, void* axpby)
{
  int i;
  if (flags & TAUCS_DOUBLE) {
    double *X = (double*)x, *Y = (double*)y, *R = (double*)axpby;
    for (i = 0; i < n; i++) R[i] = a * X[i] + b * Y[i];
  } else if (flags & TAUCS_SINGLE) {
    float *X = (float*)x, *Y = (float*)y, *R = (float*)axpby;
    for (i = 0; i < n; i++) R[i] = (float)a * X[i] + (float)b * Y[i];
  }
}

void* taucs_vec_create(int n, int flags)
{
  if (flags & TAUCS_DOUBLE) return tamd_malloc((size_t)n * sizeof(double));
  if (flags & TAUCS_SINGLE) return tamd_malloc((size_t)n * sizeof(float));
  return NULL;
}

/* pv[i] = v[p[i]] (src/taucs_vec_base.c:194-199) */
void taucs_vec_permute(int n, int flags, void* v, void* pv, int p[])
{
  int i;
  if (flags & TAUCS_DOUBLE)
    for (i = 0; i < n; i++) ((double*)pv)[i] = ((double*)v)[p[i]];
  else if (flags & TAUCS_SINGLE)
    for (i = 0; i < n; i++) ((float*)pv)[i] = ((float*)v)[p[i]];
}

/* v[p[i]] = pv[i] (src/taucs_vec_base.c:201-206) */
void taucs_vec_ipermute(int n, int flags, void* pv, void* v, int p[])
{
  int i;
  if (flags & TAUCS_DOUBLE)
    for (i = 0; i < n; i++) ((double*)v)[p[i]] = ((double*)pv)[i];
  else if (flags & TAUCS_SINGLE)
    for (i = 0; i < n; i++) ((float*)v)[p[i]] = ((float*)pv)[i];
}
