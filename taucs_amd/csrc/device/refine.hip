// FP64 residuals, norms and refinement on the device.
//
// r = b - A x is the symmetric SpMV of src/taucs_ccs_ops.c:299-344 (lower
// storage, mirrored) evaluated on a full-pattern CSR copy so each row is one
// thread's deterministic sum.  Refinement follows the reference's single
// precision precedent (progs/direct.c:656-663: FP64-accumulated residual,
// correction solved with the low-precision factor) and its preconditioned-CG
// route (src/taucs_linsolve.c:561-567: any factor as the preconditioner of
// taucs_conjugate_gradients, src/taucs_iter.c:65-222).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "kernels.h"
#include "../host/internal.h"
#include "../../../include/taucs_amd.h"

#define HIPOK(x)                                                                     \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      taucs_printf((char*)"taucs_amd: HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                   __FILE__, __LINE__);                                              \
      return -1;                                                                     \
    }                                                                                \
  } while (0)

namespace tamd {

// y(:,r) = b(:,r) - A x(:,r)     (b may be null: y = A x)
// One thread per row; A's row is read once per group of four right-hand
// sides, and each sum runs over the row in stored order.
__global__ void k_csr_residual(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                               const double* __restrict__ v, const double* __restrict__ x,
                               int ldx, const double* __restrict__ b, int ldb,
                               double* __restrict__ y, int ldy, int nrhs)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k0 = rp[i], k1 = rp[i + 1];
  for (int r0 = 0; r0 < nrhs; r0 += 4) {
    const int nr = nrhs - r0 < 4 ? nrhs - r0 : 4;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const double* xr = x + (long long)r0 * ldx;
    for (int k = k0; k < k1; k++) {
      const double a = v[k];
      const long long c = ci[k];
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (q < nr) acc[q] += a * xr[c + (long long)q * ldx];
    }
#pragma unroll
    for (int q = 0; q < 4; q++)
      if (q < nr) {
        const long long r = r0 + q;
        y[i + r * ldy] = b ? b[i + r * ldb] - acc[q] : acc[q];
      }
  }
}

// |A| row sums (for |A|_inf)
__global__ void k_csr_absrow(int n, const int* rp, const double* v, double* out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int k = rp[i]; k < rp[i + 1]; k++) acc += fabs(v[k]);
  out[i] = acc;
}

// max that propagates NaN (fmax would drop it and hide a diverged iterate)
__device__ __forceinline__ double maxnan(double x, double y) { return (x != x || x > y) ? x : y; }

// Per-column partial reductions: mode 0 sum(a*b), 1 max|a|, 2 sum(a*a).
// partial[(blk) + r*gridDim.x]
__global__ __launch_bounds__(256) void k_colreduce(int n, const double* __restrict__ a, int lda,
                                                   const double* __restrict__ b, int ldb,
                                                   int mode, double* __restrict__ partial)
{
  __shared__ double red[256];
  const int r = blockIdx.y;
  double acc = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    double av = a[i + (long long)r * lda];
    if (mode == 0) acc += av * b[i + (long long)r * ldb];
    else if (mode == 1) acc = maxnan(acc, fabs(av));
    else acc += av * av;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      red[threadIdx.x] = mode == 1 ? maxnan(red[threadIdx.x], red[threadIdx.x + off])
                                   : red[threadIdx.x] + red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x + (long long)r * gridDim.x] = red[0];
}

__global__ void k_colfinal(int nblk, int mode, const double* partial, double* out, int nrhs)
{
  for (int r = threadIdx.x; r < nrhs; r += blockDim.x) {
    double acc = 0.0;
    for (int q = 0; q < nblk; q++) {
      double v = partial[q + (long long)r * nblk];
      acc = mode == 1 ? maxnan(acc, v) : acc + v;
    }
    out[r] = acc;
  }
}

// y(:,r) += alpha[r] * x(:,r)   (sign: +1 or -1 applied to alpha)
__global__ void k_axpy_col(int n, const double* __restrict__ alpha, double sign,
                           const double* __restrict__ x, double* __restrict__ y, int nrhs)
{
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * nrhs) return;
  int r = (int)(e / n);
  y[e] += sign * alpha[r] * x[e];
}

// p = z + beta p
__global__ void k_xpby_col(int n, const double* __restrict__ z, const double* __restrict__ beta,
                           double* __restrict__ p, int nrhs)
{
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * nrhs) return;
  int r = (int)(e / n);
  p[e] = z[e] + beta[r] * p[e];
}

// out[r] = (den[r] != 0) ? num[r]/den[r] : 0, optionally masked by done[r]
__global__ void k_ratio(const double* num, const double* den, const int* done, double* out, int nrhs)
{
  for (int r = threadIdx.x; r < nrhs; r += blockDim.x)
    out[r] = (den[r] != 0.0 && !(done && done[r])) ? num[r] / den[r] : 0.0;
}

}  // namespace tamd

using namespace tamd;

struct RefineCtx {
  // the pattern the CSR structure was built from (compared by content)
  std::vector<int> pat_cp, pat_ri;
  int n = 0, nnz = 0;      // nnz: stored (lower) entries of A
  int* rp = nullptr;
  int* ci = nullptr;
  int* src = nullptr;      // CSR slot -> index of the stored entry of A
  double* v = nullptr;
  double* avals = nullptr; // device copy of A's values when they are not the factor's
  // values identity: the device values the CSR copy was gathered from
  const double* vals_from = nullptr;
  unsigned long long gen = 0;
  int nfull = 0;           // CSR entries
  bool anorm_stale = true;
  double anorm_inf = 0.0;
  int cap = 0;  // rhs capacity of the work vectors
  double *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *x = nullptr, *b = nullptr;
  double *partial = nullptr, *s1 = nullptr, *s2 = nullptr, *s3 = nullptr;
  double *mask = nullptr, *rz = nullptr, *rz_new = nullptr, *scal = nullptr;
  int* done = nullptr;
};

struct CtxEntry { void* key; RefineCtx* ctx; };
// Contexts of all factors; the list is shared by threads, a context is used
// by one thread at a time (the caller's contract for one factor).
static std::vector<CtxEntry> g_ctx;
static std::mutex g_ctx_mu;

extern "C" int tamd_dev_malloc(void** p, size_t bytes);
extern "C" void tamd_dev_free_ptr(void* p);

static void ctx_free(RefineCtx* c)
{
  if (!c) return;
  void* ps[] = {c->rp, c->ci, c->src, c->v, c->avals, c->r, c->z, c->p, c->q, c->x, c->b, c->partial,
                c->s1, c->s2, c->s3, c->mask, c->rz, c->rz_new, c->scal, c->done};
  for (void* p : ps) tamd_dev_free_ptr(p);
  delete c;
}

extern "C" void tamd_refine_forget(void* key)
{
  RefineCtx* c = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    for (size_t i = 0; i < g_ctx.size(); i++)
      if (g_ctx[i].key == key) {
        c = g_ctx[i].ctx;
        g_ctx.erase(g_ctx.begin() + i);
        break;
      }
  }
  ctx_free(c);
}

static int dmalloc(void** p, size_t bytes) { return tamd_dev_malloc(p, bytes); }

// Build (or reuse) the full-pattern CSR of A on the device.
extern "C" int tamd_dev_values_of(tamd_factor* F, const double** d_avals, const double** host_src,
                                  unsigned long long* gen);

// v[e] = vals[src[e]]
__global__ void k_gather_vals(int nnz, const int* __restrict__ src, const double* __restrict__ vals,
                              double* __restrict__ v)
{
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nnz) v[e] = vals[src[e]];
}

// The full-pattern CSR of A (built once per pattern) and its values (gathered
// on the device whenever the factor was refactored or A's values changed).
static RefineCtx* get_ctx(tamd_factor* F, const taucs_ccs_matrix* A, int nrhs, bool factored_values,
                          hipStream_t st)
{
  void* key = F;
  RefineCtx* c = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    for (auto& e : g_ctx)
      if (e.key == key) c = e.ctx;
  }
  if (c && (c->n != A->n || c->nnz != A->colptr[A->n]
            || memcmp(c->pat_cp.data(), A->colptr, sizeof(int) * (size_t)(A->n + 1)) != 0
            || memcmp(c->pat_ri.data(), A->rowind, sizeof(int) * (size_t)c->nnz) != 0)) {
    tamd_refine_forget(key);
    c = nullptr;
  }
  if (!c) {
    c = new RefineCtx();
    int n = A->n, nl = A->colptr[n];
    std::vector<int> cnt(n + 1, 0);
    for (int j = 0; j < n; j++)
      for (int k = A->colptr[j]; k < A->colptr[j + 1]; k++) {
        int i = A->rowind[k];
        cnt[i + 1]++;
        if (i != j) cnt[j + 1]++;
      }
    for (int i = 0; i < n; i++) cnt[i + 1] += cnt[i];
    int nnz = cnt[n];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1), ci(nnz), src(nnz);
    for (int j = 0; j < n; j++)
      for (int k = A->colptr[j]; k < A->colptr[j + 1]; k++) {
        int i = A->rowind[k];
        ci[fill[i]] = j; src[fill[i]++] = k;
        if (i != j) { ci[fill[j]] = i; src[fill[j]++] = k; }
      }
    c->pat_cp.assign(A->colptr, A->colptr + n + 1);
    c->pat_ri.assign(A->rowind, A->rowind + nl);
    c->n = n; c->nnz = nl;
    c->nfull = nnz;
    if (dmalloc((void**)&c->rp, sizeof(int) * (n + 1)) || dmalloc((void**)&c->ci, sizeof(int) * nnz)
        || dmalloc((void**)&c->src, sizeof(int) * nnz) || dmalloc((void**)&c->v, sizeof(double) * nnz)) {
      ctx_free(c);
      return nullptr;
    }
    if (hipMemcpyAsync(c->rp, cnt.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, st) != hipSuccess
        || hipMemcpyAsync(c->ci, ci.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, st) != hipSuccess
        || hipMemcpyAsync(c->src, src.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, st) != hipSuccess
        || hipStreamSynchronize(st) != hipSuccess) {
      ctx_free(c);
      return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    g_ctx.push_back({key, c});
  }
  // values: the factor's own device copy when the caller vouches that A's
  // values are those of the last numeric factorization, else A's host values
  // uploaded on every call (A may have changed in place, or be a new matrix at
  // a recycled address)
  const double* dvals = nullptr;
  const double* hsrc = nullptr;
  unsigned long long gen = 0;
  if (tamd_dev_values_of(F, &dvals, &hsrc, &gen) != 0) return nullptr;
  const double* from;
  if (factored_values) {
    from = dvals;
  } else {
    if (!c->avals && dmalloc((void**)&c->avals, sizeof(double) * (size_t)(c->nnz > 0 ? c->nnz : 1)))
      return nullptr;
    if (hipMemcpyAsync(c->avals, A->values.d, sizeof(double) * (size_t)c->nnz, hipMemcpyHostToDevice,
                       st) != hipSuccess)
      return nullptr;
    from = c->avals;
    gen = ~0ULL;  // always refresh
  }
  if (from != c->vals_from || gen != c->gen || gen == ~0ULL) {
    hipLaunchKernelGGL(k_gather_vals, dim3((c->nfull + 255) / 256), dim3(256), 0, st, c->nfull, c->src,
                       from, c->v);
    c->vals_from = from;
    c->gen = gen;
    c->anorm_stale = true;
  }
  if (c->cap < nrhs) {
    void** vs[] = {(void**)&c->r, (void**)&c->z, (void**)&c->p, (void**)&c->q, (void**)&c->x, (void**)&c->b};
    c->cap = 0;
    for (void** pp : vs) {
      tamd_dev_free_ptr(*pp);
      *pp = nullptr;
      if (dmalloc(pp, sizeof(double) * (size_t)c->n * nrhs)) return nullptr;
    }
    void** ss[] = {(void**)&c->partial, (void**)&c->s1,     (void**)&c->s2,   (void**)&c->s3,
                   (void**)&c->mask,    (void**)&c->rz,     (void**)&c->rz_new, (void**)&c->scal,
                   (void**)&c->done};
    for (int i = 0; i < 9; i++) {
      size_t sz = i == 0 ? sizeof(double) * 256 * (size_t)nrhs
                         : (i == 8 ? sizeof(int) : sizeof(double)) * (size_t)nrhs;
      tamd_dev_free_ptr(*ss[i]);
      *ss[i] = nullptr;
      if (dmalloc(ss[i], sz)) return nullptr;
    }
    c->cap = nrhs;
  }
  return c;
}

// out[r] = reduce over column r; blocking (returns after the stream syncs if host != null)
static void colreduce(hipStream_t st, RefineCtx* c, const double* a, const double* b, int mode,
                      double* dout, int nrhs)
{
  int nblk = std::min(256, (c->n + 255) / 256);
  if (nblk < 1) nblk = 1;
  hipLaunchKernelGGL(k_colreduce, dim3(nblk, nrhs), dim3(256), 0, st, c->n, a, c->n, b, c->n, mode,
                     c->partial);
  hipLaunchKernelGGL(k_colfinal, dim3(1), dim3(64), 0, st, nblk, mode, c->partial, dout, nrhs);
}

extern "C" int tamd_dev_stream_of(tamd_factor* F, hipStream_t* st);
extern "C" int tamd_dev_solve_device(tamd_factor* F, int nrhs, double* dX, int ldx, const double* dB, int ldb);
extern "C" int tamd_dev_sweep_error(tamd_factor* F);

// Backward errors of x: r = b - A x,  eta = |r|_inf / (|A|_inf |x|_inf + |b|_inf)
static int residual_and_error(hipStream_t st, RefineCtx* c, const double* x, const double* b,
                              double* r, int nrhs, std::vector<double>& eta,
                              std::vector<double>& relres)
{
  int n = c->n;
  hipLaunchKernelGGL(k_csr_residual, dim3((n + 255) / 256), dim3(256), 0, st, n, c->rp, c->ci, c->v, x,
                     n, b, n, r, n, nrhs);
  std::vector<double> rinf(nrhs), xinf(nrhs), binf(nrhs), r2(nrhs), b2(nrhs);
  colreduce(st, c, r, nullptr, 1, c->s1, nrhs);
  HIPOK(hipMemcpyAsync(rinf.data(), c->s1, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
  colreduce(st, c, x, nullptr, 1, c->s2, nrhs);
  HIPOK(hipMemcpyAsync(xinf.data(), c->s2, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
  colreduce(st, c, b, nullptr, 1, c->s3, nrhs);
  HIPOK(hipMemcpyAsync(binf.data(), c->s3, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
  HIPOK(hipStreamSynchronize(st));
  colreduce(st, c, r, nullptr, 2, c->s1, nrhs);
  HIPOK(hipMemcpyAsync(r2.data(), c->s1, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
  colreduce(st, c, b, nullptr, 2, c->s2, nrhs);
  HIPOK(hipMemcpyAsync(b2.data(), c->s2, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
  HIPOK(hipStreamSynchronize(st));
  eta.resize(nrhs);
  relres.resize(nrhs);
  for (int q = 0; q < nrhs; q++) {
    double den = c->anorm_inf * xinf[q] + binf[q];
    eta[q] = den > 0 ? rinf[q] / den : 0.0;
    relres[q] = b2[q] > 0 ? sqrt(r2[q]) / sqrt(b2[q]) : sqrt(r2[q]);
  }
  return 0;
}

// The refinement loop on device vectors: c->b holds B, the result is left in
// c->x.  Flags (in `method`): TAUCS_AMD_REFINE_FACTORED_VALUES,
// TAUCS_AMD_REFINE_STOP_RELRES.
static int refine_core(tamd_factor* F, RefineCtx* c, hipStream_t st, int nrhs, int method, double tol,
                       int maxit, taucs_amd_refine_info* info)
{
  const bool stop_relres = (method & TAUCS_AMD_REFINE_STOP_RELRES) != 0;
  method &= 0xff;
  int n = c->n;
  if (c->anorm_stale) {  // |A|_inf = max row sum of |a_ij|, on the device
    hipLaunchKernelGGL(k_csr_absrow, dim3((n + 255) / 256), dim3(256), 0, st, n, c->rp, c->v, c->r);
    colreduce(st, c, c->r, nullptr, 1, c->s1, 1);
    HIPOK(hipMemcpyAsync(&c->anorm_inf, c->s1, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPOK(hipStreamSynchronize(st));
    c->anorm_stale = false;
  }
  if (maxit <= 0) maxit = 100;
  std::vector<double> eta, relres;
  int it = 0;
  bool conv = false;
  const unsigned gv = (unsigned)(((long long)n * nrhs + 255) / 256);
  // Every right-hand side follows its own recurrence and stops changing once
  // its backward error reaches tol (its step is masked to zero), so a column's
  // result does not depend on which other columns are solved with it.
  std::vector<int> done(nrhs, 0), bad(nrhs, 0), stop(nrhs, 0);
  std::vector<double> mask(nrhs, 1.0), best(nrhs, 1e300);
  // A column diverges when its residual is not finite or has grown 1000x
  // above the smallest seen: a diverging iterate can grow until |A||x|
  // dwarfs the residual, and its backward error then looks small.
  auto diverging = [&](int q) {
    if (!(relres[q] <= 1e300) || !(eta[q] <= 1.0)) return true;
    best[q] = std::min(best[q], relres[q]);
    return relres[q] > 1e3 * best[q];
  };
  if (method == 0) {
    // x = M^-1 b; repeat r = b - A x, x += M^-1 r
    if (tamd_dev_solve_device(F, nrhs, c->x, n, c->b, n)) return -1;
    for (;;) {
      if (residual_and_error(st, c, c->x, c->b, c->r, nrhs, eta, relres)) return -1;
      bool all = true, settled = true;
      for (int q = 0; q < nrhs; q++) {
        if (!done[q] && !bad[q] && diverging(q)) bad[q] = 1;
        if (!bad[q] && (stop_relres ? relres[q] : eta[q]) <= tol) done[q] = 1;
        mask[q] = (done[q] || bad[q]) ? 0.0 : 1.0;
        all = all && done[q];
        settled = settled && (done[q] || bad[q]);
      }
      if (all) { conv = true; break; }
      if (settled) break;
      if (it >= maxit) break;
      if (tamd_dev_solve_device(F, nrhs, c->z, n, c->r, n)) return -1;
      HIPOK(hipMemcpyAsync(c->mask, mask.data(), sizeof(double) * nrhs, hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_axpy_col, dim3(gv), dim3(256), 0, st, n, c->mask, 1.0, c->z, c->x, nrhs);
      it++;
    }
  } else {
    // preconditioned CG, one independent recurrence per right-hand side
    HIPOK(hipMemsetAsync(c->x, 0, sizeof(double) * n * nrhs, st));
    HIPOK(hipMemcpyAsync(c->r, c->b, sizeof(double) * n * nrhs, hipMemcpyDeviceToDevice, st));
    if (tamd_dev_solve_device(F, nrhs, c->z, n, c->r, n)) return -1;
    HIPOK(hipMemcpyAsync(c->p, c->z, sizeof(double) * n * nrhs, hipMemcpyDeviceToDevice, st));
    HIPOK(hipMemsetAsync(c->done, 0, sizeof(int) * nrhs, st));
    colreduce(st, c, c->r, c->z, 0, c->rz, nrhs);
    for (;;) {
      // q = A p ; alpha = rz / (p.q)
      hipLaunchKernelGGL(k_csr_residual, dim3((n + 255) / 256), dim3(256), 0, st, n, c->rp, c->ci, c->v,
                         c->p, n, (const double*)nullptr, n, c->q, n, nrhs);
      colreduce(st, c, c->p, c->q, 0, c->s1, nrhs);
      hipLaunchKernelGGL(k_ratio, dim3(1), dim3(64), 0, st, c->rz, c->s1, c->done, c->scal, nrhs);
      hipLaunchKernelGGL(k_axpy_col, dim3(gv), dim3(256), 0, st, n, c->scal, 1.0, c->p, c->x, nrhs);
      hipLaunchKernelGGL(k_axpy_col, dim3(gv), dim3(256), 0, st, n, c->scal, -1.0, c->q, c->r, nrhs);
      it++;
      // convergence on the true residual every iteration (k small)
      if (residual_and_error(st, c, c->x, c->b, c->q, nrhs, eta, relres)) return -1;
      bool all = true, settled = true;
      for (int q = 0; q < nrhs; q++) {
        if (!done[q] && !bad[q] && diverging(q)) bad[q] = 1;
        if (!bad[q] && (stop_relres ? relres[q] : eta[q]) <= tol) done[q] = 1;
        all = all && done[q];
        settled = settled && (done[q] || bad[q]);
        stop[q] = done[q] || bad[q];  // its recurrence stops (the device mask)
      }
      if (all) { conv = true; break; }
      if (settled || it >= maxit) break;
      HIPOK(hipMemcpyAsync(c->done, stop.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, st));
      if (tamd_dev_solve_device(F, nrhs, c->z, n, c->r, n)) return -1;
      colreduce(st, c, c->r, c->z, 0, c->rz_new, nrhs);
      hipLaunchKernelGGL(k_ratio, dim3(1), dim3(64), 0, st, c->rz_new, c->rz, (const int*)nullptr, c->scal,
                         nrhs);
      hipLaunchKernelGGL(k_xpby_col, dim3(gv), dim3(256), 0, st, n, c->z, c->scal, c->p, nrhs);
      HIPOK(hipMemcpyAsync(c->rz, c->rz_new, sizeof(double) * nrhs, hipMemcpyDeviceToDevice, st));
    }
    if (residual_and_error(st, c, c->x, c->b, c->r, nrhs, eta, relres)) return -1;
  }
  if (tamd_dev_sweep_error(F)) return -1;
  if (info) {
    info->iterations = it;
    info->converged = conv;
    info->backward_error = eta.empty() ? 0 : *std::max_element(eta.begin(), eta.end());
    info->rel_residual = relres.empty() ? 0 : *std::max_element(relres.begin(), relres.end());
  }
  return 0;
}

static RefineCtx* refine_setup(tamd_factor* F, taucs_ccs_matrix* A, int nrhs, int method, hipStream_t* st)
{
  if (!F || !F->has_numeric || !A || A->n != F->n || nrhs <= 0) return nullptr;
  if ((method & 0xff) > 1) return nullptr;
  if (tamd_dev_stream_of(F, st) != 0) return nullptr;
  return get_ctx(F, A, nrhs, (method & TAUCS_AMD_REFINE_FACTORED_VALUES) != 0, *st);
}

extern "C" int taucs_amd_solve_refined(void* vL, taucs_ccs_matrix* A, int nrhs, double* X, int ldx,
                                       const double* B, int ldb, int method, double tol, int maxit,
                                       taucs_amd_refine_info* info)
{
  tamd_factor* F = (tamd_factor*)vL;
  hipStream_t st;
  RefineCtx* c = refine_setup(F, A, nrhs, method, &st);
  if (!c) return -1;
  const int n = c->n;
  HIPOK(hipMemcpy2DAsync(c->b, sizeof(double) * n, B, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                         hipMemcpyHostToDevice, st));
  if (refine_core(F, c, st, nrhs, method, tol, maxit, info) != 0) return -1;
  HIPOK(hipMemcpy2DAsync(X, sizeof(double) * ldx, c->x, sizeof(double) * n, sizeof(double) * n, nrhs,
                         hipMemcpyDeviceToHost, st));
  HIPOK(hipStreamSynchronize(st));
  return 0;
}

extern "C" int taucs_amd_solve_refined_device(void* vL, taucs_ccs_matrix* A, int nrhs, double* dX,
                                              int ldx, const double* dB, int ldb, int method, double tol,
                                              int maxit, taucs_amd_refine_info* info)
{
  tamd_factor* F = (tamd_factor*)vL;
  hipStream_t st;
  RefineCtx* c = refine_setup(F, A, nrhs, method, &st);
  if (!c) return -1;
  const int n = c->n;
  HIPOK(hipMemcpy2DAsync(c->b, sizeof(double) * n, dB, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                         hipMemcpyDeviceToDevice, st));
  if (refine_core(F, c, st, nrhs, method, tol, maxit, info) != 0) return -1;
  HIPOK(hipMemcpy2DAsync(dX, sizeof(double) * ldx, c->x, sizeof(double) * n, sizeof(double) * n, nrhs,
                         hipMemcpyDeviceToDevice, st));
  HIPOK(hipStreamSynchronize(st));
  return 0;
}
