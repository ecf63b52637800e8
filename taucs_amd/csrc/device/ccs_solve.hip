// taucs_ccs_solve_llt: solve L L^T x = b with a Cholesky factor in CCS form
// (the export of taucs_supernodal_factor_to_ccs, or any lower-triangular CCS
// matrix whose columns start with their diagonal entry), on the GPU.
// Reference: src/taucs_ccs_solve_llt.c:19-130 and :420-447 (host, column by
// column; only the double datatype is computed here).
//
// Both sweeps run sync-free, one wavefront per row (forward) or column
// (backward), with the readiness protocol of kernels.h: y and x start as the
// signaling-NaN pattern and every value is published once.
//   forward, over the rows of L (a transposed copy, CSR), in ticket order
//     i = 0, 1, ...: y_i = (b_i - sum_{k<i} L_ik y_k) / L_ii, each lane
//     awaiting the y_k of its entries;
//   backward, over the columns of L, in ticket order i = n-1, n-2, ...:
//     x_i = (y_i - sum_{j>i} L_ji x_j) / L_ii.
// A task waits only for tasks with smaller tickets, which are running, so
// the sweeps always finish; each lane sums its entries in order and the wave
// sum has a fixed tree, so the result is deterministic.  A row loads all the
// values it needs eight per lane at a time and waits only for those not yet
// published (on a dense separator almost all of them were, long before).
#include "kernels.h"
#include "taucs.h"

#include <algorithm>
#include <memory>
#include <new>
#include <thread>
#include <vector>

extern "C" {
void* taucs_amd_device_malloc(size_t bytes);
void taucs_amd_device_free(void* p);
int taucs_amd_memcpy_h2d(void* dst, const void* src, size_t bytes);
int taucs_amd_memcpy_d2h(void* dst, const void* src, size_t bytes);
}

namespace tamd {

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// The wait of one value, bounded by a larger limit than the supernodal
// sweeps use: on a large factor a row can wait for a chain of thousands of
// long rows (the root separator's), while the other waiting rows poll.
// Tickets guarantee progress; the bound only turns a fault into an error, and
// once one wait has marked the error word the others stop at their next poll.
#define CCS_SPIN_LIMIT (1 << 26)
__device__ __forceinline__ double await_long(const double* p, int* err)
{
  typedef Sentinel<double>::U U;
  U* a = reinterpret_cast<U*>(const_cast<double*>(p));
  U v = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (int it = 0; v == Sentinel<double>::bits; it++) {
    if (it > CCS_SPIN_LIMIT) {
      atomicOr(err, 1);
      break;
    }
    // another wait of this sweep already gave up: stop polling too
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    __builtin_amdgcn_s_sleep(2);
    v = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return __builtin_bit_cast(double, v);
}

// sum over p in [p0, p1) of a[p] * v[idx[p]], lane-strided: eight entries per
// lane per pass are loaded before any is awaited, so a row waits one memory
// round trip per pass plus the values still being computed, not one per entry.
// The entries are taken in the order that leaves the one computed last for
// the end: ascending along a row of L (its last off-diagonal entry is the
// previous row's), descending down a column (its first is the next column's),
// so a dense chain waits only in its final pass.
template <bool DESC>
__device__ __forceinline__ double dot_awaited(int p0, int p1, int lane, const int* __restrict__ idx,
                                              const double* __restrict__ a, const double* v,
                                              int* err)
{
  typedef Sentinel<double>::U U;
  double acc = 0.0;
  const int len = p1 - p0;
  for (int eb = lane; eb < len; eb += 64 * 8) {
    U w[8];
    double c[8];
    int k[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int e = eb + 64 * q, p = DESC ? p1 - 1 - e : p0 + e;
      k[q] = e < len ? idx[p] : 0;
      c[q] = e < len ? a[p] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 8; q++)
      w[q] = eb + 64 * q < len
                 ? __hip_atomic_load(reinterpret_cast<U*>(const_cast<double*>(v + k[q])),
                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                 : (U)0;
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (eb + 64 * q < len) {
        double x = __builtin_bit_cast(double, w[q]);
        if (w[q] == Sentinel<double>::bits) x = await_long(v + k[q], err);
        acc += c[q] * x;
      }
  }
  return acc;
}

__global__ void k_ccs_fill(double* __restrict__ a, long long n, int* __restrict__ words, int nw)
{
  const double s = __builtin_bit_cast(double, Sentinel<double>::bits);
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x)
    a[e] = s;
  if (blockIdx.x == 0)
    for (int e = threadIdx.x; e < nw; e += blockDim.x) words[e] = 0;
}

// rows of L: rp/ci/cv (columns increasing, the diagonal last)
__global__ __launch_bounds__(64) void k_ccs_fwd(int n, const int* __restrict__ rp,
                                                const int* __restrict__ ci,
                                                const double* __restrict__ cv,
                                                const double* __restrict__ b, double* y,
                                                int* ticket, int* err)
{
  __shared__ int sh;
  const int lane = threadIdx.x;
  if (lane == 0) sh = atomicAdd(ticket, 1);
  __syncthreads();
  const int i = sh;
  if (i >= n) return;
  const int p0 = rp[i], p1 = rp[i + 1] - 1;  // p1: the diagonal
  const double acc = wave_sum(dot_awaited<false>(p0, p1, lane, ci, cv, y, err));
  if (lane == 0) publish(y + i, (b[i] - acc) / cv[p1]);
}

// columns of L (the diagonal first)
__global__ __launch_bounds__(64) void k_ccs_bwd(int n, const int* __restrict__ cp,
                                                const int* __restrict__ ri,
                                                const double* __restrict__ v,
                                                const double* __restrict__ y, double* x,
                                                int* ticket, int* err)
{
  __shared__ int sh;
  const int lane = threadIdx.x;
  if (lane == 0) sh = atomicAdd(ticket, 1);
  __syncthreads();
  const int i = n - 1 - sh;
  if (i < 0) return;
  const int p0 = cp[i], p1 = cp[i + 1];
  const double acc = wave_sum(dot_awaited<true>(p0 + 1, p1, lane, ri, v, x, err));
  if (lane == 0) publish(x + i, (y[i] - acc) / v[p0]);
}

}  // namespace tamd

using namespace tamd;

extern "C" int taucs_ccs_solve_llt(void* vL, void* vx, void* vb)
{
  taucs_ccs_matrix* L = (taucs_ccs_matrix*)vL;
  double* x = (double*)vx;
  const double* b = (const double*)vb;
  if (!L || !x || !b) return -1;
  if (!(L->flags & TAUCS_TRIANGULAR)) {
    taucs_printf((char*)"taucs_ccs_solve_llt: factor matrix must be triangular\n");
    return -1;
  }
  if (!(L->flags & TAUCS_LOWER)) {
    taucs_printf((char*)"taucs_ccs_solve_llt: lower part must be represented\n");
    return -1;
  }
  if (!(L->flags & TAUCS_DOUBLE)) {
    taucs_printf((char*)"taucs_ccs_solve_llt: only double factors are supported\n");
    return -1;
  }
  const int n = L->n;
  if (n <= 0) return 0;
  const int nnz = L->colptr[n];
  const double t0 = taucs_wtime();
  // Host threads over contiguous column ranges of about equal entry counts.
  const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  std::vector<int> cb(nt + 1, n);
  cb[0] = 0;
  for (int t = 1; t < nt; t++) {
    const long long target = (long long)nnz * t / nt;
    cb[t] = (int)(std::lower_bound(L->colptr, L->colptr + n, (int)target) - L->colptr);
    cb[t] = std::max(cb[t], cb[t - 1]);
  }
  auto parallel = [&](auto&& body) {
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(body, t);
    body(0);
    for (auto& x : th) x.join();
  };
  // every column starts with its diagonal; the other rows are below it.
  // Each thread records its first bad column; the smallest one is reported.
  std::vector<int> bad(nt, -1), badkind(nt, 0);
  parallel([&](int t) {
    for (int j = cb[t]; j < cb[t + 1] && bad[t] < 0; j++) {
      const int p0 = L->colptr[j], p1 = L->colptr[j + 1];
      if (p1 <= p0 || L->rowind[p0] != j) {
        bad[t] = j;
        badkind[t] = 1;
        break;
      }
      for (int p = p0 + 1; p < p1; p++)
        if (L->rowind[p] <= j || L->rowind[p] >= n) {
          bad[t] = j;
          badkind[t] = 2;
          break;
        }
    }
  });
  for (int t = 0; t < nt; t++)
    if (bad[t] >= 0) {
      taucs_printf(badkind[t] == 1
                       ? (char*)"taucs_ccs_solve_llt: column %d does not start with its diagonal\n"
                       : (char*)"taucs_ccs_solve_llt: column %d has an entry outside the lower triangle\n",
                   bad[t]);
      return -1;
    }
  // the rows of L (transpose by counting sort; columns increase along a row,
  // so the diagonal is each row's last entry).  Thread t's entries of a row
  // follow those of threads < t, which hold smaller columns: the order is
  // the serial one.
  // ci, cv: every slot is written by the scatter below, so no zero fill
  std::vector<int> rp(n + 1, 0);
  std::unique_ptr<int[]> ci(new (std::nothrow) int[nnz > 0 ? nnz : 1]);
  std::unique_ptr<double[]> cv(new (std::nothrow) double[nnz > 0 ? nnz : 1]);
  if (!ci || !cv) {
    taucs_printf((char*)"taucs_ccs_solve_llt: out of memory\n");
    return -1;
  }
  std::vector<int> cnt((size_t)nt * n, 0);
  parallel([&](int t) {
    int* c = cnt.data() + (size_t)t * n;
    for (int p = L->colptr[cb[t]]; p < L->colptr[cb[t + 1]]; p++) c[L->rowind[p]]++;
  });
  for (int i = 0; i < n; i++) {
    int run = rp[i];
    for (int t = 0; t < nt; t++) {
      const int c = cnt[(size_t)t * n + i];
      cnt[(size_t)t * n + i] = run;  // thread t's next slot in row i
      run += c;
    }
    rp[i + 1] = run;
  }
  parallel([&](int t) {
    int* next = cnt.data() + (size_t)t * n;
    for (int j = cb[t]; j < cb[t + 1]; j++)
      for (int p = L->colptr[j]; p < L->colptr[j + 1]; p++) {
        const int q = next[L->rowind[p]]++;
        ci[q] = j;
        cv[q] = L->values.d[p];
      }
  });
  const double t1 = taucs_wtime();
  const size_t ib = sizeof(int), db = sizeof(double);
  void* buf[10] = {nullptr};
  buf[0] = taucs_amd_device_malloc(ib * (n + 1));   // colptr
  buf[1] = taucs_amd_device_malloc(ib * nnz);       // rowind
  buf[2] = taucs_amd_device_malloc(db * nnz);       // values
  buf[3] = taucs_amd_device_malloc(ib * (n + 1));   // row pointers
  buf[4] = taucs_amd_device_malloc(ib * nnz);       // column indices
  buf[5] = taucs_amd_device_malloc(db * nnz);       // row values
  buf[6] = taucs_amd_device_malloc(db * n);         // b
  buf[7] = taucs_amd_device_malloc(db * 2 * n);     // y, x
  buf[8] = taucs_amd_device_malloc(ib * 4);         // tickets, error word
  int rc = 0;
  for (int k = 0; k < 9; k++)
    if (!buf[k]) rc = -1;
  if (rc == 0)
    rc = taucs_amd_memcpy_h2d(buf[0], L->colptr, ib * (n + 1)) ||
         taucs_amd_memcpy_h2d(buf[1], L->rowind, ib * nnz) ||
         taucs_amd_memcpy_h2d(buf[2], L->values.d, db * nnz) ||
         taucs_amd_memcpy_h2d(buf[3], rp.data(), ib * (n + 1)) ||
         taucs_amd_memcpy_h2d(buf[4], ci.get(), ib * nnz) ||
         taucs_amd_memcpy_h2d(buf[5], cv.get(), db * nnz) ||
         taucs_amd_memcpy_h2d(buf[6], b, db * n);
  int words[4] = {0, 0, 0, 0};
  const double t2 = taucs_wtime();
  if (rc == 0) {
    double* y = (double*)buf[7];
    int* w = (int*)buf[8];
    const long long fill = 2LL * n;
    unsigned g = (unsigned)std::min<long long>((fill + 255) / 256, 4096);
    hipLaunchKernelGGL(k_ccs_fill, dim3(g), dim3(256), 0, 0, y, fill, w, 4);
    hipLaunchKernelGGL(k_ccs_fwd, dim3(n), dim3(64), 0, 0, n, (const int*)buf[3],
                       (const int*)buf[4], (const double*)buf[5], (const double*)buf[6], y, w,
                       w + 2);
    hipLaunchKernelGGL(k_ccs_bwd, dim3(n), dim3(64), 0, 0, n, (const int*)buf[0],
                       (const int*)buf[1], (const double*)buf[2], (const double*)y, y + n, w + 1,
                       w + 2);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = -1;
    if (rc == 0) rc = taucs_amd_memcpy_d2h(words, w, ib * 4) || taucs_amd_memcpy_d2h(x, y + n, db * n);
    if (rc == 0 && words[2]) {
      taucs_printf((char*)"taucs_ccs_solve_llt: a sweep timed out waiting for a value\n");
      rc = -1;
    }
  }
  const double t3 = taucs_wtime();
  for (int k = 0; k < 9; k++)
    if (buf[k]) taucs_amd_device_free(buf[k]);
  taucs_printf((char*)"\t\tCCS solve: checks and row copy %.3f s, upload %.3f s, sweeps %.3f s\n",
               t1 - t0, t2 - t1, t3 - t2);
  return rc ? -1 : 0;
}
