// Supernodal forward/backward solves over the level schedule, multi-RHS.
//
// Reference (one rhs, recursive): src/taucs_sn_llt.c:2597-2720 (forward:
// gather, trsm, gemm, scatter) and :2723-2841 (backward: gather, gemm^T,
// trsm^T, scatter).  Here each front owns an m x nrhs workspace W_f:
//   forward:  W_f[0:s] starts as b[cols], W_f[s:m] as 0; children add their
//             update parts at the relative rows (fixed order, no atomics);
//             the front solves L11 y = W_f[0:s] and W_f[s:m] -= L21 y,
//   backward: top-down, x[cols] = L11^-T (y - L21^T x[update rows]).
#include "kernels.h"

namespace tamd {

// W_f <- [b(cols, :); 0]
template <typename T>
__global__ void k_solve_init(const FrontDesc* __restrict__ fr, const int* __restrict__ rows,
                             const double* __restrict__ b, int ldb, T* __restrict__ W,
                             int nrhs)
{
  const FrontDesc d = fr[blockIdx.x];
  const int* rw = rows + d.rows;
  const int tot = d.m * nrhs;
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    int i = e % d.m, r = e / d.m;
    W[d.offW * nrhs + e] = i < d.s ? (T)b[rw[i] + (long long)r * ldb] : (T)0;
  }
}

// Column sweep over the panel: y_j = w_j / L_jj, then w_i -= L_ij y_j (i > j).
template <typename T>
__global__ __launch_bounds__(256) void k_solve_fwd(const FrontDesc* __restrict__ fr,
                                                   const int* __restrict__ list,
                                                   const T* __restrict__ L,
                                                   T* __restrict__ W, int nrhs)
{
  __shared__ T ycol[16];
  const FrontDesc d = fr[list[blockIdx.x]];
  const int s = d.s, m = d.m;
  const T* P = L + d.offL;
  T* w = W + d.offW * nrhs;
  for (int j = 0; j < s; j++) {
    if (threadIdx.x < nrhs) {
      T y = w[j + threadIdx.x * m] / P[j + (long long)j * m];
      w[j + threadIdx.x * m] = y;
      ycol[threadIdx.x] = y;
    }
    __syncthreads();
    const int cnt = m - j - 1;
    const T* col = P + (long long)j * m + j + 1;
    for (int e = threadIdx.x; e < cnt * nrhs; e += blockDim.x) {
      int i = e % cnt, r = e / cnt;
      w[j + 1 + i + r * m] -= col[i] * ycol[r];
    }
    __syncthreads();
  }
}

// Child update rows added into the parent's workspace.
template <typename T>
__global__ void k_solve_ext(const FrontDesc* __restrict__ fr, const int* __restrict__ list,
                            const int* __restrict__ rel, T* __restrict__ W, int nrhs)
{
  const FrontDesc c = fr[list[blockIdx.x]];
  const FrontDesc p = fr[c.parent];
  const int u = c.m - c.s;
  const int* r = rel + c.rel;
  for (int e = threadIdx.x; e < u * nrhs; e += blockDim.x) {
    int i = e % u, q = e / u;
    W[p.offW * nrhs + r[i] + (long long)q * p.m] += W[c.offW * nrhs + c.s + i + (long long)q * c.m];
  }
}

// Backward: z = y - L21^T x_upd, then x = L11^-T z by reverse column dots.
template <typename T>
__global__ __launch_bounds__(256) void k_solve_bwd(const FrontDesc* __restrict__ fr,
                                                   const int* __restrict__ list,
                                                   const int* __restrict__ rows,
                                                   const T* __restrict__ L,
                                                   T* __restrict__ W, double* __restrict__ x,
                                                   int ldx, int nrhs)
{
  __shared__ T red[256];
  const FrontDesc d = fr[list[blockIdx.x]];
  const int s = d.s, m = d.m;
  const int* rw = rows + d.rows;
  const T* P = L + d.offL;
  T* w = W + d.offW * nrhs;
  // update rows' solution from the ancestors (already final)
  for (int e = threadIdx.x; e < (m - s) * nrhs; e += blockDim.x) {
    int i = e % (m - s), r = e / (m - s);
    w[s + i + r * m] = (T)x[rw[s + i] + (long long)r * ldx];
  }
  __syncthreads();
  for (int j = s - 1; j >= 0; j--) {
    const T* col = P + (long long)j * m;
    for (int r = 0; r < nrhs; r++) {
      T acc = (T)0;
      for (int i = j + 1 + threadIdx.x; i < m; i += blockDim.x) acc += col[i] * w[i + r * m];
      red[threadIdx.x] = acc;
      __syncthreads();
      for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
      }
      if (threadIdx.x == 0) w[j + r * m] = (w[j + r * m] - red[0]) / col[j];
      __syncthreads();
    }
  }
  for (int e = threadIdx.x; e < s * nrhs; e += blockDim.x) {
    int i = e % s, r = e / s;
    x[rw[i] + (long long)r * ldx] = (double)w[i + r * m];
  }
}

#define TAMD_SINST(T)                                                                  \
  template __global__ void k_solve_init<T>(const FrontDesc*, const int*, const double*, int, T*, int); \
  template __global__ void k_solve_fwd<T>(const FrontDesc*, const int*, const T*, T*, int);  \
  template __global__ void k_solve_ext<T>(const FrontDesc*, const int*, const int*, T*, int); \
  template __global__ void k_solve_bwd<T>(const FrontDesc*, const int*, const int*, const T*, T*, \
                                          double*, int, int);
TAMD_SINST(double)
TAMD_SINST(float)

}  // namespace tamd
