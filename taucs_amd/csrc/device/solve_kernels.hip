// Supernodal forward/backward solves over the level schedule, multi-RHS.
//
// Reference (one rhs, recursive): src/taucs_sn_llt.c:2597-2720 (forward:
// gather, trsm, gemm, scatter) and :2723-2841 (backward: gather, gemm^T,
// trsm^T, scatter).  Here each front owns an m x nrhs workspace block W_f
// (column-major, at W + offW * nrhs):
//   forward:  W_f[0:s] starts as b[cols], W_f[s:m] as 0; children add their
//             update parts at the relative rows (fixed order, no atomics);
//             then, 64 columns at a time, y_blk = inv(L_kk) w_blk (the block
//             inverses come from the factorization) and every row below the
//             block is updated by many workgroups: W_f[r] -= L[r, blk] y_blk.
//             y is kept in Y, indexed by global column.
//   backward: top-down.  W_f[s:m] receives x of the update rows from the
//             ancestors; z = y - L21^T x2 is formed one wave per column
//             (k_bwd_cols); then blocks from the last to the first:
//             x_blk = inv(L_kk)^T z_blk and the update of every earlier
//             column by the block's rows (k_bwd_step).
// Right-hand sides are processed 16 at a time (nrhs <= 16 per call).
#include "kernels.h"

namespace tamd {

#define SMAXR 16

// W_f <- [b(cols, :); 0]
template <typename T>
__global__ void k_solve_init(const FrontDesc* __restrict__ fr, const int* __restrict__ rows,
                             const double* __restrict__ b, int ldb, T* __restrict__ W, int nrhs)
{
  const FrontDesc d = fr[blockIdx.x];
  const int* rw = rows + d.rows;
  const int tot = d.m * nrhs;
  T* w = W + d.offW * nrhs;
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    int i = e % d.m, r = e / d.m;
    w[e] = i < d.s ? (T)b[rw[i] + (long long)r * ldb] : (T)0;
  }
}

// ---- LDS-tiled pieces ---------------------------------------------------------
// All block operations work on 64 x 64 tiles of L staged in LDS (Lt[i + j*65],
// loaded column by column so global reads are coalesced) against small
// 64 x nrhs blocks of right-hand-side values (V[i + r*64]).  A 256-thread
// workgroup assigns thread t to tile row/column t % 64 and right-hand sides
// 4*(t/64) .. 4*(t/64)+3.

template <typename T>
__device__ __forceinline__ void load_tile(T* Lt, const T* src, long long ld, int nr, int nc)
{
  // src(i, j) = src[i + j*ld], i < nr rows, j < nc columns; zero padded
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    int i = e & 63, j = e >> 6;
    Lt[i + j * 65] = (i < nr && j < nc) ? src[i + (long long)j * ld] : (T)0;
  }
}

template <typename T>
__device__ __forceinline__ void load_inv(T* It, const T* iv)
{
  // row-major 64 x 64 inverse block -> It[i + q*65] = inv(i, q)
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    int q = e & 63, i = e >> 6;
    It[i + q * 65] = iv[i * 64 + q];
  }
}

// The same two loads issued together: all 32 global loads of a thread are in
// flight before the first LDS store (the step kernels are latency-bound).
template <typename T>
__device__ __forceinline__ void load_inv_and_tile(T* It, const T* iv, T* Lt, const T* src,
                                                  long long ld, int nr, int nc)
{
  T a[16], b[16];
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const int e = threadIdx.x + 256 * u;
    a[u] = iv[e];
    const int i = e & 63, j = e >> 6;
    b[u] = (i < nr && j < nc) ? src[i + (long long)j * ld] : (T)0;
  }
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const int e = threadIdx.x + 256 * u;
    It[(e >> 6) + (e & 63) * 65] = a[u];
    Lt[(e & 63) + (e >> 6) * 65] = b[u];
  }
}

// One 64-column block step of the forward solve.  Every task forms
// y_blk = inv(L_kk) w_blk from the block's rows of W_f (read-only in this
// step) and updates its 64 rows of the diagonal block's column range below
// the block: W_f[r] -= L[r, blk] y_blk (rows inside L11 only; the L21 rows
// get one pass with all columns in k_fwd_rect).  The task whose rows start
// right below the block also stores y_blk into Y (by global column).
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_step(const FrontDesc* __restrict__ fr,
                                                  const PanelTask* __restrict__ tasks,
                                                  const int* __restrict__ rows,
                                                  const T* __restrict__ L,
                                                  const T* __restrict__ Inv, T* __restrict__ W,
                                                  T* __restrict__ Y, int ldy, int nrhs)
{
  __shared__ T It[64 * 65];
  __shared__ T Lt[64 * 65];
  __shared__ T Wb[64 * SMAXR];
  __shared__ T Yb[64 * SMAXR];
  const PanelTask t = tasks[blockIdx.x];
  const FrontDesc d = fr[t.front];
  const int m = d.m, kb = t.kb, tid = threadIdx.x;
  T* w = W + d.offW * nrhs;
  load_inv_and_tile(It, Inv + d.offI + (long long)(t.k0 / 64) * 4096, Lt,
                    L + d.offL + t.r0 + (long long)t.k0 * m, m, t.nr, kb);
  for (int e = tid; e < 64 * nrhs; e += 256) {
    int i = e & 63, r = e >> 6;
    Wb[i + r * 64] = i < kb ? w[t.k0 + i + r * m] : (T)0;
  }
  __syncthreads();
  for (int e = tid; e < 64 * nrhs; e += 256) {
    int i = e & 63, r = e >> 6;
    T acc = (T)0;
    for (int q = 0; q <= i; q++) acc += It[i + q * 65] * Wb[q + r * 64];
    Yb[i + r * 64] = acc;
  }
  __syncthreads();
  if (t.r0 == t.k0 + kb) {
    const int* rw = rows + d.rows + t.k0;
    for (int e = tid; e < kb * nrhs; e += 256) {
      int i = e % kb, r = e / kb;
      Y[rw[i] + (long long)r * ldy] = Yb[i + r * 64];
    }
  }
  if (t.nr <= 0) return;
  const int i = tid & 63, g = tid >> 6;
  if (i < t.nr) {
    for (int q = 0; q < 4; q++) {
      int r = 4 * g + q;
      if (r >= nrhs) break;
      T acc = (T)0;
      for (int j = 0; j < kb; j++) acc += Lt[i + j * 65] * Yb[j + r * 64];
      w[t.r0 + i + r * m] -= acc;
    }
  }
}

// L21 rows of the forward solve, split over column chunks: the task's rows
// r0..r0+nr-1 (>= s) times y of columns [c0, c0+nc) (k0 = c0, kb = nc) into
// the partial block Part[slot] (64 x SMAXR); k_fwd_rect_reduce subtracts the
// partial blocks from W_f in slot order, so the sum is deterministic.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_rect(const FrontDesc* __restrict__ fr,
                                                  const PanelTask* __restrict__ tasks,
                                                  const int* __restrict__ rows,
                                                  const T* __restrict__ L,
                                                  const T* __restrict__ Y, int ldy,
                                                  T* __restrict__ Part, int nrhs)
{
  __shared__ T Lt[64 * 65];
  __shared__ T Yc[64 * SMAXR];
  const PanelTask t = tasks[blockIdx.x];
  const FrontDesc d = fr[t.front];
  const int m = d.m, tid = threadIdx.x, i = tid & 63, g = tid >> 6;
  const int* rw = rows + d.rows;
  T acc[4] = {(T)0, (T)0, (T)0, (T)0};
  for (int c0 = t.k0; c0 < t.k0 + t.kb; c0 += 64) {
    const int nc = t.k0 + t.kb - c0 < 64 ? t.k0 + t.kb - c0 : 64;
    __syncthreads();
    load_tile(Lt, L + d.offL + t.r0 + (long long)c0 * m, m, t.nr, nc);
    for (int e = tid; e < 64 * nrhs; e += 256) {
      int j = e & 63, r = e >> 6;
      Yc[j + r * 64] = j < nc ? Y[rw[c0 + j] + (long long)r * ldy] : (T)0;
    }
    __syncthreads();
    for (int q = 0; q < 4; q++) {
      int r = 4 * g + q;
      if (r >= nrhs) break;
      T a = (T)0;
      for (int j = 0; j < nc; j++) a += Lt[i + j * 65] * Yc[j + r * 64];
      acc[q] += a;
    }
  }
  T* part = Part + (long long)blockIdx.x * (64 * SMAXR);
  for (int q = 0; q < 4; q++) {
    int r = 4 * g + q;
    if (r >= nrhs) break;
    part[i + r * 64] = acc[q];
  }
}

// W_f[r0 + i] -= sum over slots k0 .. k0+kb-1 of Part[slot][i]
template <typename T>
__global__ void k_fwd_rect_reduce(const FrontDesc* __restrict__ fr,
                                  const PanelTask* __restrict__ tasks,
                                  const T* __restrict__ Part, T* __restrict__ W, int nrhs)
{
  const PanelTask t = tasks[blockIdx.x];
  const FrontDesc d = fr[t.front];
  T* w = W + d.offW * nrhs;
  for (int e = threadIdx.x; e < 64 * nrhs; e += blockDim.x) {
    int i = e & 63, r = e >> 6;
    if (i >= t.nr) continue;
    T sum = (T)0;
    for (int q = 0; q < t.kb; q++) sum += Part[(long long)(t.k0 + q) * (64 * SMAXR) + i + r * 64];
    w[t.r0 + i + r * d.m] -= sum;
  }
}

// Child update rows added into the parent's workspace (rows r0..r0+nr-1 of
// the child's update part).
template <typename T>
__global__ void k_solve_ext(const FrontDesc* __restrict__ fr, const PanelTask* __restrict__ tasks,
                            const int* __restrict__ rel, T* __restrict__ W, int nrhs)
{
  const PanelTask t = tasks[blockIdx.x];
  const FrontDesc c = fr[t.front];
  const FrontDesc p = fr[c.parent];
  const int* r = rel + c.rel;
  T* wp = W + p.offW * nrhs;
  const T* wc = W + c.offW * nrhs + c.s;
  const int i = t.r0 + threadIdx.x;
  if (threadIdx.x >= t.nr) return;
  const int pi = r[i];
  for (int q = 0; q < nrhs; q++) wp[pi + (long long)q * p.m] += wc[i + (long long)q * c.m];
}

// Backward: W_f[s + i] <- x(update row i) for rows r0..r0+nr-1 of the update part.
template <typename T>
__global__ void k_bwd_gather(const FrontDesc* __restrict__ fr, const PanelTask* __restrict__ tasks,
                             const int* __restrict__ rows, const double* __restrict__ x, int ldx,
                             T* __restrict__ W, int nrhs)
{
  const PanelTask t = tasks[blockIdx.x];
  const FrontDesc d = fr[t.front];
  const int i = t.r0 + threadIdx.x;
  if (threadIdx.x >= t.nr) return;
  const int gi = rows[d.rows + d.s + i];
  T* w = W + d.offW * nrhs + d.s;
  for (int r = 0; r < nrhs; r++) w[i + r * d.m] = (T)x[gi + (long long)r * ldx];
}

// z = y - L21^T x2, split over row chunks: columns [r0, r0+nr) of a front
// times the update rows [k0, k0+kb) (>= s) into Part[slot];
// k_bwd_cols_reduce subtracts the partials from Y in slot order.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_cols(const FrontDesc* __restrict__ fr,
                                                  const PanelTask* __restrict__ tasks,
                                                  const T* __restrict__ L,
                                                  const T* __restrict__ W,
                                                  T* __restrict__ Part, int nrhs)
{
  __shared__ T Lt[64 * 65];
  __shared__ T Xc[64 * SMAXR];
  const PanelTask t = tasks[blockIdx.x];
  const FrontDesc d = fr[t.front];
  const int m = d.m, tid = threadIdx.x, j = tid & 63, g = tid >> 6;
  const T* w = W + d.offW * nrhs;
  T acc[4] = {(T)0, (T)0, (T)0, (T)0};
  for (int i0 = t.k0; i0 < t.k0 + t.kb; i0 += 64) {
    const int ni = t.k0 + t.kb - i0 < 64 ? t.k0 + t.kb - i0 : 64;
    __syncthreads();
    load_tile(Lt, L + d.offL + i0 + (long long)t.r0 * m, m, ni, t.nr);
    for (int e = tid; e < 64 * nrhs; e += 256) {
      int ii = e & 63, r = e >> 6;
      Xc[ii + r * 64] = ii < ni ? w[i0 + ii + r * m] : (T)0;
    }
    __syncthreads();
    for (int q = 0; q < 4; q++) {
      int r = 4 * g + q;
      if (r >= nrhs) break;
      T a = (T)0;
      for (int ii = 0; ii < ni; ii++) a += Lt[ii + j * 65] * Xc[ii + r * 64];
      acc[q] += a;
    }
  }
  T* part = Part + (long long)blockIdx.x * (64 * SMAXR);
  for (int q = 0; q < 4; q++) {
    int r = 4 * g + q;
    if (r >= nrhs) break;
    part[j + r * 64] = acc[q];
  }
}

// Y[col(r0 + j)] -= sum over slots k0 .. k0+kb-1 of Part[slot][j]
template <typename T>
__global__ void k_bwd_cols_reduce(const FrontDesc* __restrict__ fr,
                                  const PanelTask* __restrict__ tasks,
                                  const int* __restrict__ rows, const T* __restrict__ Part,
                                  T* __restrict__ Y, int ldy, int nrhs)
{
  const PanelTask t = tasks[blockIdx.x];
  const FrontDesc d = fr[t.front];
  for (int e = threadIdx.x; e < 64 * nrhs; e += blockDim.x) {
    int j = e & 63, r = e >> 6;
    if (j >= t.nr) continue;
    T sum = (T)0;
    for (int q = 0; q < t.kb; q++) sum += Part[(long long)(t.k0 + q) * (64 * SMAXR) + j + r * 64];
    Y[rows[d.rows + t.r0 + j] + (long long)r * ldy] -= sum;
  }
}

// One 64-column block step of the backward solve (blocks from the last to
// the first).  Every task forms x_blk = inv(L_kk)^T z_blk from the block's
// entries of Y (read-only in this step) and updates its 64 columns
// j in [r0, r0+nr) < k0: z_j -= sum_{i in blk} L[i][j] x_i.  The task with
// r0 == 0 also stores x_blk into the solution.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_step(const FrontDesc* __restrict__ fr,
                                                  const PanelTask* __restrict__ tasks,
                                                  const int* __restrict__ rows,
                                                  const T* __restrict__ L,
                                                  const T* __restrict__ Inv, T* __restrict__ Y,
                                                  int ldy, double* __restrict__ x, int ldx,
                                                  int nrhs)
{
  __shared__ T It[64 * 65];
  __shared__ T Lt[64 * 65];
  __shared__ T Zb[64 * SMAXR];
  __shared__ T Xb[64 * SMAXR];
  const PanelTask t = tasks[blockIdx.x];
  const FrontDesc d = fr[t.front];
  const int m = d.m, kb = t.kb, tid = threadIdx.x;
  const int* rw = rows + d.rows;
  load_inv_and_tile(It, Inv + d.offI + (long long)(t.k0 / 64) * 4096, Lt,
                    L + d.offL + t.k0 + (long long)t.r0 * m, m, kb, t.nr);
  for (int e = tid; e < 64 * nrhs; e += 256) {
    int i = e & 63, r = e >> 6;
    Zb[i + r * 64] = i < kb ? Y[rw[t.k0 + i] + (long long)r * ldy] : (T)0;
  }
  __syncthreads();
  for (int e = tid; e < 64 * nrhs; e += 256) {
    int i = e & 63, r = e >> 6;
    T acc = (T)0;
    for (int q = i; q < kb; q++) acc += It[q + i * 65] * Zb[q + r * 64];
    Xb[i + r * 64] = acc;
  }
  __syncthreads();
  if (t.r0 == 0) {
    for (int e = tid; e < kb * nrhs; e += 256) {
      int i = e % kb, r = e / kb;
      x[rw[t.k0 + i] + (long long)r * ldx] = (double)Xb[i + r * 64];
    }
  }
  if (t.nr <= 0) return;
  const int j = tid & 63, g = tid >> 6;
  if (j < t.nr) {
    const int gj = rw[t.r0 + j];
    for (int q = 0; q < 4; q++) {
      int r = 4 * g + q;
      if (r >= nrhs) break;
      T acc = (T)0;
      for (int i = 0; i < kb; i++) acc += Lt[i + j * 65] * Xb[i + r * 64];
      Y[gj + (long long)r * ldy] -= acc;
    }
  }
}

#define TAMD_SINST(T)                                                                           \
  template __global__ void k_solve_init<T>(const FrontDesc*, const int*, const double*, int, T*, \
                                           int);                                                \
  template __global__ void k_fwd_step<T>(const FrontDesc*, const PanelTask*, const int*,         \
                                         const T*, const T*, T*, T*, int, int);                 \
  template __global__ void k_fwd_rect<T>(const FrontDesc*, const PanelTask*, const int*,         \
                                         const T*, const T*, int, T*, int);                     \
  template __global__ void k_fwd_rect_reduce<T>(const FrontDesc*, const PanelTask*, const T*, T*, int); \
  template __global__ void k_bwd_cols_reduce<T>(const FrontDesc*, const PanelTask*, const int*,  \
                                                const T*, T*, int, int);                        \
  template __global__ void k_solve_ext<T>(const FrontDesc*, const PanelTask*, const int*, T*, int); \
  template __global__ void k_bwd_gather<T>(const FrontDesc*, const PanelTask*, const int*,        \
                                           const double*, int, T*, int);                        \
  template __global__ void k_bwd_cols<T>(const FrontDesc*, const PanelTask*, const T*, const T*, \
                                         T*, int);                                              \
  template __global__ void k_bwd_step<T>(const FrontDesc*, const PanelTask*, const int*,         \
                                         const T*, const T*, T*, int, double*, int, int);
TAMD_SINST(double)
TAMD_SINST(float)

}  // namespace tamd
