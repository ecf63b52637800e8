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

// W_f <- [b(cols, :); 0].  One wave per task (a front's rows r0 .. r0+nr-1,
// nr <= 64), four tasks per workgroup.
template <typename T>
__global__ __launch_bounds__(256) void k_solve_init(const FrontDesc* __restrict__ fr,
                                                    const PanelTask* __restrict__ tasks, int count,
                                                    const int* __restrict__ rows,
                                                    const double* __restrict__ b, int ldb,
                                                    T* __restrict__ W, int nrhs)
{
  const int ti = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (ti >= count) return;
  const PanelTask t = tasks[ti];
  const FrontDesc d = fr[t.front];
  const int i = t.r0 + lane;
  if (lane >= t.nr) return;
  T* w = W + d.offW * nrhs + i;
  const int col = i < d.s ? rows[d.rows + i] : 0;
  for (int r = 0; r < nrhs; r++)
    w[(long long)r * d.m] = i < d.s ? (T)b[col + (long long)r * ldb] : (T)0;
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
// the partial block Part[slot] (64 x nrhs); k_fwd_rect_reduce subtracts the
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
  T* part = Part + (long long)blockIdx.x * (64LL * nrhs);
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
    for (int q = 0; q < t.kb; q++) sum += Part[(long long)(t.k0 + q) * (64LL * nrhs) + i + r * 64];
    w[t.r0 + i + r * d.m] -= sum;
  }
}

// dst[k * ldd] += src[k * lds] for the right-hand sides k < nrhs, eight at a
// time with all loads before the stores (child and parent rows are distinct
// memory, which the compiler cannot know: both lie in the workspace).
template <typename T>
__device__ __forceinline__ void ext_row(T* dst, int ldd, const T* src, int lds, int nrhs)
{
  for (int k0 = 0; k0 < nrhs; k0 += 8) {
    T a[8], b[8];
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (k0 + k < nrhs) {
        a[k] = src[(long long)(k0 + k) * lds];
        b[k] = dst[(long long)(k0 + k) * ldd];
      }
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (k0 + k < nrhs) dst[(long long)(k0 + k) * ldd] = b[k] + a[k];
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
  ext_row(wp + r[i], p.m, wc + i, c.m, nrhs);
}

// The same for parents whose children at this level are all small: one
// workgroup per parent, heads[b] = {front: first task, k0: count}, the
// children's tasks (whole update parts, nr <= blockDim) added in sibling-rank
// order, so each parent row receives its children's values in the same order
// as under the per-rank launches.
template <typename T>
__global__ void k_solve_ext_seq(const FrontDesc* __restrict__ fr, const PanelTask* __restrict__ heads,
                                const PanelTask* __restrict__ tasks, const int* __restrict__ rel,
                                T* __restrict__ W, int nrhs)
{
  const PanelTask h = heads[blockIdx.x];
  for (int q = 0; q < h.k0; q++) {
    if (q > 0) __syncthreads();
    const PanelTask t = tasks[h.front + q];
    const FrontDesc c = fr[t.front];
    const FrontDesc p = fr[c.parent];
    const int* r = rel + c.rel;
    T* wp = W + p.offW * nrhs;
    const T* wc = W + c.offW * nrhs + c.s;
    for (int i = threadIdx.x; i < t.nr; i += blockDim.x) ext_row(wp + r[i], p.m, wc + i, c.m, nrhs);
  }
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
  T* part = Part + (long long)blockIdx.x * (64LL * nrhs);
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
    for (int q = 0; q < t.kb; q++) sum += Part[(long long)(t.k0 + q) * (64LL * nrhs) + j + r * 64];
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

// ---- chained sweeps: one launch per level --------------------------------------
// The block steps of a front's triangular solve form a chain (block i needs
// the solution of blocks 0..i-1 forward, i+1..nb-1 backward).  Instead of one
// launch per block step, one launch per level runs a task per 64-row tile of
// every front, and a task reads the blocks it needs as soon as they exist.
//
// Readiness travels with the data: before a sweep, the forward solution Y
// and the solution x are filled with a signaling-NaN bit pattern (no
// arithmetic result is a signaling NaN), every value a task publishes is one
// relaxed device-scope atomic store, and a reader polls each value it needs
// with relaxed device-scope atomic loads until the pattern is gone.  Each
// value is its own flag, so no fence (and no L2 write-back or invalidation)
// is needed on either side.
//
// Forward progress does not depend on how many workgroups are resident: a
// task takes its index from a ticket counter, tasks are listed so that every
// task waits only for tasks of smaller index, and a task that holds a ticket
// is running, so the smallest unfinished ticket can always proceed.  Every
// wait is still bounded: after SPIN_LIMIT polls the value is taken as it is,
// the error word is marked and the task finishes (the solve reports failure);
// every other wait of the launch then gives up at its next poll.
// Launches whose tasks wait for nothing run with ticket == nullptr (task =
// blockIdx).


// (Sentinel<T>, publish: the readiness protocol, in kernels.h)

// Right-hand-side blocks of the chained kernels live in LDS as 64 x 16
// arrays [row][rhs] with row stride YS (padded against bank conflicts).
#define YS 17

// Vl[i*YS + r] <- V[col + r*ldv] for i < nv, r < nrhs (0 elsewhere in the
// 64 x 16 block), awaiting each value; all of a thread's loads are issued
// before it waits for any.  Thread t handles row i = t & 63 (right-hand
// sides t/64 + 4k), whose global column `col` the caller has loaded ahead.
template <typename T, typename V>
__device__ __forceinline__ void await_rows(T* Vl, const V* Vg, int ldv, int col, int nv, int nrhs,
                                           int* err)
{
  typedef typename Sentinel<V>::U U;
  U bits[4];
  U* adr[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int e = threadIdx.x + 256 * k, i = e & 63, r = e >> 6;
    adr[k] = nullptr;
    bits[k] = 0;
    if (r < nrhs && i < nv) {
      adr[k] = reinterpret_cast<U*>(const_cast<V*>(Vg + col + (long long)r * ldv));
      bits[k] = __hip_atomic_load(adr[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int e = threadIdx.x + 256 * k, i = e & 63, r = e >> 6;
    if (adr[k]) {
      int it = 0;
      while (bits[k] == Sentinel<V>::bits) {
        if (++it > SPIN_LIMIT) {
          atomicOr(err, 1);
          break;
        }
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        __builtin_amdgcn_s_sleep(1);
        bits[k] = __hip_atomic_load(adr[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    Vl[i * YS + r] = adr[k] ? (T)__builtin_bit_cast(V, bits[k]) : (T)0;
  }
}

// Y (n x nrhs, ld n) and x (rows < n of n x nrhs, ld ldx) <- the pattern;
// tickets <- 0.  The first op of every chained solve.
template <typename T>
__global__ void k_sweep_clear(T* __restrict__ Y, double* __restrict__ x, int n, int ldx, int nrhs,
                              int* __restrict__ tickets, int nt)
{
  const long long tot = (long long)n * nrhs;
  const T ys = __builtin_bit_cast(T, Sentinel<T>::bits);
  const double xs = __builtin_bit_cast(double, Sentinel<double>::bits);
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    const long long r = e / n, i = e - r * n;
    Y[e] = ys;
    x[i + r * ldx] = xs;
  }
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nt; e += gridDim.x * blockDim.x)
    tickets[e] = 0;
}

// Forward sweep of one level (fronts with m > 64).  Task (front, tile t):
// rows [64t, 64t+64) of the front's workspace W_f (m x nrhs, the right-hand
// side plus the children's updates).  Tile t < nb (nb = ceil(s/64)) owns
// diagonal block t: it subtracts L[rows, blk j] y_j for j < t as the y_j
// appear, forms y_t = inv(L_tt) w_t, publishes it in Y (by global column),
// and (last block only) updates the sub-diagonal rows that share its tile.
// Tiles t >= nb hold sub-diagonal rows only: they subtract all nb blocks and
// store the result in W_f for the parent's extend-add.
// The products run on the 16x16x4 MFMA: wave w owns rows 16w..16w+15 of the
// tile and right-hand sides 0..15 (those beyond nrhs are zero), A = L rows
// (loaded one dependency ahead), B = the awaited y_j block in LDS; the sum
// over dependencies stays in the accumulators.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_chain(const FrontDesc* __restrict__ fr,
                                                   const PanelTask* __restrict__ tasks,
                                                   const int* __restrict__ rows,
                                                   const T* __restrict__ L,
                                                   const T* __restrict__ Inv, T* __restrict__ W,
                                                   T* Y, int ldy, int* ticket, int* err, int nrhs)
{
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T It[64 * 65];
  __shared__ T Yl[2][64 * YS];
  __shared__ int sh_task;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, c16 = lane & 15, kq = lane >> 4;
  if (tid == 0) sh_task = ticket ? atomicAdd(ticket, 1) : (int)blockIdx.x;
  __syncthreads();
  const PanelTask t = tasks[sh_task];
  const FrontDesc d = fr[t.front];
  const int m = d.m, s = d.s, nb = (s + 63) >> 6, tile = t.k0, r0 = tile * 64;
  const int nr = min(64, m - r0);
  const bool isblk = tile < nb;
  const int kb = isblk ? min(64, s - r0) : 0;
  const int ndep = isblk ? tile : nb;
  T* w = W + d.offW * nrhs;
  const T* Lf = L + d.offL;
  const int* rw = rows + d.rows;  // global column of each of the front's first s rows
  const int arow = 16 * wv + c16;
  T w0[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int i = 16 * wv + Mfma<T>::drow(lane, k);
    w0[k] = (c16 < nrhs && i < nr) ? w[r0 + i + (long long)c16 * m] : (T)0;
  }
  // global columns loaded ahead: the rows this lane publishes, and the row
  // (tid & 63) of each dependency block this thread awaits
  int pcol[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int i = 16 * wv + Mfma<T>::drow(lane, k);
    pcol[k] = (isblk && i < kb) ? rw[r0 + i] : 0;
  }
  if (isblk) {
    const T* iv = Inv + d.offI + (long long)tile * 4096;
    for (int e = tid; e < 4096; e += 256) It[(e >> 6) + (e & 63) * 65] = iv[e];
  }
  acc_t D0, D1;  // even and odd k steps: two independent MFMA chains
#pragma unroll
  for (int k = 0; k < 4; k++) D0[k] = D1[k] = (T)0;
  T lv[16];
  int dcol = 0;
  auto fetch = [&](int j) {
    const int c0 = j * 64, kbj = min(64, s - c0);
#pragma unroll
    for (int st = 0; st < 16; st++) {
      const int k = 4 * st + kq;
      lv[st] = (arow < nr && k < kbj) ? Lf[r0 + arow + (long long)(c0 + k) * m] : (T)0;
    }
    dcol = (tid & 63) < kbj ? rw[c0 + (tid & 63)] : 0;
  };
  if (ndep > 0) fetch(0);
  for (int j = 0; j < ndep; j++) {
    T* yb = Yl[j & 1];
    await_rows(yb, Y, ldy, dcol, min(64, s - 64 * j), nrhs, err);
    __syncthreads();
    T cur[16];
#pragma unroll
    for (int st = 0; st < 16; st++) cur[st] = lv[st];
    if (j + 1 < ndep) fetch(j + 1);
#pragma unroll
    for (int st = 0; st < 16; st += 2) {
      D0 = Mfma<T>::step(cur[st], yb[(4 * st + kq) * YS + c16], D0);
      D1 = Mfma<T>::step(cur[st + 1], yb[(4 * st + 4 + kq) * YS + c16], D1);
    }
  }
  T acc[4];
#pragma unroll
  for (int k = 0; k < 4; k++) acc[k] = w0[k] - (D0[k] + D1[k]);
  if (!isblk) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int i = 16 * wv + Mfma<T>::drow(lane, k);
      if (c16 < nrhs && i < nr) w[r0 + i + (long long)c16 * m] = acc[k];
    }
    return;
  }
  // y = inv(L_tt) acc: A = inverse rows (zero beyond kb), B = acc (64 x 16)
  T* Ac = Yl[ndep & 1];  // not read since the last dependency's barrier
#pragma unroll
  for (int k = 0; k < 4; k++) Ac[(16 * wv + Mfma<T>::drow(lane, k)) * YS + c16] = acc[k];
  __syncthreads();
  acc_t Ya, Yb;
#pragma unroll
  for (int k = 0; k < 4; k++) Ya[k] = Yb[k] = (T)0;
#pragma unroll
  for (int st = 0; st < 16; st += 2) {
    Ya = Mfma<T>::step(It[arow + (4 * st + kq) * 65], Ac[(4 * st + kq) * YS + c16], Ya);
    Yb = Mfma<T>::step(It[arow + (4 * st + 4 + kq) * 65], Ac[(4 * st + 4 + kq) * YS + c16], Yb);
  }
  acc_t Yd;
#pragma unroll
  for (int k = 0; k < 4; k++) Yd[k] = Ya[k] + Yb[k];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int i = 16 * wv + Mfma<T>::drow(lane, k);
    if (i < kb && c16 < nrhs) publish(Y + pcol[k] + (long long)c16 * ldy, (T)Yd[k]);
  }
  if (nr > kb) {  // sub-diagonal rows in the last block's tile: acc -= L[rows, blk t] y_t
    T* Yv = Yl[(ndep + 1) & 1];  // free since the barrier above
#pragma unroll
    for (int k = 0; k < 4; k++) Yv[(16 * wv + Mfma<T>::drow(lane, k)) * YS + c16] = Yd[k];
    __syncthreads();
    acc_t D2;
#pragma unroll
    for (int k = 0; k < 4; k++) D2[k] = (T)0;
#pragma unroll
    for (int st = 0; st < 16; st++) {
      const int k = 4 * st + kq;
      const T a = (arow < nr && k < kb) ? Lf[r0 + arow + (long long)(r0 + k) * m] : (T)0;
      D2 = Mfma<T>::step(a, Yv[k * YS + c16], D2);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int i = 16 * wv + Mfma<T>::drow(lane, k);
      if (i >= kb && i < nr && c16 < nrhs) w[r0 + i + (long long)c16 * m] = acc[k] - D2[k];
    }
  }
}

// Backward sweep of one level (fronts with m > 64).  Task (front, block b,
// kb = first Part slot, nr = slot count): the update rows' part L21^T x2 of
// the block's columns comes from k_bwd_upd as nr partial blocks, summed in
// slot order; z = y_b - that - sum_{j > b} L[blk j, blk b]^T x_j with the x_j
// of the same front awaited as they are published; then x_b = inv(L_bb)^T z
// is published in x.  MFMA products as in k_fwd_chain with wave w owning the
// block's columns 16w..16w+15: A = L^T (read down the block's columns, one
// dependency ahead), B = the awaited x_j block in LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_chain(const FrontDesc* __restrict__ fr,
                                                   const PanelTask* __restrict__ tasks,
                                                   const int* __restrict__ rows,
                                                   const T* __restrict__ L,
                                                   const T* __restrict__ Inv,
                                                   const T* __restrict__ Y, int ldy,
                                                   const T* __restrict__ Part, double* x,
                                                   int ldx, int* ticket, int* err, int nrhs)
{
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T It[64 * 65];
  __shared__ T Xl[2][64 * YS];
  __shared__ int sh_task;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, c16 = lane & 15, kq = lane >> 4;
  if (tid == 0) sh_task = ticket ? atomicAdd(ticket, 1) : (int)blockIdx.x;
  __syncthreads();
  const PanelTask t = tasks[sh_task];
  const FrontDesc d = fr[t.front];
  const int m = d.m, s = d.s, nb = (s + 63) >> 6, blk = t.k0, c0 = blk * 64;
  const int kb = min(64, s - c0);
  const T* Lf = L + d.offL;
  const int* rw = rows + d.rows;
  const int acol = 16 * wv + c16;  // A-operand row = block column of this lane
  T y0[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int q = 16 * wv + Mfma<T>::drow(lane, k);
    // slot stride 64 x nrhs: only the columns c16 < nrhs exist
    const long long pst = 64LL * nrhs;
    const T* ps = Part + (long long)t.kb * pst + q + c16 * 64;
    T s0 = (T)0, s1 = (T)0, s2 = (T)0, s3 = (T)0;  // slots in four interleaved sums
    if (c16 < nrhs) {
      int u = 0;
      for (; u + 3 < t.nr; u += 4) {
        s0 += ps[(long long)u * pst];
        s1 += ps[(long long)(u + 1) * pst];
        s2 += ps[(long long)(u + 2) * pst];
        s3 += ps[(long long)(u + 3) * pst];
      }
      for (; u < t.nr; u++) s0 += ps[(long long)u * pst];
    }
    const T sum = (s0 + s1) + (s2 + s3);
    y0[k] = (c16 < nrhs && q < kb) ? Y[rw[c0 + q] + (long long)c16 * ldy] - sum : (T)0;
  }
  {
    const T* iv = Inv + d.offI + (long long)blk * 4096;
    for (int e = tid; e < 4096; e += 256) It[(e >> 6) + (e & 63) * 65] = iv[e];
  }
  const int ndep = nb - 1 - blk;  // blocks nb-1 down to blk+1
  int pcol[4];  // the columns this lane publishes, loaded ahead
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int q = 16 * wv + Mfma<T>::drow(lane, k);
    pcol[k] = q < kb ? rw[c0 + q] : 0;
  }
  acc_t D0, D1;  // even and odd k steps: two independent MFMA chains
#pragma unroll
  for (int k = 0; k < 4; k++) D0[k] = D1[k] = (T)0;
  T lv[16];
  int dcol = 0;
  auto fetch = [&](int u) {
    const int rr0 = 64 * (nb - 1 - u), nrr = min(64, s - rr0);
#pragma unroll
    for (int st = 0; st < 16; st++) {
      const int i = 4 * st + kq;
      lv[st] = (acol < kb && i < nrr) ? Lf[rr0 + i + (long long)(c0 + acol) * m] : (T)0;
    }
    dcol = (tid & 63) < nrr ? rw[rr0 + (tid & 63)] : 0;
  };
  if (ndep > 0) fetch(0);
  for (int u = 0; u < ndep; u++) {
    const int rr0 = 64 * (nb - 1 - u);
    T* xb = Xl[u & 1];
    await_rows(xb, x, ldx, dcol, min(64, s - rr0), nrhs, err);
    __syncthreads();
    T cur[16];
#pragma unroll
    for (int st = 0; st < 16; st++) cur[st] = lv[st];
    if (u + 1 < ndep) fetch(u + 1);
#pragma unroll
    for (int st = 0; st < 16; st += 2) {
      D0 = Mfma<T>::step(cur[st], xb[(4 * st + kq) * YS + c16], D0);
      D1 = Mfma<T>::step(cur[st + 1], xb[(4 * st + 4 + kq) * YS + c16], D1);
    }
  }
  T* Z = Xl[ndep & 1];  // not read since the last dependency's barrier
#pragma unroll
  for (int k = 0; k < 4; k++)
    Z[(16 * wv + Mfma<T>::drow(lane, k)) * YS + c16] = y0[k] - (D0[k] + D1[k]);
  __syncthreads();
  // x = inv(L_bb)^T z: A[q][p] = inv(p, q)
  acc_t Xa, Xb;
#pragma unroll
  for (int k = 0; k < 4; k++) Xa[k] = Xb[k] = (T)0;
#pragma unroll
  for (int st = 0; st < 16; st += 2) {
    Xa = Mfma<T>::step(It[(4 * st + kq) + acol * 65], Z[(4 * st + kq) * YS + c16], Xa);
    Xb = Mfma<T>::step(It[(4 * st + 4 + kq) + acol * 65], Z[(4 * st + 4 + kq) * YS + c16], Xb);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int q = 16 * wv + Mfma<T>::drow(lane, k);
    if (q < kb && c16 < nrhs) publish(x + pcol[k] + (long long)c16 * ldx, (double)(Xa[k] + Xb[k]));
  }
}

// Backward sweep, update rows of the fronts with m > 64, split over row
// chunks (before the chained launch): Part[slot] (64 columns x 16 rhs) =
// L[rows i0 .. i0+cnt-1, columns c0 .. c0+nc-1]^T x(rows), task (front, k0 =
// i0, kb = cnt <= BU_ROWS, r0 = c0, nr = nc); k_bwd_cols_reduce subtracts the
// slots from Y in order.  x of the update rows is final (the ancestors'
// levels wrote it) and is read through the front's row list.  MFMA as in
// k_bwd_chain: wave w owns columns 16w..16w+15, A = L^T, B = x rows in LDS.
#define BU_ROWS 128
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_upd(const FrontDesc* __restrict__ fr,
                                                 const PanelTask* __restrict__ tasks,
                                                 const int* __restrict__ rows,
                                                 const T* __restrict__ L,
                                                 const double* __restrict__ x, int ldx,
                                                 T* __restrict__ Part, int nrhs)
{
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T Xl[BU_ROWS * YS];
  const PanelTask t = tasks[blockIdx.x];
  const FrontDesc d = fr[t.front];
  const int m = d.m, i0 = t.k0, cnt = t.kb, c0 = t.r0, nc = t.nr;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, c16 = lane & 15, kq = lane >> 4;
  const int* rw = rows + d.rows;
  const int acol = 16 * wv + c16;
  const T* Lc = L + d.offL + i0 + (long long)(c0 + acol) * m;  // column c0+acol from row i0
  T a[BU_ROWS / 4];
#pragma unroll
  for (int st = 0; st < BU_ROWS / 4; st++) {
    const int i = 4 * st + kq;
    a[st] = (acol < nc && i < cnt) ? Lc[i] : (T)0;
  }
#pragma unroll
  for (int k = 0; k < BU_ROWS * 16 / 256; k++) {
    const int e = tid + 256 * k, i = e & (BU_ROWS - 1), r = e / BU_ROWS;
    Xl[i * YS + r] = (i < cnt && r < nrhs) ? (T)x[rw[i0 + i] + (long long)r * ldx] : (T)0;
  }
  __syncthreads();
  acc_t D;
#pragma unroll
  for (int k = 0; k < 4; k++) D[k] = (T)0;
  const int nst = (cnt + 3) >> 2;
#pragma unroll
  for (int st = 0; st < BU_ROWS / 4; st++)
    if (st < nst) D = Mfma<T>::step(a[st], Xl[(4 * st + kq) * YS + c16], D);
  // slot stride 64 x nrhs: the columns c16 >= nrhs are not stored
  T* part = Part + (long long)blockIdx.x * (64LL * nrhs);
  if (c16 < nrhs) {
#pragma unroll
    for (int k = 0; k < 4; k++) part[16 * wv + Mfma<T>::drow(lane, k) + c16 * 64] = D[k];
  }
}

// Fronts with m <= 64, one wave each (four per workgroup), no waits: the
// whole front's right-hand side fits one wave's rows.
//   forward:  y = inv(L11) w[0:s] into Y; w[s:m] -= L21 y for the parent.
//   backward: z = y - L21^T x(update rows); x = inv(L11)^T z.
// Per-wave LDS holds the m x nrhs values that every lane needs.
__device__ __forceinline__ void wave_sync_lds()
{
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
}

// NR: right-hand sides per pass rounded up to 1, 4, 8 or 16 (compile-time, so
// the loops over them carry no guards); columns beyond nrhs hold zeros.
template <typename T, int NR>
__global__ __launch_bounds__(256) void k_fwd_small(const FrontDesc* __restrict__ fr,
                                                   const PanelTask* __restrict__ tasks, int count,
                                                   const int* __restrict__ rows,
                                                   const T* __restrict__ L,
                                                   const T* __restrict__ Inv, T* __restrict__ W,
                                                   T* __restrict__ Y, int ldy, int nrhs)
{
  __shared__ T Ws[4][64 * NR];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ti = blockIdx.x * 4 + wv;
  if (ti >= count) return;
  const FrontDesc d = fr[tasks[ti].front];
  const int m = d.m, s = d.s;
  T* w = W + d.offW * nrhs;
  T* ws = Ws[wv];
  const int* rw = rows + d.rows;
  T v[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) {
    v[r] = (r < nrhs && lane < m) ? w[lane + r * m] : (T)0;
    ws[lane + r * 64] = v[r];
  }
  wave_sync_lds();
  const T* iv = Inv + d.offI + lane * 64;
  T y[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) y[r] = (T)0;
  for (int p = 0; p < s; p++) {
    const T a = (lane < s && p <= lane) ? iv[p] : (T)0;
#pragma unroll
    for (int r = 0; r < NR; r++) y[r] += a * ws[p + r * 64];
  }
  wave_sync_lds();
  if (lane < s) {
#pragma unroll
    for (int r = 0; r < NR; r++) {
      ws[lane + r * 64] = y[r];
      if (r < nrhs) Y[rw[lane] + (long long)r * ldy] = y[r];
    }
  }
  wave_sync_lds();
  if (m == s) return;
  const T* Lf = L + d.offL;
  for (int p = 0; p < s; p++) {
    const T a = (lane >= s && lane < m) ? Lf[lane + (long long)p * m] : (T)0;
#pragma unroll
    for (int r = 0; r < NR; r++) v[r] -= a * ws[p + r * 64];
  }
  if (lane >= s && lane < m) {
#pragma unroll
    for (int r = 0; r < NR; r++)
      if (r < nrhs) w[lane + r * m] = v[r];
  }
}

template <typename T, int NR>
__global__ __launch_bounds__(256) void k_bwd_small(const FrontDesc* __restrict__ fr,
                                                   const PanelTask* __restrict__ tasks, int count,
                                                   const int* __restrict__ rows,
                                                   const T* __restrict__ L,
                                                   const T* __restrict__ Inv,
                                                   const T* __restrict__ Y, int ldy,
                                                   double* __restrict__ x, int ldx, int nrhs)
{
  __shared__ T Ws[4][64 * NR];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ti = blockIdx.x * 4 + wv;
  if (ti >= count) return;
  const FrontDesc d = fr[tasks[ti].front];
  const int m = d.m, s = d.s;
  const int* rw = rows + d.rows;
  T* ws = Ws[wv];
  T z[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) {
    z[r] = (r < nrhs && lane < s) ? Y[rw[lane] + (long long)r * ldy] : (T)0;
    ws[lane + r * 64] =
        (r < nrhs && lane >= s && lane < m) ? (T)x[rw[lane] + (long long)r * ldx] : (T)0;
  }
  wave_sync_lds();
  const T* Lf = L + d.offL + (long long)lane * m;  // column `lane` of the panel
  for (int i = s; i < m; i++) {
    const T a = lane < s ? Lf[i] : (T)0;
#pragma unroll
    for (int r = 0; r < NR; r++) z[r] -= a * ws[i + r * 64];
  }
  if (lane < s) {
#pragma unroll
    for (int r = 0; r < NR; r++) ws[lane + r * 64] = z[r];
  }
  wave_sync_lds();
  const T* iv = Inv + d.offI + lane;  // column `lane` of the inverse: iv[p * 64]
  T xv[NR];
#pragma unroll
  for (int r = 0; r < NR; r++) xv[r] = (T)0;
  for (int p = 0; p < s; p++) {
    const T a = (lane < s && p >= lane) ? iv[p * 64] : (T)0;
#pragma unroll
    for (int r = 0; r < NR; r++) xv[r] += a * ws[p + r * 64];
  }
  if (lane < s) {
#pragma unroll
    for (int r = 0; r < NR; r++)
      if (r < nrhs) x[rw[lane] + (long long)r * ldx] = (double)xv[r];
  }
}

#define SMALL_INST(T, NR)                                                                       \
  template __global__ void k_fwd_small<T, NR>(const FrontDesc*, const PanelTask*, int, const int*, \
                                              const T*, const T*, T*, T*, int, int);            \
  template __global__ void k_bwd_small<T, NR>(const FrontDesc*, const PanelTask*, int, const int*, \
                                              const T*, const T*, const T*, int, double*, int, int);
#define TAMD_SINST(T)                                                                           \
  template __global__ void k_solve_init<T>(const FrontDesc*, const PanelTask*, int, const int*,  \
                                           const double*, int, T*, int);                        \
  template __global__ void k_fwd_step<T>(const FrontDesc*, const PanelTask*, const int*,         \
                                         const T*, const T*, T*, T*, int, int);                 \
  template __global__ void k_fwd_rect<T>(const FrontDesc*, const PanelTask*, const int*,         \
                                         const T*, const T*, int, T*, int);                     \
  template __global__ void k_fwd_rect_reduce<T>(const FrontDesc*, const PanelTask*, const T*, T*, int); \
  template __global__ void k_bwd_cols_reduce<T>(const FrontDesc*, const PanelTask*, const int*,  \
                                                const T*, T*, int, int);                        \
  template __global__ void k_solve_ext<T>(const FrontDesc*, const PanelTask*, const int*, T*, int); \
  template __global__ void k_solve_ext_seq<T>(const FrontDesc*, const PanelTask*, const PanelTask*, \
                                              const int*, T*, int);                             \
  template __global__ void k_bwd_gather<T>(const FrontDesc*, const PanelTask*, const int*,        \
                                           const double*, int, T*, int);                        \
  template __global__ void k_bwd_cols<T>(const FrontDesc*, const PanelTask*, const T*, const T*, \
                                         T*, int);                                              \
  template __global__ void k_bwd_step<T>(const FrontDesc*, const PanelTask*, const int*,         \
                                         const T*, const T*, T*, int, double*, int, int);       \
  template __global__ void k_sweep_clear<T>(T*, double*, int, int, int, int*, int);           \
  template __global__ void k_fwd_chain<T>(const FrontDesc*, const PanelTask*, const int*,        \
                                          const T*, const T*, T*, T*, int, int*, int*, int);    \
  template __global__ void k_bwd_chain<T>(const FrontDesc*, const PanelTask*, const int*,        \
                                          const T*, const T*, const T*, int, const T*, double*, \
                                          int, int*, int*, int);                                \
  template __global__ void k_bwd_upd<T>(const FrontDesc*, const PanelTask*, const int*, const T*, \
                                        const double*, int, T*, int);                          \
  SMALL_INST(T, 1) SMALL_INST(T, 4) SMALL_INST(T, 8) SMALL_INST(T, 16)
TAMD_SINST(double)
TAMD_SINST(float)

}  // namespace tamd
