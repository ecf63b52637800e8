// Numeric-factorization kernels: A assembly, small fronts, blocked panel
// POTRF/TRSM, MFMA tile GEMM/SYRK, extend-add.
//
// Reference computation per front (src/taucs_sn_llt.c:1086-1194): assemble
// A's columns into [F1;F2], dpotrf F1, dtrsm F2 <- F2 L11^-T, dsyrk
// U <- U - F2 F2^T; then extend-add U into the parent (:1200-1240).
#include "kernels.h"

namespace tamd {

template <typename T>
__global__ void k_fill(T* __restrict__ p, long long n)
{
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = (T)0;
}

// Zero the update blocks whose life starts at this level: the lower
// triangles of large blocks (one task = a column range of about 64K
// elements; every reader of an update block, the extend-add, the SYRK
// epilogue and the whole-front kernels, touches entries (i, j) with i >= j
// alone), small blocks whole as one flat range.
template <typename T>
__global__ void k_fill_ranges(const ZeroTask* __restrict__ tasks, T* __restrict__ U)
{
  const ZeroTask t = tasks[blockIdx.x];
  T* const b = U + t.off;
  for (int c = t.c0; c < t.c1; c++)
    for (int i = c + threadIdx.x; i < t.u; i += blockDim.x) b[i + (long long)c * t.u] = (T)0;
}

// One thread per A column, entries in stored order: duplicates (if any) are
// summed deterministically.
template <typename T>
__global__ void k_assemble(int ncol, const int* __restrict__ colptr,
                           const long long* __restrict__ dest,
                           const double* __restrict__ vals, T* __restrict__ L)
{
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncol) return;
  int e = colptr[j + 1];
  for (int ip = colptr[j]; ip < e; ip++) L[dest[ip]] += (T)vals[ip];
}

// One thread per A entry, for patterns whose entries all have distinct
// destinations (the host checks when it builds the schedule): the same single
// addition per destination as k_assemble, without its chain of dependent
// read-modify-writes down each column.
template <typename T>
__global__ void k_assemble_flat(long long nnz, const long long* __restrict__ dest,
                                const double* __restrict__ vals, T* __restrict__ L)
{
  const long long step = (long long)gridDim.x * blockDim.x;
  for (long long ip = (long long)blockIdx.x * blockDim.x + threadIdx.x; ip < nnz; ip += step)
    L[dest[ip]] += (T)vals[ip];
}

// Right-looking unblocked Cholesky of an nrows x ncols panel (both <= 64),
// fused with the inversion of the factored diagonal block (X = L11^{-1}).
// The panel is staged through LDS into registers: thread (wave w, lane j)
// owns rows i = w + 4q (q < 16) of column j of both A and X, so the rank-1
// updates of every step run in registers and LDS only carries broadcasts:
// the pivot column (colk), the pivot row of X (xrow) and the reciprocal of
// the pivot.  Rows beyond ncols receive the triangular solve.  On return the
// factor and the inverse are back in A and X (LDS, ld 65).
template <typename T>
__device__ __forceinline__ void lds_chol_inv(T* __restrict__ A, T* __restrict__ X,
                                             T* __restrict__ colk, T* __restrict__ xrow,
                                             int nrows, int ncols, int* bad)
{
  const int tid = threadIdx.x, j = tid & 63, w = tid >> 6;
  T a[16], x[16];
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int i = w + 4 * q;
    a[q] = (i < nrows && j < ncols) ? A[i + j * 65] : (T)0;
    x[q] = (i == j) ? (T)1 : (T)0;
  }
  __syncthreads();
  for (int k = 0; k < ncols; k++) {
    const int qk = k >> 2, wk = k & 3;
    // pivot: row k lives in wave wk, register qk, lane k
    if (w == wk && j == k) {
      T p = a[0];
#pragma unroll
      for (int q = 1; q < 16; q++)
        if (q == qk) p = a[q];
      if (!(p > (T)0)) { *bad = 1; p = (T)1; }
      const T d = sqrt(p);
      colk[64] = d;
      colk[65] = (T)1 / d;
    }
    __syncthreads();
    const T dk = colk[64], rdk = colk[65];
    if (j == k) {  // column k: scale rows below, publish
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int i = w + 4 * q;
        if (i == k) a[q] = dk;
        if (i > k && i < nrows) {
          a[q] = a[q] * rdk;
          colk[i] = a[q];
        }
      }
    }
    if (w == wk && j <= k) {  // row k of X: scale, publish
#pragma unroll
      for (int q = 0; q < 16; q++)
        if (q == qk) {
          x[q] = x[q] * rdk;
          xrow[j] = x[q];
        }
    }
    __syncthreads();
    {
      const T lj = (j > k && j < 64) ? colk[j] : (T)0;
      const T xk = j <= k ? xrow[j] : (T)0;
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int i = w + 4 * q;
        if (i > k && i < nrows) {
          const T ci = colk[i];
          if (j > k && j < ncols && j <= i) a[q] -= ci * lj;
          if (i < ncols && j <= k) x[q] -= ci * xk;
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int i = w + 4 * q;
    A[i + j * 65] = a[q];
    X[i + j * 65] = x[q];
  }
  __syncthreads();
}

// Store X (ncols x ncols, lower) row-major zero padded to 64 x 64.
template <typename T>
__device__ __forceinline__ void store_inv(const T* X, int kb, T* __restrict__ Inv)
{
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    int c = e & 63, i = e >> 6;
    Inv[i * 64 + c] = (i < kb && c < kb && c <= i) ? X[i + c * 65] : (T)0;
  }
}

// Whole front (m <= 64): POTRF + TRSM + inverse in LDS, SYRK into U.
template <typename T>
__global__ __launch_bounds__(256) void k_front_small(const FrontDesc* __restrict__ fr,
                                                     const int* __restrict__ list,
                                                     T* __restrict__ L, T* __restrict__ U,
                                                     T* __restrict__ Inv, int* __restrict__ info)
{
  __shared__ T A[64 * 65];
  __shared__ T X[64 * 65];
  __shared__ T colk[66];
  __shared__ T xrow[64];
  __shared__ int bad;
  const FrontDesc d = fr[list[blockIdx.x]];
  const int s = d.s, m = d.m, u = m - s, tid = threadIdx.x;
  T* gp = L + d.offL;
  if (tid == 0) bad = 0;
  for (int e = tid; e < 64 * s; e += 256) {
    int i = e & 63, c = e >> 6;
    if (i < m) A[i + c * 65] = gp[i + (long long)c * m];
  }
  __syncthreads();
  lds_chol_inv(A, X, colk, xrow, m, s, &bad);
  for (int e = tid; e < 64 * s; e += 256) {
    int i = e & 63, c = e >> 6;
    if (i < m && i >= c) gp[i + (long long)c * m] = A[i + c * 65];
  }
  store_inv(X, s, Inv + d.offI);
  if (u > 0) {
    T* gu = U + d.offU;
    for (int e = tid; e < u * u; e += 256) {
      int i = e % u, j = e / u;
      if (i < j) continue;
      T acc = (T)0;
      for (int k = 0; k < s; k++) acc += A[s + i + k * 65] * A[s + j + k * 65];
      gu[i + (long long)j * u] -= acc;
    }
  }
  if (tid == 0 && bad) atomicOr(info, 1);
}

// Cholesky of one kb x kb diagonal block of a large front's panel, and the
// inverse of the factored block.
template <typename T>
__global__ __launch_bounds__(256) void k_potrf_diag(const FrontDesc* __restrict__ fr,
                                                    const PanelTask* __restrict__ tasks,
                                                    T* __restrict__ L, T* __restrict__ Inv,
                                                    int* __restrict__ info)
{
  __shared__ T A[64 * 65];
  __shared__ T X[64 * 65];
  __shared__ T colk[66];
  __shared__ T xrow[64];
  __shared__ int bad;
  const PanelTask t = tasks[blockIdx.x];
  const FrontDesc d = fr[t.front];
  const int m = d.m, tid = threadIdx.x, kb = t.kb;
  T* gp = L + d.offL + t.k0 + (long long)t.k0 * m;
  if (tid == 0) bad = 0;
  for (int e = tid; e < 64 * kb; e += 256) {
    int i = e & 63, c = e >> 6;
    if (i < kb) A[i + c * 65] = gp[i + (long long)c * m];
  }
  __syncthreads();
  lds_chol_inv(A, X, colk, xrow, kb, kb, &bad);
  for (int e = tid; e < 64 * kb; e += 256) {
    int i = e & 63, c = e >> 6;
    if (i < kb && i >= c) gp[i + (long long)c * m] = A[i + c * 65];
  }
  store_inv(X, kb, Inv + d.offI + (long long)(t.k0 / 64) * 4096);
  if (tid == 0 && bad) atomicOr(info, 1);
}

// ---- one-wave Cholesky + inverse of a 64 x 64 diagonal block ----------------
// Lane i holds row i of the block in 64 registers (loops fully unrolled, so
// every register index is a constant).  Column k is published through LDS
// once per step and read back by every lane as broadcast loads (two values
// per 16-byte load): a wave's LDS operations complete in issue order, so the
// single wave needs no s_barrier, only the compiler fence below.  Updates
// are applied unmasked: entries above the diagonal take garbage and are
// never read or stored.
//   Cholesky, right-looking (src/taucs_sn_llt.c:1112-1120 calls dpotrf on
//   the same block): d = sqrt(a_kk), l_ik = a_ik / d, a_ij -= l_ik l_jk.
//   Inverse X = L^{-1}: L goes to LDS row-major; lane c forms column c by
//   forward substitution x_ic = (delta_ic - sum_{t<i} l_it x_tc) / l_ii, with
//   the l_it as broadcast loads and x_tc in the registers L's row vacated.
// A block narrower than 64 is padded with the identity, whose factor and
// inverse are the identity.
#define PW_LD 66

// v_readlane broadcast of one lane's value (lane index a constant)
template <typename T>
__device__ __forceinline__ T rlane(T v, int l);
template <>
__device__ __forceinline__ double rlane<double>(double v, int l)
{
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
template <>
__device__ __forceinline__ float rlane<float>(float v, int l)
{
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// d = sqrt(p) and rd = 1/d for a pivot p > 0.  FP32: correctly rounded.
// FP64: Newton's iteration for y = 1/sqrt(p), y <- y + y (1 - p y^2) / 2,
// twice from the hardware estimate (each step squares the relative error),
// then d = p y corrected once by the same iteration for sqrt,
// d <- d + y (p - d^2) / 2.  Both end within about one ulp, and the chain is
// half as long as a correctly rounded sqrt followed by a division: the pivot
// step is the latency of the whole block.
template <typename T>
__device__ __forceinline__ void sqrt_pair(T p, T& d, T& rd)
{
  d = sqrt(p);
  rd = (T)1 / d;
}
template <>
__device__ __forceinline__ void sqrt_pair<double>(double p, double& d, double& rd)
{
  double y = __builtin_amdgcn_rsq(p);
  double h = __builtin_fma(-p * y, y, 1.0);
  y = __builtin_fma(0.5 * y, h, y);
  h = __builtin_fma(-p * y, y, 1.0);
  y = __builtin_fma(0.5 * y, h, y);
  const double s = p * y;
  d = __builtin_fma(0.5 * y, __builtin_fma(-s, s, p), s);
  rd = y;
}

__device__ __forceinline__ void wave_lds_fence()
{
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __builtin_amdgcn_wave_barrier();
}

// Pivot growth that flags a reduced-precision diagonal block: a pivot below
// 2^-12 of the block's input diagonal entry has lost at least 12 of FP32's 24
// bits to cancellation.  Such a block is refactored in FP64.
#define GROW_TAU (1.0f / 4096.0f)

// Pivot k is eliminated when bit k of `piv` is set (wave-uniform): all 64 for
// a diagonal block, the first s columns of each front for whole fronts.
template <typename T>
__device__ __forceinline__ void wave_chol_inv(T (&r)[64], int lane, unsigned long long piv, T* col,
                                              T* Ls, T* rdiag, const T* dg, bool& bad, bool& grow)
{
#pragma clang loop unroll(full)
  for (int k = 0; k < 64; k++) {
    if ((piv >> k) & 1ULL) {
      T p = rlane(r[k], k);
      if constexpr (sizeof(T) == 4) {
        if (p < (T)GROW_TAU * dg[k]) grow = true;
      }
      if (!(p > (T)0)) { bad = true; p = (T)1; }
      T dk, rdk;
      sqrt_pair(p, dk, rdk);
      const T lik = r[k] * rdk;
      r[k] = lane == k ? dk : lik;
      col[lane] = r[k];
      rdiag[k] = rdk;
      wave_lds_fence();
      // the broadcasts go in batches of 16: all loads of a batch are issued
      // before the first is used, so their latencies overlap
#pragma clang loop unroll(full)
      for (int j0 = k + 1; j0 < 64; j0 += 16) {
        T v[16];
#pragma clang loop unroll(full)
        for (int q = 0; q < 16; q++)
          if (j0 + q < 64) v[q] = col[j0 + q];
        __builtin_amdgcn_sched_barrier(0);
#pragma clang loop unroll(full)
        for (int q = 0; q < 16; q++)
          if (j0 + q < 64) r[j0 + q] -= lik * v[q];
      }
      // the updated entries, pinned here: left free, the compiler defers each
      // entry's updates to its own pivot step and keeps every broadcast value
      // live until then (several hundred registers)
#pragma clang loop unroll(full)
      for (int j = k + 1; j < 64; j++) asm volatile("" : "+v"(r[j]));
      __builtin_amdgcn_wave_barrier();
    }
  }
}

// X = L^{-1} of the 64 x 64 block by 16 x 16 blocks, one wave.  L goes to LDS
// row-major (Ls[i*PW_LD + c]).  The four diagonal blocks are inverted by
// forward substitution (lane = block b, column c; x_ic = (delta_ic - sum_t
// l_it x_tc) / l_ii).  The off-diagonal blocks follow by block distance on
// the 16x16x4 MFMA, X_ij = -X_ii (sum_{k=j}^{i-1} L_ik X_kj).  X is kept
// transposed in the upper triangle of Ls (X(r, c) at Ls[c*PW_LD + r], r >= c;
// L's diagonal is only needed through rdiag), so no second 64 x 64 buffer is
// needed; reads of X mask the entries above its diagonal.  Registers then
// hold column `lane` of the inverse, as the stores expect.  Rows with
// rdiag = 1 and unit rows of Ls (the identity padding) give the identity.
template <typename T>
__device__ __forceinline__ void wave_inverse(T (&r)[64], int lane, T* Ls, const T* rdiag, T* Ts)
{
  typedef typename Mfma<T>::acc_t acc_t;
#pragma clang loop unroll(full)
  for (int c = 0; c < 64; c++) Ls[lane * PW_LD + c] = r[c];
  wave_lds_fence();
  const int b = lane >> 4, c16 = lane & 15, kq = lane >> 4;
  {
    T x[16];
    const T* Lb = Ls + (16 * b) * PW_LD + 16 * b;
#pragma clang loop unroll(full)
    for (int i = 0; i < 16; i++) {
      T a0 = (i == c16) ? (T)1 : (T)0, a1 = (T)0;
#pragma clang loop unroll(full)
      for (int t = 0; t < i; t += 2) {
        a0 -= Lb[i * PW_LD + t] * x[t];
        if (t + 1 < i) a1 -= Lb[i * PW_LD + t + 1] * x[t + 1];
      }
      x[i] = (a0 + a1) * rdiag[16 * b + i];
      asm volatile("" : "+v"(x[i]));  // row i done here (see wave_chol_inv)
    }
    // X(16b+i, 16b+c16) for i >= c16, at the transposed (upper) position
#pragma clang loop unroll(full)
    for (int i = 0; i < 16; i++)
      if (i >= c16) Ls[(16 * b + c16) * PW_LD + 16 * b + i] = x[i];
  }
  wave_lds_fence();
#pragma clang loop unroll(full)
  for (int d = 1; d < 4; d++) {
#pragma clang loop unroll(full)
    for (int j = 0; j + d < 4; j++) {
      const int i = j + d;
      acc_t acc;
#pragma clang loop unroll(full)
      for (int q = 0; q < 4; q++) acc[q] = (T)0;
#pragma clang loop unroll(full)
      for (int k = j; k < i; k++)
#pragma clang loop unroll(full)
        for (int st = 0; st < 4; st++) {
          const int kr = 4 * st + kq;  // row of X_kj, column of L_ik
          const T a = Ls[(16 * i + c16) * PW_LD + 16 * k + kr];
          const T xv = (k > j || kr >= c16) ? Ls[(16 * j + c16) * PW_LD + 16 * k + kr] : (T)0;
          acc = Mfma<T>::step(a, xv, acc);
        }
#pragma clang loop unroll(full)
      for (int q = 0; q < 4; q++) Ts[Mfma<T>::drow(lane, q) * 17 + c16] = acc[q];
      wave_lds_fence();
      acc_t xr;
#pragma clang loop unroll(full)
      for (int q = 0; q < 4; q++) xr[q] = (T)0;
#pragma clang loop unroll(full)
      for (int st = 0; st < 4; st++) {
        const int kr = 4 * st + kq;  // column of X_ii
        const T a = (c16 >= kr) ? Ls[(16 * i + kr) * PW_LD + 16 * i + c16] : (T)0;
        xr = Mfma<T>::step(a, Ts[kr * 17 + c16], xr);
      }
#pragma clang loop unroll(full)
      for (int q = 0; q < 4; q++)
        Ls[(16 * j + c16) * PW_LD + 16 * i + Mfma<T>::drow(lane, q)] = -xr[q];
      wave_lds_fence();
    }
  }
#pragma clang loop unroll(full)
  for (int i = 0; i < 64; i++) r[i] = i >= lane ? Ls[lane * PW_LD + i] : (T)0;
  __builtin_amdgcn_wave_barrier();
}

// FP64 fallback of a flagged FP32 diagonal block.  The block's input (not yet
// overwritten in global memory) is reloaded as FP64 into the lower triangle
// packed by rows in LDS (row i at i(i+1)/2: 2080 doubles, the footprint of the
// FP32 staging buffer Ls), factored there for its first ns pivots by a
// right-looking Cholesky (lane i owns row i; the column of pivot k is scaled,
// then each lane updates its own row), and rounded back into the FP32
// registers and reciprocal diagonal that the inverse and the stores read.
// Rows and columns >= ns keep the FP64 Schur complement.
__device__ __forceinline__ int ptri(int i, int j) { return ((i * (i + 1)) >> 1) + j; }

__device__ __forceinline__ void wave_chol_packed_f64(double* P, int lane, unsigned long long piv,
                                                     bool& bad)
{
  for (int k = 0; k < 64; k++) {
    if (!((piv >> k) & 1ULL)) continue;
    wave_lds_fence();
    double p = P[ptri(k, k)];
    if (!(p > 0.0)) { bad = true; p = 1.0; }
    const double dk = sqrt(p), rdk = 1.0 / dk;
    double lik = 0.0;
    if (lane > k) {
      lik = P[ptri(lane, k)] * rdk;
      P[ptri(lane, k)] = lik;
    }
    wave_lds_fence();
    if (lane == k) P[ptri(k, k)] = dk;
    if (lane > k)
      for (int j = k + 1; j <= lane; j++) P[ptri(lane, j)] -= lik * P[ptri(j, k)];
  }
  wave_lds_fence();
}

template <typename T>
__device__ __forceinline__ void fp64_refactor(T (&r)[64], int lane, unsigned long long piv, T* Ls,
                                              T* rdiag, bool& bad)
{
  double* P = reinterpret_cast<double*>(Ls);
  bad = false;
  wave_chol_packed_f64(P, lane, piv, bad);
#pragma unroll
  for (int c = 0; c < 64; c++) r[c] = c <= lane ? (T)P[ptri(lane, c)] : (T)0;
  rdiag[lane] = ((piv >> lane) & 1ULL) ? (T)(1.0 / P[ptri(lane, lane)]) : (T)1;
  wave_lds_fence();
}

template <typename T>
__global__ __launch_bounds__(64) void k_potrf_wave(const FrontDesc* __restrict__ fr,
                                                   const PanelTask* __restrict__ tasks,
                                                   T* __restrict__ L, T* __restrict__ Inv,
                                                   int* __restrict__ info, int fallback)
{
  __shared__ __align__(16) T col[64];
  __shared__ __align__(16) T Ls[64 * PW_LD];
  __shared__ __align__(16) T rdiag[64];
  __shared__ __align__(16) T Ts[16 * 17];
  __shared__ __align__(16) T dg[64];
  // the panel chain is the critical path: its waves win issue arbitration
  // against the trailing-update waves that share their SIMD
  __builtin_amdgcn_s_setprio(3);
  const PanelTask t = tasks[blockIdx.x];
  const FrontDesc d = fr[t.front];
  const long long m = d.m;
  const int kb = t.kb, lane = threadIdx.x;
  T* gp = L + d.offL + t.k0 + (long long)t.k0 * m;
  T r[64];
#pragma unroll
  for (int c = 0; c < 64; c++)
    r[c] = (lane < kb && c < kb) ? gp[lane + c * m] : (lane == c ? (T)1 : (T)0);
  if constexpr (sizeof(T) == 4) {
    T v = (T)1;
#pragma unroll
    for (int c = 0; c < 64; c++)
      if (c == lane) v = r[c];
    dg[lane] = v;
    wave_lds_fence();
  }
  bool bad = false, grow = false;
  wave_chol_inv<T>(r, lane, ~0ULL, col, Ls, rdiag, dg, bad, grow);
  if constexpr (sizeof(T) == 4) {
    if (fallback && (bad || grow)) {  // wave-uniform (pivots come by readlane)
      double* P = reinterpret_cast<double*>(Ls);
      const T* g0 = gp;
      asm volatile("" : "+s"(g0));
      for (int c = 0; c <= lane; c++)
        P[ptri(lane, c)] = (lane < kb && c < kb) ? (double)g0[lane + c * m] : (lane == c ? 1.0 : 0.0);
      fp64_refactor<T>(r, lane, ~0ULL, Ls, rdiag, bad);
      if (lane == 0) atomicAdd(info + 1, 1);
    }
  }
  T* gw = gp;  // a fresh copy, so the loads' addresses need not stay live
  asm volatile("" : "+s"(gw));
#pragma unroll
  for (int c = 0; c < 64; c++)
    if (lane < kb && c < kb && lane >= c) gw[lane + c * m] = r[c];
  wave_inverse<T>(r, lane, Ls, rdiag, Ts);
  T* iv = Inv + d.offI + (long long)(t.k0 / 64) * 4096;
  asm volatile("" : "+s"(iv));
#pragma unroll
  for (int i = 0; i < 64; i++) iv[i * 64 + lane] = (lane < kb && i < kb) ? r[i] : (T)0;
  if (lane == 0 && bad) atomicOr(info, 1);
}

// The diagonal block of a panel step by four waves (256 threads).  The
// one-wave kernel above is bound by its instruction stream: one wave issues
// every pivot's chain (broadcast, square root, scaling) and all 2016 rank-1
// FMAs, so FP32 runs no faster than FP64.  Here the block lives in LDS
// (column-major, A(i, c) at A[c*PM_LDA + i]) and is factored in 16-column
// panels: wave 0 factors panel p in registers (lane = row, its 16 panel
// entries; pivot by lane broadcast and sqrt_pair, the panel's own updates
// with the multipliers broadcast by readlane), then all four waves apply the
// panel to the 16 x 16 blocks right of it as rank-16 MFMA updates,
// C -= L_Qp L_Rp^T.  The inverse X = L^{-1} follows in the same LDS block,
// transposed into its upper triangle: the four 16 x 16 diagonal blocks by
// substitution (one wave, lane = (block, column)), then the off-diagonal
// blocks by distance on MFMA, X_ij = -X_ii sum_k L_ik X_kj, one block per
// wave.  Loads and stores go through LDS, coalesced.  The MFMA steps apply
// their products as one fused multiply-add after another in k order, so L
// and the inverse are bitwise those of k_potrf_wave.
#define PM_LDA 65

// C(Q, R) -= L(Q, p) L(R, p)^T on 16 x 16 blocks of A, by one wave; on a
// diagonal block only its lower triangle is written
template <typename T>
__device__ __forceinline__ void pm_block_update(T* A, int Q, int R, int p, int lane)
{
  typedef typename Mfma<T>::acc_t acc_t;
  const int c16 = lane & 15, g = lane >> 4;
  acc_t acc;
#pragma clang loop unroll(full)
  for (int e = 0; e < 4; e++) acc[e] = A[(16 * R + c16) * PM_LDA + 16 * Q + Mfma<T>::drow(lane, e)];
#pragma clang loop unroll(full)
  for (int st = 0; st < 4; st++) {
    const int k = 16 * p + 4 * st + g;
    const T a = -A[k * PM_LDA + 16 * Q + c16];
    const T b = A[k * PM_LDA + 16 * R + c16];
    acc = Mfma<T>::step(a, b, acc);
  }
#pragma clang loop unroll(full)
  for (int e = 0; e < 4; e++) {
    const int i = Mfma<T>::drow(lane, e);
    if (Q != R || i >= c16) A[(16 * R + c16) * PM_LDA + 16 * Q + i] = acc[e];
  }
}

// LDS of potrf_mw_block, in bytes: the block A (64 x PM_LDA), rdiag, dg, the
// four waves' 16 x 17 staging tiles Ts, and one flag
template <typename T>
constexpr int potrf_mw_lds() { return (int)sizeof(T) * (64 * PM_LDA + 64 + 64 + 4 * 16 * 17) + 16; }

template <typename T>
__device__ __forceinline__ void potrf_mw_block(const FrontDesc* __restrict__ fr, const PanelTask t,
                                               T* __restrict__ L, T* __restrict__ Inv,
                                               int* __restrict__ info, int fallback,
                                               unsigned char* smem)
{
  typedef typename Mfma<T>::acc_t acc_t;
  // L in the lower triangle; its inverse X, transposed, in the upper one
  // (X(r, c) at A[r*PM_LDA + c], r >= c; L's diagonal is kept in rdiag)
  // (FP32: the FP64 fallback's packed triangle, 2080 doubles, takes the
  // same 16,640 bytes); all in the caller's LDS buffer of potrf_mw_lds<T>()
  T* const A = reinterpret_cast<T*>(smem);
  T* const rdiag = A + 64 * PM_LDA;
  T* const dg = rdiag + 64;
  T (*const Ts)[16 * 17] = reinterpret_cast<T (*)[16 * 17]>(dg + 64);
  int& flagged = *reinterpret_cast<int*>(reinterpret_cast<unsigned char*>(dg + 64 + 4 * 16 * 17));
  __builtin_amdgcn_s_setprio(3);  // the panel chain is the critical path
  const FrontDesc d = fr[t.front];
  const long long m = d.m;
  const int kb = t.kb, tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  T* const gp = L + d.offL + t.k0 + (long long)t.k0 * m;
  // the lower triangle (identity padding beyond kb, zeros above), the 16
  // loads of a thread issued unconditionally and in flight together (an
  // entry outside the triangle reads the block's first entry and drops it)
  T v[16];
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int c = wave + 4 * q;
    const bool ok = lane < kb && c < kb && lane >= c;
    v[q] = *(ok ? gp + lane + c * m : gp);
  }
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int c = wave + 4 * q;
    if (!(lane < kb && c < kb && lane >= c)) v[q] = lane == c ? (T)1 : (T)0;
    A[c * PM_LDA + lane] = v[q];
    if (lane == c) dg[c] = v[q];
  }
  if (tid == 0) flagged = 0;
  __syncthreads();
  bool bad = false, grow = false;
#pragma clang loop unroll(full)
  for (int p = 0; p < 4; p++) {
    if (wave == 0) {
      const int c0 = 16 * p;
      T r[16];
#pragma unroll
      for (int c = 0; c < 16; c++) r[c] = A[(c0 + c) * PM_LDA + lane];
      T myrd = (T)1;
#pragma clang loop unroll(full)
      for (int q = 0; q < 16; q++) {
        const int k = c0 + q;
        T pv = rlane(r[q], k);
        if constexpr (sizeof(T) == 4) {
          if (pv < (T)GROW_TAU * dg[k]) grow = true;
        }
        // a nonpositive pivot is recorded beside the chain; its NaNs or
        // infinities run on through a factor that is reported as failed
        bad = bad || !(pv > (T)0);
        T dk, rdk;
        sqrt_pair(pv, dk, rdk);
        const T lik = r[q] * rdk;
        r[q] = lane == k ? dk : lik;
        myrd = lane == k ? rdk : myrd;
        // (no register barrier: 16 columns cannot spill, and the scheduler
        // can then overlap the next pivot's chain with these updates)
#pragma clang loop unroll(full)
        for (int j = q + 1; j < 16; j++) r[j] -= lik * rlane(lik, c0 + j);
      }
#pragma unroll
      for (int c = 0; c < 16; c++) A[(c0 + c) * PM_LDA + lane] = lane >= c0 + c ? r[c] : (T)0;
      if (lane >= c0 && lane < c0 + 16) rdiag[lane] = myrd;
    }
    __syncthreads();
    if (p < 3) {
      // the blocks (Q, R), Q >= R > p, round robin over the waves
      int n = 0;
#pragma clang loop unroll(full)
      for (int R = p + 1; R < 4; R++)
#pragma clang loop unroll(full)
        for (int Q = R; Q < 4; Q++) {
          if ((n & 3) == wave) pm_block_update<T>(A, Q, R, p, lane);
          n++;
        }
      __syncthreads();
    }
  }
  if constexpr (sizeof(T) == 4) {
    // a flagged block (nonpositive pivot or pivot growth) again in FP64 from
    // its input, by wave 0 in the packed triangle
    if (wave == 0 && fallback && (bad || grow) && lane == 0) flagged = 1;
    __syncthreads();
    if (flagged) {
      if (wave == 0) {
        static_assert(sizeof(T) != 4 || 64 * PM_LDA * sizeof(T) >= 2080 * sizeof(double), "");
        double* P = reinterpret_cast<double*>(A);  // the FP32 factor is discarded
        for (int c = 0; c <= lane; c++)
          P[ptri(lane, c)] = (lane < kb && c < kb) ? (double)gp[lane + c * m] : (lane == c ? 1.0 : 0.0);
        bool bad64 = false;
        wave_chol_packed_f64(P, lane, ~0ULL, bad64);
        T rd = (T)(1.0 / P[ptri(lane, lane)]);
        T v[64];
#pragma unroll
        for (int c = 0; c < 64; c++) v[c] = c <= lane ? (T)P[ptri(lane, c)] : (T)0;
        wave_lds_fence();
#pragma unroll
        for (int c = 0; c < 64; c++) A[c * PM_LDA + lane] = v[c];
        rdiag[lane] = rd;
        bad = bad64;
        if (lane == 0) atomicAdd(info + 1, 1);
      }
      __syncthreads();
    }
  }
  if (tid == 0 && bad) atomicOr(info, 1);
  // L back to the panel
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int c = wave + 4 * q;
    if (lane < kb && c < kb && lane >= c) gp[lane + c * m] = A[c * PM_LDA + lane];
  }
  __syncthreads();  // the diagonal of A is overwritten by X's next
  // inverse: the diagonal 16 x 16 blocks, lane (b, c16) holds column c16 of X_bb
  const int c16 = lane & 15, g = lane >> 4;
  if (wave == 0) {
    const int b = g;
    T x[16];
#pragma clang loop unroll(full)
    for (int i = 0; i < 16; i++) {
      asm volatile("" ::: "memory");  // one row's entries in registers at a time
      T a0 = (i == c16) ? (T)1 : (T)0, a1 = (T)0;
#pragma clang loop unroll(full)
      for (int tt = 0; tt < i; tt += 2) {
        a0 -= A[(16 * b + tt) * PM_LDA + 16 * b + i] * x[tt];
        if (tt + 1 < i) a1 -= A[(16 * b + tt + 1) * PM_LDA + 16 * b + i] * x[tt + 1];
      }
      x[i] = (a0 + a1) * rdiag[16 * b + i];
      asm volatile("" : "+v"(x[i]));
    }
    wave_lds_fence();  // every row of L_bb read before X_bb takes the upper triangle
#pragma unroll
    for (int i = 0; i < 16; i++)
      if (i >= c16) A[(16 * b + i) * PM_LDA + 16 * b + c16] = x[i];
  }
  __syncthreads();
  // off-diagonal blocks by distance: X_ij = -X_ii (sum_{k=j}^{i-1} L_ik X_kj);
  // X_kj's entries above its diagonal (k = j) and X_ii's are masked: there
  // the upper triangle holds L
#pragma clang loop unroll(full)
  for (int dd = 1; dd < 4; dd++) {
    const int j = wave, i = wave + dd;
    if (i < 4) {
      acc_t acc;
#pragma clang loop unroll(full)
      for (int e = 0; e < 4; e++) acc[e] = (T)0;
      for (int k = j; k < i; k++)
#pragma clang loop unroll(full)
        for (int st = 0; st < 4; st++) {
          const int kr = 4 * st + g;
          const T a = A[(16 * k + kr) * PM_LDA + 16 * i + c16];
          const T bv = (k > j || kr >= c16) ? A[(16 * k + kr) * PM_LDA + 16 * j + c16] : (T)0;
          acc = Mfma<T>::step(a, bv, acc);
        }
      T* ts = Ts[wave];
#pragma clang loop unroll(full)
      for (int e = 0; e < 4; e++) ts[Mfma<T>::drow(lane, e) * 17 + c16] = acc[e];
      wave_lds_fence();
      acc_t xr;
#pragma clang loop unroll(full)
      for (int e = 0; e < 4; e++) xr[e] = (T)0;
#pragma clang loop unroll(full)
      for (int st = 0; st < 4; st++) {
        const int kr = 4 * st + g;
        const T a = c16 >= kr ? A[(16 * i + c16) * PM_LDA + 16 * i + kr] : (T)0;
        xr = Mfma<T>::step(a, ts[kr * 17 + c16], xr);
      }
#pragma clang loop unroll(full)
      for (int e = 0; e < 4; e++) A[(16 * i + Mfma<T>::drow(lane, e)) * PM_LDA + 16 * j + c16] = -xr[e];
    }
    __syncthreads();
  }
  // Inv row-major, X(i, j) at Inv[i*64 + j] (zero outside the block's kb x kb)
  T* const iv = Inv + d.offI + (long long)(t.k0 / 64) * 4096;
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int idx = tid + 256 * q, i = idx >> 6, j = idx & 63;
    iv[idx] = (i < kb && j < kb && i >= j) ? A[i * PM_LDA + j] : (T)0;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_potrf_mw(const FrontDesc* __restrict__ fr,
                                                  const PanelTask* __restrict__ tasks,
                                                  T* __restrict__ L, T* __restrict__ Inv,
                                                  int* __restrict__ info, int fallback)
{
  __shared__ __align__(16) unsigned char smem[potrf_mw_lds<T>()];
  potrf_mw_block<T>(fr, tasks[blockIdx.x], L, Inv, info, fallback, smem);
}

// Whole front with m <= 64 in one wavefront: lane i holds row i of the
// m x m front [F1 .; F2 U] (panel columns from L, the update block from U,
// which already holds the children's extend-adds).  The first s steps leave
// L in the panel columns and the Schur complement U - F2 F2^T (the
// reference's dsyrk, src/taucs_sn_llt.c:1178-1186) in the update block; then
// the s x s diagonal block is inverted.  Rows and columns beyond m are
// identity padding.
template <typename T>
__global__ __launch_bounds__(64) void k_front_wave(const FrontDesc* __restrict__ fr,
                                                   const int* __restrict__ list,
                                                   T* __restrict__ L, T* __restrict__ U,
                                                   T* __restrict__ Inv, int* __restrict__ info,
                                                   int fallback)
{
  __shared__ __align__(16) T col[64];
  __shared__ __align__(16) T Ls[64 * PW_LD];
  __shared__ __align__(16) T rdiag[64];
  __shared__ __align__(16) T Ts[16 * 17];
  __shared__ __align__(16) T dg[64];
  const FrontDesc d = fr[list[blockIdx.x]];
  const int s = d.s, m = d.m, u = m - s, lane = threadIdx.x;
  T* gp = L + d.offL;
  T* gu = U + d.offU;
  // the panel columns whole, the update block's lower triangle (the rest is
  // never written), 1 on the diagonal and 0 elsewhere; every load issued
  // unconditionally (an entry outside the front reads its first entry and
  // drops it), so that all 64 are in flight together
  T r[64];
#pragma unroll
  for (int c = 0; c < 64; c++) {
    const bool ok = c < s ? lane < m : (c < m && lane >= c && lane < m);
    const T* q = c < s ? gp + lane + c * m : gu + (lane - s) + (c - s) * u;
    r[c] = *(ok ? q : gp);
  }
#pragma unroll
  for (int c = 0; c < 64; c++) {
    const bool ok = c < s ? lane < m : (c < m && lane >= c && lane < m);
    if (!ok) r[c] = lane == c ? (T)1 : (T)0;
  }
  if constexpr (sizeof(T) == 4) {
    T v = (T)1;
#pragma unroll
    for (int c = 0; c < 64; c++)
      if (c == lane) v = r[c];
    dg[lane] = v;
    wave_lds_fence();
  }
  bool bad = false, grow = false;
  wave_chol_inv<T>(r, lane, s >= 64 ? ~0ULL : (1ULL << s) - 1, col, Ls, rdiag, dg, bad, grow);
  if constexpr (sizeof(T) == 4) {
    if (fallback && (bad || grow)) {  // wave-uniform (pivots come by readlane)
      double* P = reinterpret_cast<double*>(Ls);
      const T* g0 = gp;
      asm volatile("" : "+s"(g0));
      const T* g1 = gu;
      asm volatile("" : "+s"(g1));
      for (int c = 0; c <= lane; c++) {
        double v = lane == c ? 1.0 : 0.0;
        if (c < s) {
          if (lane < m) v = (double)g0[lane + c * m];
        } else if (c < m) {
          if (lane < m) v = (double)g1[(lane - s) + (c - s) * u];
        }
        P[ptri(lane, c)] = v;
      }
      fp64_refactor<T>(r, lane, s >= 64 ? ~0ULL : (1ULL << s) - 1, Ls, rdiag, bad);
      if (lane == 0) atomicAdd(info + 1, 1);
    }
  }
  // global stores (not flat: the inverse's LDS waits below would then wait
  // for them too)
  T* gw = gp;
  T* guw = gu;
  int ln = lane;
  asm volatile("" : "+v"(ln));
#pragma unroll
  for (int c = 0; c < 64; c++) {
    if (c < s) {
      if (ln < m && ln >= c) gw[ln + c * m] = r[c];
    } else if (c < m) {
      if (ln >= c && ln < m) guw[(ln - s) + (c - s) * u] = r[c];
    }
  }
  // the inverse of the leading s x s block: rows and columns beyond s become
  // the identity (they hold the Schur complement, not L)
#pragma unroll
  for (int c = 0; c < 64; c++)
    if (ln >= s) r[c] = c == ln ? (T)1 : (T)0;
  if (ln >= s) rdiag[ln] = (T)1;
  wave_lds_fence();
  wave_inverse<T>(r, ln, Ls, rdiag, Ts);
  T* iv = Inv + d.offI;
  asm volatile("" : "+s"(iv));
  asm volatile("" : "+v"(ln));
#pragma unroll
  for (int i = 0; i < 64; i++) iv[i * 64 + ln] = (ln < s && i < s) ? r[i] : (T)0;
  if (lane == 0 && bad) atomicOr(info, 1);
}

// Row li of a small front into registers, columns 0..W-1: the panel columns
// from L, the update block's lower triangle from U, and 1 on the diagonal, 0
// elsewhere, for entries right of the diagonal and for rows outside a front
// (valid false).  Every load is issued unconditionally (an entry outside the
// row reads the front's first entry and drops it), so all W are in flight
// together rather than a few per trip of a loop bounded by li.
template <typename T, int W>
__device__ __forceinline__ void load_front_row(T (&r)[W], const T* gp, const T* gu, int li, int s,
                                               int m, int u, bool valid)
{
#pragma unroll
  for (int c = 0; c < W; c++) {
    const bool ok = valid && c <= li;
    const T* q = c < s ? gp + li + (long long)c * m : gu + (li - s) + (long long)(c - s) * u;
    r[c] = *(ok ? q : gp);
  }
#pragma unroll
  for (int c = 0; c < W; c++)
    if (!(valid && c <= li)) r[c] = li == c ? (T)1 : (T)0;
}

// The factored row back: L in the panel columns, the Schur complement in the
// update block's lower triangle (columns 0..li of a valid row).
template <typename T, int W>
__device__ __forceinline__ void store_front_row(const T (&r)[W], T* gp, T* gu, int li, int s,
                                                int m, int u, bool valid)
{
#pragma unroll
  for (int c = 0; c < W; c++)
    if (valid && c <= li) {
      if (c < s) gp[li + (long long)c * m] = r[c];
      else gu[(li - s) + (long long)(c - s) * u] = r[c];
    }
}

// Fronts with m <= 16, four to a wavefront, lane = row within its slot of 16
// lanes and each lane holding only its slot's 16 columns (a fifth of the
// registers of the 64-wide kernels, so several waves share a SIMD).  Every
// operation on a front's entries is the one k_front_wave performs on a front
// at lane 0 (pivot by lane broadcast, sqrt_pair, rank-1 update through LDS,
// the 16 x 16 diagonal-block substitution of wave_inverse), so the results
// are bitwise those of one front per wave.
template <typename T>
__global__ __launch_bounds__(64) void k_front_slot16(const FrontDesc* __restrict__ fr,
                                                     const SmallPack* __restrict__ packs,
                                                     T* __restrict__ L, T* __restrict__ U,
                                                     T* __restrict__ Inv, int* __restrict__ info,
                                                     int fallback)
{
  constexpr int NSB = sizeof(T) == 4 ? 2080 * 8 : 64 * 17 * 8;  // FP32: room for the FP64 fallback
  __shared__ __align__(16) unsigned char sbuf[NSB];
  __shared__ __align__(16) T col[64];
  __shared__ __align__(16) T rdiag[64];
  __shared__ __align__(16) T dg[64];
  T* Ls = reinterpret_cast<T*>(sbuf);  // row `lane`: the 16 columns of its slot (ld 17)
  const SmallPack* pk = packs + blockIdx.x;
  const int lane = threadIdx.x, slot = lane >> 4, base = slot << 4, li = lane & 15;
  const int fi = pk->f[slot];
  int s = 0, m = 0;
  long long offL = 0, offU = 0, offI = 0;
  if (fi >= 0) {
    const FrontDesc d = fr[fi];
    s = d.s;
    m = d.m;
    offL = d.offL;
    offU = d.offU;
    offI = d.offI;
  }
  const int u = m - s;
  T* const gp = L + offL;
  T* const gu = U + offU;
  T r[16];
  load_front_row<T, 16>(r, gp, gu, li, s, m, u, fi >= 0 && li < m);
  if constexpr (sizeof(T) == 4) {
    T v = (T)1;
#pragma unroll
    for (int c = 0; c < 16; c++)
      if (c == li) v = r[c];
    dg[lane] = v;
    wave_lds_fence();
  }
  bool bad = false, grow = false;
#pragma clang loop unroll(full)
  for (int k = 0; k < 16; k++) {
    const T p0 = rlane(r[k], k), p1 = rlane(r[k], 16 + k), p2 = rlane(r[k], 32 + k),
            p3 = rlane(r[k], 48 + k);
    T p = slot == 0 ? p0 : slot == 1 ? p1 : slot == 2 ? p2 : p3;
    const bool act = k < s;
    if (act) {
      if constexpr (sizeof(T) == 4) {
        if (p < (T)GROW_TAU * dg[base + k]) grow = true;
      }
      if (!(p > (T)0)) { bad = true; p = (T)1; }
    }
    T dk, rdk;
    sqrt_pair(p, dk, rdk);
    const T lik = act ? r[k] * rdk : (T)0;
    if (act) {
      r[k] = li == k ? dk : lik;
      if (li == k) rdiag[lane] = rdk;
    }
    col[lane] = r[k];
    wave_lds_fence();
#pragma clang loop unroll(full)
    for (int j = k + 1; j < 16; j++) r[j] -= lik * col[base + j];
#pragma clang loop unroll(full)
    for (int j = k + 1; j < 16; j++) asm volatile("" : "+v"(r[j]));  // as in wave_chol_inv
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_sched_barrier(0);
  }
  if constexpr (sizeof(T) == 4) {
    // a flagged slot anywhere in the wave: every front of the wave in FP64
    if (fallback && __builtin_amdgcn_read_exec() && __any_sync(~0ULL, bad || grow)) {
      double* P = reinterpret_cast<double*>(sbuf);
      unsigned long long piv = 0;
      for (int q = 0; q < 4; q++) {
        const int f = pk->f[q];
        if (f >= 0) piv |= ((1ULL << fr[f].s) - 1) << (16 * q);
      }
      for (int c = 0; c <= lane; c++) P[ptri(lane, c)] = c == lane ? 1.0 : 0.0;
      if (fi >= 0 && li < m) {
        for (int cl = 0; cl <= li; cl++) {
          if (cl < s) P[ptri(lane, base + cl)] = (double)gp[li + (long long)cl * m];
          else P[ptri(lane, base + cl)] = (double)gu[(li - s) + (long long)(cl - s) * u];
        }
      }
      bool bad64 = false;
      wave_chol_packed_f64(P, lane, piv, bad64);
#pragma unroll
      for (int c = 0; c < 16; c++) r[c] = c <= li ? (T)P[ptri(lane, base + c)] : (T)0;
      rdiag[lane] = li < s ? (T)(1.0 / P[ptri(lane, lane)]) : (T)1;
      wave_lds_fence();
      bad = bad64;
      if (lane == 0) atomicAdd(info + 1, 1);
    }
  }
  store_front_row<T, 16>(r, gp, gu, li, s, m, u, fi >= 0 && li < m);
  // inverse of each front's leading s x s block (its other rows: identity)
  if (li >= s) {
#pragma unroll
    for (int c = 0; c < 16; c++) r[c] = c == li ? (T)1 : (T)0;
    rdiag[lane] = (T)1;
  }
  wave_lds_fence();
#pragma unroll
  for (int c = 0; c < 16; c++) Ls[lane * 17 + c] = r[c];
  wave_lds_fence();
  const T* Lb = Ls + base * 17;
  T x[16];
#pragma clang loop unroll(full)
  for (int i = 0; i < 16; i++) {
    // a compiler-only fence per row: the row's entries are loaded where they
    // are used, so the 120 entries never occupy registers at once
    asm volatile("" ::: "memory");
    T a0 = (i == li) ? (T)1 : (T)0, a1 = (T)0;
#pragma clang loop unroll(full)
    for (int t = 0; t < i; t += 2) {
      a0 -= Lb[i * 17 + t] * x[t];
      if (t + 1 < i) a1 -= Lb[i * 17 + t + 1] * x[t + 1];
    }
    x[i] = (a0 + a1) * rdiag[base + i];
    asm volatile("" : "+v"(x[i]));
  }
  if (fi >= 0 && li < s) {
    T* iv = Inv + offI;
#pragma unroll
    for (int i = 0; i < 16; i++)
      if (i < s) iv[i * 64 + li] = i >= li ? x[i] : (T)0;
  }
  if (bad) atomicOr(info, 1);
}

// Fronts with 16 < m <= 32, two to a wavefront (lanes 0-31 and 32-63), each
// lane holding only its slot's 32 columns.  The Cholesky steps are those of
// k_front_slot16; the inverse follows wave_inverse on the wave's block-
// diagonal matrix: the four 16 x 16 diagonal blocks (two per front) by
// substitution, then only the block pairs (1,0) and (3,2), each a front's own
// off-diagonal block, by the same MFMA products.  A front in either slot goes
// through exactly the operations k_front_wave applies to it at lane 0, so the
// results are bitwise those of one front per wave.
template <typename T>
__global__ __launch_bounds__(64) void k_front_slot32(const FrontDesc* __restrict__ fr,
                                                     const SmallPack* __restrict__ packs,
                                                     T* __restrict__ L, T* __restrict__ U,
                                                     T* __restrict__ Inv, int* __restrict__ info,
                                                     int fallback)
{
  typedef typename Mfma<T>::acc_t acc_t;
  // each slot's 32 x 32 block in its own rows (row stride SL, columns local
  // to the slot); FP32 also needs room for the FP64 fallback's triangle
  constexpr int SL = 33;
  constexpr int NSB = sizeof(T) == 4 ? 2080 * 8 : 64 * SL * (int)sizeof(T);
  __shared__ __align__(16) unsigned char sbuf[NSB];
  __shared__ __align__(16) T col[64];
  __shared__ __align__(16) T rdiag[64];
  __shared__ __align__(16) T dg[64];
  __shared__ __align__(16) T Ts[16 * 17];
  T* Ls = reinterpret_cast<T*>(sbuf);  // row r, slot-local column c at Ls[r * SL + c]
  const SmallPack* pk = packs + blockIdx.x;
  const int lane = threadIdx.x, slot = lane >> 5, base = slot << 5, li = lane & 31;
  const int fi = pk->f[slot];
  int s = 0, m = 0;
  long long offL = 0, offU = 0, offI = 0;
  if (fi >= 0) {
    const FrontDesc d = fr[fi];
    s = d.s;
    m = d.m;
    offL = d.offL;
    offU = d.offU;
    offI = d.offI;
  }
  const int u = m - s;
  T* const gp = L + offL;
  T* const gu = U + offU;
  T r[32];
  load_front_row<T, 32>(r, gp, gu, li, s, m, u, fi >= 0 && li < m);
  if constexpr (sizeof(T) == 4) {
    T v = (T)1;
#pragma unroll
    for (int c = 0; c < 32; c++)
      if (c == li) v = r[c];
    dg[lane] = v;
    wave_lds_fence();
  }
  bool bad = false, grow = false;
#pragma clang loop unroll(full)
  for (int k = 0; k < 32; k++) {
    const T p0 = rlane(r[k], k), p1 = rlane(r[k], 32 + k);
    T p = slot == 0 ? p0 : p1;
    const bool act = k < s;
    if (act) {
      if constexpr (sizeof(T) == 4) {
        if (p < (T)GROW_TAU * dg[base + k]) grow = true;
      }
      if (!(p > (T)0)) { bad = true; p = (T)1; }
    }
    T dk, rdk;
    sqrt_pair(p, dk, rdk);
    const T lik = act ? r[k] * rdk : (T)0;
    if (act) {
      r[k] = li == k ? dk : lik;
      if (li == k) rdiag[lane] = rdk;
    }
    col[lane] = r[k];
    wave_lds_fence();
#pragma clang loop unroll(full)
    for (int j0 = k + 1; j0 < 32; j0 += 16) {
      T v[16];
#pragma clang loop unroll(full)
      for (int q = 0; q < 16; q++)
        if (j0 + q < 32) v[q] = col[base + j0 + q];
      __builtin_amdgcn_sched_barrier(0);
#pragma clang loop unroll(full)
      for (int q = 0; q < 16; q++)
        if (j0 + q < 32) r[j0 + q] -= lik * v[q];
    }
#pragma clang loop unroll(full)
    for (int j = k + 1; j < 32; j++) asm volatile("" : "+v"(r[j]));  // as in wave_chol_inv
    __builtin_amdgcn_wave_barrier();
  }
  if constexpr (sizeof(T) == 4) {
    // a flagged slot anywhere in the wave: every front of the wave in FP64
    if (fallback && __any_sync(~0ULL, bad || grow)) {
      double* P = reinterpret_cast<double*>(sbuf);
      unsigned long long piv = 0;
      for (int q = 0; q < 2; q++) {
        const int f = pk->f[q];
        if (f >= 0) piv |= ((1ULL << fr[f].s) - 1) << (32 * q);
      }
      for (int c = 0; c <= lane; c++) P[ptri(lane, c)] = c == lane ? 1.0 : 0.0;
      if (fi >= 0 && li < m) {
        for (int cl = 0; cl <= li; cl++) {
          if (cl < s) P[ptri(lane, base + cl)] = (double)gp[li + (long long)cl * m];
          else P[ptri(lane, base + cl)] = (double)gu[(li - s) + (long long)(cl - s) * u];
        }
      }
      bool bad64 = false;
      wave_chol_packed_f64(P, lane, piv, bad64);
#pragma unroll
      for (int c = 0; c < 32; c++) r[c] = c <= li ? (T)P[ptri(lane, base + c)] : (T)0;
      rdiag[lane] = li < s ? (T)(1.0 / P[ptri(lane, lane)]) : (T)1;
      wave_lds_fence();
      bad = bad64;
      if (lane == 0) atomicAdd(info + 1, 1);
    }
  }
  store_front_row<T, 32>(r, gp, gu, li, s, m, u, fi >= 0 && li < m);
  // inverse of each front's leading s x s block (its other rows: identity)
  if (li >= s) {
#pragma unroll
    for (int c = 0; c < 32; c++) r[c] = c == li ? (T)1 : (T)0;
    rdiag[lane] = (T)1;
  }
  wave_lds_fence();
#pragma unroll
  for (int c = 0; c < 32; c++) Ls[lane * SL + c] = r[c];
  wave_lds_fence();
  const int b = lane >> 4, c16 = lane & 15, kq = lane >> 4;
  {
    T x[16];
    const T* Lb = Ls + (16 * b) * SL + 16 * (b & 1);
#pragma clang loop unroll(full)
    for (int i = 0; i < 16; i++) {
      asm volatile("" ::: "memory");  // one row's entries in registers at a time
      T a0 = (i == c16) ? (T)1 : (T)0, a1 = (T)0;
#pragma clang loop unroll(full)
      for (int t = 0; t < i; t += 2) {
        a0 -= Lb[i * SL + t] * x[t];
        if (t + 1 < i) a1 -= Lb[i * SL + t + 1] * x[t + 1];
      }
      x[i] = (a0 + a1) * rdiag[16 * b + i];
      asm volatile("" : "+v"(x[i]));
    }
#pragma clang loop unroll(full)
    for (int i = 0; i < 16; i++)
      if (i >= c16) Ls[(16 * b + c16) * SL + 16 * (b & 1) + i] = x[i];
  }
  wave_lds_fence();
  // each front's off-diagonal block: X_10 = -X_11 (L_10 X_00), pairs (1,0), (3,2)
#pragma clang loop unroll(full)
  for (int q = 0; q < 2; q++) {
    const int i = 2 * q + 1, j = 2 * q;
    acc_t acc;
#pragma clang loop unroll(full)
    for (int e = 0; e < 4; e++) acc[e] = (T)0;
#pragma clang loop unroll(full)
    for (int st = 0; st < 4; st++) {
      const int kr = 4 * st + kq;
      const T a = Ls[(16 * i + c16) * SL + 16 * (j & 1) + kr];
      const T xv = kr >= c16 ? Ls[(16 * j + c16) * SL + 16 * (j & 1) + kr] : (T)0;
      acc = Mfma<T>::step(a, xv, acc);
    }
#pragma clang loop unroll(full)
    for (int e = 0; e < 4; e++) Ts[Mfma<T>::drow(lane, e) * 17 + c16] = acc[e];
    wave_lds_fence();
    acc_t xr;
#pragma clang loop unroll(full)
    for (int e = 0; e < 4; e++) xr[e] = (T)0;
#pragma clang loop unroll(full)
    for (int st = 0; st < 4; st++) {
      const int kr = 4 * st + kq;
      const T a = (c16 >= kr) ? Ls[(16 * i + kr) * SL + 16 * (i & 1) + c16] : (T)0;
      xr = Mfma<T>::step(a, Ts[kr * 17 + c16], xr);
    }
#pragma clang loop unroll(full)
    for (int e = 0; e < 4; e++)
      Ls[(16 * j + c16) * SL + 16 * (i & 1) + Mfma<T>::drow(lane, e)] = -xr[e];
    wave_lds_fence();
  }
  // lane holds column `lane` of its front's inverse: X(base + i, lane)
  if (fi >= 0 && li < s) {
    T* iv = Inv + offI;
    for (int i = 0; i < s; i++)
      iv[i * 64 + li] = i >= li ? Ls[lane * SL + i] : (T)0;
  }
  if (bad) atomicOr(info, 1);
}

// Fronts with 32 < m <= 48, one to a wavefront: each lane holds 48 columns of
// its row and the LDS staging is 48 wide, 20 KB in FP64 against the 37 KB of
// the 64-row kernel, so twice as many waves fit on a CU.  Lanes 48..63 carry
// no row.  Every operation on the front is the one k_front_wave applies to it
// (pivot by lane broadcast and sqrt_pair for the first s columns, rank-1
// updates in column order, the FP64 fallback over the packed 64-row
// triangle, the diagonal 16 x 16 substitutions and the MFMA products of the
// block pairs (1,0), (2,1), (2,0); the pairs with block 3 only meet its
// identity padding there), so the results are bitwise those of k_front_wave.
template <typename T>
__global__ __launch_bounds__(64) void k_front_w48(const FrontDesc* __restrict__ fr,
                                                  const int* __restrict__ list,
                                                  T* __restrict__ L, T* __restrict__ U,
                                                  T* __restrict__ Inv, int* __restrict__ info,
                                                  int fallback)
{
  typedef typename Mfma<T>::acc_t acc_t;
  constexpr int SL = 49;
  constexpr int NSB = sizeof(T) == 4 ? 2080 * 8 : 48 * SL * (int)sizeof(T);
  __shared__ __align__(16) unsigned char sbuf[NSB];
  __shared__ __align__(16) T col[64];
  __shared__ __align__(16) T rdiag[64];
  __shared__ __align__(16) T dg[64];
  __shared__ __align__(16) T Ts[16 * 17];
  T* Ls = reinterpret_cast<T*>(sbuf);  // row r < 48, column c < 48 at Ls[r * SL + c]
  const FrontDesc d = fr[list[blockIdx.x]];
  const int s = d.s, m = d.m, u = m - s, lane = threadIdx.x;
  const bool act = lane < 48;
  T* const gp = L + d.offL;
  T* const gu = U + d.offU;
  T r[48];  // lanes 48-63: rows of zeros
  load_front_row<T, 48>(r, gp, gu, lane, s, m, u, act && lane < m);
  if constexpr (sizeof(T) == 4) {
    T v = (T)1;
#pragma unroll
    for (int c = 0; c < 48; c++)
      if (c == lane) v = r[c];
    dg[lane] = v;
    wave_lds_fence();
  }
  bool bad = false, grow = false;
#pragma clang loop unroll(full)
  for (int k = 0; k < 48; k++) {
    if (k < s) {
      T p = rlane(r[k], k);
      if constexpr (sizeof(T) == 4) {
        if (p < (T)GROW_TAU * dg[k]) grow = true;
      }
      if (!(p > (T)0)) { bad = true; p = (T)1; }
      T dk, rdk;
      sqrt_pair(p, dk, rdk);
      const T lik = r[k] * rdk;
      r[k] = lane == k ? dk : lik;
      col[lane] = r[k];
      rdiag[k] = rdk;
      wave_lds_fence();
#pragma clang loop unroll(full)
      for (int j0 = k + 1; j0 < 48; j0 += 16) {
        T v[16];
#pragma clang loop unroll(full)
        for (int q = 0; q < 16; q++)
          if (j0 + q < 48) v[q] = col[j0 + q];
        __builtin_amdgcn_sched_barrier(0);
#pragma clang loop unroll(full)
        for (int q = 0; q < 16; q++)
          if (j0 + q < 48) r[j0 + q] -= lik * v[q];
      }
#pragma clang loop unroll(full)
      for (int j = k + 1; j < 48; j++) asm volatile("" : "+v"(r[j]));  // as in wave_chol_inv
      __builtin_amdgcn_wave_barrier();
    }
  }
  if constexpr (sizeof(T) == 4) {
    if (fallback && (bad || grow)) {  // wave-uniform (pivots come by readlane)
      double* P = reinterpret_cast<double*>(sbuf);
      const unsigned long long piv = (1ULL << s) - 1;
      for (int c = 0; c <= lane; c++) {
        double v = lane == c ? 1.0 : 0.0;
        if (c < s) {
          if (lane < m) v = (double)gp[lane + (long long)c * m];
        } else if (c < m) {
          if (lane < m) v = (double)gu[(lane - s) + (long long)(c - s) * u];
        }
        P[ptri(lane, c)] = v;
      }
      bool bad64 = false;
      wave_chol_packed_f64(P, lane, piv, bad64);
#pragma unroll
      for (int c = 0; c < 48; c++) r[c] = c <= lane ? (T)P[ptri(lane, c)] : (T)0;
      rdiag[lane] = ((piv >> lane) & 1ULL) ? (T)(1.0 / P[ptri(lane, lane)]) : (T)1;
      wave_lds_fence();
      bad = bad64;
      if (lane == 0) atomicAdd(info + 1, 1);
    }
  }
  store_front_row<T, 48>(r, gp, gu, lane, s, m, u, act && lane < m);
  // inverse of the leading s x s block (its other rows: identity)
  if (lane >= s) {
#pragma unroll
    for (int c = 0; c < 48; c++) r[c] = c == lane ? (T)1 : (T)0;
    rdiag[lane] = (T)1;
  }
  wave_lds_fence();
  if (act) {
#pragma unroll
    for (int c = 0; c < 48; c++) Ls[lane * SL + c] = r[c];
  }
  wave_lds_fence();
  const int b = lane >> 4, c16 = lane & 15, kq = lane >> 4;
  if (b < 3) {
    T x[16];
    const T* Lb = Ls + (16 * b) * SL + 16 * b;
#pragma clang loop unroll(full)
    for (int i = 0; i < 16; i++) {
      asm volatile("" ::: "memory");  // one row's entries in registers at a time
      T a0 = (i == c16) ? (T)1 : (T)0, a1 = (T)0;
#pragma clang loop unroll(full)
      for (int t = 0; t < i; t += 2) {
        a0 -= Lb[i * SL + t] * x[t];
        if (t + 1 < i) a1 -= Lb[i * SL + t + 1] * x[t + 1];
      }
      x[i] = (a0 + a1) * rdiag[16 * b + i];
      asm volatile("" : "+v"(x[i]));
    }
#pragma clang loop unroll(full)
    for (int i = 0; i < 16; i++)
      if (i >= c16) Ls[(16 * b + c16) * SL + 16 * b + i] = x[i];
  }
  wave_lds_fence();
  // off-diagonal blocks by distance: X_ij = -X_ii (sum_{k=j}^{i-1} L_ik X_kj)
#pragma clang loop unroll(full)
  for (int dd = 1; dd < 3; dd++) {
#pragma clang loop unroll(full)
    for (int j = 0; j + dd < 3; j++) {
      const int i = j + dd;
      acc_t acc;
#pragma clang loop unroll(full)
      for (int e = 0; e < 4; e++) acc[e] = (T)0;
#pragma clang loop unroll(full)
      for (int k = j; k < i; k++)
#pragma clang loop unroll(full)
        for (int st = 0; st < 4; st++) {
          const int kr = 4 * st + kq;
          const T a = Ls[(16 * i + c16) * SL + 16 * k + kr];
          const T xv = (k > j || kr >= c16) ? Ls[(16 * j + c16) * SL + 16 * k + kr] : (T)0;
          acc = Mfma<T>::step(a, xv, acc);
        }
#pragma clang loop unroll(full)
      for (int e = 0; e < 4; e++) Ts[Mfma<T>::drow(lane, e) * 17 + c16] = acc[e];
      wave_lds_fence();
      acc_t xr;
#pragma clang loop unroll(full)
      for (int e = 0; e < 4; e++) xr[e] = (T)0;
#pragma clang loop unroll(full)
      for (int st = 0; st < 4; st++) {
        const int kr = 4 * st + kq;
        const T a = (c16 >= kr) ? Ls[(16 * i + kr) * SL + 16 * i + c16] : (T)0;
        xr = Mfma<T>::step(a, Ts[kr * 17 + c16], xr);
      }
#pragma clang loop unroll(full)
      for (int e = 0; e < 4; e++)
        Ls[(16 * j + c16) * SL + 16 * i + Mfma<T>::drow(lane, e)] = -xr[e];
      wave_lds_fence();
    }
  }
  // the whole 64 x 64 inverse block: X(i, lane) in the transposed upper
  // triangle, zero outside the leading s x s
  T* iv = Inv + d.offI;
  for (int i = 0; i < 64; i++)
    iv[i * 64 + lane] = (lane < s && i < s && i >= lane) ? Ls[lane * SL + i] : (T)0;
  if (lane == 0 && bad) atomicOr(info, 1);
}

// ---- MFMA tile GEMM ---------------------------------------------------------
// 64x64 tile, 4 waves (2x2), each wave a 32x32 sub-tile of 2x2 16x16 MFMA
// blocks.  K is staged through LDS in chunks of 16.  The MFMA is issued as
// D = B_tile * A_tile^T so that a lane's four results are four consecutive
// columns of one row: the epilogue's read-modify-write of C walks down
// columns (16 consecutive rows per lane group).
//
// (Mfma<T>: the 16x16x4 MFMA step and its lane layout, in kernels.h)

#define TK 16
#define TPAD 4

// K is consumed in chunks of 16 (measured on gfx950 against 8, 32 and 64 for
// K = 256 and 512 trailing updates: the small chunk keeps the register and
// LDS footprint low enough for 4 (FP64) / 8 (FP32) waves per SIMD, 8-15%
// faster than chunks of 32): the next chunk is fetched from global memory
// into registers while the MFMAs run on the current chunk in LDS.
// K-major LDS staging: a chunk of TK k-values is stored as As[row][k], and
// the MFMA's k index is assigned so that lane group g (lane / 16) takes the
// consecutive k = g*TK/4 .. g*TK/4 + TK/4 - 1 over the TK/4 steps of the chunk
// (any assignment is valid as long as A and B use the same one).  Each lane
// then reads its fragments for the whole chunk as contiguous 16-byte loads.
// The next chunk is fetched into registers while the MFMAs run.  NBUF = 2:
// it is stored into a second LDS buffer and one barrier per chunk separates
// the phases; NBUF = 1 (half the LDS, so more resident workgroups): stored
// after a second barrier.
#define KP (TK + 2)
template <typename T, int NBUF>
__device__ __forceinline__ void gemm_tile(const GemmTask* __restrict__ tasks, const T* Lsrc, T* L,
                                          T* U, const T* __restrict__ Inv)
{
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ __align__(16) T As[NBUF][64][KP];
  __shared__ __align__(16) T Bs[NBUF][64][KP];
  const GemmTask t = tasks[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const T* A = Lsrc + t.a_off;
  const bool binv = (t.flags & 4) != 0;
  const T* B = binv ? Inv + t.b_off : Lsrc + t.b_off;
  acc_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) acc[a][b][r] = (T)0;

  const int li = tid & 63, lk = tid >> 6;  // loader: row li, k = lk + 4q
  const int lam = min(li, t.M - 1), lbn = min(li, t.N - 1);
  T ra[TK / 4], rb[TK / 4];
  // the chunk's loads, unconditional (entries outside the operands read a
  // clamped in-range one), so that they stay in flight during the MFMAs;
  // stash zeroes those entries as it stores the chunk into LDS (the empty
  // asm there keeps the compiler from sinking a load into that select)
  auto fetch = [&](int kc) {
#pragma unroll
    for (int q = 0; q < TK / 4; q++) {
      const int kcl = min(kc + lk + 4 * q, t.K - 1);
      ra[q] = A[lam + (long long)kcl * t.ld];
      rb[q] = binv ? B[lbn * 64 + kcl] : B[lbn + (long long)kcl * t.ld];
    }
  };
  auto stash = [&](int buf, int kc) {
#pragma unroll
    for (int q = 0; q < TK / 4; q++) {
      asm volatile("" : "+v"(ra[q]));
      asm volatile("" : "+v"(rb[q]));
    }
#pragma unroll
    for (int q = 0; q < TK / 4; q++) {
      const bool kin = kc + lk + 4 * q < t.K;
      As[buf][li][lk + 4 * q] = (kin && li < t.M) ? ra[q] : (T)0;
      Bs[buf][li][lk + 4 * q] = (kin && li < t.N) ? rb[q] : (T)0;
    }
  };
  fetch(0);
  stash(0, 0);
  __syncthreads();
  const int c16 = lane & 15, g = lane >> 4;
  const int nch = (t.K + TK - 1) / TK;
  for (int ch = 0; ch < nch; ch++) {
    const int buf = NBUF == 2 ? (ch & 1) : 0;
    if (ch + 1 < nch) fetch((ch + 1) * TK);
    T fa0[TK / 4], fa1[TK / 4], fb0[TK / 4], fb1[TK / 4];
#pragma unroll
    for (int s2 = 0; s2 < TK / 4; s2++) {
      fa0[s2] = As[buf][wr + c16][g * (TK / 4) + s2];
      fa1[s2] = As[buf][wr + 16 + c16][g * (TK / 4) + s2];
      fb0[s2] = Bs[buf][wc + c16][g * (TK / 4) + s2];
      fb1[s2] = Bs[buf][wc + 16 + c16][g * (TK / 4) + s2];
    }
#pragma unroll
    for (int s2 = 0; s2 < TK / 4; s2++) {
      acc[0][0] = Mfma<T>::step(fb0[s2], fa0[s2], acc[0][0]);
      acc[0][1] = Mfma<T>::step(fb1[s2], fa0[s2], acc[0][1]);
      acc[1][0] = Mfma<T>::step(fb0[s2], fa1[s2], acc[1][0]);
      acc[1][1] = Mfma<T>::step(fb1[s2], fa1[s2], acc[1][1]);
    }
    if (NBUF == 1) __syncthreads();
    if (ch + 1 < nch) stash(NBUF == 2 ? (buf ^ 1) : 0, (ch + 1) * TK);
    __syncthreads();
  }

  // the C entries loaded together (clamped to the tile, dropped outside it)
  T* C = ((t.flags & 1) ? U : L) + t.c_off;
  const bool diag = (t.flags & 2) != 0;
  T cv[2][2][4];
  if (!binv) {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = min(wr + a * 16 + (lane & 15), t.M - 1);
          const int j = min(wc + b * 16 + Mfma<T>::drow(lane, r), t.N - 1);
          cv[a][b][r] = C[i + (long long)j * t.ldc];
        }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int r = 0; r < 4; r++) asm volatile("" : "+v"(cv[a][b][r]));
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int i = wr + a * 16 + (lane & 15);
        int j = wc + b * 16 + Mfma<T>::drow(lane, r);
        if (i < t.M && j < t.N && (!diag || i >= j)) {
          if (binv) C[i + (long long)j * t.ldc] = acc[a][b][r];
          else C[i + (long long)j * t.ldc] = cv[a][b][r] - acc[a][b][r];
        }
      }
}

// The same tile shaped for latency (launches with few tiles, where each
// workgroup has its CU to itself and gemm_tile's one-chunk prefetch pays a
// global round trip per 16-deep chunk): K in groups of four chunks, each
// group's loads issued together (branch-free, out-of-range entries read a
// clamped in-range one and are dropped), so the tile pays one round trip per
// 64 of K, and the next group's loads are issued once the current group's
// last two chunks are in LDS, while they are multiplied.  The chunks, their
// LDS layout and the MFMA order are those of gemm_tile: the results are
// bitwise equal.
template <typename T>
__device__ __forceinline__ void gemm_tile_short(const GemmTask* __restrict__ tasks, const T* Lsrc,
                                                T* L, T* U, const T* __restrict__ Inv)
{
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ __align__(16) T As[2][64][KP];
  __shared__ __align__(16) T Bs[2][64][KP];
  const GemmTask t = tasks[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const T* A = Lsrc + t.a_off;
  const bool binv = (t.flags & 4) != 0;
  const T* B = binv ? Inv + t.b_off : Lsrc + t.b_off;
  acc_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) acc[a][b][r] = (T)0;

  const int li = tid & 63, lk = tid >> 6;  // loader: row li, k = 16 ch + lk + 4q
  const int lam = min(li, t.M - 1), lbn = min(li, t.N - 1);
  T ra[4][TK / 4], rb[4][TK / 4];
  // the group's loads only (clamped, unconditional); the entries outside
  // the operands are zeroed as they are stored into LDS, where the empty asm
  // keeps the compiler from sinking a load into that select
  auto fetch4 = [&](int k0) {
#pragma unroll
    for (int ch = 0; ch < 4; ch++)
#pragma unroll
      for (int q = 0; q < TK / 4; q++) {
        const int kc = min(k0 + ch * TK + lk + 4 * q, t.K - 1);
        ra[ch][q] = A[lam + (long long)kc * t.ld];
        rb[ch][q] = binv ? B[lbn * 64 + kc] : B[lbn + (long long)kc * t.ld];
      }
  };
  fetch4(0);
  const int c16 = lane & 15, g = lane >> 4;
  const int nch = (t.K + TK - 1) / TK;
  for (int c0 = 0; c0 < nch; c0 += 4) {
#pragma unroll
  for (int h = 0; h < 2; h++) {
    if (c0 + 2 * h >= nch) break;
    if (c0 > 0 || h) __syncthreads();  // the previous chunks' fragments are read
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int q = 0; q < TK / 4; q++) {
        asm volatile("" : "+v"(ra[2 * h + c][q]));
        asm volatile("" : "+v"(rb[2 * h + c][q]));
      }
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int q = 0; q < TK / 4; q++) {
        const bool kin = (c0 + 2 * h + c) * TK + lk + 4 * q < t.K;
        As[c][li][lk + 4 * q] = (kin && li < t.M) ? ra[2 * h + c][q] : (T)0;
        Bs[c][li][lk + 4 * q] = (kin && li < t.N) ? rb[2 * h + c][q] : (T)0;
      }
    if (h == 1 && c0 + 4 < nch) fetch4((c0 + 4) * TK);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 2; c++) {
      if (c0 + 2 * h + c >= nch) break;
      T fa0[TK / 4], fa1[TK / 4], fb0[TK / 4], fb1[TK / 4];
#pragma unroll
      for (int s2 = 0; s2 < TK / 4; s2++) {
        fa0[s2] = As[c][wr + c16][g * (TK / 4) + s2];
        fa1[s2] = As[c][wr + 16 + c16][g * (TK / 4) + s2];
        fb0[s2] = Bs[c][wc + c16][g * (TK / 4) + s2];
        fb1[s2] = Bs[c][wc + 16 + c16][g * (TK / 4) + s2];
      }
#pragma unroll
      for (int s2 = 0; s2 < TK / 4; s2++) {
        acc[0][0] = Mfma<T>::step(fb0[s2], fa0[s2], acc[0][0]);
        acc[0][1] = Mfma<T>::step(fb1[s2], fa0[s2], acc[0][1]);
        acc[1][0] = Mfma<T>::step(fb0[s2], fa1[s2], acc[1][0]);
        acc[1][1] = Mfma<T>::step(fb1[s2], fa1[s2], acc[1][1]);
      }
    }
  }
  }

  // the C entries loaded together (clamped to the tile, dropped outside it)
  T* C = ((t.flags & 1) ? U : L) + t.c_off;
  const bool diag = (t.flags & 2) != 0;
  T cv[2][2][4];
  if (!binv) {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = min(wr + a * 16 + (lane & 15), t.M - 1);
          const int j = min(wc + b * 16 + Mfma<T>::drow(lane, r), t.N - 1);
          cv[a][b][r] = C[i + (long long)j * t.ldc];
        }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int r = 0; r < 4; r++) asm volatile("" : "+v"(cv[a][b][r]));
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int i = wr + a * 16 + (lane & 15);
        int j = wc + b * 16 + Mfma<T>::drow(lane, r);
        if (i < t.M && j < t.N && (!diag || i >= j)) {
          if (binv) C[i + (long long)j * t.ldc] = acc[a][b][r];
          else C[i + (long long)j * t.ldc] = cv[a][b][r] - acc[a][b][r];
        }
      }
}

// Trailing update / SYRK tiles (C -= A B^T) and the panel solves (C = A
// inv(L_kk)^T, flag 4) share the tile code; two symbols keep their launches
// apart in profiles.
template <typename T>
__global__ __launch_bounds__(256) void k_gemm_tile(const GemmTask* __restrict__ tasks,
                                                   const T* Lsrc, T* L, T* U,
                                                   const T* __restrict__ Inv)
{
  gemm_tile<T, 1>(tasks, Lsrc, L, U, Inv);
}
// The latency-shaped tile: updates with K <= 64 (the trailing update of a
// front's last outer panel when it is at most 64 columns wide) and the main
// chain's launches with few tiles.
template <typename T>
__global__ __launch_bounds__(256) void k_gemm_short(const GemmTask* __restrict__ tasks,
                                                    const T* Lsrc, T* L, T* U,
                                                    const T* __restrict__ Inv)
{
  __builtin_amdgcn_s_setprio(2);
  gemm_tile_short<T>(tasks, Lsrc, L, U, Inv);
}

// One 64-row tile of a panel step, left-looking inside the outer panel (one
// launch per step on the panel chain, instead of a solve launch and then an
// update launch): with X the panel's columns already solved,
//   C  <- C - Xr Xk^T   the tile's updates from the outer panel's earlier
//                       blocks (Xr: the tile's rows, Xk: the block's rows), in
//                       one accumulation over their kpre columns;
//   X  <- C inv(L_kk)^T the solve by the block inverse (C staged through LDS as
//                       the A operand), stored in place;
//   Cd <- Cd - X X^T    for a tile whose first row starts a later diagonal
//                       block of the outer panel: that block's update from
//                       this step (lower part), so that its POTRF finds it
//                       complete.
// The other tiles of the later blocks receive their updates when their own
// step comes.  The workgroup of the tile that starts the next step's block
// then factors that block (potrf_mw_block, as k_potrf_mw would in a launch of
// its own), so an outer panel's steps after the first need one launch each.
// MFMA layout and 16-deep K chunks as in gemm_tile_short: wave w holds the
// 32 x 32 sub-tile (wr, wc), element r of acc[a][b] at row wr + 16a + c16 and
// column wc + 16b + drow(lane, r).
template <typename T>
__global__ __launch_bounds__(256) void k_panel_solve(const SolveTask* __restrict__ tasks, T* L,
                                                     T* Inv, const FrontDesc* __restrict__ fr,
                                                     int* __restrict__ info, int fallback)
{
  typedef typename Mfma<T>::acc_t acc_t;
  // one LDS buffer: the staging halves As, Bs, then the next block's POTRF
  constexpr int STAGE = 2 * 2 * 64 * KP * (int)sizeof(T);
  __shared__ __align__(16) unsigned char smem[STAGE > potrf_mw_lds<T>() ? STAGE : potrf_mw_lds<T>()];
  T (*const As)[64][KP] = reinterpret_cast<T (*)[64][KP]>(smem);
  T (*const Bs)[64][KP] = As + 2;
  __builtin_amdgcn_s_setprio(2);
  const SolveTask t = tasks[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  const int c16 = lane & 15, g = lane >> 4;
  const int li = tid & 63, lk = tid >> 6;  // loader: row li, k = 16 ch + lk + 4q
  const long long ld = t.ld;
  // issued first: the tile (accumulator layout)
  acc_t C[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = wr + 16 * a + c16, j = wc + 16 * b + Mfma<T>::drow(lane, r);
        C[a][b][r] = (i < t.M && j < t.N) ? L[t.c_off + i + j * ld] : (T)0;
      }
  // one 64 x 64 x 64 product of LDS-staged halves into acc (fa rows wr..,
  // fb rows wc.. of Bsrc, chunks c of the half)
  auto mma_half = [&](acc_t (&acc)[2][2], T (*Asrc)[64][KP], T (*Bsrc)[64][KP]) {
#pragma unroll
    for (int c = 0; c < 2; c++) {
      T fa0[TK / 4], fa1[TK / 4], fb0[TK / 4], fb1[TK / 4];
#pragma unroll
      for (int s2 = 0; s2 < TK / 4; s2++) {
        fa0[s2] = Asrc[c][wr + c16][g * (TK / 4) + s2];
        fa1[s2] = Asrc[c][wr + 16 + c16][g * (TK / 4) + s2];
        fb0[s2] = Bsrc[c][wc + c16][g * (TK / 4) + s2];
        fb1[s2] = Bsrc[c][wc + 16 + c16][g * (TK / 4) + s2];
      }
#pragma unroll
      for (int s2 = 0; s2 < TK / 4; s2++) {
        acc[0][0] = Mfma<T>::step(fb0[s2], fa0[s2], acc[0][0]);
        acc[0][1] = Mfma<T>::step(fb1[s2], fa0[s2], acc[0][1]);
        acc[1][0] = Mfma<T>::step(fb0[s2], fa1[s2], acc[1][0]);
        acc[1][1] = Mfma<T>::step(fb1[s2], fa1[s2], acc[1][1]);
      }
    }
  };
  // a matrix in the accumulator layout into As, half h (its columns 32h..32h+31
  // are held by the waves with wc == 32h), as the A operand's chunks
  auto stage_acc = [&](const acc_t (&M)[2][2], int h) {
    if (wc == 32 * h) {
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
          for (int r = 0; r < 4; r++) As[b][wr + 16 * a + c16][Mfma<T>::drow(lane, r)] = M[a][b][r];
    }
  };
  acc_t acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) acc[a][b][r] = (T)0;
  // the inverse in the loader layout: loaded into rb once the last earlier
  // block is staged (at once without earlier blocks)
  T ra[4][TK / 4], rb[4][TK / 4];
  auto fetch_inv = [&]() {
#pragma unroll
    for (int ch = 0; ch < 4; ch++)
#pragma unroll
      for (int q = 0; q < TK / 4; q++) {
        const int k = ch * TK + lk + 4 * q;
        rb[ch][q] = (li < t.N && k < t.N) ? Inv[t.inv_off + li * 64 + k] : (T)0;
      }
  };
  // C -= Xr Xk^T over the earlier blocks, 64 columns at a time, the next
  // block's loads issued before the current block's second half
  if (t.kpre == 0) fetch_inv();
  if (t.kpre > 0) {
    auto fetch = [&](int k0) {
#pragma unroll
      for (int ch = 0; ch < 4; ch++)
#pragma unroll
        for (int q = 0; q < TK / 4; q++) {
          const long long k = k0 + ch * TK + lk + 4 * q;
          ra[ch][q] = li < t.M ? L[t.xr_off + li + k * ld] : (T)0;
          rb[ch][q] = li < t.N ? L[t.xk_off + li + k * ld] : (T)0;
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < t.kpre; k0 += 64) {
#pragma unroll
      for (int h = 0; h < 2; h++) {
        __syncthreads();  // the previous half's fragments are read
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
          for (int q = 0; q < TK / 4; q++) {
            As[c][li][lk + 4 * q] = ra[2 * h + c][q];
            Bs[c][li][lk + 4 * q] = rb[2 * h + c][q];
          }
        if (h == 1) {
          if (k0 + 64 < t.kpre) fetch(k0 + 64);
          else fetch_inv();
        }
        __syncthreads();
        mma_half(acc, As, Bs);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) C[a][b][r] -= acc[a][b][r];
  // X = C inv^T
  acc_t X[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) X[a][b][r] = (T)0;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    __syncthreads();
    stage_acc(C, h);
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
      for (int q = 0; q < TK / 4; q++) Bs[c][li][lk + 4 * q] = rb[2 * h + c][q];
    __syncthreads();
    mma_half(X, As, Bs);
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = wr + 16 * a + c16, j = wc + 16 * b + Mfma<T>::drow(lane, r);
        if (i < t.M && j < t.N) L[t.c_off + i + j * ld] = X[a][b][r];
      }
  if (t.Md == 0) return;
  // Cd -= X X^T (both operands X, staged once into As).  Cd's entries are
  // loaded first, together (clamped to the block, dropped outside it), so
  // that their round trip overlaps the product; no other workgroup of the
  // launch writes this block.
  T dv[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = min(wr + 16 * a + c16, t.Md - 1);
        const int j = min(wc + 16 * b + Mfma<T>::drow(lane, r), t.Md - 1);
        dv[a][b][r] = L[t.d_off + i + j * ld];
      }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) asm volatile("" : "+v"(dv[a][b][r]));
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) acc[a][b][r] = (T)0;
#pragma unroll
  for (int h = 0; h < 2; h++) {
    __syncthreads();
    stage_acc(X, h);
    __syncthreads();
    mma_half(acc, As, As);
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = wr + 16 * a + c16, j = wc + 16 * b + Mfma<T>::drow(lane, r);
        if (i < t.Md && j < t.Md && i >= j) L[t.d_off + i + j * ld] = dv[a][b][r] - acc[a][b][r];
      }
  if (t.pf_front < 0) return;
  __syncthreads();  // the block's update is stored, and As read, before the POTRF
  potrf_mw_block<T>(fr, PanelTask{t.pf_front, t.pf_k0, t.Md, 0, 0}, L, Inv, info, fallback, smem);
}

// 128 x 128 tile for the large trailing updates and SYRKs: 8 waves (2 x 4),
// each a 64 x 32 sub-tile of 4 x 2 MFMA blocks.  A 64 x 64 sub-tile per wave
// (4 waves) reuses each LDS fragment more, but its 128 accumulator registers
// leave room for two waves per SIMD only; with 64 x 32 four waves share a
// SIMD and keep its MFMA pipe busier (measured on a K = 512 trailing update
// of a 6144-row front: 45.0 against 37.9 TFLOP/s FP64, the same bits).  K is
// staged in chunks of 16 with the next chunk prefetched into registers.
// Same task format (M, N <= 128) and epilogue as the 64 x 64 tile.
#define TK2 16
#define TPAD2 4
template <typename T>
__global__ __launch_bounds__(512) void k_gemm_tile128(const GemmTask* __restrict__ tasks,
                                                      const T* Lsrc, T* L, T* U)
{
  typedef typename Mfma<T>::acc_t acc_t;
  __shared__ T As[TK2][128 + TPAD2];
  __shared__ T Bs[TK2][128 + TPAD2];
  const GemmTask t = tasks[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 2) * 64, wc = (wave & 3) * 32;
  const T* A = Lsrc + t.a_off;
  const T* B = Lsrc + t.b_off;
  acc_t acc[4][2];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) acc[a][b][r] = (T)0;

  const int li = tid & 127, lk = tid >> 7;  // loader: row li, k = lk + 4q
  T ra[TK2 / 4], rb[TK2 / 4];
  auto fetch = [&](int kc) {
#pragma unroll
    for (int q = 0; q < TK2 / 4; q++) {
      const int k = kc + lk + 4 * q;
      const bool kin = k < t.K;
      ra[q] = (kin && li < t.M) ? A[li + (long long)k * t.ld] : (T)0;
      rb[q] = (kin && li < t.N) ? B[li + (long long)k * t.ld] : (T)0;
    }
  };
  fetch(0);
  for (int kc = 0; kc < t.K; kc += TK2) {
#pragma unroll
    for (int q = 0; q < TK2 / 4; q++) {
      As[lk + 4 * q][li] = ra[q];
      Bs[lk + 4 * q][li] = rb[q];
    }
    __syncthreads();
    if (kc + TK2 < t.K) fetch(kc + TK2);
#pragma unroll
    for (int kk = 0; kk < TK2; kk += 4) {
      const int kr = kk + (lane >> 4), c16 = lane & 15;
      T af[4], bf[2];
#pragma unroll
      for (int a = 0; a < 4; a++) af[a] = As[kr][wr + 16 * a + c16];
#pragma unroll
      for (int b = 0; b < 2; b++) bf[b] = Bs[kr][wc + 16 * b + c16];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = Mfma<T>::step(bf[b], af[a], acc[a][b]);
    }
    __syncthreads();
  }

  // C -= acc, eight entries a lane at a time with their loads in flight
  // together (clamped to the tile, dropped outside it; the empty asm keeps
  // the compiler from sinking each load into its store's condition; eight
  // keep the kernel within the 128 registers of four waves per SIMD)
  T* C = ((t.flags & 1) ? U : L) + t.c_off;
  const bool diag = (t.flags & 2) != 0;
#pragma unroll
  for (int a0 = 0; a0 < 4; a0++) {
    T cv[1][2][4];
#pragma unroll
    for (int a = 0; a < 1; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = min(wr + (a0 + a) * 16 + (lane & 15), t.M - 1);
          const int j = min(wc + b * 16 + Mfma<T>::drow(lane, r), t.N - 1);
          cv[a][b][r] = C[i + (long long)j * t.ldc];
        }
#pragma unroll
    for (int a = 0; a < 1; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int r = 0; r < 4; r++) asm volatile("" : "+v"(cv[a][b][r]));
#pragma unroll
    for (int a = 0; a < 1; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = wr + (a0 + a) * 16 + (lane & 15);
          const int j = wc + b * 16 + Mfma<T>::drow(lane, r);
          if (i < t.M && j < t.N && (!diag || i >= j))
            C[i + (long long)j * t.ldc] = cv[a][b][r] - acc[a0 + a][b][r];
        }
  }
}

// Same tile, BF16 inputs with FP32 accumulation (v_mfma_f32_16x16x32_bf16):
// the update variant of the FP32 factor.  Operands are rounded to bf16 as
// they are staged into LDS; K chunk 32.
__global__ __launch_bounds__(256) void k_gemm_tile_bf16(const GemmTask* __restrict__ tasks,
                                                        const float* Lsrc, float* L,
                                                        float* U)
{
  __shared__ __bf16 As[64][32 + 8];  // [row][k]
  __shared__ __bf16 Bs[64][32 + 8];
  const GemmTask t0 = tasks[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
  // a task larger than 64 x 64 (the 128-tile schedules) runs as 64 x 64 pieces
  for (int si = 0; si < t0.M; si += 64)
    for (int sj = 0; sj < t0.N; sj += 64) {
  const bool tdiag = (t0.flags & 2) != 0;
  if (tdiag && sj > si) continue;
  GemmTask t = t0;
  t.a_off += si;
  t.b_off += sj;
  t.c_off += si + (long long)sj * t0.ldc;
  t.M = t0.M - si < 64 ? t0.M - si : 64;
  t.N = t0.N - sj < 64 ? t0.N - sj : 64;
  t.flags = (t0.flags & 1) | ((tdiag && si == sj) ? 2 : 0);
  const float* A = Lsrc + t.a_off;
  const float* B = Lsrc + t.b_off;
  f4_t acc[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int r = 0; r < 4; r++) acc[a][b][r] = 0.f;
  const int li = tid & 63, lk = tid >> 6;
  for (int kc = 0; kc < t.K; kc += 32) {
    for (int q = 0; q < 8; q++) {
      int k = lk + 4 * q;
      bool kin = kc + k < t.K;
      As[li][k] = (__bf16)((kin && li < t.M) ? A[li + (long long)(kc + k) * t.ld] : 0.f);
      Bs[li][k] = (__bf16)((kin && li < t.N) ? B[li + (long long)(kc + k) * t.ld] : 0.f);
    }
    __syncthreads();
    {
      const int c16 = lane & 15, kg = (lane >> 4) * 8;
      bf16x8_t a0, a1, b0, b1;
      for (int e = 0; e < 8; e++) {
        a0[e] = As[wr + c16][kg + e];
        a1[e] = As[wr + 16 + c16][kg + e];
        b0[e] = Bs[wc + c16][kg + e];
        b1[e] = Bs[wc + 16 + c16][kg + e];
      }
      acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b0, a0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b1, a0, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b0, a1, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b1, a1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
  float* C = ((t.flags & 1) ? U : L) + t.c_off;
  const bool diag = (t.flags & 2) != 0;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int r = 0; r < 4; r++) {
        int i = wr + a * 16 + (lane & 15);
        int j = wc + b * 16 + 4 * (lane >> 4) + r;
        if (i < t.M && j < t.N && (!diag || i >= j))
          C[i + (long long)j * t.ldc] -= acc[a][b][r];
      }
  __syncthreads();
    }
}

// Extend-add: child update block entries (i >= j) added into the parent's
// panel or update block at the relative positions (src/taucs_sn_llt.c:1200-1240).
// The child's update rows are ordered by their parent positions, so
// r[i] > r[j] for i > j: entry (i, j) goes to parent (r[i], r[j]), below the
// diagonal, and one child column lands in one parent column.  Each wave
// takes every fourth column of the task; its lanes walk the column.
template <typename T>
__device__ __forceinline__ void extend_add_cols(const FrontDesc* __restrict__ fr, const EATask& t,
                                                const int* __restrict__ rel, T* __restrict__ L,
                                                T* __restrict__ U)
{
  const FrontDesc c = fr[t.child];
  const FrontDesc p = fr[c.parent];
  const int uc = c.m - c.s, sp = p.s, mp = p.m, up = mp - sp;
  const T* Uc = U + c.offU;
  const int* r = rel + c.rel;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (t.j1 - t.j0 < 4) {
    // fewer columns than waves (the long columns of a large child): all 256
    // threads walk each column, four rows a thread per pass, loads first
    for (int j = t.j0; j < t.j1; j++) {
      const int pj = r[j];
      const T* src = Uc + (long long)j * uc;
      T* base = pj < sp ? L : U;
      const long long col0 = pj < sp ? p.offL + (long long)pj * mp
                                     : p.offU + (long long)(pj - sp) * up - sp;
      for (int i0 = j + (int)threadIdx.x; i0 < uc; i0 += 1024) {
        T sv[4], bv[4];
        int ri[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int i = min(i0 + 256 * q, uc - 1);  // rows past the end: read, dropped
          sv[q] = src[i];
          ri[q] = r[i];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) bv[q] = base[col0 + ri[q]];
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (i0 + 256 * q < uc) base[col0 + ri[q]] = bv[q] + sv[q];
      }
    }
    return;
  }
  for (int j = t.j0 + wave; j < t.j1; j += 4) {
    const int pj = r[j];
    const T* src = Uc + (long long)j * uc;
    // parent column pj: panel column (rows 0..mp) or update-block column (rows sp..mp)
    T* base = pj < sp ? L : U;
    const long long col0 = pj < sp ? p.offL + (long long)pj * mp
                                   : p.offU + (long long)(pj - sp) * up - sp;
    // four rows per lane per pass, all loads before the stores: the rows of a
    // column map to distinct parent rows, and child and parent blocks never
    // overlap, so the batch is free of conflicts (the compiler cannot know
    // that: both may point into the update arena).  The loads are issued
    // unconditionally (a row past the end reads the column's last row and
    // drops it), so they are in flight together.
    for (int i0 = j + lane; i0 < uc; i0 += 256) {
      T sv[4], bv[4];
      int ri[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int i = min(i0 + 64 * q, uc - 1);  // rows past the end: read, dropped
        sv[q] = src[i];
        ri[q] = r[i];
      }
#pragma unroll
      for (int q = 0; q < 4; q++) bv[q] = base[col0 + ri[q]];
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (i0 + 64 * q < uc) base[col0 + ri[q]] = bv[q] + sv[q];
    }
  }
}

// A whole child whose u x u block has at most 2048 entries, in one pass of
// the workgroup: the block's entries as one flat index, eight a thread, the
// lower ones (i >= j) added into the parent with all loads before the stores
// (the entries go to distinct parent positions).
#define EA_FLAT_MAX 2048
template <typename T>
__device__ __forceinline__ void extend_add_flat(const FrontDesc* __restrict__ fr, const EATask& t,
                                                const int* __restrict__ rel, T* __restrict__ L,
                                                T* __restrict__ U)
{
  const FrontDesc c = fr[t.child];
  const FrontDesc p = fr[c.parent];
  const int uc = c.m - c.s, sp = p.s, mp = p.m, up = mp - sp, tot = uc * uc;
  const T* Uc = U + c.offU;
  const int* r = rel + c.rel;
  // every load issued unconditionally (an entry that is not added reads
  // entry (0, 0) and drops it), so that they are in flight together
  T sv[8], bv[8];
  int ri[8], rj[8];
  bool on[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int e = (int)threadIdx.x + 256 * q, j = e / uc, i = e - j * uc;
    on[q] = e < tot && i >= j;
    sv[q] = Uc[on[q] ? e : 0];
    ri[q] = r[on[q] ? i : 0];
    rj[q] = r[on[q] ? j : 0];
  }
  T* dst[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int pi = ri[q], pj = rj[q];
    dst[q] = pj < sp ? L + (p.offL + (long long)pj * mp + pi)
                     : U + (p.offU + (long long)(pj - sp) * up + (pi - sp));
    bv[q] = *dst[q];
  }
#pragma unroll
  for (int q = 0; q < 8; q++)
    if (on[q]) *dst[q] = bv[q] + sv[q];
}

template <typename T>
__global__ __launch_bounds__(256) void k_extend_add(const FrontDesc* __restrict__ fr,
                                                    const EATask* __restrict__ tasks,
                                                    const int* __restrict__ rel,
                                                    T* __restrict__ L, T* __restrict__ U)
{
  const EATask t = tasks[blockIdx.x];
  const FrontDesc c = fr[t.child];
  const int uc = c.m - c.s;
  if (t.j0 == 0 && t.j1 == uc && uc * uc <= EA_FLAT_MAX)
    extend_add_flat(fr, t, rel, L, U);
  else
    extend_add_cols(fr, t, rel, L, U);
}

// One parent per workgroup: heads[b] = {first, count} names the tasks (whole
// children, in sibling-rank order) added one after the other, so each parent
// entry receives its children's values in the same order as under the
// per-rank launches.
template <typename T>
__global__ __launch_bounds__(256) void k_extend_add_seq(const FrontDesc* __restrict__ fr,
                                                        const EATask* __restrict__ heads,
                                                        const EATask* __restrict__ tasks,
                                                        const int* __restrict__ rel,
                                                        T* __restrict__ L, T* __restrict__ U)
{
  const EATask h = heads[blockIdx.x];
  for (int q = 0; q < h.j0; q++) {
    if (q > 0) __syncthreads();
    const EATask t = tasks[h.child + q];
    const FrontDesc c = fr[t.child];
    const int uc = c.m - c.s;
    if (t.j0 == 0 && t.j1 == uc && uc * uc <= EA_FLAT_MAX)
      extend_add_flat(fr, t, rel, L, U);
    else
      extend_add_cols(fr, t, rel, L, U);
  }
}

#define TAMD_INST(T)                                                                   \
  template __global__ void k_fill<T>(T*, long long);                                   \
  template __global__ void k_fill_ranges<T>(const ZeroTask*, T*);                       \
  template __global__ void k_assemble_flat<T>(long long, const long long*, const double*, T*); \
  template __global__ void k_assemble<T>(int, const int*, const long long*,             \
                                         const double*, T*);                            \
  template __global__ void k_front_small<T>(const FrontDesc*, const int*, T*, T*, T*, int*); \
  template __global__ void k_potrf_diag<T>(const FrontDesc*, const PanelTask*, T*, T*, int*); \
  template __global__ void k_potrf_wave<T>(const FrontDesc*, const PanelTask*, T*, T*, int*, int); \
  template __global__ void k_potrf_mw<T>(const FrontDesc*, const PanelTask*, T*, T*, int*, int);   \
  template __global__ void k_front_wave<T>(const FrontDesc*, const int*, T*, T*, T*, int*, int); \
  template __global__ void k_front_slot16<T>(const FrontDesc*, const SmallPack*, T*, T*, T*, int*, int); \
  template __global__ void k_front_slot32<T>(const FrontDesc*, const SmallPack*, T*, T*, T*, int*, int); \
  template __global__ void k_front_w48<T>(const FrontDesc*, const int*, T*, T*, T*, int*, int);          \
  template __global__ void k_gemm_tile<T>(const GemmTask*, const T*, T*, T*, const T*);  \
  template __global__ void k_panel_solve<T>(const SolveTask*, T*, T*, const FrontDesc*, int*, int); \
  template __global__ void k_gemm_tile128<T>(const GemmTask*, const T*, T*, T*);          \
  template __global__ void k_gemm_short<T>(const GemmTask*, const T*, T*, T*, const T*);  \
  template __global__ void k_extend_add<T>(const FrontDesc*, const EATask*, const int*, T*, T*); \
  template __global__ void k_extend_add_seq<T>(const FrontDesc*, const EATask*, const EATask*,   \
                                               const int*, T*, T*);
TAMD_INST(double)
TAMD_INST(float)
template __global__ void k_fill<int>(int*, long long);

}  // namespace tamd
