// Device kernels of the MI355X multifrontal Cholesky (gfx950, wave64).
//
// Storage (all column-major, one arena per kind):
//   L  panels: supernode f owns m_f x s_f at L + offL, leading dim m_f = s_f+u_f;
//              rows 0..s-1 are the diagonal block (F1), rows s..m-1 the
//              sub-diagonal block (F2) of src/taucs_sn_llt.c:1086-1194.
//   U  update (contribution) blocks: u_f x u_f at U + offU, lower part used;
//      blocks of disjoint lifetimes share space, each is zeroed (ZeroTask)
//      at the start of the level where its first writer runs.
//   W  solve workspace: m_f x nrhs per front.
// Work is described by task arrays built once on the host per level schedule.
// A kernel reads its task by blockIdx, except the chained solve sweeps, whose
// tasks take a ticket and wait (bounded) for values that earlier tickets of
// the same launch publish.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

struct FrontDesc {
  long long offL;   // panel offset (elements)
  long long offU;   // update block offset
  long long offW;   // solve workspace offset (elements per rhs column: m)
  long long offI;   // inverses of the 64x64 diagonal blocks (row-major, 4096 each)
  int s, m;         // columns, rows (m = s + u)
  int parent;       // index into the front table, -1 at a root
  int rel;          // offset of this front's update rows in the rel array
  int rows;         // offset of this front's row list in the rows array
  int pad;
};

struct GemmTask {   // C[i,j] -= sum_k A[i,k] * B[j,k] on one 64x64 tile
  long long a_off;  // A(i,k) = L[a_off + i + k*ld]
  long long b_off;  // B(j,k) = L[b_off + j + k*ld]  (flag 4: Inv[b_off + j*64 + k])
  long long c_off;  // C(i,j) = C[c_off + i + j*ldc]
  int ld, ldc;
  int M, N, K;
  int flags;        // 1: C lives in U (else in L); 2: diagonal tile, lower part only;
                    // 4: B is a row-major inverse block and C = A B^T is stored
                    //    (the panel triangular solve X <- X L_kk^{-T}, in place)
};

struct SolveTask {  // one 64-row tile of a panel step (k_panel_solve)
  long long c_off;   // the tile: C(i, j) = L[c_off + i + j*ld], i < M, j < N (the step's block)
  long long inv_off; // the block inverse, row-major: Inv[inv_off + j*64 + k]
  long long xr_off;  // the tile's rows in the outer panel's earlier blocks: L[xr_off + i + k*ld]
  long long xk_off;  // the block's own rows in those blocks: L[xk_off + j + k*ld]
  long long d_off;   // the diagonal block starting at the tile's first row: L[d_off + i + j*ld]
  int ld, M, N;
  int kpre;          // columns of the earlier blocks (k < kpre; 0 at an outer panel's first step)
  int Md;            // order of that diagonal block when it lies in the outer panel, else 0
  int pf_front;      // the next step's diagonal block, factored by this tile's workgroup once
  int pf_k0;         // it is updated: front, first column (pf_front < 0: none)
};

struct PanelTask {  // one diagonal block, or one row block of the panel
  int front;
  int k0, kb;       // columns k0 .. k0+kb-1 of the panel
  int r0, nr;       // rows (or columns) r0 .. r0+nr-1
};

struct SmallPack {  // fronts with m <= width, one per `width`-lane slot of a wavefront
  int f[4];         // front indices (-1: empty slot)
  int width;        // 16, 32 or 64
  int pad[3];
};

struct ZeroTask {   // rows c..u-1 of columns c0..c1-1 of the u x u block at U[off] <- 0
                    // (u = n, c0 = 0, c1 = 1: the n elements from off)
  long long off;
  int u, c0, c1, pad;
};

struct EATask {     // extend-add columns j0..j1-1 of a child's update block
  int child;
  int j0, j1;
  int pad;
};

#define TAMD_OP_ASSEMBLE 1
#define TAMD_OP_SMALL    2
#define TAMD_OP_POTRF    3
#define TAMD_OP_TRSM     4
#define TAMD_OP_GEMM     5
#define TAMD_OP_EXTEND   6
#define TAMD_OP_ZERO     7

// Lane layout of v_mfma_*_16x16x4 (checked on gfx950): A operand lane l holds
// A[l%16][l/16], B operand lane l holds B[l/16][l%16].  Result register r of
// lane l is D[l/16 + 4r][l%16] for f64 and D[4(l/16) + r][l%16] for f32.
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef d4_t acc_t;
  static __device__ acc_t step(double a, double b, acc_t c)
  {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ int drow(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <> struct Mfma<float> {
  typedef f4_t acc_t;
  static __device__ acc_t step(float a, float b, acc_t c)
  {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// ---- readiness by value (the chained solves) ------------------------------
// A buffer is filled with a signaling-NaN bit pattern before a sweep (no
// arithmetic result is a signaling NaN); every value is then published with
// one relaxed device-scope atomic store and read with relaxed device-scope
// atomic loads until the pattern is gone, so each value is its own flag.
#define SPIN_LIMIT (1 << 20)

template <typename T> struct Sentinel;
template <> struct Sentinel<double> {
  typedef unsigned long long U;
  static constexpr U bits = 0x7FF0DEAD0000BEEFull;
};
template <> struct Sentinel<float> {
  typedef unsigned int U;
  static constexpr U bits = 0x7F80BEEFu;
};

template <typename T>
__device__ __forceinline__ void publish(T* p, T v)
{
  typedef typename Sentinel<T>::U U;
  __hip_atomic_store(reinterpret_cast<U*>(p), __builtin_bit_cast(U, v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// one value, awaited (bounded: after SPIN_LIMIT polls the error word is marked
// and the pattern is returned as it is; once any wait of the launch has marked
// the error word, every other wait gives up at its next poll)
template <typename T>
__device__ __forceinline__ T await_value(const T* p, int* err)
{
  typedef typename Sentinel<T>::U U;
  U* a = reinterpret_cast<U*>(const_cast<T*>(p));
  U v = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  int it = 0;
  while (v == Sentinel<T>::bits) {
    if (++it > SPIN_LIMIT) {
      atomicOr(err, 1);
      break;
    }
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
    __builtin_amdgcn_s_sleep(1);
    v = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return __builtin_bit_cast(T, v);
}
