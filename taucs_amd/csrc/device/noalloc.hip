// Allocation-free supernodal solve.
//
// taucs_amd_supernodal_solve_many_device allocates its sweep workspace (front
// blocks W, the forward solution Y, split-K partial blocks, tickets and the
// error word) on first use and replays the sweeps as a hipGraph it builds.
// Callers that must not allocate device memory (code under graph capture,
// tight serving loops, harnesses that forbid hipMalloc) use the variant here
// instead: the same kernels, launched eagerly on the factor's stream, with the
// workspace carved out of one file-scope device array.
//
// The array is per GPU (each device holds its own copy at its own address) and
// lives until the process ends.  Host code never passes the array's name to a
// kernel: its device address is looked up once per device with
// hipGetSymbolAddress and only that address is used.  Every call on one GPU
// shares the array, so calls that overlap on one GPU race (different factors
// included); calls on one GPU must be serialized by the caller.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../host/internal.h"

extern "C" int tamd_dev_solve_dims(tamd_factor* F, long long* W_per_rhs, long long* part_slots,
                                   long long* sync_n, int* elem);
extern "C" int tamd_dev_solve_eager(tamd_factor* F, void* W, void* Y, void* Part, int* sync, int nr,
                                    double* dX, int ldx, const double* dB, int ldb);
extern "C" int tamd_dev_stream_of(tamd_factor* F, hipStream_t* st);

// 64 MiB.  A sweep needs, per FP64 right-hand side, 8 bytes per front row and
// per unknown plus 512 bytes per split-K slot of its widest level.
#define TAMD_NOALLOC_SCRATCH_BYTES (64ull << 20)
#define TAMD_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[TAMD_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[TAMD_NOALLOC_MAX_DEVICES];

// Device address of this GPU's scratch array, looked up once per device with
// that device current.
static unsigned char* scratch_of_current_device()
{
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= TAMD_NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      taucs_printf((char*)"taucs_amd: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
                   hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  return g_scratch_ptr[dev];
}

__global__ void k_clear_words(int* __restrict__ w, long long n)
{
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    w[i] = 0;
}

static inline long long up256(long long b) { return (b + 255) & ~255LL; }

// Bytes of scratch one sweep of nr right-hand sides takes: W, Y, the split-K
// partial blocks (64 x nr each), then tickets and the error word, each
// 256-byte aligned.
static long long sweep_bytes(long long W_per_rhs, long long n, long long part_slots,
                             long long sync_n, int elem, int nr)
{
  return up256((long long)elem * W_per_rhs * nr) + up256((long long)elem * n * nr)
         + up256((long long)elem * 64 * nr * part_slots) + up256(4 * (sync_n + 1));
}

extern "C" long long taucs_amd_noalloc_scratch_bytes(void) { return (long long)TAMD_NOALLOC_SCRATCH_BYTES; }

extern "C" int taucs_amd_supernodal_solve_many_device_noalloc(void* vL, int nrhs, double* dX, int ldx,
                                                              const double* dB, int ldb)
{
  tamd_factor* F = (tamd_factor*)vL;
  if (!F || !F->dev || !F->has_numeric) {
    taucs_printf((char*)"taucs_amd: allocation-free solve needs a numeric factor\n");
    return -1;
  }
  const long long n = F->n;
  if (nrhs < 0 || (nrhs > 0 && (!dX || !dB || ldx < n || ldb < n))) {
    taucs_printf((char*)"taucs_amd: allocation-free solve: bad arguments\n");
    return -1;
  }
  if (nrhs == 0) return 0;
  long long W_per_rhs = 0, part_slots = 0, sync_n = 0;
  int elem = 0;
  hipStream_t st = nullptr;
  if (tamd_dev_solve_dims(F, &W_per_rhs, &part_slots, &sync_n, &elem) != 0
      || tamd_dev_stream_of(F, &st) != 0)
    return -1;
  // right-hand sides per sweep: at most 16 (the sweep kernels' limit) and as
  // many as the scratch holds
  int chunk = 0;
  for (int nr = 16; nr >= 1; nr--)
    if (sweep_bytes(W_per_rhs, n, part_slots, sync_n, elem, nr) <= (long long)TAMD_NOALLOC_SCRATCH_BYTES) {
      chunk = nr;
      break;
    }
  if (chunk == 0) {
    taucs_printf((char*)"taucs_amd: allocation-free solve: one right-hand side needs %lld bytes of "
                        "workspace, above the %llu-byte scratch limit; use "
                        "taucs_amd_supernodal_solve_many_device\n",
                 sweep_bytes(W_per_rhs, n, part_slots, sync_n, elem, 1),
                 (unsigned long long)TAMD_NOALLOC_SCRATCH_BYTES);
    return -1;
  }
  unsigned char* base = scratch_of_current_device();
  if (!base) return -1;
  unsigned char* W = base;
  unsigned char* Y = W + up256((long long)elem * W_per_rhs * chunk);
  unsigned char* Part = Y + up256((long long)elem * n * chunk);
  int* sync = (int*)(Part + up256((long long)elem * 64 * chunk * part_slots));
  for (int r0 = 0; r0 < nrhs; r0 += chunk) {
    const int nr = nrhs - r0 < chunk ? nrhs - r0 : chunk;
    // the array keeps the previous call's contents: the tickets start at zero
    // in every sweep, the error word once per call (a time-out in any sweep
    // of this call is reported)
    const long long nw = r0 == 0 ? sync_n + 1 : sync_n;
    if (nw > 0) {
      const unsigned blocks = (unsigned)(nw < 1024 * 256 ? (nw + 255) / 256 : 1024);
      hipLaunchKernelGGL(k_clear_words, dim3(blocks), dim3(256), 0, st, sync, nw);
      if (hipGetLastError() != hipSuccess) return -1;
    }
    if (tamd_dev_solve_eager(F, W, Y, Part, sync, nr, dX + (size_t)r0 * ldx, ldx,
                             dB + (size_t)r0 * ldb, ldb) != 0)
      return -1;
  }
  if (hipStreamSynchronize(st) != hipSuccess) return -1;
  int e = 0;
  if (hipMemcpy(&e, sync + sync_n, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  if (e) {
    taucs_printf((char*)"taucs_amd: a solve sweep timed out waiting for a block\n");
    return -1;
  }
  return 0;
}
