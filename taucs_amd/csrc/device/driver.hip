// Host driver of the device numeric phase: builds the task lists of the level
// schedule once per symbolic factor, owns the HBM arenas, launches the
// factorization and the solves on the factor's stream.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <tuple>
#include <utility>
#include <unordered_map>
#include <vector>

#include "kernels.h"
#include "../host/internal.h"
#include "../../../include/taucs_amd.h"

namespace tamd {
template <typename T> __global__ void k_fill(T*, long long);
template <typename T> __global__ void k_fill_ranges(const ZeroTask*, T*);
template <typename T> __global__ void k_assemble(int, const int*, const long long*, const double*, T*);
template <typename T> __global__ void k_assemble_flat(long long, const long long*, const double*, T*);
template <typename T> __global__ void k_front_small(const FrontDesc*, const int*, T*, T*, T*, int*);
template <typename T> __global__ void k_potrf_diag(const FrontDesc*, const PanelTask*, T*, T*, int*);
template <typename T> __global__ void k_potrf_wave(const FrontDesc*, const PanelTask*, T*, T*, int*, int);
template <typename T> __global__ void k_potrf_mw(const FrontDesc*, const PanelTask*, T*, T*, int*, int);
template <typename T> __global__ void k_front_wave(const FrontDesc*, const int*, T*, T*, T*, int*, int);
template <typename T> __global__ void k_front_slot16(const FrontDesc*, const SmallPack*, T*, T*, T*, int*, int);
template <typename T> __global__ void k_front_slot32(const FrontDesc*, const SmallPack*, T*, T*, T*, int*, int);
template <typename T> __global__ void k_front_w48(const FrontDesc*, const int*, T*, T*, T*, int*, int);
template <typename T> __global__ void k_gemm_tile(const GemmTask*, const T*, T*, T*, const T*);
template <typename T> __global__ void k_panel_solve(const SolveTask*, T*, T*, const FrontDesc*, int*, int);
template <typename T> __global__ void k_gemm_tile128(const GemmTask*, const T*, T*, T*);
template <typename T> __global__ void k_gemm_short(const GemmTask*, const T*, T*, T*, const T*);
__global__ void k_gemm_tile_bf16(const GemmTask*, const float*, float*, float*);
template <typename T> __global__ void k_extend_add(const FrontDesc*, const EATask*, const int*, T*, T*);
template <typename T> __global__ void k_extend_add_seq(const FrontDesc*, const EATask*, const EATask*, const int*, T*, T*);
template <typename T> __global__ void k_solve_init(const FrontDesc*, const PanelTask*, int, const int*, const double*, int, T*, int);
template <typename T> __global__ void k_fwd_step(const FrontDesc*, const PanelTask*, const int*, const T*, const T*, T*, T*, int, int);
template <typename T> __global__ void k_fwd_rect(const FrontDesc*, const PanelTask*, const int*, const T*, const T*, int, T*, int);
template <typename T> __global__ void k_fwd_rect_reduce(const FrontDesc*, const PanelTask*, const T*, T*, int);
template <typename T> __global__ void k_bwd_cols_reduce(const FrontDesc*, const PanelTask*, const int*, const T*, T*, int, int);
template <typename T> __global__ void k_solve_ext(const FrontDesc*, const PanelTask*, const int*, T*, int);
template <typename T> __global__ void k_solve_ext_seq(const FrontDesc*, const PanelTask*, const PanelTask*, const int*, T*, int);
template <typename T> __global__ void k_bwd_gather(const FrontDesc*, const PanelTask*, const int*, const double*, int, T*, int);
template <typename T> __global__ void k_bwd_cols(const FrontDesc*, const PanelTask*, const T*, const T*, T*, int);
template <typename T> __global__ void k_bwd_step(const FrontDesc*, const PanelTask*, const int*, const T*, const T*, T*, int, double*, int, int);
template <typename T> __global__ void k_sweep_clear(T*, double*, int, int, int, int*, int);
template <typename T> __global__ void k_fwd_chain(const FrontDesc*, const PanelTask*, const int*, const T*, const T*, T*, T*, int, int*, int*, int);
template <typename T> __global__ void k_bwd_chain(const FrontDesc*, const PanelTask*, const int*, const T*, const T*, const T*, int, const T*, double*, int, int*, int*, int);
template <typename T> __global__ void k_bwd_upd(const FrontDesc*, const PanelTask*, const int*, const T*, const double*, int, T*, int);
template <typename T, int NR> __global__ void k_fwd_small(const FrontDesc*, const PanelTask*, int, const int*, const T*, const T*, T*, T*, int, int);
template <typename T, int NR> __global__ void k_bwd_small(const FrontDesc*, const PanelTask*, int, const int*, const T*, const T*, const T*, int, double*, int, int);
}  // namespace tamd

using namespace tamd;

static int g_profiling = 0;
extern "C" void taucs_amd_set_profiling(int on) { g_profiling = on; }

#define HIPCHECK(x)                                                                  \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      taucs_printf((char*)"taucs_amd: HIP error %s at %s:%d\n", hipGetErrorString(e_), \
                   __FILE__, __LINE__);                                              \
      return -1;                                                                     \
    }                                                                                \
  } while (0)

struct Op {
  int kind;
  int start, count;  // range in the task array of that kind
  int level;
  int lane = 0;      // 1: may run on the side stream (the lookahead's trailing update)
  bool nojoin = false;  // touches only the current outer panel: need not wait for the side stream
  int tile = 64;        // GEMM tile edge of the op's tasks (64 or 128)
  bool grouped = false; // extend-add: the tasks are group heads (k_extend_add_seq)
  bool shortk = false;  // GEMM of 64-tiles whose tasks all have K <= 64 (k_gemm_short)
};

// solve op kinds
enum { SV_INIT, SV_FSTEP, SV_FRECT, SV_FRED, SV_EXT, SV_GATHER, SV_BCOLS, SV_BRED, SV_BSTEP,
       SV_CLEAR, SV_FCHAIN, SV_BCHAIN, SV_FSMALL, SV_BSMALL, SV_BUPD, SV_EXTSEQ };

struct tamd_dev {
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;    // second stream of the panel lookahead
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int elem = 8;  // bytes per factor element
  // structure (uploaded once)
  FrontDesc* d_fronts = nullptr;
  int* d_rows = nullptr;
  int* d_rel = nullptr;
  int* d_colptr = nullptr;
  bool assemble_flat = false;  // every A entry has its own destination in L
  long long* d_adest = nullptr;
  int* d_small = nullptr;        // front indices for k_front_small / k_front_wave
  SmallPack* d_packs = nullptr;  // fronts packed 4 (m <= 16) or 2 (m <= 32) to a wave
  PanelTask* d_potrf = nullptr;
  SolveTask* d_trsm = nullptr;  // panel steps: left-looking updates and solves of 64-row tiles
  GemmTask* d_gemm = nullptr;
  EATask* d_ea = nullptr;
  ZeroTask* d_zero = nullptr;
  PanelTask* d_sv_tasks = nullptr;  // solve tasks
  std::vector<Op> sv_ops;           // block-step sweeps (one launch per block step)
  std::vector<Op> sv_chain_ops;     // chained sweeps (one launch per level); Op.tile = ticket
  int solve_chain = 1;
  int* d_sync = nullptr;            // the chained launches' tickets, then the error word
  long long sync_n = 0;             // tickets (cleared at the start of every sweep)
  int* d_info = nullptr;            // [0] a pivot failed, [1] blocks refactored in FP64
  int fp64_fallbacks = 0;           // of the last numeric factorization
  std::vector<Op> ops;
  std::vector<double> op_flops, op_bytes;  // algorithmic work of each op
  std::vector<float> op_ms;                // device time of each op (profiling runs)
  std::vector<int> lvl_ptr;      // fronts of level l: [lvl_ptr[l], lvl_ptr[l+1])
  int n_fronts = 0, ncol = 0, nnz = 0, n_levels = 0;
  int panel_w = 256;              // outer panel width of large fronts (multiple of 64)
  int tile_big = 64;              // GEMM tile edge of the trailing updates and SYRKs
  long long W_per_rhs = 0;
  size_t struct_bytes = 0;       // device bytes of the schedule and A values
  double gemm_flops = 0, assembly_bytes = 0;
  int gemm_launches = 0;
  // numeric arenas (precision-dependent)
  int prec = -1;
  void* L = nullptr;
  void* U = nullptr;
  void* Inv = nullptr;           // inverses of the 64x64 diagonal blocks
  long long I_elems = 0;
  double* d_avals = nullptr;
  const double* num_vals = nullptr;   // host values of the last numeric factorization
  unsigned long long num_gen = 0;     // numeric factorizations so far
  long long L_elems = 0, U_elems = 0;
  // solve buffers
  void* W = nullptr;
  void* Y = nullptr;             // forward solution, n x W_rhs, by global column
  void* Part = nullptr;          // split-K partial blocks of the solve (64 x 16 each)
  long long part_slots = 0;
  int W_rhs = 0;
  double* d_x = nullptr;
  double* d_b = nullptr;
  int xb_rhs = 0;
  taucs_amd_numeric_timing timing;
  // captured launch sequences (replayed while the buffers they name live)
  hipGraphExec_t num_graph = nullptr;
  int num_graph_key = -1;
  int num_graph_launches = 0;
  struct SolveGraph {
    hipGraphExec_t exec = nullptr;
    int nrhs = 0, prec = -1, ldx = 0, ldb = 0;
    const void* x = nullptr;
    const void* b = nullptr;
    unsigned long long used = 0;
  };
  SolveGraph sv[4];
  unsigned long long sv_clock = 0;
};

static int g_graphs = 1;
static int g_solve_chain = 1;
extern "C" void taucs_amd_set_solve_kernel(int chained) { g_solve_chain = chained ? 1 : 0; }
extern "C" void taucs_amd_set_graphs(int on) { g_graphs = on; }
static int g_panel_w = 256;
extern "C" void taucs_amd_set_panel_width(int w)
{
  w = (w + 63) / 64 * 64;
  g_panel_w = w < 64 ? 64 : (w > 1024 ? 1024 : w);
}
static int g_tile_big = 64;
// With 128-tiles selected, a launch with fewer 128-tiles than this uses
// 64-tiles instead (four times as many workgroups for the 256 CUs).
#ifndef TILE128_MIN
#define TILE128_MIN 512
#endif
extern "C" void taucs_amd_set_gemm_tile(int edge) { g_tile_big = edge == 128 ? 128 : 64; }
static int g_lookahead = 1;
extern "C" void taucs_amd_set_lookahead(int on) { g_lookahead = on; }
// Diagonal-block kernel of the block that starts each outer panel: 2 = four
// waves, 16-column panels with MFMA updates (default); 1 = one wave with
// register broadcasts; 0 = 256 threads through LDS (kept for comparison).  2
// and 1 give bitwise equal factors.  The later blocks of an outer panel are
// factored by the four-wave code inside the previous step's k_panel_solve.
static int g_potrf_wave = 2;
extern "C" void taucs_amd_set_potrf_kernel(int kind) { g_potrf_wave = kind >= 0 && kind <= 2 ? kind : 2; }
// Small fronts several to a wavefront: m <= 16 four per wave (k_front_slot16),
// m <= 32 two per wave (k_front_slot32), and m <= 48 one per wave with
// 48-wide staging (k_front_w48, half the LDS of k_front_wave); results are
// bitwise those of k_front_wave.  Applies to schedules built afterwards
// (default on).
static int g_front_packing = 1;
extern "C" void taucs_amd_set_front_packing(int on) { g_front_packing = on ? 1 : 0; }
// Reduced-precision factors: a diagonal block whose FP32 Cholesky meets a
// nonpositive pivot or a pivot below 2^-12 of its input diagonal entry is
// refactored in FP64 (the diagonal-block and whole-front kernels; default on).
static int g_fp64_fallback = 1;
extern "C" void taucs_amd_set_fp64_fallback(int on) { g_fp64_fallback = on ? 1 : 0; }

static void drop_solve_graphs(tamd_dev* D)
{
  for (auto& g : D->sv) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g = tamd_dev::SolveGraph();
  }
}

static void drop_graphs(tamd_dev* D)
{
  if (D->num_graph) (void)hipGraphExecDestroy(D->num_graph);
  D->num_graph = nullptr;
  D->num_graph_key = -1;
  drop_solve_graphs(D);
}

// Every device allocation of the library goes through dev_malloc/free_ptr,
// which keep a ledger of live bytes (for admission by HBM budget) under one
// mutex.  Graphs are built node by node (never by stream capture), so
// threads working on different factors never constrain each other's HIP
// calls.
static std::mutex g_mem_mu;
static std::unordered_map<void*, size_t> g_live;
static size_t g_live_bytes = 0;

static int dev_malloc(void** p, size_t bytes)
{
  if (tamd_alloc_should_fail()) {
    *p = nullptr;
    taucs_printf((char*)"taucs_amd: device allocation failed (induced)\n");
    return -1;
  }
  const size_t b = bytes ? bytes : 1;
  std::lock_guard<std::mutex> lk(g_mem_mu);
  hipError_t e = hipMalloc(p, b);
  if (e != hipSuccess) {
    *p = nullptr;
    taucs_printf((char*)"taucs_amd: hipMalloc(%zu) failed: %s\n", bytes, hipGetErrorString(e));
    return -1;
  }
  g_live[*p] = b;
  g_live_bytes += b;
  return 0;
}

static void free_ptr(void* p)
{
  if (!p) return;
  std::lock_guard<std::mutex> lk(g_mem_mu);
  (void)hipFree(p);
  auto it = g_live.find(p);
  if (it != g_live.end()) {
    g_live_bytes -= it->second;
    g_live.erase(it);
  }
}

extern "C" int tamd_dev_malloc(void** p, size_t bytes) { return dev_malloc(p, bytes); }
extern "C" void tamd_dev_free_ptr(void* p) { free_ptr(p); }

extern "C" long long taucs_amd_device_bytes_in_use(void)
{
  std::lock_guard<std::mutex> lk(g_mem_mu);
  return (long long)g_live_bytes;
}

template <typename V>
static int upload(V** dptr, const std::vector<V>& h, hipStream_t s)
{
  if (dev_malloc((void**)dptr, sizeof(V) * (h.size() ? h.size() : 1)) != 0) return -1;
  if (!h.empty()) HIPCHECK(hipMemcpyAsync(*dptr, h.data(), sizeof(V) * h.size(), hipMemcpyHostToDevice, s));
  return 0;
}

extern "C" int tamd_dev_available(void)
{
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n > 0;
}

extern "C" void tamd_dev_free_numeric(tamd_factor* F)
{
  tamd_dev* D = F ? F->dev : nullptr;
  if (!D) return;
  if (D->stream) (void)hipStreamSynchronize(D->stream);
  drop_graphs(D);
  free_ptr(D->L);
  free_ptr(D->U);
  free_ptr(D->Inv);
  free_ptr(D->W);
  free_ptr(D->Y);
  free_ptr(D->Part);
  free_ptr(D->d_sync);
  D->d_sync = nullptr;
  D->L = D->U = D->Inv = D->W = D->Y = D->Part = nullptr;
  D->W_rhs = 0;
  D->prec = -1;
}

extern "C" void tamd_refine_forget(void* key);

extern "C" void tamd_dev_free(tamd_factor* F)
{
  tamd_dev* D = F ? F->dev : nullptr;
  if (!D) return;
  tamd_refine_forget(F);
  tamd_dev_free_numeric(F);
  free_ptr(D->d_fronts); free_ptr(D->d_rows); free_ptr(D->d_rel); free_ptr(D->d_colptr);
  free_ptr(D->d_adest); free_ptr(D->d_small); free_ptr(D->d_packs); free_ptr(D->d_potrf);
  free_ptr(D->d_trsm);
  free_ptr(D->d_gemm); free_ptr(D->d_ea); free_ptr(D->d_sv_tasks); free_ptr(D->d_zero);
  free_ptr(D->d_info); free_ptr(D->d_avals); free_ptr(D->d_x); free_ptr(D->d_b);
  if (D->ev_fork) (void)hipEventDestroy(D->ev_fork);
  if (D->ev_join) (void)hipEventDestroy(D->ev_join);
  if (D->side) (void)hipStreamDestroy(D->side);
  if (D->stream) (void)hipStreamDestroy(D->stream);
  delete D;
  F->dev = nullptr;
}

// ---- schedule construction --------------------------------------------------

// Algorithmic flops of a tile: a diagonal tile only needs its lower part.
static double tile_flops(const GemmTask& g)
{
  double entries = (g.flags & 2) ? (double)g.N * g.M - 0.5 * g.N * (g.N - 1) : (double)g.M * g.N;
  return 2.0 * g.K * entries;
}

static const int SMALL_M = 64;   // fronts with m <= 64 run whole in one wave
static const int NB = 64;        // panel block width of large fronts

// One independent system of a (possibly batched) device schedule.
struct SysView {
  const tamd_symbolic* S;
  const tamd_plan* P;
  long long xoff;             // first row of this system in the concatenated vectors
};

// Builds the device schedule of one or more independent systems.  Fronts of
// level l of every system share the launches of level l; each system keeps
// its own arenas (offsets), row numbering (shifted by xoff) and tree.
static int dev_build(tamd_dev* D, const std::vector<SysView>& sys, bool dry = false, int panel_w = 0,
                     int tile = 0)
{
  D->panel_w = panel_w > 0 ? panel_w : g_panel_w;
  D->tile_big = tile > 0 ? tile : g_tile_big;
  int nlev = 0, nf = 0;
  long long nnz_tot = 0;
  for (const SysView& v : sys) {
    nlev = std::max(nlev, v.P->n_levels);
    nf += v.P->n_live;
    nnz_tot += v.P->a_nnz;
  }
  D->n_fronts = nf;
  std::vector<FrontDesc> fr(nf);
  std::vector<int> sibr(nf), rows, rel;
  std::vector<int> lvl(nlev + 1, 0);
  std::vector<long long> adest;
  adest.reserve(nnz_tot);
  // global index of (system, plan position), grouped by level across systems
  std::vector<std::vector<int>> gidx(sys.size());
  {
    int g = 0;
    for (int l = 0; l < nlev; l++) {
      lvl[l] = g;
      for (size_t q = 0; q < sys.size(); q++) {
        const tamd_plan* P = sys[q].P;
        if (gidx[q].empty()) gidx[q].assign(P->n_live, -1);
        if (l >= P->n_levels) continue;
        for (int k = P->level_ptr[l]; k < P->level_ptr[l + 1]; k++) gidx[q][k] = g++;
      }
    }
    lvl[nlev] = g;
  }
  long long Lbase = 0, Ubase = 0, offW = 0, offI = 0;
  for (size_t q = 0; q < sys.size(); q++) {
    const tamd_symbolic* S = sys[q].S;
    const tamd_plan* P = sys[q].P;
    std::vector<int> idx_of(P->n_sn + 1, -1);
    for (int k = 0; k < P->n_live; k++) idx_of[P->order[k]] = gidx[q][k];
    const int relbase = (int)rel.size();
    for (int k = 0; k < P->n_live; k++) {
      int sn = P->order[k];
      FrontDesc d;
      memset(&d, 0, sizeof(d));
      d.offL = Lbase + P->off_L[sn];
      d.offU = Ubase + P->off_U[sn];
      d.s = S->sn_size[sn];
      d.m = S->sn_up_size[sn];
      d.parent = P->sn_parent[sn] < P->n_sn ? idx_of[P->sn_parent[sn]] : -1;
      d.rel = relbase + P->rel_ptr[sn];
      fr[gidx[q][k]] = d;
      sibr[gidx[q][k]] = P->sib_rank[sn];
    }
    int nrel = P->rel_ptr[P->n_sn] > 0 ? P->rel_ptr[P->n_sn] : 0;
    rel.insert(rel.end(), P->rel, P->rel + nrel);
    for (int e = 0; e < P->a_nnz; e++) adest.push_back(Lbase + P->a_dest[e]);
    Lbase += P->L_elems;
    Ubase += P->U_elems;
  }
  // row lists, workspace and inverse offsets in global (level-grouped) order
  for (int g = 0; g < nf; g++) fr[g].rows = -1;
  for (size_t q = 0; q < sys.size(); q++) {
    const tamd_symbolic* S = sys[q].S;
    const tamd_plan* P = sys[q].P;
    for (int k = 0; k < P->n_live; k++) {
      int sn = P->order[k], g = gidx[q][k];
      fr[g].rows = (int)rows.size();
      for (int e = 0; e < S->sn_up_size[sn]; e++) rows.push_back((int)(sys[q].xoff + P->drow[sn][e]));
    }
  }
  for (int g = 0; g < nf; g++) {
    fr[g].offW = offW;
    fr[g].offI = offI;
    offW += fr[g].m;
    offI += (long long)((fr[g].s + 63) / 64) * 4096;
  }
  D->W_per_rhs = offW;
  D->I_elems = offI;
  D->L_elems = Lbase;
  D->U_elems = Ubase;
  D->nnz = (int)nnz_tot;
  D->n_levels = nlev;

  // factorization task lists, level by level
  std::vector<int> small;
  std::vector<SmallPack> packs;
  std::vector<PanelTask> potrf;
  std::vector<SolveTask> trsm;
  std::vector<GemmTask> gemm;
  std::vector<EATask> ea;
  std::vector<ZeroTask> zero;
  // update blocks by the level that zeroes them (global U offsets): a leaf
  // front's own block (its first writer is its own front op or SYRK), or a
  // parent's block (first written by that level's extend-add)
  std::vector<std::vector<ZeroTask>> zero_at(nlev), zero_par_at(nlev);
  {
    long long ub = 0;
    for (size_t q = 0; q < sys.size(); q++) {
      const tamd_symbolic* S = sys[q].S;
      const tamd_plan* P = sys[q].P;
      for (int k = 0; k < P->n_live; k++) {
        int sn = P->order[k];
        long long u = S->sn_up_size[sn] - S->sn_size[sn];
        if (u <= 0) continue;
        const long long off = ub + P->off_U[sn];
        auto& za = (P->u_start[sn] == P->sn_level[sn] ? zero_at : zero_par_at)[P->u_start[sn]];
        if (u <= 128) {  // a small block whole, as one flat range (column 0 of a u*u x 1 block)
          za.push_back({off, (int)(u * u), 0, 1, 0});
          continue;
        }
        // the lower triangle, by column ranges of about 64K elements
        long long cnt = 0;
        for (int c = 0, c0 = 0; c < (int)u; c++) {
          cnt += u - c;
          if (cnt >= 65536 || c == (int)u - 1) {
            za.push_back({off, (int)u, c0, c + 1, 0});
            c0 = c + 1;
            cnt = 0;
          }
        }
      }
      ub += P->U_elems;
    }
  }
  D->lvl_ptr = lvl;
  D->ops.clear();
  D->ops.push_back({TAMD_OP_ASSEMBLE, 0, 0, -1});
  D->gemm_flops = 0;
  D->assembly_bytes = 12.0 * (double)nnz_tot;
  for (int l = 0; l < nlev; l++) {
    int b0 = lvl[l], b1 = lvl[l + 1];
    if (!zero_at[l].empty()) {
      D->ops.push_back({TAMD_OP_ZERO, (int)zero.size(), (int)zero_at[l].size(), l});
      zero.insert(zero.end(), zero_at[l].begin(), zero_at[l].end());
    }
    // parents' blocks are first written by this level's extend-add (which
    // joins the side chain): their zeroing overlaps the level's fronts
    if (!zero_par_at[l].empty()) {
      Op o = {TAMD_OP_ZERO, (int)zero.size(), (int)zero_par_at[l].size(), l};
      o.lane = 1;
      D->ops.push_back(o);
      zero.insert(zero.end(), zero_par_at[l].begin(), zero_par_at[l].end());
    }
    int st = (int)small.size();
    int maxouter = 0;
    const int W = D->panel_w;
    std::vector<int> c16, c32, c48;
    for (int k = b0; k < b1; k++) {
      if (g_front_packing && fr[k].m <= 16) c16.push_back(k);
      else if (g_front_packing && fr[k].m <= 32) c32.push_back(k);
      else if (g_front_packing && fr[k].m <= 48) c48.push_back(k);
      else if (fr[k].m <= SMALL_M) small.push_back(k);
      else maxouter = std::max(maxouter, (fr[k].s + W - 1) / W);
    }
    // own blocks only: not the parents' blocks zeroed beside them
    for (int cls = 0; cls < 2; cls++) {
      const std::vector<int>& cf = cls == 0 ? c16 : c32;
      const int per = cls == 0 ? 4 : 2;
      if (cf.empty()) continue;
      const int p0 = (int)packs.size();
      for (size_t i = 0; i < cf.size(); i += per) {
        SmallPack pk;
        for (int q = 0; q < 4; q++) pk.f[q] = (q < per && i + q < cf.size()) ? cf[i + q] : -1;
        pk.width = cls == 0 ? 16 : 32;
        pk.pad[0] = pk.pad[1] = pk.pad[2] = 0;
        packs.push_back(pk);
      }
      Op o = {TAMD_OP_SMALL, p0, (int)packs.size() - p0, l};
      o.nojoin = true;
      o.tile = cls == 0 ? 16 : 32;
      D->ops.push_back(o);
    }
    if ((int)small.size() > st) {
      Op o = {TAMD_OP_SMALL, st, (int)small.size() - st, l};
      o.nojoin = true;
      D->ops.push_back(o);
    }
    if (!c48.empty()) {  // one front per wavefront, 48-wide staging
      Op o = {TAMD_OP_SMALL, (int)small.size(), (int)c48.size(), l};
      o.nojoin = true;
      o.tile = 48;
      D->ops.push_back(o);
      small.insert(small.end(), c48.begin(), c48.end());
    }
    // Two-level blocking: outer panels of W columns, each factored in 64-column
    // steps (POTRF of the diagonal block, TRSM of the rows below by its
    // inverse, update of the panel's remaining columns with K = 64); then one
    // trailing update of the columns right of the panel with K = W, so the
    // trailing matrix is read and written once per W columns.
    // Lookahead: that update is split into the next outer panel's columns
    // (main stream) and the rest (side stream); the next panel's own steps
    // touch only its columns and overlap the rest.
    auto tile = [&](const FrontDesc& d, int r0, int c0, int k0, int kw, int ts) {
      GemmTask g;
      g.a_off = d.offL + r0 + (long long)k0 * d.m;
      g.b_off = d.offL + c0 + (long long)k0 * d.m;
      g.c_off = d.offL + r0 + (long long)c0 * d.m;
      g.ld = d.m;
      g.ldc = d.m;
      g.M = std::min(ts, d.m - r0);
      g.N = std::min(ts, d.s - c0);
      g.K = kw;
      g.flags = (r0 == c0) ? 2 : 0;
      D->gemm_flops += tile_flops(g);
      return g;
    };
    std::vector<GemmTask> rest;
    for (int os = 0; os < maxouter; os++) {
      for (int j = 0; j < W / NB; j++) {
        int p0 = (int)potrf.size(), t0 = (int)trsm.size();
        for (int k = b0; k < b1; k++) {
          const FrontDesc& d = fr[k];
          if (d.m <= SMALL_M) continue;
          const int K0 = os * W, k0 = K0 + j * NB;
          if (k0 >= d.s) continue;
          const int kb = std::min(NB, d.s - k0), pend = std::min(K0 + W, d.s);
          // an outer panel's first block waits for the trailing update; the
          // later ones are factored at the end of the previous step's launch
          if (k0 == K0) potrf.push_back({k, k0, kb, 0, 0});
          // one task per 64-row tile below the block (k_panel_solve): its
          // updates from the outer panel's earlier blocks, its solve, and for
          // a tile that starts a later diagonal block of the outer panel that
          // block's update from this step
          for (int r0 = k0 + kb; r0 < d.m; r0 += 64) {
            SolveTask st;
            st.c_off = d.offL + r0 + (long long)k0 * d.m;
            st.inv_off = d.offI + (long long)(k0 / 64) * 4096;
            st.xr_off = d.offL + r0 + (long long)K0 * d.m;
            st.xk_off = d.offL + k0 + (long long)K0 * d.m;
            st.d_off = d.offL + r0 + (long long)r0 * d.m;
            st.ld = d.m;
            st.M = std::min(64, d.m - r0);
            st.N = kb;
            st.kpre = k0 - K0;
            st.Md = r0 < pend ? std::min(64, d.s - r0) : 0;
            st.pf_front = r0 == k0 + kb && r0 < pend ? k : -1;
            st.pf_k0 = r0;
            trsm.push_back(st);
          }
        }
        Op o;
        if ((int)potrf.size() > p0) {
          o = {TAMD_OP_POTRF, p0, (int)potrf.size() - p0, l};
          o.nojoin = true;
          D->ops.push_back(o);
        }
        if ((int)trsm.size() > t0) {
          o = {TAMD_OP_TRSM, t0, (int)trsm.size() - t0, l};
          o.nojoin = true;
          D->ops.push_back(o);
        }
      }
      int g0 = (int)gemm.size();
      rest.clear();
      // the tile edge of each launch: 128 only where the launch still has
      // enough 128-tiles to fill the GPU, else 64
      long long n_main = 0, n_rest = 0;
      if (D->tile_big == 128)
        for (int k = b0; k < b1; k++) {
          const FrontDesc& d = fr[k];
          if (d.m <= SMALL_M || os * W >= d.s) continue;
          for (int c0 = os * W + W; c0 < d.s; c0 += 128)
            (c0 < os * W + 2 * W ? n_main : n_rest) += (d.m - c0 + 127) / 128;
        }
      const int ts_main = D->tile_big == 128 && n_main >= TILE128_MIN ? 128 : 64;
      const int ts_rest = D->tile_big == 128 && n_rest >= TILE128_MIN ? 128 : 64;
      for (int k = b0; k < b1; k++) {
        const FrontDesc& d = fr[k];
        if (d.m <= SMALL_M) continue;
        const int K0 = os * W;
        if (K0 >= d.s) continue;
        const int kw = std::min(W, d.s - K0);
        // main: the tiles that start within the next outer panel's columns
        // (with W below the tile edge they reach beyond it); rest: the others
        int c0 = K0 + W;
        for (; c0 < d.s && c0 < K0 + 2 * W; c0 += ts_main)
          for (int r0 = c0; r0 < d.m; r0 += ts_main) gemm.push_back(tile(d, r0, c0, K0, kw, ts_main));
        for (; c0 < d.s; c0 += ts_rest)
          for (int r0 = c0; r0 < d.m; r0 += ts_rest) rest.push_back(tile(d, r0, c0, K0, kw, ts_rest));
      }
      if ((int)gemm.size() > g0) {
        Op o = {TAMD_OP_GEMM, g0, (int)gemm.size() - g0, l};
        o.tile = ts_main;
        D->ops.push_back(o);
      }
      if (!rest.empty()) {
        Op o = {TAMD_OP_GEMM, (int)gemm.size(), (int)rest.size(), l};
        o.lane = 1;
        o.tile = ts_rest;
        D->ops.push_back(o);
        gemm.insert(gemm.end(), rest.begin(), rest.end());
      }
    }
    // SYRK of the update blocks, K = s
    {
      int g0 = (int)gemm.size();
      long long n128 = 0;
      for (int k = b0; k < b1; k++) {
        const long long u = fr[k].m - fr[k].s;
        if (fr[k].m > SMALL_M && u > 0) n128 += ((u + 127) / 128) * ((u + 127) / 128 + 1) / 2;
      }
      const int ts = D->tile_big == 128 && n128 >= TILE128_MIN ? 128 : 64;
      for (int k = b0; k < b1; k++) {
        const FrontDesc& d = fr[k];
        int u = d.m - d.s;
        if (d.m <= SMALL_M || u == 0) continue;
        for (int j0 = 0; j0 < u; j0 += ts)
          for (int i0 = j0; i0 < u; i0 += ts) {
            GemmTask g;
            g.a_off = d.offL + d.s + i0;
            g.b_off = d.offL + d.s + j0;
            g.c_off = d.offU + i0 + (long long)j0 * u;
            g.ld = d.m;
            g.ldc = u;
            g.M = std::min(ts, u - i0);
            g.N = std::min(ts, u - j0);
            g.K = d.s;
            g.flags = 1 | (i0 == j0 ? 2 : 0);
            gemm.push_back(g);
            D->gemm_flops += tile_flops(g);
          }
      }
      if ((int)gemm.size() > g0) {
        Op o = {TAMD_OP_GEMM, g0, (int)gemm.size() - g0, l};
        o.tile = ts;
        D->ops.push_back(o);
      }
    }
    // extend-add of this level's fronts into their parents.  A parent whose
    // children at this level are all small (at most EA_CHUNK entries each, at
    // most EA_GROUP in all) gets one workgroup that adds them in sibling-rank
    // order (k_extend_add_seq, one launch for all such parents); the children
    // of the other parents go in one launch per sibling rank, split into
    // column chunks of about EA_CHUNK entries.  Either way a parent entry
    // receives its children's values in rank order.  (Chunk and group sizes
    // measured on mesh3d 60^3: 16K-entry chunks 4.9 ms, 1K 3.2 ms; grouping
    // pays only where parents gather many tiny children.)
    {
      const long long EA_GROUP = 8192, EA_CHUNK = 1024;
      std::vector<long long> pwork(nf, 0), pmax(nf, 0);
      for (int k = b0; k < b1; k++) {
        const long long u = fr[k].m - fr[k].s;
        if (fr[k].parent < 0 || u == 0) continue;
        pwork[fr[k].parent] += u * (u + 1) / 2;
        pmax[fr[k].parent] = std::max(pmax[fr[k].parent], u * (u + 1) / 2);
      }
      auto grouped = [&](int p) { return pwork[p] <= EA_GROUP && pmax[p] <= EA_CHUNK; };
      std::vector<std::vector<int>> kids(nf);
      for (int k = b0; k < b1; k++)
        if (fr[k].parent >= 0 && fr[k].m > fr[k].s && grouped(fr[k].parent))
          kids[fr[k].parent].push_back(k);
      std::vector<EATask> heads, body;
      for (int p = 0; p < nf; p++) {
        if (kids[p].empty()) continue;
        std::sort(kids[p].begin(), kids[p].end(), [&](int a, int b) { return sibr[a] < sibr[b]; });
        heads.push_back({(int)body.size(), (int)kids[p].size(), 0, 0});
        for (int k : kids[p]) {
          const int u = fr[k].m - fr[k].s;
          D->assembly_bytes += 24.0 * ((double)u * (u + 1) / 2);
          body.push_back({k, 0, u, 0});
        }
      }
      if (!heads.empty()) {
        const int hb = (int)ea.size();
        const int bb = hb + (int)heads.size();
        for (EATask& h : heads) h.child += bb;  // heads index the bodies in the same array
        ea.insert(ea.end(), heads.begin(), heads.end());
        ea.insert(ea.end(), body.begin(), body.end());
        Op o{TAMD_OP_EXTEND, hb, (int)heads.size(), l};
        o.grouped = true;
        D->ops.push_back(o);
      }
      int maxrank = -1;
      for (int k = b0; k < b1; k++)
        if (fr[k].parent >= 0) maxrank = std::max(maxrank, sibr[k]);
      for (int rk = 0; rk <= maxrank; rk++) {
        int e0 = (int)ea.size();
        for (int k = b0; k < b1; k++) {
          const FrontDesc& d = fr[k];
          int u = d.m - d.s;
          if (d.parent < 0 || u == 0 || sibr[k] != rk) continue;
          if (grouped(d.parent)) continue;
          D->assembly_bytes += 24.0 * ((double)u * (u + 1) / 2);
          // chunks of about EA_CHUNK entries
          int j = 0;
          while (j < u) {
            int j1 = j;
            long long acc = 0;
            while (j1 < u && acc < EA_CHUNK) { acc += u - j1; j1++; }
            ea.push_back({k, j, j1, 0});
            j = j1;
          }
        }
        if ((int)ea.size() > e0) D->ops.push_back({TAMD_OP_EXTEND, e0, (int)ea.size() - e0, l});
      }
    }
  }

  // solve schedule: forward bottom-up, backward top-down, 64-column blocks;
  // every op is a list of PanelTasks (front, k0, kb, r0, nr)
  std::vector<PanelTask> svt;
  D->sv_ops.clear();
  for (int k = 0; k < nf; k++)  // k_solve_init: 64-row pieces of every front
    for (int r0 = 0; r0 < fr[k].m; r0 += 64) svt.push_back({k, 0, 0, r0, std::min(64, fr[k].m - r0)});
  const int n_init = (int)svt.size();
  D->sv_ops.push_back({SV_INIT, 0, n_init, -1});
  auto push_tasks = [&](int kind, size_t t0, int l) {
    if (svt.size() > t0) D->sv_ops.push_back({kind, (int)t0, (int)(svt.size() - t0), l});
  };
  const int SB = 64, CH = 512, CHB = 128;  // CHB: rows per k_bwd_upd task (its BU_ROWS)
  D->part_slots = 0;
  for (int l = 0; l < nlev; l++) {
    int b0 = lvl[l], b1 = lvl[l + 1];
    int maxb = 0;
    for (int k = b0; k < b1; k++) maxb = std::max(maxb, (fr[k].s + SB - 1) / SB);
    for (int step = 0; step < maxb; step++) {
      size_t t0 = svt.size();
      for (int k = b0; k < b1; k++) {
        int k0 = step * SB;
        if (k0 >= fr[k].s) continue;
        int kb = std::min(SB, fr[k].s - k0);
        if (k0 + kb >= fr[k].s) svt.push_back({k, k0, kb, k0 + kb, 0});  // stores y only
        for (int r0 = k0 + kb; r0 < fr[k].s; r0 += 64)
          svt.push_back({k, k0, kb, r0, std::min(64, fr[k].s - r0)});
      }
      push_tasks(SV_FSTEP, t0, l);
    }
    {  // L21 rows: split over column chunks of CH, then a deterministic reduce
      size_t t0 = svt.size();
      std::vector<PanelTask> red;
      for (int k = b0; k < b1; k++)
        for (int r0 = fr[k].s; r0 < fr[k].m; r0 += 64) {
          int nr = std::min(64, fr[k].m - r0), slot0 = (int)(svt.size() - t0), cnt = 0;
          for (int c0 = 0; c0 < fr[k].s; c0 += CH, cnt++)
            svt.push_back({k, c0, std::min(CH, fr[k].s - c0), r0, nr});
          red.push_back({k, slot0, cnt, r0, nr});
        }
      D->part_slots = std::max(D->part_slots, (long long)(svt.size() - t0));
      push_tasks(SV_FRECT, t0, l);
      t0 = svt.size();
      svt.insert(svt.end(), red.begin(), red.end());
      push_tasks(SV_FRED, t0, l);
    }
    int maxrank = -1;
    for (int k = b0; k < b1; k++)
      if (fr[k].parent >= 0 && fr[k].m > fr[k].s) maxrank = std::max(maxrank, sibr[k]);
    for (int rk = 0; rk <= maxrank; rk++) {
      size_t t0 = svt.size();
      for (int k = b0; k < b1; k++) {
        int u = fr[k].m - fr[k].s;
        if (fr[k].parent < 0 || u == 0 || sibr[k] != rk) continue;
        for (int r0 = 0; r0 < u; r0 += 256) svt.push_back({k, 0, 0, r0, std::min(256, u - r0)});
      }
      push_tasks(SV_EXT, t0, l);
    }
  }
  for (int l = nlev - 1; l >= 0; l--) {
    int b0 = lvl[l], b1 = lvl[l + 1];
    size_t t0 = svt.size();
    for (int k = b0; k < b1; k++) {
      int u = fr[k].m - fr[k].s;
      for (int r0 = 0; r0 < u; r0 += 256) svt.push_back({k, 0, 0, r0, std::min(256, u - r0)});
    }
    push_tasks(SV_GATHER, t0, l);
    {  // z -= L21^T x2: 64 columns per task, split over row chunks of CH
      t0 = svt.size();
      std::vector<PanelTask> red;
      for (int k = b0; k < b1; k++) {
        if (fr[k].m == fr[k].s) continue;
        for (int c0 = 0; c0 < fr[k].s; c0 += 64) {
          int nc = std::min(64, fr[k].s - c0), slot0 = (int)(svt.size() - t0), cnt = 0;
          for (int i0 = fr[k].s; i0 < fr[k].m; i0 += CH, cnt++)
            svt.push_back({k, i0, std::min(CH, fr[k].m - i0), c0, nc});
          red.push_back({k, slot0, cnt, c0, nc});
        }
      }
      D->part_slots = std::max(D->part_slots, (long long)(svt.size() - t0));
      push_tasks(SV_BCOLS, t0, l);
      t0 = svt.size();
      svt.insert(svt.end(), red.begin(), red.end());
      push_tasks(SV_BRED, t0, l);
    }
    int maxb = 0;
    for (int k = b0; k < b1; k++) maxb = std::max(maxb, (fr[k].s + SB - 1) / SB);
    for (int q = 0; q < maxb; q++) {  // q-th block from the end of each front
      t0 = svt.size();
      for (int k = b0; k < b1; k++) {
        int nbk = (fr[k].s + SB - 1) / SB, blk = nbk - 1 - q;
        if (blk < 0) continue;
        int k0 = blk * SB, kb = std::min(SB, fr[k].s - k0);
        if (k0 == 0) svt.push_back({k, 0, kb, 0, 0});  // stores x only
        for (int c0 = 0; c0 < k0; c0 += 64) svt.push_back({k, k0, kb, c0, std::min(64, k0 - c0)});
      }
      push_tasks(SV_BSTEP, t0, l);
    }
  }

  // chained sweeps: per level, the fronts with m <= 64 one wave each
  // (k_fwd_small / k_bwd_small); the others through k_fwd_chain /
  // k_bwd_chain, in one ticketed launch whose task order puts every task
  // after the tasks it waits for (forward: the block tiles round-robin by
  // block index, then the sub-diagonal tiles; backward: the blocks from the
  // last one), single-block fronts of the backward without tickets.  Op.tile
  // holds the launch's ticket index (-1: none).
  D->sv_chain_ops.clear();
  // b is read into the workspaces before x is cleared: x may alias b
  D->sv_chain_ops.push_back({SV_INIT, 0, n_init, -1});
  D->sv_chain_ops.push_back({SV_CLEAR, 0, 0, -1});
  int ntickets = 0;
  auto push_chain = [&](int kind, size_t t0, int l, bool ticketed) {
    if (svt.size() <= t0) return;
    Op o{kind, (int)t0, (int)(svt.size() - t0), l};
    o.tile = ticketed ? ntickets++ : -1;
    D->sv_chain_ops.push_back(o);
  };
  auto nblk = [&](int k) { return (fr[k].s + SB - 1) / SB; };
  for (int l = 0; l < nlev; l++) {
    const int b0 = lvl[l], b1 = lvl[l + 1];
    size_t t0 = svt.size();
    for (int k = b0; k < b1; k++)
      if (fr[k].m <= 64) svt.push_back({k, 0, 0, 0, 0});
    push_chain(SV_FSMALL, t0, l, false);
    t0 = svt.size();
    int maxb = 0;
    for (int k = b0; k < b1; k++)
      if (fr[k].m > 64) maxb = std::max(maxb, nblk(k));
    for (int b = 0; b < maxb; b++)
      for (int k = b0; k < b1; k++)
        if (fr[k].m > 64 && b < nblk(k)) svt.push_back({k, b, 0, 0, 0});
    for (int k = b0; k < b1; k++)
      if (fr[k].m > 64)
        for (int tt = nblk(k); tt < (fr[k].m + 63) / 64; tt++) svt.push_back({k, tt, 0, 0, 0});
    push_chain(SV_FCHAIN, t0, l, true);
    // children's updates into the parents: parents whose children here all
    // have at most 256 update rows get one workgroup each (k_solve_ext_seq,
    // children in sibling-rank order); the others by sibling rank
    std::vector<int> pmaxu(nf, 0);
    for (int k = b0; k < b1; k++)
      if (fr[k].parent >= 0) pmaxu[fr[k].parent] = std::max(pmaxu[fr[k].parent], fr[k].m - fr[k].s);
    auto ext_grouped = [&](int p) { return pmaxu[p] <= 256; };
    {
      std::vector<std::vector<int>> kids(nf);
      for (int k = b0; k < b1; k++)
        if (fr[k].parent >= 0 && fr[k].m > fr[k].s && ext_grouped(fr[k].parent))
          kids[fr[k].parent].push_back(k);
      std::vector<PanelTask> heads, body;
      for (int p = 0; p < nf; p++) {
        if (kids[p].empty()) continue;
        std::sort(kids[p].begin(), kids[p].end(), [&](int a, int b) { return sibr[a] < sibr[b]; });
        heads.push_back({(int)body.size(), (int)kids[p].size(), 0, 0, 0});
        for (int k : kids[p]) body.push_back({k, 0, 0, 0, fr[k].m - fr[k].s});
      }
      if (!heads.empty()) {
        t0 = svt.size();
        const int bb = (int)(t0 + heads.size());
        for (PanelTask& h : heads) h.front += bb;  // heads index the bodies in the same array
        svt.insert(svt.end(), heads.begin(), heads.end());
        push_chain(SV_EXTSEQ, t0, l, false);
        svt.insert(svt.end(), body.begin(), body.end());
      }
    }
    int maxrank = -1;
    for (int k = b0; k < b1; k++)
      if (fr[k].parent >= 0 && fr[k].m > fr[k].s && !ext_grouped(fr[k].parent))
        maxrank = std::max(maxrank, sibr[k]);
    for (int rk = 0; rk <= maxrank; rk++) {
      t0 = svt.size();
      for (int k = b0; k < b1; k++) {
        int u = fr[k].m - fr[k].s;
        if (fr[k].parent < 0 || u == 0 || sibr[k] != rk || ext_grouped(fr[k].parent)) continue;
        for (int r0 = 0; r0 < u; r0 += 256) svt.push_back({k, 0, 0, r0, std::min(256, u - r0)});
      }
      push_chain(SV_EXT, t0, l, false);
    }
  }
  for (int l = nlev - 1; l >= 0; l--) {
    const int b0 = lvl[l], b1 = lvl[l + 1];
    size_t t0 = svt.size();
    // fronts with m > 64: the update rows' part L21^T x2 first, split over
    // row chunks (k_bwd_upd, one Part slot per chunk: block b of front k owns
    // slots sl0[k] + b * scnt[k] ..); the chained launch sums its slots and
    // follows the diagonal block chain
    std::vector<int> sl0(nf, 0), scnt(nf, 0);
    {
      t0 = svt.size();
      for (int k = b0; k < b1; k++) {
        if (fr[k].m <= 64 || fr[k].m == fr[k].s) continue;
        sl0[k] = (int)(svt.size() - t0);
        scnt[k] = (fr[k].m - fr[k].s + CHB - 1) / CHB;
        for (int c0 = 0; c0 < fr[k].s; c0 += 64) {
          const int nc = std::min(64, fr[k].s - c0);
          for (int i0 = fr[k].s; i0 < fr[k].m; i0 += CHB)
            svt.push_back({k, i0, std::min(CHB, fr[k].m - i0), c0, nc});
        }
      }
      D->part_slots = std::max(D->part_slots, (long long)(svt.size() - t0));
      push_chain(SV_BUPD, t0, l, false);
    }
    t0 = svt.size();
    for (int k = b0; k < b1; k++)
      if (fr[k].m <= 64) svt.push_back({k, 0, 0, 0, 0});
    push_chain(SV_BSMALL, t0, l, false);
    t0 = svt.size();
    for (int k = b0; k < b1; k++)
      if (fr[k].m > 64 && nblk(k) == 1) svt.push_back({k, 0, sl0[k], 0, scnt[k]});
    push_chain(SV_BCHAIN, t0, l, false);
    t0 = svt.size();
    int maxb = 0;
    for (int k = b0; k < b1; k++)
      if (nblk(k) > 1) maxb = std::max(maxb, nblk(k));
    for (int q = 0; q < maxb; q++)
      for (int k = b0; k < b1; k++)
        if (nblk(k) > 1 && q < nblk(k)) {
          const int b = nblk(k) - 1 - q;
          svt.push_back({k, b, sl0[k] + b * scnt[k], 0, scnt[k]});
        }
    push_chain(SV_BCHAIN, t0, l, true);
  }
  D->sync_n = ntickets;
  D->solve_chain = g_solve_chain;

  // concatenated A pattern: column j of system q is global column xoff_q + j
  long long ncol = 0;
  for (const SysView& v : sys) ncol = std::max(ncol, v.xoff + (long long)v.S->n);
  std::vector<int> colptr(ncol + 1, 0);
  {
    long long base = 0;
    for (const SysView& v : sys) {
      for (int j = 0; j <= v.S->n; j++) colptr[v.xoff + j] = (int)(base + v.P->a_colptr[j]);
      base += v.P->a_nnz;
    }
  }
  D->ncol = (int)ncol;
  // A's entries can only share a destination within one column (a duplicated
  // (row, column) pair); without one, assembly runs one thread per entry
  {
    bool dup = false;
    std::vector<long long> cd;
    for (long long j = 0; j < ncol && !dup; j++) {
      const long long* c = adest.data() + colptr[j];
      const int cnt = colptr[j + 1] - colptr[j];
      if (cnt <= 16) {  // short columns pairwise, in place
        for (int a = 1; a < cnt && !dup; a++)
          for (int b = 0; b < a; b++)
            if (c[a] == c[b]) dup = true;
      } else {
        cd.assign(c, c + cnt);
        std::sort(cd.begin(), cd.end());
        for (size_t q = 1; q < cd.size(); q++)
          if (cd[q] == cd[q - 1]) dup = true;
      }
    }
    D->assemble_flat = !dup;
  }
  // GEMM launches of 64-tiles with K <= 64 use the one-round-trip tile
  for (Op& op : D->ops)
    if (op.kind == TAMD_OP_GEMM && op.tile == 64) {
      bool sk = true;
      for (int q = op.start; q < op.start + op.count && sk; q++) sk = gemm[q].K <= 64;
      op.shortk = sk;
    }
  // algorithmic work per op: tile flops (lower half of diagonal tiles), and
  // extend-add / assembly bytes (read child + read-modify-write parent)
  D->op_flops.assign(D->ops.size(), 0.0);
  D->op_bytes.assign(D->ops.size(), 0.0);
  for (size_t o = 0; o < D->ops.size(); o++) {
    const Op& op = D->ops[o];
    double f = 0, b = 0;
    if (op.kind == TAMD_OP_GEMM)
      for (int q = op.start; q < op.start + op.count; q++) f += tile_flops(gemm[q]);
    else if (op.kind == TAMD_OP_TRSM)
      for (int q = op.start; q < op.start + op.count; q++) {
        const SolveTask& st = trsm[q];
        f += 2.0 * st.M * st.N * (st.N + st.kpre) + (double)st.Md * (st.Md + 1) * st.N;
      }
    else if (op.kind == TAMD_OP_POTRF)
      f = op.count * (2.0 / 3.0) * 64.0 * 64.0 * 64.0;
    else if (op.kind == TAMD_OP_EXTEND)
      for (int q = op.start; q < op.start + op.count; q++) {
        const int q0 = op.grouped ? ea[q].child : q, q1 = op.grouped ? q0 + ea[q].j0 : q + 1;
        for (int e = q0; e < q1; e++) {
          const int uc = fr[ea[e].child].m - fr[ea[e].child].s;
          for (int j = ea[e].j0; j < ea[e].j1; j++) b += 24.0 * (uc - j);
        }
      }
    else if (op.kind == TAMD_OP_ASSEMBLE)
      b = 20.0 * (double)nnz_tot;
    else if (op.kind == TAMD_OP_ZERO)
      for (int q = op.start; q < op.start + op.count; q++) {
        const double u = zero[q].u, c0 = zero[q].c0, c1 = zero[q].c1;
        b += 8.0 * ((c1 - c0) * u - 0.5 * (c1 * (c1 - 1) - c0 * (c0 - 1)));
      }
    D->op_flops[o] = f;
    D->op_bytes[o] = b;
  }
  D->struct_bytes = sizeof(FrontDesc) * fr.size() + sizeof(int) * (rows.size() + rel.size())
                    + sizeof(long long) * adest.size()
                    + sizeof(int) * (small.size() + colptr.size() + 1)
                    + sizeof(SmallPack) * packs.size()
                    + sizeof(PanelTask) * (potrf.size() + svt.size())
                    + sizeof(SolveTask) * trsm.size() + sizeof(GemmTask) * gemm.size()
                    + sizeof(EATask) * ea.size()
                    + sizeof(ZeroTask) * zero.size()
                    + sizeof(double) * (size_t)nnz_tot;
  if (dry) return 0;
  if (upload(&D->d_fronts, fr, D->stream) || upload(&D->d_rows, rows, D->stream)
      || upload(&D->d_rel, rel, D->stream) || upload(&D->d_adest, adest, D->stream)
      || upload(&D->d_small, small, D->stream) || upload(&D->d_packs, packs, D->stream)
      || upload(&D->d_potrf, potrf, D->stream)
      || upload(&D->d_trsm, trsm, D->stream) || upload(&D->d_gemm, gemm, D->stream)
      || upload(&D->d_ea, ea, D->stream) || upload(&D->d_sv_tasks, svt, D->stream)
      || upload(&D->d_zero, zero, D->stream)
      || upload(&D->d_colptr, colptr, D->stream)
      || dev_malloc((void**)&D->d_info, 2 * sizeof(int))
      || dev_malloc((void**)&D->d_avals, sizeof(double) * (nnz_tot > 0 ? nnz_tot : 1)))
    return -1;
  if (hipStreamSynchronize(D->stream) != hipSuccess) return -1;
  return 0;
}

// The factor's own stream plus the side stream of the lookahead, both
// non-blocking and at the default priority.
static int create_streams(tamd_dev* D)
{
  if (hipStreamCreateWithFlags(&D->stream, hipStreamNonBlocking) != hipSuccess
      || hipStreamCreateWithFlags(&D->side, hipStreamNonBlocking) != hipSuccess
      || hipEventCreateWithFlags(&D->ev_fork, hipEventDisableTiming) != hipSuccess
      || hipEventCreateWithFlags(&D->ev_join, hipEventDisableTiming) != hipSuccess)
    return -1;
  return 0;
}

// Device schedule of `count` independent systems sharing the launches of
// every level; the systems' vectors are concatenated at offsets xoff[q].
extern "C" int tamd_dev_batch_setup(tamd_factor* FB, int count, tamd_factor* const* F,
                                    const long long* xoff)
{
  if (!tamd_dev_available()) {
    taucs_printf((char*)"taucs_amd: no GPU device available; the numeric phase runs only on the GPU\n");
    return -1;
  }
  tamd_dev* D = new tamd_dev();
  FB->dev = D;
  if (create_streams(D) != 0) {
    tamd_dev_free(FB);
    return -1;
  }
  std::vector<SysView> v(count);
  for (int q = 0; q < count; q++) {
    v[q].S = &F[q]->sym;
    v[q].P = &F[q]->plan;
    v[q].xoff = xoff[q];
  }
  if (dev_build(D, v, false, FB->panel_w, FB->gemm_tile) != 0) {
    tamd_dev_free(FB);
    return -1;
  }
  return 0;
}

extern "C" int tamd_dev_setup(tamd_factor* F)
{
  if (!tamd_dev_available()) {
    taucs_printf((char*)"taucs_amd: no GPU device available; the numeric phase runs only on the GPU\n");
    return -1;
  }
  tamd_dev* D = new tamd_dev();
  F->dev = D;
  if (create_streams(D) != 0) {
    tamd_dev_free(F);
    return -1;
  }
  std::vector<SysView> one(1);
  one[0].S = &F->sym;
  one[0].P = &F->plan;
  one[0].xoff = 0;
  if (dev_build(D, one, false, F->panel_w, F->gemm_tile) != 0) {
    tamd_dev_free(F);
    return -1;
  }
  return 0;
}

// Device bytes a factor will hold after a numeric factorization in
// `precision` and solves of up to `nrhs` right-hand sides from host buffers
// (plus the refinement context when `refine` is set), computed on the host
// from the same schedule construction, without touching the device.
extern "C" long long taucs_amd_factor_device_bytes(void* vL, int precision, int nrhs, int refine)
{
  tamd_factor* F = (tamd_factor*)vL;
  if (!F || nrhs < 0) return -1;
  tamd_dev D;
  std::vector<SysView> one(1);
  one[0].S = &F->sym;
  one[0].P = &F->plan;
  one[0].xoff = 0;
  if (dev_build(&D, one, true, F->panel_w, F->gemm_tile) != 0) return -1;
  const long long elem = precision == TAMD_PREC_FP64 ? 8 : 4;
  const long long n = F->n, cap = std::max(nrhs, 16);
  long long b = (long long)D.struct_bytes + sizeof(int);
  b += elem * (std::max(D.L_elems, 1LL) + std::max(D.U_elems, 1LL) + std::max(D.I_elems, 1LL));
  b += elem * (D.W_per_rhs * cap + n * cap + 64LL * 16 * std::max(D.part_slots, 1LL));
  b += (long long)sizeof(int) * (D.sync_n + 1);
  b += 2LL * 8 * n * std::max(nrhs, 1);
  if (refine) {
    const long long full = 2LL * F->plan.a_nnz;
    b += 4 * (n + 1) + 12 * std::max(full, 1LL);
    b += 6LL * 8 * n * std::max(nrhs, 1) + 8LL * (256 + 8) * std::max(nrhs, 1);
  }
  return b;
}

// ---- numeric factorization ------------------------------------------------

static inline unsigned grid_for(long long n, int block)
{
  long long g = (n + block - 1) / block;
  if (g > 65535LL * 64) g = 65535LL * 64;
  return (unsigned)(g > 0 ? g : 1);
}

// Issues kernels either on streams (eager) or as kernel nodes of a hipGraph
// built explicitly (no stream capture, so other threads' HIP calls never
// interfere with it).  Lane 0 is the main chain; lane 1 ops fork from the
// main chain's last op; `join` makes a main op wait for the side chain too.
struct Emitter {
  hipStream_t main_st = nullptr, side_st = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipGraph_t g = nullptr;
  hipGraphNode_t last_main = nullptr, last_side = nullptr;
  bool side_pending = false;
  int count = 0;
  unsigned shmem = 0;  // dynamic LDS of the next launch (then reset)

  template <typename... P, size_t... I>
  static void ptrs_of(std::tuple<P...>& t, void** p, std::index_sequence<I...>)
  {
    ((p[I] = (void*)&std::get<I>(t)), ...);
  }

  template <typename... P, typename... A>
  int kernel(int lane, bool join, void (*k)(P...), dim3 grid, dim3 block, A... a)
  {
    std::tuple<typename std::decay<P>::type...> vals(a...);
    void* p[sizeof...(P) > 0 ? sizeof...(P) : 1];
    ptrs_of(vals, p, std::index_sequence_for<P...>{});
    count++;
    const unsigned lds = shmem;
    shmem = 0;
    if (g) {
      hipGraphNode_t deps[2];
      int nd = 0;
      if (lane == 1) {
        if (last_main) deps[nd++] = last_main;
        if (last_side) deps[nd++] = last_side;
      } else {
        if (last_main) deps[nd++] = last_main;
        if (join && last_side) deps[nd++] = last_side;
      }
      hipKernelNodeParams kp;
      memset(&kp, 0, sizeof(kp));
      kp.func = (void*)k;
      kp.gridDim = grid;
      kp.blockDim = block;
      kp.sharedMemBytes = lds;
      kp.kernelParams = p;
      kp.extra = nullptr;
      hipGraphNode_t node;
      HIPCHECK(hipGraphAddKernelNode(&node, g, deps, nd, &kp));
      if (lane == 1) {
        last_side = node;
      } else {
        last_main = node;
        if (join) last_side = nullptr;
      }
      return 0;
    }
    hipStream_t st = main_st;
    if (lane == 1 && side_st) {
      HIPCHECK(hipEventRecord(ev_fork, main_st));
      HIPCHECK(hipStreamWaitEvent(side_st, ev_fork, 0));
      st = side_st;
    } else if (join && side_pending) {
      HIPCHECK(hipStreamWaitEvent(main_st, ev_join, 0));
      side_pending = false;
    }
    HIPCHECK(hipLaunchKernel((const void*)k, grid, block, p, lds, st));
    if (lane == 1 && side_st) {
      HIPCHECK(hipEventRecord(ev_join, side_st));
      side_pending = true;
    }
    return 0;
  }

  // eager mode: the main stream waits for the side chain
  int finish()
  {
    if (!g && side_pending) {
      HIPCHECK(hipStreamWaitEvent(main_st, ev_join, 0));
      side_pending = false;
    }
    return 0;
  }
};

#define EMIT(...)                            \
  do {                                       \
    if (E.kernel(__VA_ARGS__) != 0) return -1; \
  } while (0)

// LDS per CU is 160 KiB.  Side-lane tiles: FP64 two per CU (waves of 104
// registers, leaving 304 on each SIMD for the 288-register POTRF wave and
// 48 KiB of LDS for its 37 KiB); FP32 four per CU (64-register waves, POTRF
// 152 registers and 18.5 KiB).
template <typename T> constexpr int SIDE_WGS = sizeof(T) == 8 ? 2 : 4;
#define SIDE_PAD_MAX_TILES_FP64 4096
#ifndef LAT_TILES_MAX
#define LAT_TILES_MAX 512
#endif
template <typename T>
static unsigned side_lds_pad()
{
  const unsigned tile_lds = 2u * 64 * (16 + 2) * sizeof(T);  // As + Bs of the 64x64 tile
  const unsigned per_wg = (160u * 1024 - (sizeof(T) == 8 ? 48u : 24u) * 1024) / SIDE_WGS<T>;
  return per_wg > tile_lds ? per_wg - tile_lds : 0;
}

template <typename T>
static int launch_numeric(tamd_dev* D, Emitter& E, bool bf16upd, bool prof)
{
  T* L = (T*)D->L;
  T* U = (T*)D->U;
  T* Iv = (T*)D->Inv;
  hipEvent_t a0 = nullptr, a1 = nullptr;
  if (prof) {
    HIPCHECK(hipEventCreate(&a0));
    HIPCHECK(hipEventCreate(&a1));
    D->op_ms.assign(D->ops.size(), 0.0f);
  }
  EMIT(0, true, k_fill<T>, dim3(grid_for(D->L_elems, 256)), dim3(256), L, D->L_elems);
  EMIT(0, true, k_fill<int>, dim3(1), dim3(64), D->d_info, 2LL);
  // Side-lane ops (the lookahead's trailing update) may overlap the next
  // panel's POTRF/TRSM, which touch other columns; every other op joins.
  const bool la = !prof && g_lookahead && D->side;
  for (const Op& op : D->ops) {
    const int lane = la ? op.lane : 0;
    const bool join = !op.nojoin;
    if (prof) HIPCHECK(hipEventRecord(a0, E.main_st));
    switch (op.kind) {
      case TAMD_OP_ZERO:
        EMIT(lane, join, k_fill_ranges<T>, dim3(op.count), dim3(256),
             (const ZeroTask*)(D->d_zero + op.start), U);
        break;
      case TAMD_OP_ASSEMBLE:
        if (D->assemble_flat)
          EMIT(lane, join, k_assemble_flat<T>, dim3(grid_for(D->nnz, 256)), dim3(256),
               (long long)D->nnz, (const long long*)D->d_adest, (const double*)D->d_avals, L);
        else
          EMIT(lane, join, k_assemble<T>, dim3(grid_for(D->ncol, 128)), dim3(128), D->ncol,
               (const int*)D->d_colptr, (const long long*)D->d_adest, (const double*)D->d_avals, L);
        break;
      case TAMD_OP_SMALL:
        if (op.tile == 16)
          EMIT(lane, join, k_front_slot16<T>, dim3(op.count), dim3(64), (const FrontDesc*)D->d_fronts,
               (const SmallPack*)(D->d_packs + op.start), L, U, Iv, D->d_info, g_fp64_fallback);
        else if (op.tile == 32)
          EMIT(lane, join, k_front_slot32<T>, dim3(op.count), dim3(64), (const FrontDesc*)D->d_fronts,
               (const SmallPack*)(D->d_packs + op.start), L, U, Iv, D->d_info, g_fp64_fallback);
        else if (op.tile == 48)
          EMIT(lane, join, k_front_w48<T>, dim3(op.count), dim3(64), (const FrontDesc*)D->d_fronts,
               (const int*)(D->d_small + op.start), L, U, Iv, D->d_info, g_fp64_fallback);
        else if (g_potrf_wave)
          EMIT(lane, join, k_front_wave<T>, dim3(op.count), dim3(64), (const FrontDesc*)D->d_fronts,
               (const int*)(D->d_small + op.start), L, U, Iv, D->d_info, g_fp64_fallback);
        else
          EMIT(lane, join, k_front_small<T>, dim3(op.count), dim3(256), (const FrontDesc*)D->d_fronts,
               (const int*)(D->d_small + op.start), L, U, Iv, D->d_info);
        break;
      case TAMD_OP_POTRF:
        if (g_potrf_wave == 2)
          EMIT(lane, join, k_potrf_mw<T>, dim3(op.count), dim3(256), (const FrontDesc*)D->d_fronts,
               (const PanelTask*)(D->d_potrf + op.start), L, Iv, D->d_info, g_fp64_fallback);
        else if (g_potrf_wave)
          EMIT(lane, join, k_potrf_wave<T>, dim3(op.count), dim3(64), (const FrontDesc*)D->d_fronts,
               (const PanelTask*)(D->d_potrf + op.start), L, Iv, D->d_info, g_fp64_fallback);
        else
          EMIT(lane, join, k_potrf_diag<T>, dim3(op.count), dim3(256), (const FrontDesc*)D->d_fronts,
               (const PanelTask*)(D->d_potrf + op.start), L, Iv, D->d_info);
        break;
      case TAMD_OP_TRSM:
        EMIT(lane, join, k_panel_solve<T>, dim3(op.count), dim3(256),
             (const SolveTask*)(D->d_trsm + op.start), L, Iv, (const FrontDesc*)D->d_fronts,
             D->d_info, g_fp64_fallback);
        break;
      case TAMD_OP_GEMM:
        if (bf16upd)
          EMIT(lane, join, k_gemm_tile_bf16, dim3(op.count), dim3(256),
               (const GemmTask*)(D->d_gemm + op.start), (const float*)L, (float*)L, (float*)U);
        else if (op.tile == 128)
          EMIT(lane, join, k_gemm_tile128<T>, dim3(op.count), dim3(512),
               (const GemmTask*)(D->d_gemm + op.start), (const T*)L, L, U);
        // the latency-shaped tile (the same bits) where each workgroup has a
        // CU to itself: K <= 64, or a main-chain launch of at most
        // LAT_TILES_MAX tiles (the next outer panel's columns of the top
        // fronts: K = W in 4 round trips per 256 instead of 16)
        else if (lane == 0 && (op.shortk || op.count <= LAT_TILES_MAX))
          EMIT(lane, join, k_gemm_short<T>, dim3(op.count), dim3(256),
               (const GemmTask*)(D->d_gemm + op.start), (const T*)L, L, U, (const T*)Iv);
        else {
          // The lookahead's trailing update reserves LDS it does not use, so
          // that at most SIDE_WGS<T> of its workgroups share a CU: the panel
          // chain's POTRF, TRSM and inner updates then find room beside it
          // instead of waiting for the whole update to drain.  In FP64 only
          // while the update is short enough for the chain to be the longer
          // branch: a long update at half occupancy costs more than the
          // overlap saves (measured: mesh3d 60^3 root, 2.3K tiles, gains 5%;
          // 80^3 root, 20K tiles, loses 4%).
          if (lane == 1 && (sizeof(T) == 4 || op.count <= SIDE_PAD_MAX_TILES_FP64))
            E.shmem = side_lds_pad<T>();
          EMIT(lane, join, k_gemm_tile<T>, dim3(op.count), dim3(256),
               (const GemmTask*)(D->d_gemm + op.start), (const T*)L, L, U, (const T*)Iv);
        }
        break;
      case TAMD_OP_EXTEND:
        if (op.grouped)
          EMIT(lane, join, k_extend_add_seq<T>, dim3(op.count), dim3(256),
               (const FrontDesc*)D->d_fronts, (const EATask*)(D->d_ea + op.start),
               (const EATask*)D->d_ea, (const int*)D->d_rel, L, U);
        else
          EMIT(lane, join, k_extend_add<T>, dim3(op.count), dim3(256), (const FrontDesc*)D->d_fronts,
               (const EATask*)(D->d_ea + op.start), (const int*)D->d_rel, L, U);
        break;
    }
    if (prof) {
      float ms = 0;
      HIPCHECK(hipEventRecord(a1, E.main_st));
      HIPCHECK(hipEventSynchronize(a1));
      HIPCHECK(hipEventElapsedTime(&ms, a0, a1));
      D->op_ms[&op - D->ops.data()] = ms;
      if (op.kind == TAMD_OP_ASSEMBLE || op.kind == TAMD_OP_EXTEND || op.kind == TAMD_OP_ZERO)
        D->timing.assembly_ms += ms;
      else
        D->timing.front_ms += ms;
      if (op.kind == TAMD_OP_GEMM) {
        D->timing.gemm_ms += ms;
        D->timing.gemm_launches++;
      }
    }
  }
  if (E.finish() != 0) return -1;
  if (prof) {
    (void)hipEventDestroy(a0);
    (void)hipEventDestroy(a1);
  }
  D->timing.launches = E.count;
  return 0;
}

static Emitter eager_emitter(tamd_dev* D)
{
  Emitter E;
  E.main_st = D->stream;
  E.side_st = D->side;
  E.ev_fork = D->ev_fork;
  E.ev_join = D->ev_join;
  return E;
}

// Runs the whole numeric schedule.  Without profiling the schedule is built
// once into a hipGraph (kernel nodes only; the lookahead's two chains become
// graph branches) and replayed afterwards.
template <typename T>
static int run_numeric(tamd_dev* D, bool bf16upd)
{
  hipStream_t st = D->stream;
  const bool prof = g_profiling != 0;
  memset(&D->timing, 0, sizeof(D->timing));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (prof) {
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, st));
  }
  const int key = (int)sizeof(T) * 2 + (bf16upd ? 1 : 0) + (g_lookahead ? 16 : 0)
                  + 32 * g_potrf_wave + (g_fp64_fallback ? 128 : 0);
  if (prof || !g_graphs) {
    Emitter E = eager_emitter(D);
    if (launch_numeric<T>(D, E, bf16upd, prof) != 0) return -1;
  } else {
    if (!D->num_graph || D->num_graph_key != key) {
      if (D->num_graph) (void)hipGraphExecDestroy(D->num_graph);
      D->num_graph = nullptr;
      Emitter E;
      E.main_st = st;
      HIPCHECK(hipGraphCreate(&E.g, 0));
      int rc = launch_numeric<T>(D, E, bf16upd, false);
      hipError_t ie = rc == 0 ? hipGraphInstantiate(&D->num_graph, E.g, nullptr, nullptr, 0)
                              : hipErrorUnknown;
      (void)hipGraphDestroy(E.g);
      if (ie != hipSuccess) {
        D->num_graph = nullptr;
        taucs_printf((char*)"taucs_amd: building the numeric graph failed\n");
        return -1;
      }
      D->num_graph_key = key;
      D->num_graph_launches = E.count;
    }
    HIPCHECK(hipGraphLaunch(D->num_graph, st));
    D->timing.launches = D->num_graph_launches;
  }
  HIPCHECK(hipGetLastError());
  if (prof) {
    float ms = 0;
    HIPCHECK(hipEventRecord(e1, st));
    HIPCHECK(hipEventSynchronize(e1));
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    D->timing.total_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  D->timing.gemm_flops = D->gemm_flops;
  D->timing.assembly_bytes = D->assembly_bytes;
  int info[2] = {0, 0};
  HIPCHECK(hipMemcpyAsync(info, D->d_info, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
  HIPCHECK(hipStreamSynchronize(st));
  D->fp64_fallbacks = info[1];
  if (info[0]) {
    taucs_printf((char*)"\t\tLL^T Factorization: Matrix is not positive definite.\n");
    return -1;
  }
  return 0;
}

static int ensure_numeric_buffers(tamd_factor* F)
{
  tamd_dev* D = F->dev;
  int prec = F->precision;
  int elem = prec == TAMD_PREC_FP64 ? 8 : 4;
  if (D->prec != prec) {
    tamd_dev_free_numeric(F);
    if (dev_malloc(&D->L, (size_t)elem * (D->L_elems > 0 ? D->L_elems : 1)) != 0
        || dev_malloc(&D->U, (size_t)elem * (D->U_elems > 0 ? D->U_elems : 1)) != 0
        || dev_malloc(&D->Inv, (size_t)elem * (D->I_elems > 0 ? D->I_elems : 1)) != 0) {
      tamd_dev_free_numeric(F);
      return -1;
    }
    D->prec = prec;
    D->elem = elem;
  }
  return 0;
}

// Numeric factorization from the concatenated A values (host pointer).
extern "C" int tamd_dev_numeric_values(tamd_factor* F, const double* vals)
{
  tamd_dev* D = F->dev;
  if (!D || ensure_numeric_buffers(F) != 0) return -1;
  HIPCHECK(hipMemcpyAsync(D->d_avals, vals, sizeof(double) * (size_t)D->nnz, hipMemcpyHostToDevice,
                          D->stream));
  D->num_vals = vals;
  D->num_gen++;
  if (F->precision == TAMD_PREC_FP64) return run_numeric<double>(D, false);
  return run_numeric<float>(D, F->precision == TAMD_PREC_BF16UPD);
}

// The device copy of the values of the last numeric factorization, the host
// array they came from, and a counter that changes with every factorization.
extern "C" int tamd_dev_values_of(tamd_factor* F, const double** d_avals, const double** host_src,
                                  unsigned long long* gen)
{
  tamd_dev* D = F ? F->dev : nullptr;
  if (!D || !D->d_avals) return -1;
  *d_avals = D->d_avals;
  *host_src = D->num_vals;
  *gen = D->num_gen;
  return 0;
}

// Numeric factorization from values already in device memory (same order as
// the permuted matrix's stored entries; for a batch, the concatenation).
extern "C" int tamd_dev_numeric_device(tamd_factor* F, const double* d_vals)
{
  tamd_dev* D = F->dev;
  if (!D || ensure_numeric_buffers(F) != 0) return -1;
  HIPCHECK(hipMemcpyAsync(D->d_avals, d_vals, sizeof(double) * (size_t)D->nnz, hipMemcpyDeviceToDevice,
                          D->stream));
  D->num_vals = nullptr;  // no host array holds these values
  D->num_gen++;
  if (F->precision == TAMD_PREC_FP64) return run_numeric<double>(D, false);
  return run_numeric<float>(D, F->precision == TAMD_PREC_BF16UPD);
}

extern "C" int tamd_dev_numeric(tamd_factor* F, const taucs_ccs_matrix* A)
{
  return tamd_dev_numeric_values(F, A->values.d);
}

// Per-op record of the numeric schedule: kind (TAMD_OP_*), level, tasks,
// algorithmic flops and bytes, and the device time of the last profiled run.
extern "C" int taucs_amd_get_op_profile(void* vL, int cap, int* kind, int* level, int* count,
                                        double* flops, double* bytes, double* ms)
{
  tamd_factor* F = (tamd_factor*)vL;
  if (!F || !F->dev) return -1;
  tamd_dev* D = F->dev;
  const int n = (int)D->ops.size();
  for (int o = 0; o < n && o < cap; o++) {
    if (kind) kind[o] = D->ops[o].kind;
    if (level) level[o] = D->ops[o].level;
    if (count) count[o] = D->ops[o].count;
    if (flops) flops[o] = D->op_flops[o];
    if (bytes) bytes[o] = D->op_bytes[o];
    if (ms) ms[o] = o < (int)D->op_ms.size() ? D->op_ms[o] : 0.0;
  }
  return n;
}

extern "C" int taucs_amd_get_op_tiles(void* vL, int cap, int* tile, int* shortk)
{
  tamd_factor* F = (tamd_factor*)vL;
  if (!F || !F->dev) return -1;
  tamd_dev* D = F->dev;
  const int n = (int)D->ops.size();
  for (int o = 0; o < n && o < cap; o++) {
    const bool g = D->ops[o].kind == TAMD_OP_GEMM;
    if (tile) tile[o] = g ? D->ops[o].tile : 0;
    if (shortk) shortk[o] = g && D->ops[o].shortk ? 1 : 0;
  }
  return n;
}

extern "C" int taucs_amd_factor_fp64_fallbacks(void* vL)
{
  tamd_factor* F = (tamd_factor*)vL;
  if (!F || !F->dev) return -1;
  return F->dev->fp64_fallbacks;
}

extern "C" int taucs_amd_get_numeric_timing(void* vL, taucs_amd_numeric_timing* t)
{
  tamd_factor* F = (tamd_factor*)vL;
  if (!F || !F->dev || !t) return -1;
  *t = F->dev->timing;
  return 0;
}

// ---- solves ------------------------------------------------------------------

// The workspace of one sweep: front blocks W (W_per_rhs x nrhs), the forward
// solution Y (n x nrhs), split-K partial blocks, then the chained launches'
// tickets followed by the error word.
struct SolveBufs {
  void* W;
  void* Y;
  void* Part;
  int* sync;
};

template <typename T>
static int emit_solve(tamd_dev* D, Emitter& E, const SolveBufs& sb, int n, int nrhs, double* dX,
                      int ldx, const double* dB, int ldb)
{
  T* W = (T*)sb.W;
  T* Y = (T*)sb.Y;
  T* Part = (T*)sb.Part;
  const T* L = (const T*)D->L;
  const T* Iv = (const T*)D->Inv;
  const FrontDesc* fr = D->d_fronts;
  const int* rows = D->d_rows;
  int* tickets = sb.sync;
  int* err = sb.sync + D->sync_n;
  for (const Op& op : D->solve_chain ? D->sv_chain_ops : D->sv_ops) {
    const PanelTask* tk = D->d_sv_tasks + op.start;
    const dim3 grid(op.count);
    switch (op.kind) {
      case SV_INIT:
        EMIT(0, true, k_solve_init<T>, dim3((op.count + 3) / 4), dim3(256), fr, tk, op.count, rows,
             dB, ldb, W, nrhs);
        break;
      case SV_FSTEP:
        EMIT(0, true, k_fwd_step<T>, grid, dim3(256), fr, tk, rows, L, Iv, W, Y, n, nrhs);
        break;
      case SV_FRECT:
        EMIT(0, true, k_fwd_rect<T>, grid, dim3(256), fr, tk, rows, L, (const T*)Y, n, Part, nrhs);
        break;
      case SV_FRED:
        EMIT(0, true, k_fwd_rect_reduce<T>, grid, dim3(256), fr, tk, (const T*)Part, W, nrhs);
        break;
      case SV_EXT:
        EMIT(0, true, k_solve_ext<T>, grid, dim3(256), fr, tk, (const int*)D->d_rel, W, nrhs);
        break;
      case SV_GATHER:
        EMIT(0, true, k_bwd_gather<T>, grid, dim3(256), fr, tk, rows, (const double*)dX, ldx, W, nrhs);
        break;
      case SV_BCOLS:
        EMIT(0, true, k_bwd_cols<T>, grid, dim3(256), fr, tk, L, (const T*)W, Part, nrhs);
        break;
      case SV_BRED:
        EMIT(0, true, k_bwd_cols_reduce<T>, grid, dim3(256), fr, tk, rows, (const T*)Part, Y, n, nrhs);
        break;
      case SV_BSTEP:
        EMIT(0, true, k_bwd_step<T>, grid, dim3(256), fr, tk, rows, L, Iv, Y, n, dX, ldx, nrhs);
        break;
      case SV_CLEAR:
        EMIT(0, true, k_sweep_clear<T>, dim3(std::min(grid_for((long long)n * nrhs, 256), 4096u)),
             dim3(256), Y, dX, n, ldx, nrhs, tickets, (int)D->sync_n);
        break;
      case SV_FCHAIN:
        EMIT(0, true, k_fwd_chain<T>, grid, dim3(256), fr, tk, rows, L, Iv, W, Y, n,
             op.tile >= 0 ? tickets + op.tile : (int*)nullptr, err, nrhs);
        break;
      case SV_BCHAIN:
        EMIT(0, true, k_bwd_chain<T>, grid, dim3(256), fr, tk, rows, L, Iv, (const T*)Y, n,
             (const T*)Part, dX, ldx, op.tile >= 0 ? tickets + op.tile : (int*)nullptr, err, nrhs);
        break;
      case SV_FSMALL: {
        auto k = nrhs <= 1   ? k_fwd_small<T, 1>
                 : nrhs <= 4 ? k_fwd_small<T, 4>
                 : nrhs <= 8 ? k_fwd_small<T, 8>
                             : k_fwd_small<T, 16>;
        EMIT(0, true, k, dim3((op.count + 3) / 4), dim3(256), fr, tk, op.count, rows, L, Iv, W, Y, n,
             nrhs);
        break;
      }
      case SV_EXTSEQ:
        EMIT(0, true, k_solve_ext_seq<T>, grid, dim3(256), fr, tk, (const PanelTask*)D->d_sv_tasks,
             (const int*)D->d_rel, W, nrhs);
        break;
      case SV_BUPD:
        EMIT(0, true, k_bwd_upd<T>, grid, dim3(256), fr, tk, rows, L, (const double*)dX, ldx, Part, nrhs);
        break;
      case SV_BSMALL: {
        auto k = nrhs <= 1   ? k_bwd_small<T, 1>
                 : nrhs <= 4 ? k_bwd_small<T, 4>
                 : nrhs <= 8 ? k_bwd_small<T, 8>
                             : k_bwd_small<T, 16>;
        EMIT(0, true, k, dim3((op.count + 3) / 4), dim3(256), fr, tk, op.count, rows, L, Iv,
             (const T*)Y, n, dX, ldx, nrhs);
        break;
      }
    }
  }
  return 0;
}

template <typename T>
static int run_solve(tamd_factor* F, int nrhs, double* dX, int ldx, const double* dB, int ldb)
{
  tamd_dev* D = F->dev;
  hipStream_t st = D->stream;
  const int n = F->n;
  if (!D->W || D->W_rhs < nrhs) {
    free_ptr(D->W);
    free_ptr(D->Y);
    D->W = D->Y = nullptr;
    int cap = std::max(nrhs, 16);
    free_ptr(D->Part);
    D->Part = nullptr;
    drop_solve_graphs(D);
    if (dev_malloc(&D->W, sizeof(T) * (size_t)D->W_per_rhs * cap) != 0
        || dev_malloc(&D->Y, sizeof(T) * (size_t)n * cap) != 0
        || dev_malloc(&D->Part, sizeof(T) * 64 * 16 * (size_t)std::max(D->part_slots, 1LL)) != 0)
      return -1;
    D->W_rhs = cap;
  }
  if (!D->d_sync) {
    if (dev_malloc((void**)&D->d_sync, sizeof(int) * (size_t)(D->sync_n + 1)) != 0) return -1;
    HIPCHECK(hipMemsetAsync(D->d_sync, 0, sizeof(int) * (size_t)(D->sync_n + 1), st));
  }
  const SolveBufs sb = {D->W, D->Y, D->Part, D->d_sync};
  if (!g_graphs) {
    Emitter E = eager_emitter(D);
    if (emit_solve<T>(D, E, sb, n, nrhs, dX, ldx, dB, ldb) != 0) return -1;
    HIPCHECK(hipGetLastError());
    return 0;
  }
  for (auto& g : D->sv)
    if (g.exec && g.nrhs == nrhs && g.prec == D->prec && g.x == dX && g.b == dB && g.ldx == ldx
        && g.ldb == ldb) {
      g.used = ++D->sv_clock;
      HIPCHECK(hipGraphLaunch(g.exec, st));
      return 0;
    }
  tamd_dev::SolveGraph* slot = &D->sv[0];
  for (auto& g : D->sv)
    if (g.used < slot->used) slot = &g;
  if (slot->exec) (void)hipGraphExecDestroy(slot->exec);
  *slot = tamd_dev::SolveGraph();
  Emitter E;
  E.main_st = st;
  HIPCHECK(hipGraphCreate(&E.g, 0));
  int rc = emit_solve<T>(D, E, sb, n, nrhs, dX, ldx, dB, ldb);
  hipError_t ie = rc == 0 ? hipGraphInstantiate(&slot->exec, E.g, nullptr, nullptr, 0)
                          : hipErrorUnknown;
  (void)hipGraphDestroy(E.g);
  if (ie != hipSuccess) {
    slot->exec = nullptr;
    taucs_printf((char*)"taucs_amd: building the solve graph failed\n");
    return -1;
  }
  slot->nrhs = nrhs;
  slot->prec = D->prec;
  slot->x = dX;
  slot->b = dB;
  slot->ldx = ldx;
  slot->ldb = ldb;
  slot->used = ++D->sv_clock;
  HIPCHECK(hipGraphLaunch(slot->exec, st));
  return 0;
}

static int solve_device(tamd_factor* F, int nrhs, double* dX, int ldx, const double* dB, int ldb)
{
  // at most 16 right-hand sides per sweep (a front's workspace block is
  // m x nrhs at offset offW * nrhs)
  tamd_dev* D = F->dev;
  for (int r0 = 0; r0 < nrhs; r0 += 16) {
    int nr = std::min(16, nrhs - r0);
    int rc = (D->prec == TAMD_PREC_FP64)
                 ? run_solve<double>(F, nr, dX + (size_t)r0 * ldx, ldx, dB + (size_t)r0 * ldb, ldb)
                 : run_solve<float>(F, nr, dX + (size_t)r0 * ldx, ldx, dB + (size_t)r0 * ldb, ldb);
    if (rc) return rc;
  }
  return 0;
}

extern "C" int tamd_dev_stream_of(tamd_factor* F, hipStream_t* st)
{
  if (!F || !F->dev) return -1;
  *st = F->dev->stream;
  return 0;
}

// A chained sweep whose wait timed out leaves a mark in the error word (its
// results are not valid); read and cleared after the caller's synchronization.
extern "C" int tamd_dev_sweep_error(tamd_factor* F)
{
  tamd_dev* D = F ? F->dev : nullptr;
  if (!D || !D->d_sync) return 0;
  int e = 0;
  HIPCHECK(hipMemcpy(&e, D->d_sync + D->sync_n, sizeof(int), hipMemcpyDeviceToHost));
  if (!e) return 0;
  HIPCHECK(hipMemsetAsync(D->d_sync + D->sync_n, 0, sizeof(int), D->stream));
  HIPCHECK(hipStreamSynchronize(D->stream));
  taucs_printf((char*)"taucs_amd: a solve sweep timed out waiting for a block\n");
  return -1;
}

extern "C" int tamd_dev_solve_device(tamd_factor* F, int nrhs, double* dX, int ldx, const double* dB,
                                     int ldb)
{
  return solve_device(F, nrhs, dX, ldx, dB, ldb);
}

extern "C" int taucs_amd_supernodal_solve_many_device(void* vL, int nrhs, double* dX, int ldx,
                                                      const double* dB, int ldb)
{
  tamd_factor* F = (tamd_factor*)vL;
  if (!F || !F->dev || !F->has_numeric) return -1;
  int rc = solve_device(F, nrhs, dX, ldx, dB, ldb);
  if (rc) return rc;
  HIPCHECK(hipStreamSynchronize(F->dev->stream));
  return tamd_dev_sweep_error(F);
}

// Internal hooks of the allocation-free solve (noalloc.hip): the workspace
// dimensions of one sweep, and one sweep launched eagerly on the factor's
// stream with workspace the caller provides.
extern "C" int tamd_dev_solve_dims(tamd_factor* F, long long* W_per_rhs, long long* part_slots,
                                   long long* sync_n, int* elem)
{
  tamd_dev* D = F ? F->dev : nullptr;
  if (!D) return -1;
  *W_per_rhs = D->W_per_rhs;
  *part_slots = std::max(D->part_slots, 1LL);
  *sync_n = D->sync_n;
  *elem = D->elem;
  return 0;
}

extern "C" int tamd_dev_solve_eager(tamd_factor* F, void* W, void* Y, void* Part, int* sync, int nr,
                                    double* dX, int ldx, const double* dB, int ldb)
{
  tamd_dev* D = F ? F->dev : nullptr;
  if (!D || !D->L || nr <= 0 || nr > 16) return -1;
  const SolveBufs sb = {W, Y, Part, sync};
  Emitter E = eager_emitter(D);
  const int rc = D->prec == TAMD_PREC_FP64
                     ? emit_solve<double>(D, E, sb, F->n, nr, dX, ldx, dB, ldb)
                     : emit_solve<float>(D, E, sb, F->n, nr, dX, ldx, dB, ldb);
  if (rc != 0) return -1;
  HIPCHECK(hipGetLastError());
  return 0;
}

extern "C" int tamd_dev_solve(tamd_factor* F, int nrhs, double* X, int ldx, const double* B, int ldb)
{
  tamd_dev* D = F->dev;
  int n = F->n;
  if (!D || !D->L) return -1;
  if (!D->d_x || D->xb_rhs < nrhs) {
    free_ptr(D->d_x);
    free_ptr(D->d_b);
    D->d_x = D->d_b = nullptr;
    if (dev_malloc((void**)&D->d_x, sizeof(double) * (size_t)n * nrhs) != 0
        || dev_malloc((void**)&D->d_b, sizeof(double) * (size_t)n * nrhs) != 0)
      return -1;
    D->xb_rhs = nrhs;
  }
  HIPCHECK(hipMemcpy2DAsync(D->d_b, sizeof(double) * n, B, sizeof(double) * ldb, sizeof(double) * n,
                            nrhs, hipMemcpyHostToDevice, D->stream));
  if (solve_device(F, nrhs, D->d_x, n, D->d_b, n) != 0) return -1;
  HIPCHECK(hipMemcpy2DAsync(X, sizeof(double) * ldx, D->d_x, sizeof(double) * n, sizeof(double) * n,
                            nrhs, hipMemcpyDeviceToHost, D->stream));
  HIPCHECK(hipStreamSynchronize(D->stream));
  return tamd_dev_sweep_error(F);
}

extern "C" int tamd_dev_download_blocks(tamd_factor* F, double* Lhost)
{
  tamd_dev* D = F->dev;
  if (!D || !D->L) return -1;
  size_t n = (size_t)D->L_elems;
  if (D->prec == TAMD_PREC_FP64) {
    HIPCHECK(hipMemcpyAsync(Lhost, D->L, sizeof(double) * n, hipMemcpyDeviceToHost, D->stream));
    HIPCHECK(hipStreamSynchronize(D->stream));
  } else {
    std::vector<float> tmp(n);
    HIPCHECK(hipMemcpyAsync(tmp.data(), D->L, sizeof(float) * n, hipMemcpyDeviceToHost, D->stream));
    HIPCHECK(hipStreamSynchronize(D->stream));
    for (size_t i = 0; i < n; i++) Lhost[i] = tmp[i];
  }
  return 0;
}

// Digest of the factor's stored bits: the sum modulo 2^64 of a mix of every
// 32-bit word of the L arena with its position.  Integer sums do not depend
// on the reduction order, so equal factors give equal digests; a change of
// any bit changes the digest (up to 2^-64).  Used to compare factors of
// different schedules bit for bit without copying them out.
__global__ __launch_bounds__(256) void k_digest(const unsigned int* __restrict__ w, long long nw,
                                                unsigned long long* out)
{
  unsigned long long acc = 0;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < nw; i += (long long)gridDim.x * 256) {
    unsigned long long z = ((unsigned long long)(i + 1) << 32 | w[i]) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    acc += z ^ (z >> 31);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, acc);
}

extern "C" int taucs_amd_factor_digest(void* vL, unsigned long long* digest)
{
  tamd_factor* F = (tamd_factor*)vL;
  tamd_dev* D = F ? F->dev : nullptr;
  if (!D || !D->L || !F->has_numeric || !digest) return -1;
  unsigned long long* d_out = nullptr;
  if (dev_malloc((void**)&d_out, sizeof(unsigned long long)) != 0) return -1;
  const long long nw = D->L_elems * (long long)D->elem / 4;
  int rc = 0;
  if (hipMemsetAsync(d_out, 0, sizeof(unsigned long long), D->stream) != hipSuccess) rc = -1;
  if (!rc) {
    hipLaunchKernelGGL(k_digest, dim3(2048), dim3(256), 0, D->stream, (const unsigned int*)D->L, nw,
                       d_out);
    if (hipGetLastError() != hipSuccess
        || hipMemcpyAsync(digest, d_out, sizeof(unsigned long long), hipMemcpyDeviceToHost, D->stream)
               != hipSuccess
        || hipStreamSynchronize(D->stream) != hipSuccess)
      rc = -1;
  }
  tamd_dev_free_ptr(d_out);
  return rc;
}

// ---- timing for the schedule autotuner ---------------------------------------
// reps + 1 numeric factorizations from host values; the first (which builds
// the graph) is not counted; *best_ms = the fastest of the others, by device
// events around each replay on the factor's stream.
extern "C" int tamd_dev_time_numeric(tamd_factor* F, const double* vals, int reps, double* best_ms)
{
  double t = 1e300;
  if (!F->dev) return -1;
  for (int r = 0; r < reps + 1; r++) {
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, F->dev->stream));
    const int rc = tamd_dev_numeric_values(F, vals);
    HIPCHECK(hipEventRecord(e1, F->dev->stream));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != 0) return -1;
    if (r > 0 || reps == 0) t = std::min(t, (double)ms);
  }
  *best_ms = t;
  return 0;
}

// ---- device buffers for callers that keep right-hand sides in HBM ----------
extern "C" void* taucs_amd_device_malloc(size_t bytes)
{
  void* p = nullptr;
  if (dev_malloc(&p, bytes) != 0) return nullptr;
  return p;
}
extern "C" void taucs_amd_device_free(void* p) { free_ptr(p); }
extern "C" int taucs_amd_memcpy_h2d(void* dst, const void* src, size_t bytes)
{
  return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}
extern "C" int taucs_amd_memcpy_d2d(void* dst, const void* src, size_t bytes)
{
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) == hipSuccess ? 0 : -1;
}
extern "C" int taucs_amd_memcpy_d2h(void* dst, const void* src, size_t bytes)
{
  return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}
extern "C" int taucs_amd_device_synchronize(void)
{
  return hipDeviceSynchronize() == hipSuccess ? 0 : -1;
}
