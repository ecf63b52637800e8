/*
 * Internal host-side types of the MI355X multifrontal Cholesky.
 *
 * The factor object keeps the reference's opaque-handle contract
 * (src/taucs_sn_llt.c:603-623: `flags` first so generic entry points can
 * dispatch on the datatype), but its numeric blocks live in GPU memory and it
 * carries the level schedule the device numeric phase runs on.
 */
#ifndef TAMD_INTERNAL_H
#define TAMD_INTERNAL_H

#include <stdint.h>
#include "../../../include/taucs.h"

#ifdef __cplusplus
extern "C" {
#endif

#define tamd_malloc(x)    taucs_malloc_stub(x)
#define tamd_calloc(x, y) taucs_calloc_stub(x, y)
#define tamd_realloc(x, y) taucs_realloc_stub(x, y)
#define tamd_free(x)      taucs_free_stub(x)

int tamd_alloc_should_fail(void);

/* Precision of the numeric factor held on the device. */
enum {
  TAMD_PREC_FP64    = 0, /* plain FP64 fronts (the drop-in) */
  TAMD_PREC_FP32    = 1, /* FP32 fronts, FP32 MFMA updates */
  TAMD_PREC_BF16UPD = 2  /* FP32 fronts, BF16-input MFMA for trailing updates */
};

/* Result of the symbolic analysis (src/taucs_sn_llt.c:3663-4001).
 * Supernode slots that amalgamation merged away keep size 0 and are unlinked
 * from the tree, exactly as in the reference; n_sn is the slot count and the
 * virtual root is slot n_sn. */
typedef struct {
  int    n;
  int    n_sn;
  int*   sn_size;      /* [n_sn+1] */
  int*   sn_up_size;   /* [n_sn+1] */
  int**  sn_struct;    /* [n_sn]   first sn_size entries are the columns */
  int*   first_child;  /* [n_sn+1] */
  int*   next_child;   /* [n_sn+1] */
  int    etree_depth;
  double fund_nnz, fund_flops;    /* fundamental (pre-amalgamation) */
  double relax_nnz, relax_flops;  /* after amalgamation */
} tamd_symbolic;

/* Level schedule and HBM layout derived from the symbolic structure. */
typedef struct {
  int      n_sn;          /* slots (same as symbolic) */
  int      n_live;
  int      n_levels;
  int*     sn_parent;     /* [n_sn] parent slot, n_sn for tree roots, -1 dead */
  int*     sn_level;      /* [n_sn] height-based level, -1 dead */
  int*     order;         /* [n_live] live slots grouped by level */
  int*     level_ptr;     /* [n_levels+1] */
  int64_t* off_L;         /* [n_sn] element offset of the (s+u) x s panel */
  int64_t* off_U;         /* [n_sn] element offset of the u x u update block */
  int*     u_start;       /* [n_sn] level whose start zeroes the update block: the
                             lowest child level (children extend-add into it), else
                             the front's own level; it is dead after its own level */
  int64_t  L_elems;
  int64_t  U_elems;
  int**    drow;          /* [n_sn] device row order of each live front: its s columns as
                             in sn_struct, then its update rows sorted by their position in
                             the parent's device order (so extend-add indices increase) */
  int*     rel_ptr;       /* [n_sn+1] offsets into rel */
  int*     rel;           /* child update row i -> position in parent front */
  int*     sib_rank;      /* [n_sn] rank among same-level siblings */
  int      max_front;     /* max s+u */
  /* A-assembly map: PAPT entry k goes to L[a_dest[k]] */
  int      a_nnz;
  int64_t* a_dest;
  int*     a_colptr;      /* [n+1] column pointers of the pattern the map was built for */
  uint64_t a_pattern_hash;
} tamd_plan;

struct tamd_dev;

typedef struct tamd_factor {
  int            flags;      /* TAUCS_DOUBLE; must stay the first member */
  char           uplo;       /* 'l' */
  int            n;
  int            precision;  /* TAMD_PREC_* */
  tamd_symbolic  sym;
  tamd_plan      plan;
  struct tamd_dev* dev;      /* device-resident numeric state */
  int            has_numeric;
  double         t_symbolic, t_numeric; /* seconds */
  int            panel_w;    /* outer panel width of the schedule (0: library default) */
  int            gemm_tile;  /* tile edge of the large GEMMs, 64 or 128 (0: library default) */
} tamd_factor;

/* ---- host symbolic layer (symbolic.c, mmd.c) ---- */
int  tamd_mmd_order(int n, const int* colptr, const int* rowind, int* perm, int* invperm);
int  tamd_etree(const taucs_ccs_matrix* A, int* parent);
int  tamd_symbolic_analyze(const taucs_ccs_matrix* A, int max_depth, tamd_symbolic* S);
void tamd_symbolic_free(tamd_symbolic* S);
int  tamd_plan_build(const tamd_symbolic* S, const taucs_ccs_matrix* A, tamd_plan* P);
int  tamd_plan_amap(const tamd_symbolic* S, const taucs_ccs_matrix* A, tamd_plan* P);
void tamd_plan_free(tamd_plan* P);
uint64_t tamd_pattern_hash(const taucs_ccs_matrix* A);
int  tamd_dev_time_numeric(struct tamd_factor* F, const double* vals, int reps, double* best_ms);

/* ---- device numeric layer (the HIP sources under device/) ---- */
int  tamd_dev_available(void);
int  tamd_dev_setup(tamd_factor* F);
int  tamd_dev_numeric(tamd_factor* F, const taucs_ccs_matrix* A);
int  tamd_dev_numeric_device(tamd_factor* F, const double* d_vals);
int  tamd_dev_solve(tamd_factor* F, int nrhs, double* X, int ldx, const double* B, int ldb);
int  tamd_dev_download_blocks(tamd_factor* F, double* Lhost);
void tamd_dev_free_numeric(tamd_factor* F);
void tamd_dev_free(tamd_factor* F);

#ifdef __cplusplus
}
#endif
#endif
