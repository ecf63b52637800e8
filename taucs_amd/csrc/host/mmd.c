/*
 * Multiple minimum external degree ordering.
 *
 * The permutation must equal the one TAUCS obtains from Liu's GENMMD
 * (external/src/genmmd.c:56-236, called from src/taucs_ccs_order.c:262-442
 * with delta = 1, maxint = 2^31-1).  Minimum degree is decided by tie-breaks
 * -- which node heads a degree list, the order in which reachable sets are
 * walked, where absorbed element lists are recycled in the quotient graph --
 * so this implementation performs the same sequence of steps as Liu's
 * published algorithm (J.W.H. Liu, "Modification of the minimum-degree
 * algorithm by multiple elimination", ACM TOMS 11(2), 1985):
 *
 *   init   -- degree lists seeded in node order, each insert at the head
 *   elim   -- reachable set of the minimum-degree node built in place in its
 *             own adjacency, overflowing into absorbed elements' storage
 *   update -- external degrees recomputed per new element, two-neighbour
 *             nodes first (they expose indistinguishable pairs), then the rest
 *   number -- indistinguishable nodes numbered right after their
 *             representative, in increasing node index
 *
 * Node ids are 1-based inside this file so that 0 can serve as the list and
 * adjacency terminator and negative ids as links, as the algorithm requires.
 * Everything is iterative; no recursion.
 */
#include <limits.h>
#include <string.h>
#include "internal.h"

typedef struct {
  int  n;
  int* xadj;    /* [n+2], 1-based positions into adj */
  int* adj;     /* [nadj+1], 1-based */
  int* dhead;   /* [n+2]  degree -> first node of the list */
  int* dforw;   /* next in degree list; <0: -number (eliminated) or -rep (merged) */
  int* dbakw;   /* prev in list, or -deg at list head; 0: needs update; -MAXI: out */
  int* qsize;   /* supernode size, 0 for merged */
  int* llist;   /* scratch linked lists */
  int* marker;
  int  tag;
} mmd_state;

#define MAXI INT_MAX

static void mmd_reset_markers(mmd_state* s)
{
  int i;
  for (i = 1; i <= s->n; i++)
    if (s->marker[i] < MAXI) s->marker[i] = 0;
}

static void mmd_list_push(mmd_state* s, int node, int deg)
{
  int head = s->dhead[deg];
  s->dforw[node] = head;
  s->dbakw[node] = -deg;
  if (head > 0) s->dbakw[head] = node;
  s->dhead[deg] = node;
}

/* Eliminate `md` and rewrite the quotient graph (the "elim" step). */
static void mmd_eliminate(mmd_state* s, int md)
{
  int* adj = s->adj;
  int* xadj = s->xadj;
  int tag = s->tag;
  int i, j, elmnt, rloc, rlmt, istop;

  s->marker[md] = tag;
  istop = xadj[md + 1] - 1;
  elmnt = 0;
  rloc = xadj[md];
  rlmt = istop;

  /* Uneliminated neighbours stay in md's list; eliminated ones (elements)
     are chained on llist, most recent first. */
  for (i = xadj[md]; i <= istop; i++) {
    int nb = adj[i];
    if (nb == 0) break;
    if (s->marker[nb] >= tag) continue;
    s->marker[nb] = tag;
    if (s->dforw[nb] < 0) {
      s->llist[nb] = elmnt;
      elmnt = nb;
    } else {
      adj[rloc++] = nb;
    }
  }

  /* Merge in the nodes reachable through each adjacent element. */
  while (elmnt > 0) {
    int link = elmnt;
    int scanning = 1;
    adj[rlmt] = -elmnt;
    while (scanning) {
      int jstop = xadj[link + 1] - 1;
      scanning = 0;
      for (j = xadj[link]; j <= jstop; j++) {
        int node = adj[j];
        if (node < 0) { link = -node; scanning = 1; break; }
        if (node == 0) break;
        if (s->marker[node] >= tag || s->dforw[node] < 0) continue;
        s->marker[node] = tag;
        while (rloc >= rlmt) {          /* recycle absorbed storage */
          int nxt = -adj[rlmt];
          rloc = xadj[nxt];
          rlmt = xadj[nxt + 1] - 1;
        }
        adj[rloc++] = node;
      }
    }
    elmnt = s->llist[elmnt];
  }
  if (rloc <= rlmt) adj[rloc] = 0;

  /* Visit every reachable node. */
  {
    int link = md;
    int again = 1;
    while (again) {
      int stop = xadj[link + 1] - 1;
      again = 0;
      for (i = xadj[link]; i <= stop; i++) {
        int rn = adj[i];
        int pv, q, jstrt, jstop, nq;
        if (rn < 0) { link = -rn; again = 1; break; }
        if (rn == 0) break;

        /* take rn out of its degree list */
        pv = s->dbakw[rn];
        if (pv != 0 && pv != -MAXI) {
          int nx = s->dforw[rn];
          if (nx > 0) s->dbakw[nx] = pv;
          if (pv > 0) s->dforw[pv] = nx;
          if (pv < 0) s->dhead[-pv] = nx;
        }

        /* drop neighbours that became inactive */
        jstrt = xadj[rn];
        jstop = xadj[rn + 1] - 1;
        q = jstrt;
        for (j = jstrt; j <= jstop; j++) {
          int nb = adj[j];
          if (nb == 0) break;
          if (s->marker[nb] >= tag) continue;
          adj[q++] = nb;
        }
        nq = q - jstrt;
        if (nq <= 0) {
          /* no other neighbour: rn is absorbed into md (mass elimination) */
          s->qsize[md] += s->qsize[rn];
          s->qsize[rn] = 0;
          s->marker[rn] = MAXI;
          s->dforw[rn] = -md;
          s->dbakw[rn] = -MAXI;
        } else {
          /* flag for degree update; md becomes an element neighbour */
          s->dforw[rn] = nq + 1;
          s->dbakw[rn] = 0;
          adj[q++] = md;
          if (q <= jstop) adj[q] = 0;
        }
      }
    }
  }
}

/* Sum of qsize over the nodes of element `el` not yet marked with `tag`,
   marking them. */
static int mmd_element_weight(mmd_state* s, int el, int tag)
{
  int link = el, j, deg = 0, again = 1;
  while (again) {
    int stop = s->xadj[link + 1] - 1;
    again = 0;
    for (j = s->xadj[link]; j <= stop; j++) {
      int node = s->adj[j];
      if (node < 0) { link = -node; again = 1; break; }
      if (node == 0) return deg;
      if (s->marker[node] < tag) {
        s->marker[node] = tag;
        deg += s->qsize[node];
      }
    }
  }
  return deg;
}

/* Recompute external degrees after a round of multiple elimination. */
static void mmd_update(mmd_state* s, int ehead, int delta, int* mdeg)
{
  int* adj = s->adj;
  int* xadj = s->xadj;
  int mdeg0 = *mdeg + delta;
  int elmnt = ehead;

  while (elmnt > 0) {
    int mtag = s->tag + mdeg0;
    int q2head = 0, qxhead = 0, deg0 = 0;
    int link, i, pass;

    if (mtag >= MAXI) {
      s->tag = 1;
      mmd_reset_markers(s);
      mtag = s->tag + mdeg0;
    }

    /* Split the element's nodes into two-neighbour (q2) and other (qx)
       lists; deg0 is the element's total size. */
    link = elmnt;
    for (;;) {
      int stop = xadj[link + 1] - 1, more = 0;
      for (i = xadj[link]; i <= stop; i++) {
        int en = adj[i];
        if (en < 0) { link = -en; more = 1; break; }
        if (en == 0) break;
        if (s->qsize[en] == 0) continue;
        deg0 += s->qsize[en];
        s->marker[en] = mtag;
        if (s->dbakw[en] != 0) continue;
        if (s->dforw[en] == 2) { s->llist[en] = q2head; q2head = en; }
        else                   { s->llist[en] = qxhead; qxhead = en; }
      }
      if (!more) break;
    }

    for (pass = 0; pass < 2; pass++) {
      int en = pass == 0 ? q2head : qxhead;
      while (en > 0) {
        int deg;
        if (s->dbakw[en] != 0) { en = s->llist[en]; continue; }
        s->tag++;
        deg = deg0;

        if (pass == 0) {
          /* the one neighbour besides elmnt */
          int nb = adj[xadj[en]];
          if (nb == elmnt) nb = adj[xadj[en] + 1];
          if (s->dforw[nb] >= 0) {
            deg += s->qsize[nb];
          } else {
            /* walk the second element; nodes already carrying the
               element's mark are indistinguishable or outmatched */
            int lk = nb, again = 1, j;
            while (again) {
              int stop = xadj[lk + 1] - 1;
              again = 0;
              for (j = xadj[lk]; j <= stop; j++) {
                int node = adj[j];
                if (node == en) continue;
                if (node < 0) { lk = -node; again = 1; break; }
                if (node == 0) break;
                if (s->qsize[node] == 0) continue;
                if (s->marker[node] < s->tag) {
                  s->marker[node] = s->tag;
                  deg += s->qsize[node];
                  continue;
                }
                if (s->dbakw[node] != 0) continue;
                if (s->dforw[node] == 2) {
                  s->qsize[en] += s->qsize[node];
                  s->qsize[node] = 0;
                  s->marker[node] = MAXI;
                  s->dforw[node] = -en;
                  s->dbakw[node] = -MAXI;
                } else if (s->dbakw[node] == 0) {
                  s->dbakw[node] = -MAXI;
                }
              }
            }
          }
        } else {
          int stop = xadj[en + 1] - 1;
          for (i = xadj[en]; i <= stop; i++) {
            int nb = adj[i];
            if (nb == 0) break;
            if (s->marker[nb] >= s->tag) continue;
            s->marker[nb] = s->tag;
            if (s->dforw[nb] >= 0) deg += s->qsize[nb];
            else deg += mmd_element_weight(s, nb, s->tag);
          }
        }

        deg = deg - s->qsize[en] + 1;
        mmd_list_push(s, en, deg);
        if (deg < *mdeg) *mdeg = deg;
        en = s->llist[en];
      }
    }
    s->tag = mtag;
    elmnt = s->llist[elmnt];
  }
}

/* Final numbering: merged nodes follow their representative. */
static void mmd_number(mmd_state* s, int* perm1, int* invp1)
{
  int n = s->n, node;
  for (node = 1; node <= n; node++)
    perm1[node] = s->qsize[node] > 0 ? -invp1[node] : invp1[node];
  for (node = 1; node <= n; node++) {
    int root, f, num;
    if (perm1[node] > 0) continue;
    root = node;
    while (perm1[root] <= 0) root = -perm1[root];
    num = perm1[root] + 1;
    invp1[node] = -num;
    perm1[root] = num;
    f = node;
    for (;;) {
      int nf = -perm1[f];
      if (nf <= 0) break;
      perm1[f] = -root;
      f = nf;
    }
  }
  for (node = 1; node <= n; node++) {
    int num = -invp1[node];
    invp1[node] = num;
    perm1[num] = node;
  }
}

/*
 * Order the symmetric matrix whose lower triangle is (colptr, rowind).
 * The adjacency is assembled exactly as src/taucs_ccs_order.c:348-390 does:
 * columns scanned in stored order, each off-diagonal entry appended to both
 * endpoints.  perm[i] = old index at new position i; invperm its inverse.
 */
int tamd_mmd_order(int n, const int* colptr, const int* rowind, int* perm, int* invperm)
{
  mmd_state s;
  int nnz = colptr[n];
  int64_t nadj;
  int i, j, ip, num, delta = 1;
  int *len, *next, *dforw, *dbakw, *block;

  if (n <= 0) return 0;
  memset(&s, 0, sizeof(s));
  s.n = n;
  nadj = 2 * (int64_t)nnz - n;
  if (nadj < 1) nadj = 1;

  block = (int*)tamd_malloc(sizeof(int) * (size_t)(7 * (n + 4) + nadj + 2));
  if (!block) return -1;
  s.xadj   = block;
  s.dhead  = s.xadj + (n + 4);
  dforw    = s.dhead + (n + 4);
  dbakw    = dforw + (n + 4);
  s.qsize  = dbakw + (n + 4);
  s.llist  = s.qsize + (n + 4);
  s.marker = s.llist + (n + 4);
  s.adj    = s.marker + (n + 4);
  s.dforw = dforw;
  s.dbakw = dbakw;

  /* adjacency: len in dhead, fill pointer in qsize (both reset below) */
  len = s.dhead;
  next = s.qsize;
  for (i = 0; i <= n + 1; i++) len[i] = 0;
  for (j = 0; j < n; j++)
    for (ip = colptr[j]; ip < colptr[j + 1]; ip++) {
      i = rowind[ip];
      if (i != j) { len[i + 1]++; len[j + 1]++; }
    }
  s.xadj[1] = 1;
  for (i = 1; i <= n; i++) s.xadj[i + 1] = s.xadj[i] + len[i];
  for (i = 1; i <= n; i++) next[i] = s.xadj[i];
  for (j = 0; j < n; j++)
    for (ip = colptr[j]; ip < colptr[j + 1]; ip++) {
      i = rowind[ip];
      if (i != j) {
        s.adj[next[i + 1]++] = j + 1;
        s.adj[next[j + 1]++] = i + 1;
      }
    }

  /* init: every node in the list of its degree (+1), pushed in node order */
  for (i = 0; i < n + 4; i++) s.dhead[i] = 0;
  for (i = 1; i <= n; i++) { s.qsize[i] = 1; s.marker[i] = 0; s.llist[i] = 0; }
  for (i = 1; i <= n; i++) mmd_list_push(&s, i, s.xadj[i + 1] - s.xadj[i] + 1);

  num = 1;
  /* isolated nodes first, in list order */
  {
    int md = s.dhead[1];
    while (md > 0) {
      int nx = dforw[md];
      s.marker[md] = MAXI;
      dforw[md] = -num;
      num++;
      md = nx;
    }
  }

  if (num <= n) {
    int mdeg = 2;
    s.tag = 1;
    s.dhead[1] = 0;
    for (;;) {
      int mdlmt, ehead = 0, finished = 0;
      while (mdeg <= n && s.dhead[mdeg] <= 0) mdeg++;
      if (mdeg > n) { tamd_free(block); return -1; } /* corrupt lists */
      mdlmt = mdeg + delta;
      for (;;) {
        int md = s.dhead[mdeg], nx;
        if (md <= 0) {
          mdeg++;
          if (mdeg > mdlmt) break;
          continue;
        }
        nx = dforw[md];
        s.dhead[mdeg] = nx;
        if (nx > 0) dbakw[nx] = -mdeg;
        dforw[md] = -num;
        if (num + s.qsize[md] > n) { finished = 1; break; }
        s.tag++;
        if (s.tag >= MAXI) { s.tag = 1; mmd_reset_markers(&s); }
        mmd_eliminate(&s, md);
        num += s.qsize[md];
        s.llist[md] = ehead;
        ehead = md;
      }
      if (finished || num > n) break;
      mmd_update(&s, ehead, delta, &mdeg);
    }
  }

  /* number: dbakw becomes perm (1-based), dforw the inverse */
  mmd_number(&s, dbakw, dforw);
  for (i = 0; i < n; i++) perm[i] = dbakw[i + 1] - 1;
  for (i = 0; i < n; i++) invperm[perm[i]] = i;
  tamd_free(block);
  return 0;
}
