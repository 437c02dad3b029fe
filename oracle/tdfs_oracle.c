/*
 * tdfs_oracle.c -- CPU restatement of the T-DFS counting path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (tdfs_amd/libtdfs.so) never links or calls it.
 *
 * It restates, in plain C, the reference algorithm of lyuheng/tdfs:
 *   - pattern parsing           src/pattern.h:110-143
 *   - Aut(P), multiplicity      src/pattern.h:279-310 (brute force over k!)
 *     (the reference obtains the same group from bliss + closure,
 *      src/pattern.h:637-663; the output below is a set function of the group)
 *   - symmetry-breaking pairs   src/pattern.h:665-719 (GetConditions rule)
 *   - sym01 (partial_ori[0][0]) src/pattern.h:312-340
 *   - backward neighbours       src/pattern.h:342-358
 *   - result-reuse levels       src/pattern.h:370-413 and the unlabeled veto
 *                               src/pattern.h:459-480
 *   - condition list            src/pattern.h:416-448
 *   - level-0 edge filter       src/gpu_match.cu:526-558
 *   - candidate sets            src/gpu_match.cu:260-494, 571-779
 *   - count and multiplicity    cu_test.cu:100-107
 *   - binary CSR format         src/graph.h:136-193
 *
 * Counting semantics (raw count, before multiplicity):
 *   for every directed edge (r, c) passing the level-0 filter, M(0)=r, M(1)=c;
 *   for u = 2..k-1: C(u) = AND_{b in BN(u)} N(M(b)) AND F(u);
 *   recurse on each x in C(u) for u < k-1; add |C(k-1)| at the leaf.
 *
 * Parity pins (all in tests/): closed forms (trace(A^3), K_n falling
 * factorials, stars, C4), a brute-force injective-map counter in this file
 * (orc_bruteforce_count) that shares no code with the DFS, networkx-generated
 * fixtures, bliss 0.73 built from the reference sources (oracle/_ref), and
 * invariants (reuse on == reuse off; raw x |Aut| == #monomorphisms).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MAXK 7

enum { ORC_LT = 0, ORC_GT = 1, ORC_NE = 2 }; /* src/pattern.h:21 */

typedef struct {
  int32_t k;
  int32_t labeled;
  int32_t adj[ORC_MAXK][ORC_MAXK];
  int32_t vlabel[ORC_MAXK];
  int32_t deg[ORC_MAXK];
  int32_t num_bn[ORC_MAXK];
  int32_t bn[ORC_MAXK][ORC_MAXK];
  int32_t cond_op[ORC_MAXK][ORC_MAXK]; /* [u][j], j < u */
  int32_t shared_lvl[ORC_MAXK];
  int32_t num_bn_sh[ORC_MAXK];
  int32_t bn_sh[ORC_MAXK][ORC_MAXK];
  int32_t sym01;
  int32_t npairs;
  int32_t pairs[64][2];
  int64_t mult;
} orc_plan;

/* ------------------------------------------------------------------ */
/* .g parsing: src/pattern.h:110-143.                                   */
/* Returns 0 on success, negative on malformed input.                  */
/* ------------------------------------------------------------------ */
int orc_parse_pattern(const char *text, int labeled, orc_plan *p) {
  memset(p, 0, sizeof(*p));
  p->labeled = labeled ? 1 : 0;
  const char *s = text;
  int k = 0;
  while (*s) {
    const char *eol = strchr(s, '\n');
    size_t len = eol ? (size_t)(eol - s) : strlen(s);
    char line[256];
    if (len >= sizeof(line)) return -1;
    memcpy(line, s, len);
    line[len] = 0;
    s = eol ? eol + 1 : s + len;
    if (line[0] == 'v') {
      int id, lab;
      if (sscanf(line + 1, "%d %d", &id, &lab) != 2) return -1;
      if (id != k || k >= ORC_MAXK) return -2;
      p->vlabel[k] = labeled ? lab : 1; /* pattern.h:124-129 */
      k++;
    } else if (line[0] == 'e') {
      int a, b, lab;
      if (sscanf(line + 1, "%d %d %d", &a, &b, &lab) != 3) return -1;
      if (a < 0 || b < 0 || a >= k || b >= k || a == b) return -3;
      if (lab != 1) return -4; /* pattern.h:296 only recognises label 1 */
      p->adj[a][b] = p->adj[b][a] = 1;
    }
  }
  p->k = k;
  if (k < 3) return -5;
  return 0;
}

/* ------------------------------------------------------------------ */
/* Planner                                                             */
/* ------------------------------------------------------------------ */
static int next_perm(int *a, int n) { /* lexicographic next permutation */
  int i = n - 2;
  while (i >= 0 && a[i] >= a[i + 1]) i--;
  if (i < 0) return 0;
  int j = n - 1;
  while (a[j] <= a[i]) j--;
  int t = a[i]; a[i] = a[j]; a[j] = t;
  for (int l = i + 1, r = n - 1; l < r; l++, r--) { t = a[l]; a[l] = a[r]; a[r] = t; }
  return 1;
}

/* Every automorphism of the (unlabeled, uncoloured) pattern graph.
 * src/pattern.h:279-310; bliss is also given no vertex colours
 * (src/pattern.h:581-594). */
static int all_automorphisms(const orc_plan *p, int (*out)[ORC_MAXK]) {
  int k = p->k, a[ORC_MAXK], n = 0;
  for (int i = 0; i < k; i++) a[i] = i;
  do {
    int ok = 1;
    for (int i = 0; i < k && ok; i++)
      for (int j = 0; j < k; j++)
        if (p->adj[i][j] != p->adj[a[i]][a[j]]) { ok = 0; break; }
    if (ok) { memcpy(out[n], a, sizeof(int) * ORC_MAXK); n++; }
  } while (next_perm(a, k));
  return n;
}

/* GetConditions: src/pattern.h:665-707.  Orbits are keyed by their minimum;
 * the first orbit (smallest representative) with more than one element gives
 * n0; for every permutation the member of orbit\{n0} with the smallest image
 * is paired with n0; then the group is restricted to the stabiliser of n0. */
static void get_conditions(orc_plan *p, int (*aut)[ORC_MAXK], int naut) {
  int k = p->k;
  int live[5040];
  for (int i = 0; i < naut; i++) live[i] = 1;
  int npairs = 0;
  int pairs[5040 * 2][2];
  for (;;) {
    /* orbit bitmasks over the live permutations */
    unsigned orb[ORC_MAXK];
    for (int v = 0; v < k; v++) {
      orb[v] = 0;
      for (int i = 0; i < naut; i++) if (live[i]) orb[v] |= 1u << aut[i][v];
    }
    /* first orbit (by smallest representative) of size > 1 */
    int n0 = -1;
    unsigned cls = 0;
    for (int rep = 0; rep < k && n0 < 0; rep++) {
      /* rep must be the minimum of its own orbit */
      unsigned o = orb[rep];
      if (o == 0) continue;
      int mn = __builtin_ctz(o);
      if (mn != rep) continue;
      if (__builtin_popcount(o) > 1) { n0 = rep; cls = o; }
    }
    if (n0 < 0) break;
    for (int i = 0; i < naut; i++) {
      if (!live[i]) continue;
      int best = -1;
      for (int x = 0; x < k; x++) {
        if (!(cls >> x & 1u) || x == n0) continue;
        if (best < 0 || aut[i][x] < aut[i][best]) best = x;
      }
      pairs[npairs][0] = n0;
      pairs[npairs][1] = best;
      npairs++;
    }
    for (int i = 0; i < naut; i++)
      if (live[i] && aut[i][n0] != n0) live[i] = 0;
  }
  /* sort + unique (src/pattern.h:703-704) */
  for (int i = 1; i < npairs; i++) {
    int a0 = pairs[i][0], a1 = pairs[i][1], j = i - 1;
    while (j >= 0 && (pairs[j][0] > a0 || (pairs[j][0] == a0 && pairs[j][1] > a1))) {
      pairs[j + 1][0] = pairs[j][0]; pairs[j + 1][1] = pairs[j][1]; j--;
    }
    pairs[j + 1][0] = a0; pairs[j + 1][1] = a1;
  }
  p->npairs = 0;
  for (int i = 0; i < npairs; i++) {
    if (p->npairs > 0 && p->pairs[p->npairs - 1][0] == pairs[i][0] &&
        p->pairs[p->npairs - 1][1] == pairs[i][1])
      continue;
    p->pairs[p->npairs][0] = pairs[i][0];
    p->pairs[p->npairs][1] = pairs[i][1];
    p->npairs++;
  }
}

/* Full plan: src/pattern.h:90-101 minus the dead STMatch slot code. */
int orc_plan_build(orc_plan *p) {
  int k = p->k;
  if (k < 3 || k > ORC_MAXK) return -5;
  static __thread int aut[5040][ORC_MAXK];
  int naut = all_automorphisms(p, aut);
  p->mult = naut;
  get_conditions(p, aut, naut);

  /* pattern degrees: src/pattern.h:189-195 */
  for (int i = 0; i < k; i++) {
    p->deg[i] = 0;
    for (int j = 0; j < k; j++) p->deg[i] += p->adj[i][j] > 0;
  }
  /* sym01 = 1 in orbit(0): src/pattern.h:312-340 */
  p->sym01 = 0;
  for (int i = 0; i < naut; i++) if (aut[i][0] == 1) p->sym01 = 1;

  /* backward neighbours, ascending: src/pattern.h:342-358 */
  for (int i = 0; i < k; i++) {
    p->num_bn[i] = 0;
    if (i == 0) continue;
    for (int j = 0; j < i; j++)
      if (p->adj[i][j]) p->bn[i][p->num_bn[i]++] = j;
  }
  /* result reuse: src/pattern.h:370-413 */
  int nbr_bits[ORC_MAXK] = {0};
  p->shared_lvl[0] = p->shared_lvl[1] = -1;
  for (int i = 1; i < k; i++) {
    nbr_bits[i] = 0;
    for (int j = 0; j < i; j++) if (p->adj[j][i]) nbr_bits[i] |= 1 << j;
    if (i >= 2) p->shared_lvl[i] = -1;
    for (int j = i - 1; j >= 1; j--) {
      if ((nbr_bits[i] & nbr_bits[j]) == nbr_bits[j] && p->num_bn[i] >= 2 &&
          p->num_bn[j] >= 2 && p->vlabel[i] == p->vlabel[j]) {
        p->shared_lvl[i] = j;
        break;
      }
      p->shared_lvl[i] = -1;
    }
  }
  for (int i = 0; i < k; i++) {
    p->num_bn_sh[i] = 0;
    int dep = p->shared_lvl[i];
    if (dep == -1) continue;
    for (int j = i - 1; j >= 0; j--)
      if ((nbr_bits[i] >> j & 1) && !(nbr_bits[dep] >> j & 1)) p->bn_sh[i][p->num_bn_sh[i]++] = j;
  }
  /* condition list: src/pattern.h:416-448 (order_ from SetConditions) */
  for (int u = 0; u < k; u++)
    for (int j = 0; j < k; j++) {
      p->cond_op[u][j] = ORC_NE;
      if (j >= u) continue;
      for (int q = 0; q < p->npairs; q++) {
        if (p->pairs[q][0] == u && p->pairs[q][1] == j) { p->cond_op[u][j] = ORC_LT; break; }
        if (p->pairs[q][1] == u && p->pairs[q][0] == j) { p->cond_op[u][j] = ORC_GT; break; }
      }
    }
  /* unlabeled reuse veto: src/pattern.h:459-480 */
  if (!p->labeled)
    for (int i = 0; i < k; i++) {
      int dep = p->shared_lvl[i];
      if (dep == -1) continue;
      for (int j = 0; j < dep; j++)
        if (p->cond_op[dep][j] != p->cond_op[i][j]) { p->shared_lvl[i] = -1; break; }
    }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Graph (borrowed pointers)                                           */
/* ------------------------------------------------------------------ */
typedef struct {
  int32_t n;
  int64_t nnz;
  const int64_t *rowptr;
  const int32_t *colidx;
  const int32_t *labels; /* may be NULL (unlabeled) */
  int32_t *src;          /* COO row of each directed edge (owned) */
  int32_t *maxnbrdeg;    /* max_{v in N(c)} deg(v) (owned) */
  int32_t dmax;
} orc_graph;

static void graph_prepare(orc_graph *g) {
  g->src = (int32_t *)malloc(sizeof(int32_t) * (g->nnz > 0 ? g->nnz : 1));
  g->maxnbrdeg = (int32_t *)malloc(sizeof(int32_t) * (g->n > 0 ? g->n : 1));
  g->dmax = 0;
  for (int32_t v = 0; v < g->n; v++) {
    int64_t d = g->rowptr[v + 1] - g->rowptr[v];
    if (d > g->dmax) g->dmax = (int32_t)d;
    int32_t m = -1;
    for (int64_t e = g->rowptr[v]; e < g->rowptr[v + 1]; e++) {
      g->src[e] = v;
      int32_t w = g->colidx[e];
      int32_t dw = (int32_t)(g->rowptr[w + 1] - g->rowptr[w]);
      if (dw > m) m = dw;
    }
    g->maxnbrdeg[v] = m;
  }
}

static void graph_release(orc_graph *g) {
  free(g->src);
  free(g->maxnbrdeg);
}

/* ------------------------------------------------------------------ */
/* DFS                                                                 */
/* ------------------------------------------------------------------ */
typedef struct {
  const orc_graph *g;
  const orc_plan *p;
  int reuse;
  int32_t *cand[ORC_MAXK]; /* cand[u]: C(u) (full list, kept for reuse) */
  int32_t csize[ORC_MAXK];
  int32_t M[ORC_MAXK];
  uint64_t count;
  uint64_t balg;
  uint64_t nodes[ORC_MAXK]; /* extensions performed per pattern vertex u */
  int32_t *emit;            /* enumeration: k ids per embedding (NULL: count only) */
  uint64_t emit_cap, emitted;
} orc_worker;

static inline int64_t deg_of(const orc_graph *g, int32_t v) { return g->rowptr[v + 1] - g->rowptr[v]; }

static inline int contains(const int32_t *a, int64_t n, int32_t x) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < n && a[lo] == x;
}

/* F(u): src/gpu_match.cu:277-326 */
static inline int filter_ok(const orc_worker *w, int u, int32_t x) {
  const orc_plan *p = w->p;
  if (!p->labeled) {
    for (int j = 0; j < u; j++) {
      int32_t mj = w->M[j];
      switch (p->cond_op[u][j]) {
        case ORC_LT: if (!(x < mj)) return 0; break;
        case ORC_GT: if (!(x > mj)) return 0; break;
        default: if (x == mj) return 0; break;
      }
    }
    return 1;
  }
  if (w->g->labels[x] != p->vlabel[u]) return 0;
  for (int j = 0; j < u; j++) if (w->M[j] == x) return 0;
  return 1;
}

/* Builds C(u) given M(0..u-1); returns |C(u)|.  Writes the list only when
 * u < k-1 (the leaf only counts: src/gpu_match.cu:331,409,487). */
static int64_t extend(orc_worker *w, int u) {
  const orc_plan *p = w->p;
  const orc_graph *g = w->g;
  int leaf = (u == p->k - 1);
  int dep = w->reuse ? p->shared_lvl[u] : -1;
  const int32_t *lists[ORC_MAXK];
  int64_t sizes[ORC_MAXK];
  int nl = 0;
  uint64_t bytes = 0;
  if (dep != -1) { /* result reuse: src/gpu_match.cu:733-776 */
    lists[nl] = w->cand[dep];
    sizes[nl] = w->csize[dep];
    bytes += 4ull * (uint64_t)w->csize[dep];
    nl++;
    for (int i = 0; i < p->num_bn_sh[u]; i++) {
      int32_t t = w->M[p->bn_sh[u][i]];
      lists[nl] = g->colidx + g->rowptr[t];
      sizes[nl] = deg_of(g, t);
      bytes += 4ull * (uint64_t)sizes[nl] + 16;
      nl++;
    }
  } else { /* src/gpu_match.cu:635-731 */
    for (int i = 0; i < p->num_bn[u]; i++) {
      int32_t t = w->M[p->bn[u][i]];
      lists[nl] = g->colidx + g->rowptr[t];
      sizes[nl] = deg_of(g, t);
      bytes += 4ull * (uint64_t)sizes[nl] + 16;
      nl++;
    }
  }
  /* probe = smallest list (set semantics: the choice does not change C(u)) */
  int pi = 0;
  for (int i = 1; i < nl; i++) if (sizes[i] < sizes[pi]) pi = i;
  if (p->labeled) bytes += (uint64_t)sizes[pi];
  int64_t n = 0;
  int32_t *out = w->cand[u];
  for (int64_t e = 0; e < sizes[pi]; e++) {
    int32_t x = lists[pi][e];
    if (!filter_ok(w, u, x)) continue;
    int ok = 1;
    for (int i = 0; i < nl && ok; i++) if (i != pi) ok = contains(lists[i], sizes[i], x);
    if (!ok) continue;
    if (!leaf) out[n] = x;
    if (leaf && w->emit) { /* embedding (M(0..k-2), x), in plan order */
      if (w->emitted < w->emit_cap) {
        int32_t *e = w->emit + w->emitted * (uint64_t)p->k;
        for (int j = 0; j < u; j++) e[j] = w->M[j];
        e[u] = x;
      }
      w->emitted++;
    }
    n++;
  }
  if (!leaf) bytes += 4ull * (uint64_t)n;
  w->balg += bytes;
  w->nodes[u]++;
  return n;
}

static void dfs(orc_worker *w, int u) {
  int64_t n = extend(w, u);
  if (u == w->p->k - 1) { w->count += (uint64_t)n; return; }
  w->csize[u] = (int32_t)n;
  for (int64_t i = 0; i < n; i++) {
    w->M[u] = w->cand[u][i];
    dfs(w, u + 1);
  }
}

/* Level-0 filter for directed edge e: src/gpu_match.cu:526-558.
 * "exists v in N(c) with deg(v) >= dP(2)" is max_{v in N(c)} deg(v) >= dP(2). */
static int edge_ok(const orc_worker *w, int64_t e) {
  const orc_graph *g = w->g;
  const orc_plan *p = w->p;
  int32_t r = g->src[e], c = g->colidx[e];
  if (!p->labeled) {
    if (p->sym01 && !(r < c)) return 0;
  } else {
    if (g->labels[r] != p->vlabel[0] || g->labels[c] != p->vlabel[1]) return 0;
  }
  if (deg_of(g, r) < p->deg[0] || deg_of(g, c) < p->deg[1]) return 0;
  return g->maxnbrdeg[c] >= p->deg[2];
}

static int edge_pre_ok(const orc_worker *w, int64_t e) { /* order/label part only */
  const orc_graph *g = w->g;
  const orc_plan *p = w->p;
  int32_t r = g->src[e], c = g->colidx[e];
  if (!p->labeled) return !(p->sym01 && !(r < c));
  return g->labels[r] == p->vlabel[0] && g->labels[c] == p->vlabel[1];
}

typedef struct {
  orc_worker w;
  const int64_t *lo, *hi, *pre; /* ranges and prefix sums of their lengths */
  int nranges;
  int64_t total, chunk;
  int64_t *cursor;
  pthread_mutex_t *mu;
  pthread_t th;
} orc_thread;

static void run_edge(orc_worker *w, int64_t e) {
  w->balg += 8; /* colidx + row id */
  if (!edge_pre_ok(w, e)) return;
  w->balg += 16; /* rowptr pair reads of the degree prune */
  if (!edge_ok(w, e)) return;
  w->nodes[1]++;
  w->M[0] = w->g->src[e];
  w->M[1] = w->g->colidx[e];
  dfs(w, 2);
}

static void *thread_main(void *arg) {
  orc_thread *t = (orc_thread *)arg;
  orc_worker *w = &t->w;
  for (;;) {
    int64_t s;
    pthread_mutex_lock(t->mu);
    s = *t->cursor;
    *t->cursor += t->chunk;
    pthread_mutex_unlock(t->mu);
    if (s >= t->total) break;
    int64_t s1 = s + t->chunk < t->total ? s + t->chunk : t->total;
    /* virtual positions [s, s1) -> (range, offset) */
    int r = 0;
    while (r + 1 < t->nranges && t->pre[r + 1] <= s) r++;
    for (int64_t v = s; v < s1; v++) {
      while (v >= t->pre[r + 1]) r++;
      run_edge(w, t->lo[r] + (v - t->pre[r]));
    }
  }
  return NULL;
}

/* Counts over the union of directed-edge ranges [lo[i], hi[i]) of the CSR
 * (ranges clipped to [0, nnz); overlapping ranges are counted twice).
 * out[0] = raw count (pre-multiplicity, cu_test.cu:103)
 * out[1] = algorithmic bytes
 * out[2+u], u = 0..6: extensions computing C(u) (u = 1: accepted level-0 edges)
 * Returns 0, or a negative code. */
/* A graph prepared once (row ids, max neighbour degree) for many counts. */
void *orc_prepare(int32_t n, int64_t nnz, const int64_t *rowptr, const int32_t *colidx,
                  const int32_t *labels) {
  orc_graph *g = (orc_graph *)calloc(1, sizeof(orc_graph));
  g->n = n;
  g->nnz = nnz;
  g->rowptr = rowptr;
  g->colidx = colidx;
  g->labels = labels;
  graph_prepare(g);
  return g;
}

void orc_release(void *h) {
  if (!h) return;
  graph_release((orc_graph *)h);
  free(h);
}

static int count_prepared(const orc_graph *gp, const orc_plan *p, const int64_t *lo_in,
                          const int64_t *hi_in, int nranges, int nthreads, int reuse, uint64_t *out);

int orc_count_ranges(int32_t n, int64_t nnz, const int64_t *rowptr, const int32_t *colidx,
                     const int32_t *labels, const orc_plan *p, const int64_t *lo_in,
                     const int64_t *hi_in, int nranges, int nthreads, int reuse, uint64_t *out) {
  if (p->labeled && !labels) return -1;
  if (nranges < 1) return -2;
  orc_graph *g = (orc_graph *)orc_prepare(n, nnz, rowptr, colidx, labels);
  int rc = count_prepared(g, p, lo_in, hi_in, nranges, nthreads, reuse, out);
  orc_release(g);
  return rc;
}

int orc_count_prepared(const void *h, const orc_plan *p, const int64_t *lo_in, const int64_t *hi_in,
                       int nranges, int nthreads, int reuse, uint64_t *out) {
  const orc_graph *g = (const orc_graph *)h;
  if (p->labeled && !g->labels) return -1;
  if (nranges < 1) return -2;
  return count_prepared(g, p, lo_in, hi_in, nranges, nthreads, reuse, out);
}

static int count_prepared(const orc_graph *gp, const orc_plan *p, const int64_t *lo_in,
                          const int64_t *hi_in, int nranges, int nthreads, int reuse, uint64_t *out) {
  const orc_graph g = *gp;
  const int64_t nnz = g.nnz;
  int64_t *lo = (int64_t *)malloc(sizeof(int64_t) * (size_t)nranges);
  int64_t *hi = (int64_t *)malloc(sizeof(int64_t) * (size_t)nranges);
  int64_t *pre = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nranges + 1));
  pre[0] = 0;
  for (int i = 0; i < nranges; i++) {
    int64_t a = lo_in[i] < 0 ? 0 : lo_in[i];
    int64_t b = (hi_in[i] < 0 || hi_in[i] > nnz) ? nnz : hi_in[i];
    if (a > b) a = b;
    lo[i] = a;
    hi[i] = b;
    pre[i + 1] = pre[i] + (b - a);
  }
  if (nthreads < 1) nthreads = 1;
  orc_thread *ts = (orc_thread *)calloc((size_t)nthreads, sizeof(orc_thread));
  int64_t cursor = 0;
  pthread_mutex_t mu = PTHREAD_MUTEX_INITIALIZER;
  for (int i = 0; i < nthreads; i++) {
    orc_thread *t = &ts[i];
    t->w.g = &g;
    t->w.p = p;
    t->w.reuse = reuse;
    for (int u = 0; u < ORC_MAXK; u++)
      t->w.cand[u] = (int32_t *)malloc(sizeof(int32_t) * (size_t)(g.dmax > 0 ? g.dmax : 1));
    t->lo = lo;
    t->hi = hi;
    t->pre = pre;
    t->nranges = nranges;
    t->total = pre[nranges];
    /* small chunks when the union is short next to the thread count (hub
       edges make single chunks slow: fine chunks keep the threads balanced) */
    t->chunk = pre[nranges] / ((int64_t)nthreads * 64);
    if (t->chunk < 1) t->chunk = 1;
    if (t->chunk > 64) t->chunk = 64;
    t->cursor = &cursor;
    t->mu = &mu;
    if (nthreads > 1) pthread_create(&t->th, NULL, thread_main, t);
  }
  if (nthreads == 1) thread_main(&ts[0]);
  memset(out, 0, sizeof(uint64_t) * 9);
  for (int i = 0; i < nthreads; i++) {
    orc_thread *t = &ts[i];
    if (nthreads > 1) pthread_join(t->th, NULL);
    out[0] += t->w.count;
    out[1] += t->w.balg;
    for (int u = 0; u < ORC_MAXK; u++) out[2 + u] += t->w.nodes[u];
    for (int u = 0; u < ORC_MAXK; u++) free(t->w.cand[u]);
  }
  free(ts);
  free(lo);
  free(hi);
  free(pre);
  return 0;
}

int orc_count(int32_t n, int64_t nnz, const int64_t *rowptr, const int32_t *colidx,
              const int32_t *labels, const orc_plan *p, int64_t edge_lo, int64_t edge_hi,
              int nthreads, int reuse, uint64_t *out) {
  return orc_count_ranges(n, nnz, rowptr, colidx, labels, p, &edge_lo, &edge_hi, 1, nthreads, reuse,
                          out);
}

/* Enumeration (the counting DFS, single-threaded, writing every embedding):
 * out receives up to cap embeddings of k vertex ids each, in plan order
 * (M(0..k-1)), in the DFS's edge order; *total = the raw count (all
 * embeddings, also those beyond cap).  Returns 0, or a negative code. */
int orc_enumerate(int32_t n, int64_t nnz, const int64_t *rowptr, const int32_t *colidx,
                  const int32_t *labels, const orc_plan *p, int32_t *out, uint64_t cap, uint64_t *total) {
  if (p->labeled && !labels) return -1;
  orc_graph *g = (orc_graph *)orc_prepare(n, nnz, rowptr, colidx, labels);
  orc_worker w;
  memset(&w, 0, sizeof(w));
  w.g = g;
  w.p = p;
  w.reuse = 1;
  w.emit = out;
  w.emit_cap = out ? cap : 0;
  for (int u = 0; u < ORC_MAXK; u++) w.cand[u] = (int32_t *)malloc(sizeof(int32_t) * (size_t)(g->dmax > 0 ? g->dmax : 1));
  for (int64_t e = 0; e < nnz; e++) run_edge(&w, e);
  for (int u = 0; u < ORC_MAXK; u++) free(w.cand[u]);
  *total = w.count;
  orc_release(g);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Independent brute force: number of injective maps phi: V(P) -> V(G) */
/* with phi(a)phi(b) in E(G) for every pattern edge ab (and labels     */
/* equal when labeled).  No plan, no symmetry breaking, no pruning.    */
/* For tiny graphs only.                                               */
/* ------------------------------------------------------------------ */
static int adj_query(const int64_t *rowptr, const int32_t *colidx, int32_t a, int32_t b) {
  for (int64_t e = rowptr[a]; e < rowptr[a + 1]; e++) if (colidx[e] == b) return 1;
  return 0;
}

static uint64_t bf_rec(int32_t n, const int64_t *rowptr, const int32_t *colidx, const int32_t *labels,
                       const int32_t (*padj)[ORC_MAXK], const int32_t *plab, int k, int labeled,
                       int depth, int32_t *phi) {
  if (depth == k) return 1;
  uint64_t tot = 0;
  for (int32_t v = 0; v < n; v++) {
    int ok = 1;
    for (int j = 0; j < depth && ok; j++) if (phi[j] == v) ok = 0;
    if (ok && labeled && labels[v] != plab[depth]) ok = 0;
    for (int j = 0; j < depth && ok; j++)
      if (padj[depth][j] && !adj_query(rowptr, colidx, phi[j], v)) ok = 0;
    if (!ok) continue;
    phi[depth] = v;
    tot += bf_rec(n, rowptr, colidx, labels, padj, plab, k, labeled, depth + 1, phi);
  }
  return tot;
}

uint64_t orc_bruteforce_count(int32_t n, const int64_t *rowptr, const int32_t *colidx,
                              const int32_t *labels, const orc_plan *p) {
  int32_t phi[ORC_MAXK];
  return bf_rec(n, rowptr, colidx, labels, (const int32_t(*)[ORC_MAXK])p->adj, p->vlabel, p->k,
                p->labeled, 0, phi);
}

/* ------------------------------------------------------------------ */
/* Binary CSR reader: src/graph.h:136-193.                             */
/* ------------------------------------------------------------------ */
int orc_read_meta(const char *prefix, int64_t *meta4) {
  char path[4096];
  snprintf(path, sizeof(path), "%s.meta.txt", prefix);
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  long long a, b, c, d;
  int ok = fscanf(f, "%lld %lld %lld %lld", &a, &b, &c, &d) == 4;
  fclose(f);
  if (!ok) return -2;
  meta4[0] = a; meta4[1] = b; meta4[2] = c; meta4[3] = d;
  return 0;
}

static int read_file(const char *path, void *buf, size_t bytes) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  size_t got = fread(buf, 1, bytes, f);
  fclose(f);
  return got == bytes ? 0 : -2;
}

/* rowptr int64[n+1] (on-disk uint64), colidx int32[nnz] (on-disk uint32),
 * labels int32[n] (on-disk uint32; NULL to skip). */
int orc_read_graph(const char *prefix, int64_t n, int64_t nnz, int64_t *rowptr, int32_t *colidx,
                   int32_t *labels) {
  char path[4096];
  snprintf(path, sizeof(path), "%s.vertex.bin", prefix);
  if (read_file(path, rowptr, sizeof(int64_t) * (size_t)(n + 1))) return -1;
  snprintf(path, sizeof(path), "%s.edge.bin", prefix);
  if (read_file(path, colidx, sizeof(int32_t) * (size_t)nnz)) return -2;
  if (labels) {
    snprintf(path, sizeof(path), "%s.label.bin", prefix);
    if (read_file(path, labels, sizeof(int32_t) * (size_t)n)) return -3;
  }
  return 0;
}

size_t orc_plan_sizeof(void) { return sizeof(orc_plan); }
