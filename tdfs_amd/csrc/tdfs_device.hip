// tdfs_device.hip -- the T-DFS counting path on MI355X (gfx950), and the
// device half of the libtdfs C ABI.
//
// Design (one persistent launch per call, any number of patterns):
//   * One DFS per 64-lane wavefront.  The wave-uniform DFS state (path M,
//     iterator / end / length per stack level) lives in LDS; candidate lists
//     live in a per-wave HBM workspace sized to the graph's maximum degree, so
//     no list is ever truncated (the reference caps lists at 4096,
//     src/config.h:12).
//   * Level 0: waves claim chunks of directed edges from one HBM cursor and
//     filter them 64 at a time (src/gpu_match.cu:496-569 does this on one lane);
//     survivors are ballot-compacted into the wave's edge list.
//   * Deeper levels: C(u) = filter(probe) AND membership in every other list,
//     one pass, 64 candidates per step, ballot + mbcnt compaction
//     (src/gpu_match.cu:260-494).  The probe is the shortest list; the
//     symmetry-breaking bounds narrow it by binary search before any lane work.
//   * Result reuse: C(u) starts from C(shared_lvl[u]) (src/gpu_match.cu:733-776).
//   * Hoisted intersections: the inputs of C(u) that do not involve M(u-1)
//     are intersected into H(u) once the vertices they depend on are fixed,
//     and shared by every candidate below.
//   * Leaf: counted for 64 parents at once (flattened (parent, candidate)
//     pairs, or one shared short list probed against every parent's list).
//   * Labeled graphs keep each adjacency list sorted by (label, id), so an
//     extension reads only the label-l(u) segment; hubs (degree >= 256) get
//     neighbour bitmaps; a degree-ordered copy of the graph is counted on
//     whenever the count cannot depend on the numbering.
//   * Load balance: a wave that has held one task longer than the timeout while
//     other waves idle splits the remaining iterations of its shallowest open
//     level into tasks on an HBM ring (src/gpu_match.cu:863-908, src/queue.h);
//     a receiver rebuilds the prefix's candidate lists, so reuse stays exact.
//   * Termination: a wave leaves only when the cursors are spent, the ring is
//     empty and no wave holds work.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <vector>

#include "tdfs.h"
#include "tdfs_internal.h"

using tdfs::fail;

namespace {

constexpr int kWave = 64;
constexpr int kMaxK = TDFS_MAX_K;
constexpr int kChunkMax = 1024;      // level-0 edges per cursor grab, upper bound
constexpr int kWavesPerBlock = 4;    // 256-thread workgroups
constexpr int kBlock = kWavesPerBlock * kWave;
constexpr long long kSpinMax = 1ll << 24;  // bound on waits for a slot handshake
constexpr int kMaxSegLabels = 64;          // label segments are built for at most this many labels
constexpr int kTile = 512;                 // merge-path diagonals per round (8 per lane)
constexpr int kHashLog = 10;               // leaf hash: 1,024 LDS slots per wave
constexpr int kHash = 1 << kHashLog;
constexpr int kProf = 16;
constexpr int kXcds = 8;                   // XCDs of an MI355X: interleaved level-0 cursors per launch                  // profile words per launch (8 cycle / wave figures, 8 counters)
constexpr int kHashMax = kHash / 2;        // shared lists of at most this many elements are hashed

// Per-pattern plan in the device's layout; read with wave-uniform loads.
struct DevPlan {
  int k, labeled, sym01, pad_;
  int label[8];
  int deg[8];
  int num_bn[8];
  int bn[8][8];
  int gt_mask[8];  // bit j: M(u) > M(j)
  int lt_mask[8];  // bit j: M(u) < M(j)
  int dep[8];      // reuse source vertex, -1 if none
  int num_bn_sh[8];
  int bn_sh[8][8];
  int same_label_mask[8];  // bit j (j < u): l(j) == l(u)
  // Hoisted partial intersection H(u) (u >= 3), index [0] with result reuse,
  // [1] without: the inputs of C(u) that do not involve M(u-1) (C(dep) under
  // reuse, and the adjacency lists hx[..][u][0 .. hx_n-1]).  They depend on
  // M(0..hfix) only, so H(u) is built, filtered by the conditions against
  // M(0..hfix), as soon as M(hfix) is fixed, and serves every M(hfix+1..u-1)
  // below it.  hoist 1: C(u) = H(u) ∩ N(M(u-1)); 2: C(u) = H(u); both then
  // filtered against all of M(0..u-1).  0: not hoisted.
  int hoist[2][8];
  int hfix[2][8];
  int hx_n[2][8];
  int hx[2][8][8];
  // level-0 directed edges of this plan: [e_lo, e_lo + e_span) (per-plan
  // cursors; the interleaved cursor uses the launch's common range)
  long long e_lo, e_span;
  // [0] with result reuse, [1] without: 1 when no input of the leaf C(k-1)
  // involves M(k-2), so the leaf is counted for all its siblings from one
  // intersection (leaf_memo)
  int leaf_free[2];
  // bit w (2 <= w <= k-2): vertex w is free -- no later vertex has it as a
  // backward neighbour or shares its label (labeled plans), so no later
  // candidate set depends on M(w) and every M(w) has the same subtree count
  int free_mask;
  // bit j (j <= k-2): M(j) lies in the leaf's base set (the intersection of
  // N(M(b)), b in BN(k-1)) in every embedding -- j is adjacent to every such b
  int leaf_in;
  // tail: the last tail_t vertices (tail_t >= 2, 0: none) have one backward
  // neighbour set, are pairwise non-adjacent and either are totally ordered
  // by the symmetry conditions and otherwise bound only by injectivity
  // (unlabeled: their completions are the tail_t-subsets of C(k - tail_t))
  // or are labeled (their completions: for each label, the ordered tuples of
  // distinct elements of the candidates of the label's first tail vertex, as
  // many as the tail has vertices of that label)
  int tail_t;
  int tail_gsz[8];  // tail vertex v first of its label in the tail: the tail's vertices of that label; else 0
  // 1: the last three vertices form a triangle over one set of earlier
  // neighbours, totally ordered and otherwise bound only by injectivity
  // (unlabeled): a path's completions are the triangles of G[C(k-3)];
  // 2: the same, labeled, with three distinct labels: the triangles across
  // the candidates of the three labels
  int tri_tail;
  // leaf factor (counting kernel): lf_a >= 1 when this plan is a pattern's
  // first k vertices and its dropped last vertex (kept at index k: label,
  // backward neighbours, conditions) has candidates fixed by M(0..lf_a) up to
  // lf_excl later vertices that always lie among them and are excluded; each
  // iteration at level lf_a - 1 stands for that many completions of the leaf
  // per count of the rest; -1: none
  int lf_a;
  int lf_excl;
  // tx_w >= 2: the vertex right above the tail, which no tail vertex has as a
  // backward neighbour and which the tail's group led by tx_lead must differ
  // from (its label; unlabeled: the whole tail, NE only); -1: none
  int tx_w;
  int tx_lead;
};

// Modes of extend(): C(u) from the plan's inputs (kFull) or from H(u)
// (kHoisted), or H(u) itself (kBuildH); kMemo: the leaf's candidates from the
// plan's inputs filtered against M(0..u-2) only (leaf_memo).
enum { kFull = 0, kHoisted = 1, kBuildH = 2, kMemo = 3 };

// A donated unit of work: prefix M(0..p-1) and the iteration range [a, b) of
// C(p).  fresh = 1 marks level-0 donations whose first extension was never
// performed by the donor (statistics only).
struct __attribute__((aligned(64))) TaskRec {
  int pid, p, a, b;
  int fresh, pad0;
  // the task's count stands for this many (the product of 1 + the skipped
  // iterations of every free level it lies under)
  unsigned long long mul;
  int M[8];
};
static_assert(sizeof(TaskRec) == 64, "one task per 64-byte slot");

// Control words, each on its own 128-byte line.
struct Ctrl {
  unsigned long long cursor;
  long long pad0[15];
  unsigned long long q_tail;
  long long pad1[15];
  unsigned long long q_head;
  long long pad2[15];
  long long q_count;  // reserved - released (producer space accounting)
  long long pad3[15];
  long long q_avail;  // published - claimed (consumer accounting)
  long long pad4[15];
  int busy;
  int pad5[31];
  int idle;
  int pad6[31];
  int error;
  int pad7[31];
  unsigned long long progress;  // bumped by busy waves while they advance (idle deadline)
  long long pad8[15];
};

struct KParams {
  // graph
  const long long *rowptr;
  const int *colidx;
  const int *degv;
  const int *src;
  const int *maxnbr;
  const unsigned short *labels;  // dense label ids (the upload maps the caller's values)
  const int *colidx_lab;      // each list re-sorted by (label, id); null when unlabeled
  const int *loff;            // per vertex: nlab+1 offsets of its label segments
  int nlab;
  const int *hub_idx;         // per vertex: row of its neighbour bitmap, or -1
  const unsigned *hub_bits;   // hub_words 32-bit words per hub
  long long hub_words;
  // plans and the level-0 cursor space
  const DevPlan *plans;
  int nplans;
  int chunk;
  long long edge_lo, edge_hi;
  unsigned long long span;      // edges per plan: edge_hi - edge_lo
  unsigned long long *cursors;  // per plan (128-byte stride): next level-0 offset
  int interleave;               // 1: cursors over (chunk, plan); 0: per-plan cursors
  int xcd_groups;               // interleaved: cursors, one per group of chunks (8: one per XCD)
  // per-wave workspace: level-0 pairs, then C(2..k-2), then H(3..k-1) at hoff
  int *wbuf;
  long long wstride;
  long long hoff;
  int dstride;
  // task ring
  TaskRec *q_data;
  unsigned long long *q_seq;
  unsigned long long q_mask;
  // results: counts[pid]; stats[pid*4 + {0 alg bytes, 1 extensions, 2 donated}]
  unsigned long long *counts;
  unsigned long long *stats;
  unsigned long long *prof;  // kProf per launch: cycles in filter, extend, leaf, idle, rebuild, donate, total;
                             // waves; then the operation counters of the statistics kernel (WaveLds::ctr)
  Ctrl *ctrl;
  // tuning
  long long timeout_ticks;  // < 0: never donate
  long long deadline_ticks; // a wave idle this long with no progress reported gives up (error); <= 0: never
  long long prog_ticks;     // busy waves report progress at most this often
  int split_chunk;
  int donate_levels;
  int reuse;
  int leaf_batch;
  int leaf_memo;             // 1: leaves that do not involve M(k-2) are counted for all siblings at once
  int leaf_hash;             // bit 0: a short shared leaf input may be hashed in LDS for membership tests;
                             // bit 1: a dense one may be held as an LDS bitmap
  int ext_hash;              // 1: an extension hashes one short non-probe input in LDS for membership tests
  int tails;                 // 1: closing vertices counted without enumeration (timed kernel)
  long long limit_ticks;     // > 0: work still pending this long after a wave started is dropped
  int variant;               // host side only: 0 default kernel, 4 / 8 the kernel compiled for that occupancy
  int leaf_pair_cap;         // (owner, candidate) pairs per leaf batch
  int steal;                 // 1: idle waves take work from sibling waves of their workgroup (LDS atomics)
  int merge_min;             // two-input extensions whose inputs both hold at least this many candidates
                             // intersect by merge path (0: never)
  int hoist_min;             // build H(u) for a level of at least this many candidates (0: never)
  // per-plan completion (serving): pend[pid * 16] counts the plan's level-0
  // chunks and ring tasks in flight; when it drops to zero with the plan's
  // cursor spent, the plan's count and flags go to done_host[pid * 2 + {0, 1}]
  // (host-mapped memory) so the host can collect it before the launch ends
  unsigned long long *pend;
  unsigned long long *done_host;
  // enumeration (emit_n non-null): embeddings as rows of k ids
  int *emit;                    // rows buffer (null when emit_cap is 0)
  unsigned long long emit_cap;  // rows
  unsigned long long *emit_n;   // rows produced (also past emit_cap); null: count only
};

struct WaveLds {
  // per stack level: the untaken iterations [it, end) of the level's list
  // and the list's generation, packed as (gen << 48) | (it << 24) | end; the
  // owner claims iterations from the front and an idle sibling wave (same
  // workgroup) may take the back half, both by compare-and-swap
  unsigned long long ie[8];
  int M[8];
  int len[8];
  int pid;      // plan of the work in progress
  int readers;  // sibling waves copying this wave's lists (the owner does not overwrite them meanwhile)
  int hv[8];  // H(u) built for the current M(0..hfix(u))
  int hl[8];  // |H(u)|
  // wave-uniform input lists of the extension in progress
  const int *ulp[8];
  const unsigned *uhb[8];
  int uln[8];
  int *tile;  // merge-path variant: this wave's 2 (kTile + 4) LDS ints (the next kTile + 1 of both lists)
  // this wave's kHash LDS words holding the window [tab_lo, tab_hi) of the
  // leaf's shared input list tab_list, as a hash table (tab_kind 1: slots,
  // -1 empty) or a bitmap (tab_kind 2: bit y - tab_base of word
  // (y - tab_base) / 32 set for every element y in [tab_base, tab_end), path
  // vertices cleared), or the memoised leaf's set (tab_kind 3: sorted, tab_lo
  // elements); tab_kind 4: the path's leaf uses no table, 0: nothing held.
  // Valid while `epoch` (bumped by every level_set) equals tab_epoch
  int *htab;
  unsigned epoch, tab_epoch;
  int tab_kind;
  const int *tab_list;
  int tab_lo, tab_hi, tab_base, tab_end;
  // statistics kernel only (lane 0): leaf batches, their owners, their units
  // (pairs, words or elements tested), their set-up cycles; extensions, their
  // probe elements, their membership tests by binary search (per 64-lane
  // round and input); memoised-leaf calls
  unsigned long long ctr[8];
  // free levels: the count when the level's first (measured) iteration
  // began, and the level's other iterations (claimed with it)
  unsigned long long mbase[8];
  int mrest[8];
};

template <bool STATS>
__device__ __forceinline__ void count_op(WaveLds &S, int i, unsigned long long v) {
  if (STATS && (threadIdx.x & (kWave - 1)) == 0) S.ctr[i] += v;
}

// LDS-typed views of a wave's table (ds_* instructions instead of flat ones)
typedef __attribute__((address_space(3))) int lds_int;
typedef __attribute__((address_space(3))) unsigned lds_uint;
__device__ __forceinline__ lds_int *lds_i(int *p) { return (lds_int *)p; }
__device__ __forceinline__ lds_uint *lds_u(int *p) { return (lds_uint *)p; }

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ unsigned long long uni64(unsigned long long x) {
  unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)x);
  unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(x >> 32));
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ unsigned mbcnt64(unsigned long long m) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// ---------------------------------------------------------------------------
// Stack-level iteration ranges.  The owner wave and, with stealing on, idle
// sibling waves of its workgroup update them with LDS atomics (workgroup
// scope); without stealing, lane 0 of the owner reads and writes them
// plainly.  Every helper is called by all lanes and returns wave-uniform
// values.
// ---------------------------------------------------------------------------
constexpr int kIeMax = (1 << 24) - 1;  // list lengths and positions fit 24 bits
__device__ __forceinline__ unsigned long long ie_make(unsigned long long gen, int it, int end) {
  return (gen << 48) | ((unsigned long long)(unsigned)it << 24) | (unsigned long long)(unsigned)end;
}
__device__ __forceinline__ int ie_it(unsigned long long w) { return (int)((w >> 24) & kIeMax); }
__device__ __forceinline__ int ie_end(unsigned long long w) { return (int)(w & kIeMax); }
__device__ __forceinline__ unsigned long long ie_gen(unsigned long long w) { return w >> 48; }

__device__ __forceinline__ unsigned long long ie_read(WaveLds &S, int l) {
  unsigned long long w = 0;
  if (lane_id() == 0) w = __hip_atomic_load(&S.ie[l], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
  return uni64(w);
}

// Starts level l on a new list: iterations [it, end), next generation.
// The list and the path it depends on are written before (release).
__device__ __forceinline__ void level_set(WaveLds &S, int l, int it, int end, bool shared) {
  if (shared) __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the lists are in L2 before a sibling may copy them
  if (lane_id() == 0) {
    const unsigned long long w = __hip_atomic_load(&S.ie[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_store(&S.ie[l], ie_make(ie_gen(w) + 1, it, end), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    S.epoch += 1;  // the path changed: a hashed leaf list is stale
  }
}

// Claims up to `want` iterations from the front of level l: returns the
// first, n = how many (0: none left).
__device__ __forceinline__ int level_claim(WaveLds &S, int l, int want, int &n, bool shared) {
  int first = 0, got = 0;
  if (lane_id() == 0) {
    unsigned long long w = __hip_atomic_load(&S.ie[l], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    for (;;) {
      const int it = ie_it(w), end = ie_end(w);
      if (it >= end) break;
      const int c = min(want, end - it);
      const unsigned long long nw = ie_make(ie_gen(w), it + c, end);
      if (!shared) {
        __hip_atomic_store(&S.ie[l], nw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      } else if (!__hip_atomic_compare_exchange_strong(&S.ie[l], &w, nw, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP)) {
        continue;
      }
      first = it;
      got = c;
      break;
    }
  }
  n = uni(got);
  return uni(first);
}

// Returns the back part [from, to) of the owner's last claim of level l
// (to is the level's front: only the owner moves it, and thieves take only
// from the back of [to, end)).
__device__ __forceinline__ void level_unclaim(WaveLds &S, int l, int from, int to, bool shared) {
  if (lane_id() == 0) {
    unsigned long long w = __hip_atomic_load(&S.ie[l], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    for (;;) {
      const unsigned long long nw = ie_make(ie_gen(w), from, max(ie_end(w), to));
      if (!shared) {
        __hip_atomic_store(&S.ie[l], nw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        break;
      }
      if (__hip_atomic_compare_exchange_strong(&S.ie[l], &w, nw, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                               __HIP_MEMORY_SCOPE_WORKGROUP))
        break;
    }
  }
}

// Takes iterations from the back of level l for donation: all of [it, end)
// (all = true) or at most `most` of them; returns the first taken, n = how many.
__device__ __forceinline__ int level_take_back(WaveLds &S, int l, bool all, int most, int &n, bool shared) {
  int first = 0, got = 0;
  if (lane_id() == 0) {
    unsigned long long w = __hip_atomic_load(&S.ie[l], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    for (;;) {
      const int it = ie_it(w), end = ie_end(w);
      if (it >= end) break;
      const int c = all ? end - it : min(most, end - it);
      const unsigned long long nw = ie_make(ie_gen(w), it, end - c);
      if (!shared) {
        __hip_atomic_store(&S.ie[l], nw, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      } else if (!__hip_atomic_compare_exchange_strong(&S.ie[l], &w, nw, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                                       __HIP_MEMORY_SCOPE_WORKGROUP)) {
        continue;
      }
      first = end - c;
      got = c;
      break;
    }
  }
  n = uni(got);
  return uni(first);
}

// Before the owner overwrites a list: waits until no sibling is copying its
// lists.  Returns false if the wait did not end (error flag set by caller).
__device__ __forceinline__ bool wait_readers(WaveLds &S, bool shared) {
  if (!shared) return true;
  int ok = 1;
  if (lane_id() == 0) {
    long long spins = 0;
    while (__hip_atomic_load(&S.readers, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != 0) {
      if (++spins > (1ll << 22)) {
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  return uni(ok) != 0;
}

template <typename T>
__device__ __forceinline__ T ld_acq(T *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_rel(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
// Counters are relaxed: ordering of the ring payload is carried only by the
// per-slot sequence numbers (release on publish, acquire on take).  A
// seq_cst or acq_rel RMW at agent scope would write back and invalidate the
// XCD's L2 on every call.
template <typename T>
__device__ __forceinline__ T fadd(T *p, T v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T ld_rlx(T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Adds, then waits until the add has been performed: a later atomic of this
// wave (to another address) cannot overtake it.
template <typename T>
__device__ __forceinline__ void fadd_done(T *p, T v) {
  (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
}


// First index i in sorted a[0, n) with a[i] >= x (uniform or per-lane).
__device__ __forceinline__ int lower_bound_i(const int *a, int n, int x) {
  int lo = 0;
  while (n > 0) {
    const int half = n >> 1;
    if (a[lo + half] < x) {
      lo += half + 1;
      n -= half + 1;
    } else {
      n = half;
    }
  }
  return lo;
}

// The same for a search every lane of the wave makes with the same (a, n,
// x), all lanes active: each round loads 64 evenly spaced elements of the
// remaining range at once and keeps the block between the last sample below
// x and the next, so a search costs about log64(n) + 1 dependent loads
// instead of log2(n).
__device__ __forceinline__ int lower_bound_wave(const int *a, int n, int x) {
  const int lane = lane_id();
  int lo = 0, len = n;  // the answer lies in [lo, lo + len]
  while (len > kWave) {
    const int step = (len + kWave - 1) / kWave;
    const int p = lo + lane * step;
    const int c = __popcll(__ballot(p < lo + len && a[p] < x));
    if (c == 0) return lo;  // a[lo] >= x
    const int nlo = lo + (c - 1) * step + 1;
    len = min(lo + c * step, lo + len) - nlo;
    lo = nlo;
  }
  return lo + __popcll(__ballot(lane < len && a[lo + lane] < x));
}

// Several searches at once: the lanes of each group of G (aligned) search
// sorted a[0, n) for the key they share (G-ary: each round samples G
// elements of the group's range); all lanes call, and each gets its group's
// first index i with a[i] >= key.  64 / G independent searches cost about
// log_G(n) + 1 dependent loads together.
template <int G>
__device__ __forceinline__ int lower_bound_group(const int *a, int n, int key) {
  const int lane = lane_id();
  const int sub = lane & (G - 1);
  const int shift = lane & ~(G - 1);
  const unsigned long long gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
  int lo = 0, len = n;  // the group's answer lies in [lo, lo + len]
  while (__ballot(len > G)) {
    if (len > G) {
      const int step = (len + G - 1) / G;
      const int p = lo + sub * step;
      const unsigned long long m = __ballot(p < lo + len && a[p] < key);
      const int c = __popcll((m >> shift) & gmask);
      if (c == 0) {
        len = 0;  // a[lo] >= key
      } else {
        const int nlo = lo + (c - 1) * step + 1;
        len = min(lo + c * step, lo + len) - nlo;
        lo = nlo;
      }
    }
  }
  const unsigned long long m = __ballot(sub < len && a[lo + sub] < key);
  return lo + __popcll((m >> shift) & gmask);
}

// Membership of x in sorted a[0, n): branch-free halving, trip count depends
// only on n.
__device__ __forceinline__ bool contains(const int *a, int n, int x) {
  if (n <= 0) return false;
  const int *base = a;
  while (n > 1) {
    const int half = n >> 1;
    base = (base[half - 1] < x) ? base + half : base;
    n -= half;
  }
  return *base == x;
}

// LDS hash tables (open addressing, linear probing, multiplicative hash).
// A table of 2^hlog slots holds at most half as many elements.
__device__ __forceinline__ unsigned hash_slot(int y, int hlog = kHashLog) {
  return ((unsigned)y * 2654435761u) >> (32 - hlog);
}
__device__ __forceinline__ bool hash_has(const lds_int *tab, int y, int hlog = kHashLog) {
  unsigned h = hash_slot(y, hlog);
  const unsigned mask = (1u << hlog) - 1u;
  for (;;) {
    const int v = tab[h];
    if (v == y) return true;
    if (v < 0) return false;
    h = (h + 1) & mask;
  }
}

// Fills this wave's LDS table with a hash of a[0, n) (n <= 2^(hlog-1)).
// All lanes call it; the table is complete when it returns.
__device__ __forceinline__ void hash_build(lds_int *tab, const int *a, int n, int hlog) {
  const int lane = lane_id();
  const int slots = 1 << hlog;
  for (int i = lane; i < slots; i += kWave) tab[i] = -1;
  __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  const unsigned mask = (unsigned)slots - 1u;
  for (int idx = lane; idx < n; idx += kWave) {
    const int y = a[idx];
    unsigned h = hash_slot(y, hlog);
    for (;;) {
      int old = -1;
      __hip_atomic_compare_exchange_strong(tab + h, &old, y, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WAVEFRONT);
      if (old < 0 || old == y) break;
      h = (h + 1) & mask;
    }
  }
  __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
}

// Membership through a hub's dense neighbour bitmap: one load.
__device__ __forceinline__ bool bit_test(const unsigned *bits, int x) { return (bits[x >> 5] >> (x & 31)) & 1u; }

__device__ __forceinline__ const unsigned *hub_row(const KParams &P, int v) {
  const int h = P.hub_idx[v];
  return h >= 0 ? P.hub_bits + (long long)h * P.hub_words : nullptr;
}

// Adjacency of t, restricted to its label-`lab` segment when lab >= 0 (the
// segment is sorted by id, so it serves both as a probe and for membership).
__device__ __forceinline__ const int *adj_list(const KParams &P, int t, int lab, int &len) {
  if (lab >= 0) {
    if (lab >= P.nlab) {  // no vertex carries this label
      len = 0;
      return P.colidx_lab;
    }
    const int *o = P.loff + (long long)t * (P.nlab + 1);
    const int a = o[lab];
    len = o[lab + 1] - a;
    return P.colidx_lab + P.rowptr[t] + a;
  }
  len = P.degv[t];
  return P.colidx + P.rowptr[t];
}

// Writes the leaf embeddings (M(0..u-1), x) of the lanes set in `mask` to
// consecutive rows of the enumeration buffer (one atomic per wave).
__device__ __forceinline__ void emit_rows(const KParams &P, const WaveLds &S, int u, int x, bool ok,
                                          unsigned long long mask) {
  const unsigned c = (unsigned)__popcll(mask);
  if (c == 0) return;
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(P.emit_n, (unsigned long long)c);
  base = __shfl(base, 0);
  const unsigned long long pos = base + mbcnt64(mask);
  if (ok && pos < P.emit_cap) {
    int *row = P.emit + pos * (unsigned long long)(u + 1);
    for (int j = 0; j < u; j++) row[j] = S.M[j];
    row[u] = x;
  }
}

// ---------------------------------------------------------------------------
// Merge-path intersection of two sorted, duplicate-free lists A[0, na) and
// B[0, nb): the merge of A and B (ties take A first) is cut into rounds of
// kTile steps.  Each round stages the next kTile + 1 elements of both lists
// in LDS with coalesced loads; lane l finds where diagonal 8l crosses the
// staged lists (co-rank: binary search in LDS) and performs merge steps
// 8l .. 8l+7 there.  Taking A[a] while B's head equals it is a match, so
// every common element is found exactly once, by one lane.  Matches passing
// keep(x) are counted and, when out is non-null, written in ascending order
// (count, wave scan, then the same steps again writing).  Returns the count.
// ---------------------------------------------------------------------------
template <typename Keep>
__device__ __forceinline__ unsigned merge_intersect(WaveLds &S, const int *A, int na, const int *B, int nb, int *out,
                                                    Keep keep) {
  const int lane = lane_id();
  constexpr int per = kTile / kWave;
  int i0 = 0, j0 = 0;
  unsigned n = 0;
  int *ta = S.tile, *tb = S.tile + kTile + 4;
  auto corank = [&](int d) {  // A elements among the first d merge steps of the staged lists
    int lo = max(0, d - kTile), hi = min(d, kTile);
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (ta[mid] <= tb[d - mid - 1]) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  while (i0 < na && j0 < nb) {
    for (int t = lane; t <= kTile; t += kWave) {
      ta[t] = i0 + t < na ? A[i0 + t] : INT_MAX;
      tb[t] = j0 + t < nb ? B[j0 + t] : INT_MAX;
    }
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // the staged tile is written before any lane reads it
    const int a0 = corank(lane * per);
    auto steps = [&](int *dst) -> int {  // this lane's merge steps; writes matches to dst when non-null
      int a = a0, b = lane * per - a0, c = 0;
      for (int st = 0; st < per; st++) {
        const int x = ta[a], y = tb[b];
        if (x <= y) {
          if (x == y && i0 + a < na && j0 + b < nb && keep(x)) {
            if (dst) dst[c] = x;
            c++;
          }
          a++;
        } else {
          b++;
        }
      }
      return c;
    };
    const int c = steps(nullptr);
    int incl = c;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int t = __shfl_up(incl, d);
      if (lane >= d) incl += t;
    }
    if (out && c > 0) steps(out + n + (incl - c));
    n += (unsigned)__shfl(incl, kWave - 1);
    const int aend = corank(kTile);
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // every lane is done with the tile before it is restaged
    i0 += aend;
    j0 += kTile - aend;
  }
  return n;
}

// ---------------------------------------------------------------------------
// Candidate generation for pattern vertex u (2 <= u <= k-1) given M(0..u-1).
// Returns |C(u)|; writes C(u) (ascending) to out unless out == nullptr.
//
// mode kFull: the plan's inputs -- C(dep) and N(M(b)) for b in BN_sh(u) under
// result reuse, else N(M(b)) for b in BN(u).  kHoisted: H(u) (and
// N(M(u-1)) when hoist = 1).  kBuildH: writes H(u), filtering only against
// M(0..hfix(u)).  Statistics always count the plan's inputs (the merge
// model), whichever lists are read.  The inputs are wave-uniform; their
// descriptors live in the wave's LDS slot (S.ulp / S.uhb / S.uln), which
// keeps them out of the vector registers.
// ---------------------------------------------------------------------------
template <bool STATS, bool MERGE = false>
__device__ __forceinline__ unsigned extend(const KParams &P, const DevPlan *__restrict__ pl, WaveLds &S, int u,
                           int *__restrict__ out, const int *lbase, const int *hbase, int mode, bool count_stats,
                           unsigned long long &balg, unsigned long long &next, int lab_as = -1,
                           const int *xl = nullptr, int xn = 0) {
  // lab_as >= 0 (labeled plans, mode kFull): the candidates of label lab_as
  // over u's inputs, distinct from the path vertices of that label; xl (a
  // sorted list of xn vertices of l(u)): one more input, intersected too
  count_op<STATS>(S, 4, 1);
  const int lane = lane_id();
  const int ri = P.reuse ? 0 : 1;
  const int dep = (P.reuse && lab_as < 0) ? pl->dep[u] : -1;
  const int ulab = lab_as >= 0 ? lab_as : pl->label[u];
  unsigned long long bytes = 0;
  // labeled plans read only the label-l(u) segment of each adjacency list
  const int seg = (pl->labeled && P.colidx_lab) ? ulab : -1;
  int minfull = INT_MAX;  // statistics: shortest full input list
  if (STATS && count_stats) {
    if (dep >= 0) {
      const int l = S.len[dep - 1];
      bytes += 4ull * (unsigned)l;
      minfull = l;
      const int nsh = pl->num_bn_sh[u];
      for (int i = 0; i < nsh; i++) {
        const int d = P.degv[S.M[pl->bn_sh[u][i]]];
        bytes += 4ull * (unsigned)d + 16;
        minfull = min(minfull, d);
      }
    } else {
      const int nb = pl->num_bn[u];
      for (int i = 0; i < nb; i++) {
        const int d = P.degv[S.M[pl->bn[u][i]]];
        bytes += 4ull * (unsigned)d + 16;
        minfull = min(minfull, d);
      }
    }
  }
  int nl = 0;
  auto put_adj = [&](int b) {  // input nl: adjacency (segment) of M(b)
    const int t = S.M[b];
    int len;
    const int *p = adj_list(P, t, seg, len);
    S.ulp[nl] = p;
    S.uhb[nl] = hub_row(P, t);
    S.uln[nl] = len;
    nl++;
  };
  auto put_list = [&](const int *p, int len) {  // input nl: a workspace list
    S.ulp[nl] = p;
    S.uhb[nl] = nullptr;
    S.uln[nl] = len;
    nl++;
  };
  if (mode == kHoisted) {
    put_list(hbase + (long long)(u - 3) * P.dstride, S.hl[u]);
    if (pl->hoist[ri][u] == 1) put_adj(u - 1);
  } else if (mode == kBuildH) {
    if (dep >= 0) put_list(lbase + (long long)(dep - 2) * P.dstride, S.len[dep - 1]);
    const int nx = pl->hx_n[ri][u];
    for (int i = 0; i < nx; i++) put_adj(pl->hx[ri][u][i]);
  } else if (dep >= 0) {
    put_list(lbase + (long long)(dep - 2) * P.dstride, S.len[dep - 1]);  // C(dep): list of level dep-1
    const int nsh = pl->num_bn_sh[u];
    for (int i = 0; i < nsh; i++) put_adj(pl->bn_sh[u][i]);
  } else {
    const int nb = pl->num_bn[u];
    for (int i = 0; i < nb; i++) put_adj(pl->bn[u][i]);
  }
  if (xl) put_list(xl, xn);
  // probe = shortest list
  int pidx = 0;
  int pn = S.uln[0];
  for (int i = 1; i < nl; i++) {
    const int l = S.uln[i];
    if (l < pn) {
      pn = l;
      pidx = i;
    }
  }
  const int *pp = S.ulp[pidx];
  const bool leaf = (u == pl->k - 1) && mode != kBuildH && mode != kMemo;
  const bool labeled = pl->labeled != 0;
  if (STATS && count_stats) {
    if (labeled) bytes += (unsigned)minfull;
  }
  const bool label_check = labeled && seg < 0;
  // every probe element carries l(u): segments, C(dep) (same label as u) or H(u)
  const bool label_known = !labeled || seg >= 0 || dep >= 0 || mode == kHoisted;
  // filter bounds; H(u) is filtered only against M(0..hfix(u))
  const int jlim = mode == kBuildH ? pl->hfix[ri][u] + 1 : (mode == kMemo ? u - 1 : u);
  int lo = -1, hi = INT_MAX;
  int nev[6];
  int nemask = 0;
  const int gt = pl->gt_mask[u], lt = pl->lt_mask[u];
  // over label segments every candidate carries l(u): only path vertices of
  // that label can coincide with one
  int same = (labeled && seg >= 0) ? pl->same_label_mask[u] : 0x7f;
  if (labeled && seg >= 0 && lab_as >= 0) {
    same = 0;
    for (int j = 0; j < u; j++) same |= (pl->label[j] == lab_as) << j;
  }
#pragma unroll
  for (int j = 0; j < 6; j++) {
    nev[j] = -1;
    if (j < jlim) {
      const int m = S.M[j];
      if (!labeled && ((gt >> j) & 1)) lo = max(lo, m);
      else if (!labeled && ((lt >> j) & 1)) hi = min(hi, m);
      else if ((same >> j) & 1) {
        nev[j] = m;
        nemask |= 1 << j;
      }
    }
  }
  // A probe of at most 64 elements is loaded once: its window comes from
  // ballots over the lanes, and the probe loop takes its elements from them.
  const bool small = pn <= kWave;
  int xr = 0;
  int start = 0, stop = pn;
  if (small) {
    xr = lane < pn ? pp[lane] : INT_MAX;
    if (lo >= 0) start = __popcll(__ballot(lane < pn && xr <= lo));
    if (hi != INT_MAX) stop = __popcll(__ballot(lane < pn && xr < hi));
  } else {
    if (lo >= 0) start = lower_bound_wave(pp, pn, lo + 1);
    if (hi != INT_MAX) stop = lower_bound_wave(pp, pn, hi);
  }
  unsigned n = 0;
  // merge path: two inputs, both long, neither more than 8x the other
  int mo = -1, ms = 0, me = 0;
  if (MERGE && P.merge_min > 0 && nl == 2 && stop - start >= P.merge_min && !P.emit_n) {
    const int oi = 1 - pidx;
    const int *ob = S.ulp[oi];
    const int on = S.uln[oi];
    ms = lo >= 0 ? lower_bound_wave(ob, on, lo + 1) : 0;
    me = hi != INT_MAX ? lower_bound_wave(ob, on, hi) : on;
    if (me - ms >= P.merge_min && me - ms <= 8 * (stop - start)) mo = oi;
  }
  if (start < stop && mo >= 0) {
    const int lab = ulab;
    n = merge_intersect(S, pp + start, stop - start, S.ulp[mo] + ms, me - ms, out, [&](int x) {
      bool ok = true;
#pragma unroll
      for (int j = 0; j < 6; j++)
        if ((nemask >> j) & 1) ok = ok && (x != nev[j]);
      if (label_check && ok) ok = (int)P.labels[x] == lab;
      return ok;
    });
  } else if (start < stop) {
    if ((leaf || (out == nullptr && mode != kBuildH && mode != kMemo)) && nl == 1 && label_known && !P.emit_n) {
      // Count only (the leaf, a tail or a free vertex) over one list:
      // |list ∩ (lo,hi)| minus the excluded path vertices that lie in that
      // window and in the list.
      n = (unsigned)(stop - start);
#pragma unroll
      for (int j = 0; j < 6; j++) {
        if ((nemask >> j) & 1) {
          const int v = nev[j];
          if (v > lo && v < hi) {
            if (small) {
              if (__ballot(lane >= start && lane < stop && xr == v)) n--;
            } else {
              const int pos = lower_bound_wave(pp + start, stop - start, v);
              if (pos < stop - start && pp[start + pos] == v) n--;
            }
          }
        }
      }
    } else {
      const int lab = ulab;
      // the longest other input without a neighbour bitmap, when it has at
      // most kHashMax elements, goes into this wave's LDS hash table (sized to
      // it): one coalesced pass over it instead of a binary search per probe
      int hidx = -1, hlog = 0;
      if (P.ext_hash) {
        int best = 7;  // a search of at most 3 loads is kept
        for (int i = 0; i < nl; i++) {
          const int l = S.uln[i];
          if (i != pidx && !S.uhb[i] && l <= kHashMax && l > best) {
            best = l;
            hidx = i;
          }
        }
        if (hidx >= 0) {
          hlog = 6;
          while ((1 << (hlog - 1)) < best) hlog++;
          hash_build(lds_i(S.htab), S.ulp[hidx], best, hlog);
          if (lane == 0) S.tab_kind = 0;  // a leaf's table held there is gone
        }
      }
      const lds_int *tab = lds_i(S.htab);
      if (STATS) {
        int ns = 0;
        for (int i = 0; i < nl; i++) ns += (i != pidx && i != hidx && !S.uhb[i]) ? 1 : 0;
        count_op<STATS>(S, 5, (unsigned)(stop - start));
        count_op<STATS>(S, 6, (unsigned long long)ns * (unsigned)((stop - start + kWave - 1) / kWave));
      }
      for (int base = start; base < stop; base += kWave) {
        const int idx = base + lane;
        bool ok = idx < stop;
        const int xs = small ? __shfl(xr, min(idx, kWave - 1)) : 0;  // all lanes take part in the shuffle
        int x = 0;
        if (ok) {
          x = small ? xs : pp[idx];
#pragma unroll
          for (int j = 0; j < 6; j++)
            if ((nemask >> j) & 1) ok = ok && (x != nev[j]);
          if (label_check && ok) ok = (int)P.labels[x] == lab;
#pragma unroll 1
          for (int i = 0; i < nl; i++) {
            if (i != pidx && ok) {
              const unsigned *h = S.uhb[i];
              ok = i == hidx ? hash_has(tab, x, hlog) : (h ? bit_test(h, x) : contains(S.ulp[i], S.uln[i], x));
            }
          }
        }
        const unsigned long long mask = __ballot(ok);
        if (out) {
          const unsigned pos = n + mbcnt64(mask);
          if (ok) out[pos] = x;
        } else if (leaf && P.emit_n) {
          emit_rows(P, S, u, x, ok, mask);
        }
        n += (unsigned)__popcll(mask);
      }
    }
  }
  if (STATS && count_stats) {
    if (!leaf) bytes += 4ull * n;
    balg += bytes;
    next += 1;
  }
  return n;
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
  return v;
}

// Count-only leaf for elements [i0, i1) (at most 64) of stack level L = k-3:
// element x fixes M(k-2) (for L = 0 the edge pair fixes M(0), M(1)) and
// contributes |C(k-1)|.  Lane l owns element i0 + l and builds that
// element's window and probe; the probe candidates of all owners are then
// flattened into one pair space (wave prefix scan) and each lane tests one
// (owner, candidate) pair per step, finding its owner by a binary search over
// the scan with ds_bpermute.  Inputs that do not depend on the owner (C(dep),
// H(u), N(M(b)) for b <= L) are wave-uniform and kept in LDS; at most two
// depend on it (N(M(L+1)), or N(M(0)) and N(M(1)) at L = 0) and stay in
// registers.  Same result and statistics as calling extend() for each element
// in turn (with H(k-1) when it is built).
template <bool STATS>
__device__ __forceinline__ unsigned long long leaf_batch(const KParams &P, const DevPlan *__restrict__ pl, WaveLds &S, int L,
                                         int i0, int i1, const int *L0r, const int *L0c, const int *lbase,
                                         const int *hbase, unsigned long long &balg, unsigned long long &next,
                                         int &taken) {
  const long long t_setup = STATS ? clock64() : 0;
  const int lane = lane_id();
  const int k = pl->k;
  const int u = k - 1;
  const bool labeled = pl->labeled != 0;
  const int idx = i0 + lane;
  const bool act = idx < i1;
  // the owner's own vertices: M(L+1) for L > 0; M(0), M(1) for L = 0 (u = 2)
  int xa = 0, xb = 0;
  if (L == 0) {
    xa = act ? L0r[idx] : 0;
    xb = act ? L0c[idx] : 0;
  } else {
    xa = act ? lbase[(long long)(L - 1) * P.dstride + idx] : 0;
  }
  auto own = [&](int b) { return L == 0 ? b <= 1 : b == L + 1; };
  auto pick = [&](int b) -> int {  // by value: no address of xa / xb is formed
    if (L == 0) {
      if (b == 0) return xa;
      return xb;
    }
    if (b == L + 1) return xa;
    const int m = S.M[b];
    return m;
  };
  const int ri = P.reuse ? 0 : 1;
  const int dep = P.reuse ? pl->dep[u] : -1;
  // H(u) serves every owner of the batch (L > 0: M(0..u-2) are shared)
  const bool hoisted = L > 0 && S.hv[u] != 0;
  unsigned long long bytes = 0;
  const int seg = (labeled && P.colidx_lab) ? pl->label[u] : -1;
  int minfull = INT_MAX;
  if (STATS) {
    if (dep >= 0) {
      const int l = S.len[dep - 1];
      bytes += 4ull * (unsigned)l;
      minfull = l;
      const int nsh = pl->num_bn_sh[u];
      for (int i = 0; i < nsh; i++) {
        const int d = P.degv[pick(pl->bn_sh[u][i])];
        bytes += 4ull * (unsigned)d + 16;
        minfull = min(minfull, d);
      }
    } else {
      const int nb = pl->num_bn[u];
      for (int i = 0; i < nb; i++) {
        const int d = P.degv[pick(pl->bn[u][i])];
        bytes += 4ull * (unsigned)d + 16;
        minfull = min(minfull, d);
      }
    }
  }
  // inputs: nu wave-uniform ones in LDS; the owner's own in two fixed slots
  // (A: N(M(L+1)), or N(M(0)) at L = 0; B: N(M(1)) at L = 0)
  int nu = 0;
  bool hasA = false, hasB = false;
  const int *vp0 = nullptr, *vp1 = nullptr;
  int vh0 = -1, vh1 = -1;  // hub bitmap rows (-1: none)
  int vn0 = 0, vn1 = 0;
  auto add_adj = [&](int b) {
    const int t = pick(b);
    int len;
    const int *p = adj_list(P, t, seg, len);
    if (own(b)) {
      const int h = P.hub_idx[t];
      if (L == 0 && b == 1) {
        vp1 = p;
        vh1 = h;
        vn1 = len;
        hasB = true;
      } else {
        vp0 = p;
        vh0 = h;
        vn0 = len;
        hasA = true;
      }
    } else {
      S.ulp[nu] = p;
      S.uhb[nu] = hub_row(P, t);
      S.uln[nu] = len;
      nu++;
    }
  };
  auto add_list = [&](const int *p, int len) {
    S.ulp[nu] = p;
    S.uhb[nu] = nullptr;
    S.uln[nu] = len;
    nu++;
  };
  if (hoisted) {
    add_list(hbase + (long long)(u - 3) * P.dstride, S.hl[u]);
    if (pl->hoist[ri][u] == 1) add_adj(u - 1);
  } else if (dep >= 0) {
    add_list(lbase + (long long)(dep - 2) * P.dstride, S.len[dep - 1]);
    const int nsh = pl->num_bn_sh[u];
    for (int i = 0; i < nsh; i++) add_adj(pl->bn_sh[u][i]);
  } else {
    const int nb = pl->num_bn[u];
    for (int i = 0; i < nb; i++) add_adj(pl->bn[u][i]);
  }
  const int nl = nu + (hasA ? 1 : 0) + (hasB ? 1 : 0);
  // One shared input U and the owner's list.  The window of U (bounds of the
  // path before the owner) goes into this wave's LDS table once per path
  // (kept while the path is unchanged): as a bitmap when its elements span at
  // most 32 kHash ids (the dense core of a degree-ordered power-law graph
  // sits in one such span), else as a hash table when it holds at most
  // kHashMax elements.  Membership in U is then an LDS lookup instead of a
  // chain of dependent global loads, and with the bitmap an owner that has a
  // neighbour bitmap intersects 32 ids per load (word AND + popcount).
  // The decision, the window and the table depend on the path only: they
  // are made by the path's first batch and kept in LDS for the others
  // (tab_kind 4: no table for this path).
  bool use_hash = false, use_bits = false;
  int hash_n = 0, bbase = 0, bend = 0;
  if (P.leaf_hash && L > 0 && nu == 1 && hasA && !hasB) {
    const int *U = S.ulp[0];
    int kind = 0;
    if (S.tab_epoch == S.epoch && S.tab_list == U) kind = S.tab_kind;
    if (kind == 1 || kind == 2 || kind == 4) {
      use_hash = kind == 1;
      use_bits = kind == 2;
      hash_n = S.tab_hi - S.tab_lo;
      bbase = S.tab_base;
      bend = S.tab_end;
    } else {
      const int gt0 = pl->gt_mask[u], lt0 = pl->lt_mask[u];
      int ulo = -1, uhi = INT_MAX;
#pragma unroll
      for (int j = 0; j < 6; j++) {
        if (j < u - 1 && !labeled) {
          if ((gt0 >> j) & 1) ulo = max(ulo, S.M[j]);
          else if ((lt0 >> j) & 1) uhi = min(uhi, S.M[j]);
        }
      }
      const int nU = S.uln[0];
      // U of at most 64 elements is loaded once (window by ballots, ends and
      // bitmap from the lanes)
      const bool smallU = nU <= kWave;
      const int yU = smallU && lane < nU ? U[lane] : INT_MAX;
      int us, ue;
      if (smallU) {
        us = ulo >= 0 ? __popcll(__ballot(lane < nU && yU <= ulo)) : 0;
        ue = uhi != INT_MAX ? __popcll(__ballot(lane < nU && yU < uhi)) : nU;
      } else {
        us = ulo >= 0 ? lower_bound_wave(U, nU, ulo + 1) : 0;
        ue = uhi != INT_MAX ? lower_bound_wave(U, nU, uhi) : nU;
      }
      kind = 4;
      if (P.leaf_hash & 2) {
        int words = 0;
        bbase = 0;
        if (ue > us) {
          const int first = smallU ? __shfl(yU, us) : U[us];
          const int last = smallU ? __shfl(yU, ue - 1) : U[ue - 1];
          bbase = first & ~31;
          words = ((last - bbase) >> 5) + 1;
        }
        // dense enough that a word holds an element per 128 ids on average (a
        // sparse U is cheaper hashed, or probed element by element)
        if (words <= kHash && words <= 4 * (ue - us) + 1) {
          use_bits = true;
          kind = 2;
          bend = bbase + 32 * words;
          lds_uint *bits = lds_u(S.htab);
          for (int i = lane; i < words; i += kWave) bits[i] = 0u;
          __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
          const bool lab_test = labeled && seg < 0;  // U may hold other labels (no segments)
          const int lu = pl->label[u];
          auto set_bit = [&](int y) {
            if (!lab_test || (int)P.labels[y] == lu)
              __hip_atomic_fetch_or(bits + ((y - bbase) >> 5), 1u << (y & 31), __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_WAVEFRONT);
          };
          if (smallU) {
            if (lane >= us && lane < ue) set_bit(yU);  // lane i holds U[i]
          } else {
            for (int idx = us + lane; idx < ue; idx += kWave) set_bit(U[idx]);
          }
          __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
          // injectivity: no path vertex before the owner is a candidate
          if (lane < u - 1) {
            const int m = S.M[lane];
            if (m >= bbase && m < bend)
              __hip_atomic_fetch_and(bits + ((m - bbase) >> 5), ~(1u << (m & 31)), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_WAVEFRONT);
          }
          __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        }
      }
      // a hash is worth it when most parents' own lists are short next to U
      // (else U is the better probe: bitmap or binary search in each
      // parent's list); judged on the path's first batch
      const int nact = __popcll(__ballot(act));
      const int nshort = __popcll(__ballot(act && vn0 <= 4 * (ue - us)));
      if (!use_bits && (P.leaf_hash & 1) && ue - us <= kHashMax && 2 * nshort >= nact) {
        use_hash = true;
        kind = 1;
        hash_n = max(0, ue - us);
        hash_build(lds_i(S.htab), U + us, ue - us, kHashLog);
      }
      if (lane == 0) {
        S.tab_epoch = S.epoch;
        S.tab_kind = kind;
        S.tab_list = U;
        S.tab_lo = us;
        S.tab_hi = ue;
        S.tab_base = bbase;
        S.tab_end = bend;
      }
    }
  }
  if (use_bits) {
    // Each owner x counts the elements of U's window in [a, b): the bitmap's
    // span cut by x's own bound (GT: above x, LT: below x).  Units: the
    // bitmap words of [a, b) when x has a neighbour bitmap and they are no
    // more than x's list elements (row word AND LDS word, popcount), else the
    // elements of x's list in [a, b) (one LDS bit each).  The units of all
    // owners are flattened as in the pair path below.
    taken = 0;
    if (STATS && labeled) bytes += (unsigned)minfull;
    const int ju = u - 1;
    const bool c_gt = !labeled && ((pl->gt_mask[u] >> ju) & 1);
    const bool c_lt = !labeled && ((pl->lt_mask[u] >> ju) & 1);
    int a = bbase, b = bend;
    if (c_gt) a = max(a, xa + 1);
    if (c_lt) b = min(b, xa);
    if (!act || b < a) b = a;
    const int nw = b > a ? ((b - 1) >> 5) - (a >> 5) + 1 : 0;
    const bool wm = act && vh0 >= 0 && nw <= vn0;
    int st = 0, cnt = 0;
    if (wm) {
      st = a >> 5;
      cnt = nw;
    } else if (act && b > a) {
      if (vn0 > kWave) {
        st = lower_bound_i(vp0, vn0, a);
        cnt = lower_bound_i(vp0, vn0, b) - st;
      } else {
        cnt = vn0;  // short list: bounds tested per element
      }
    }
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int t = __shfl_up(incl, d);
      if (lane >= d) incl += t;
    }
    int tk = __popcll(__ballot(act && incl <= P.leaf_pair_cap));
    if (tk == 0) tk = 1;
    taken = tk;
    const bool mine = lane < tk;
    if (!mine) cnt = 0;
    const int T = __shfl(incl, tk - 1);
    if (STATS) {
      balg += wave_sum(mine ? bytes : 0ull);
      next += (unsigned)tk;
      count_op<STATS>(S, 2, (unsigned)T);
      count_op<STATS>(S, 3, clock64() - t_setup);
    }
    const lds_uint *bits = lds_u(S.htab);
    const int bw = bbase >> 5;
    unsigned long long found = 0;
    for (int q0 = 0; q0 < T; q0 += kWave) {
      const int q = q0 + lane;
      int o = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1) {
        const int w = __shfl(incl, o + step - 1);
        if (w <= q) o += step;
      }
      if (q >= T) o = 0;
      const int i = q - (__shfl(incl, o) - __shfl(cnt, o));
      const int ao = __shfl(a, o), bo = __shfl(b, o), so = __shfl(st, o);
      const bool wmo = __shfl((int)wm, o) != 0;
      const int ho = __shfl(vh0, o);
      const int *lpo = reinterpret_cast<const int *>(__shfl(reinterpret_cast<unsigned long long>(vp0), o));
      if (q < T) {
        if (wmo) {
          const int w = so + i;
          unsigned v = P.hub_bits[(long long)ho * P.hub_words + w] & bits[w - bw];
          const int s = ao - 32 * w, e = bo - 32 * w;
          if (s > 0) v &= ~0u << s;
          if (e < 32) v &= (1u << e) - 1u;
          found += (unsigned)__popc(v);
        } else {
          const int y = lpo[so + i];
          if (y >= ao && y < bo && ((bits[(y >> 5) - bw] >> (y & 31)) & 1u)) found++;
        }
      }
    }
    return wave_sum(found);
  }
  if (!use_hash && L > 0 && nu == 1 && hasA && !hasB && S.uln[0] <= 2 * kWave) {
    // One shared input U (C(dep), H(u) or N(M(b)), b <= L) of at most 128
    // elements and the owner's N(M(L+1)).  U is the probe for every owner
    // (set semantics: the probe does not change C(u)): lane j holds U[j mod
    // g] (and U[64 + j] when |U| > 64), filtered once against everything but
    // the owner's vertex, and 64 / g owners are tested per step (g = |U|
    // rounded up to a power of two, at most 64) -- no pair scan, no owner
    // search.
    const int *U = S.ulp[0];
    const int nU = S.uln[0];
    taken = i1 - i0;
    if (STATS) {
      if (labeled) bytes += (unsigned)minfull;
      balg += wave_sum(act ? bytes : 0ull);
      next += (unsigned)__popcll(__ballot(act));
    }
    const int ju = u - 1;  // the owner's vertex
    const int gt = pl->gt_mask[u], lt = pl->lt_mask[u];
    const int same = (labeled && seg >= 0) ? pl->same_label_mask[u] : 0x7f;
    const bool c_gt = !labeled && ((gt >> ju) & 1);
    const bool c_lt = !labeled && ((lt >> ju) & 1);
    const bool c_ne = !c_gt && !c_lt && ((same >> ju) & 1);
    auto shared_ok = [&](int y) -> bool {  // conditions against M(0..u-2) and the label
      bool ok = true;
#pragma unroll 1
      for (int i = 0; i < ju; i++) {
        const int m = S.M[i];
        if (!labeled && ((gt >> i) & 1)) ok = ok && y > m;
        else if (!labeled && ((lt >> i) & 1)) ok = ok && y < m;
        else if ((same >> i) & 1) ok = ok && y != m;
      }
      if (labeled && seg < 0 && ok) ok = (int)P.labels[y] == pl->label[u];
      return ok;
    };
    int g = 1;
    while (g < nU && g < kWave) g <<= 1;
    const int per = kWave / g;
    const int j = lane & (g - 1);
    const int slot = lane / g;
    const int y = j < nU ? U[j] : 0;
    const bool yv = j < nU && shared_ok(y);
    const bool two = nU > kWave;  // then g = 64: one owner per step, two elements per lane
    const int y2 = two && kWave + lane < nU ? U[kWave + lane] : 0;
    const bool yv2 = two && kWave + lane < nU && shared_ok(y2);
    const int owners = i1 - i0;
    if (STATS) {
      count_op<STATS>(S, 2, (unsigned long long)owners * (unsigned)nU);
      count_op<STATS>(S, 3, clock64() - t_setup);
    }
    unsigned long long found = 0;
    for (int o0 = 0; o0 < owners; o0 += per) {
      const int o = o0 + slot;
      const bool ov = o < owners;
      const int oo = ov ? o : 0;
      const int x = __shfl(xa, oo);
      const int h = __shfl(vh0, oo);
      const int *lp = reinterpret_cast<const int *>(__shfl(reinterpret_cast<unsigned long long>(vp0), oo));
      const int ln = __shfl(vn0, oo);
      auto test = [&](int yy, bool v) -> bool {
        bool ok = v && ov;
        if (c_gt) ok = ok && yy > x;
        if (c_lt) ok = ok && yy < x;
        if (c_ne) ok = ok && yy != x;
        if (ok) ok = h >= 0 ? bit_test(P.hub_bits + (long long)h * P.hub_words, yy) : contains(lp, ln, yy);
        return ok;
      };
      found += (unsigned)__popcll(__ballot(test(y, yv)));
      if (two) found += (unsigned)__popcll(__ballot(test(y2, yv2)));
    }
    return found;
  }
  // probe = shortest input: index i < nu (uniform) or 8 / 9 (own)
  int pidx = 0, pn = INT_MAX;
  for (int i = 0; i < nu; i++) {
    const int l = S.uln[i];
    if (l < pn) {
      pn = l;
      pidx = i;
    }
  }
  const int *pp = nu > 0 ? S.ulp[pidx] : nullptr;
  if (hasA && vn0 < pn) {
    pn = vn0;
    pidx = 8;
    pp = vp0;
  }
  if (hasB && vn1 < pn) {
    pn = vn1;
    pidx = 9;
    pp = vp1;
  }
  if (use_hash && pidx != 8 && vn0 <= 4 * hash_n) {  // with U hashed, the owner's list is the cheaper probe
    pn = vn0;
    pidx = 8;
    pp = vp0;
  }
  if (STATS && labeled) bytes += (unsigned)minfull;
  const bool label_check = labeled && seg < 0;
  int lo = -1, hi = INT_MAX;
  int nev[6];
  int nemask = 0;
  const int gt = pl->gt_mask[u], lt = pl->lt_mask[u];
  const int same = (labeled && seg >= 0) ? pl->same_label_mask[u] : 0x7f;
#pragma unroll
  for (int j = 0; j < 6; j++) {
    nev[j] = -1;
    if (j < u) {
      const int m = pick(j);
      if (!labeled && ((gt >> j) & 1)) lo = max(lo, m);
      else if (!labeled && ((lt >> j) & 1)) hi = min(hi, m);
      else if ((same >> j) & 1) {
        nev[j] = m;
        nemask |= 1 << j;
      }
    }
  }
  int start = 0, stop = 0;
  if (act) {
    stop = pn;
    if (lo >= 0) start = lower_bound_i(pp, pn, lo + 1);
    if (hi != INT_MAX) stop = lower_bound_i(pp, pn, hi);
    if (stop < start) stop = start;
  }
  // every probe element carries l(u): segments, C(dep) (same label as u) or H(u)
  const bool label_known = !labeled || seg >= 0 || dep >= 0 || hoisted;
  if (nl == 1 && label_known) {
    // |list ∩ (lo,hi)| minus excluded path vertices inside the window
    taken = i1 - i0;
    if (STATS) {
      balg += wave_sum(act ? bytes : 0ull);
      next += (unsigned)__popcll(__ballot(act));
    }
    unsigned c = (unsigned)(stop - start);
    if (act) {
#pragma unroll
      for (int j = 0; j < 6; j++) {
        if ((nemask >> j) & 1) {
          const int v = nev[j];
          if (v > lo && v < hi && stop > start) {
            const int pos = lower_bound_i(pp + start, stop - start, v);
            if (pos < stop - start && pp[start + pos] == v) c--;
          }
        }
      }
    }
    return wave_sum(act ? (unsigned long long)c : 0ull);
  }
  int s = stop - start;
  int incl = s;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int t = __shfl_up(incl, d);
    if (lane >= d) incl += t;
  }
  // Take the longest prefix of owners whose pairs fit in leaf_pair_cap (at
  // least one owner), so that a batch stays short enough for the donation
  // check between batches to keep the tail balanced.
  int tk = __popcll(__ballot(act && incl <= P.leaf_pair_cap));
  if (tk == 0) tk = 1;
  taken = tk;
  const bool mine = lane < tk;
  if (!mine) s = 0;
  const int T = __shfl(incl, tk - 1);
  if (STATS) {
    balg += wave_sum(mine ? bytes : 0ull);
    next += (unsigned)tk;
    count_op<STATS>(S, 2, (unsigned)T);
    count_op<STATS>(S, 3, clock64() - t_setup);
  }
  const int lab = pl->label[u];
  unsigned long long found = 0;
  for (int q0 = 0; q0 < T; q0 += kWave) {
    const int q = q0 + lane;
    const bool v = q < T;
    int o = 0;
#pragma unroll
    for (int step = 32; step >= 1; step >>= 1) {
      const int w = __shfl(incl, o + step - 1);
      if (w <= q) o += step;
    }
    if (!v) o = 0;
    const int excl_o = __shfl(incl, o) - __shfl(s, o);
    const int *ppo = reinterpret_cast<const int *>(__shfl(reinterpret_cast<unsigned long long>(pp), o));
    const int st = __shfl(start, o);
    const int po = __shfl(pidx, o);
    bool ok = v;
    int y = 0;
    if (ok) y = ppo[st + (q - excl_o)];
#pragma unroll
    for (int j = 0; j < 6; j++) {
      if ((nemask >> j) & 1) {
        const int ne = own(j) ? __shfl(nev[j], o) : nev[j];
        ok = ok && (y != ne);
      }
    }
    if (label_check && ok) ok = (int)P.labels[y] == lab;
#pragma unroll 1
    for (int i = 0; i < nu; i++) {
      if (i != po && ok) {
        const unsigned *h = S.uhb[i];
        ok = use_hash ? hash_has(lds_i(S.htab), y) : (h ? bit_test(h, y) : contains(S.ulp[i], S.uln[i], y));
      }
    }
    if (hasA) {
      const int *p = reinterpret_cast<const int *>(__shfl(reinterpret_cast<unsigned long long>(vp0), o));
      const int h = __shfl(vh0, o);
      const int n = __shfl(vn0, o);
      if (po != 8 && ok) ok = h >= 0 ? bit_test(P.hub_bits + (long long)h * P.hub_words, y) : contains(p, n, y);
    }
    if (hasB) {
      const int *p = reinterpret_cast<const int *>(__shfl(reinterpret_cast<unsigned long long>(vp1), o));
      const int h = __shfl(vh1, o);
      const int n = __shfl(vn1, o);
      if (po != 9 && ok) ok = h >= 0 ? bit_test(P.hub_bits + (long long)h * P.hub_words, y) : contains(p, n, y);
    }
    found += (unsigned)__popcll(__ballot(ok));
  }
  return found;
}

// Leaf counts of the owners [i0, i1) of stack level L = k-3 (L >= 1) of a
// plan whose leaf inputs do not involve the owner's vertex M(k-2) = x
// (leaf_free).  Then C(k-1) = I minus what x itself excludes, where I is the
// intersection of the leaf's inputs over the window, labels and injectivity
// of M(0..k-3) -- the same for every owner.  I is computed once per call:
// as a window of one input list less the path vertices found in it (one
// input whose elements all carry the leaf's label), or materialised into the
// unused H(k-1) slot by extend (kMemo).  Each owner then costs one binary
// search: elements of I above (GT) or below (LT) x, or I less x itself when
// x falls under the leaf's injectivity (NE) and lies in I.  Same counts and
// statistics as leaf_batch.
template <bool STATS, bool MERGE>
__device__ __forceinline__ unsigned long long leaf_memo(const KParams &P, const DevPlan *__restrict__ pl, WaveLds &S,
                                                        int L, int i0, int i1, const int *lbase, int *hbase,
                                                        unsigned long long &balg, unsigned long long &next) {
  count_op<STATS>(S, 7, 1);
  const int lane = lane_id();
  const int u = pl->k - 1;
  const int ju = u - 1;  // the owner's vertex
  const bool labeled = pl->labeled != 0;
  const int ri = P.reuse ? 0 : 1;
  const int dep = P.reuse ? pl->dep[u] : -1;
  const bool hoisted = S.hv[u] != 0;  // then C(k-1) = H(k-1) (the owner is no input)
  const int seg = (labeled && P.colidx_lab) ? pl->label[u] : -1;
  const int owners = i1 - i0;
  if (STATS) {
    // per owner, the merge-model bytes of the reference's inputs (all fixed by the path)
    unsigned long long bytes = 0;
    int minfull = INT_MAX;
    if (dep >= 0) {
      const int l = S.len[dep - 1];
      bytes += 4ull * (unsigned)l;
      minfull = l;
      for (int i = 0; i < pl->num_bn_sh[u]; i++) {
        const int d = P.degv[S.M[pl->bn_sh[u][i]]];
        bytes += 4ull * (unsigned)d + 16;
        minfull = min(minfull, d);
      }
    } else {
      for (int i = 0; i < pl->num_bn[u]; i++) {
        const int d = P.degv[S.M[pl->bn[u][i]]];
        bytes += 4ull * (unsigned)d + 16;
        minfull = min(minfull, d);
      }
    }
    if (labeled) bytes += (unsigned)minfull;
    balg += bytes * (unsigned long long)owners;
    next += (unsigned long long)owners;
  }
  // the leaf's relation to the owner
  const int gt = pl->gt_mask[u], lt = pl->lt_mask[u];
  const int same = (labeled && seg >= 0) ? pl->same_label_mask[u] : 0x7f;
  const bool c_gt = !labeled && ((gt >> ju) & 1);
  const bool c_lt = !labeled && ((lt >> ju) & 1);
  const bool c_ne = !c_gt && !c_lt && ((same >> ju) & 1);
  // I as a sorted list base[0, len) less `nex` excluded path vertices ex[] that occur in it
  const int *base = nullptr;
  int len = 0;
  int ex[6];    // path vertices excluded from I (-1: none), at fixed positions
  int nex = 0;
#pragma unroll
  for (int j = 0; j < 6; j++) ex[j] = -1;
  int lo = -1, hi = INT_MAX;
#pragma unroll
  for (int j = 0; j < 6; j++) {
    if (j < ju) {
      const int m = S.M[j];
      if (!labeled && ((gt >> j) & 1)) lo = max(lo, m);
      else if (!labeled && ((lt >> j) & 1)) hi = min(hi, m);
    }
  }
  // I less its excluded vertices, staged ascending in this wave's LDS table
  // when it fits (searched there, and kept for the later calls of the same
  // path: I depends on M(0..k-3) only)
  // The leaf's input is one raw adjacency list N(M(b0)) (no reuse, no
  // hoisting): the plan then tells which path vertices lie in it in every
  // embedding (adjacent to b0) -- no search for them -- and an owner test
  // "x in I" is a window compare (owner adjacent to b0) or, with b0's
  // vertex a hub, one bitmap load.
  const bool raw = !hoisted && dep < 0 && pl->num_bn[u] == 1 && (!labeled || seg >= 0);
  const bool owner_in = raw && ((pl->leaf_in >> ju) & 1);
  // owners need no search in I
  const bool fast_ne = c_ne && (owner_in || (raw && P.hub_idx[uni(S.M[pl->bn[u][0]])] >= 0));
  const bool staged = S.tab_kind == 3 && S.tab_epoch == S.epoch;
  if (!staged) {
    int nin = 0;  // inputs of the leaf
    int single_len = 0;
    const int *single = nullptr;
    if (hoisted) {
      single = hbase + (long long)(u - 3) * P.dstride;
      single_len = S.hl[u];
      nin = 1;
    } else if (dep >= 0) {
      single = lbase + (long long)(dep - 2) * P.dstride;
      single_len = S.len[dep - 1];
      nin = 1 + pl->num_bn_sh[u];
    } else {
      nin = pl->num_bn[u];
      if (nin == 1) single = adj_list(P, S.M[pl->bn[u][0]], seg, single_len);
    }
    const bool label_known = !labeled || seg >= 0 || dep >= 0 || hoisted;
    if (nin == 1 && label_known) {
      int start = 0, stop = single_len;
      if (lo >= 0) start = lower_bound_wave(single, single_len, lo + 1);
      if (hi != INT_MAX) stop = lower_bound_wave(single, single_len, hi);
      if (stop < start) stop = start;
      base = single + start;
      len = stop - start;
#pragma unroll
      for (int j = 0; j < 6; j++) {
        if (j < ju && ((same >> j) & 1) && !(!labeled && (((gt | lt) >> j) & 1))) {
          const int v = S.M[j];
          if (v > lo && v < hi) {
            bool in;
            if (raw && ((pl->leaf_in >> j) & 1)) {
              in = true;  // adjacent to b0 in the pattern
            } else if (raw && pl->bn[u][0] == j) {
              in = false;  // b0 itself: no self-loops
            } else {
              const int pos = lower_bound_wave(base, len, v);
              in = pos < len && base[pos] == v;
            }
            if (in) {
              ex[j] = v;
              nex++;
            }
          }
        }
      }
    } else {
      int *buf = hbase + (long long)(u - 3) * P.dstride;  // H(k-1) is not built: its slot is free
      unsigned long long db = 0, dn = 0;
      len = (int)extend<false, MERGE>(P, pl, S, u, buf, lbase, hbase, kMemo, false, db, dn);
      base = buf;
    }
    if (len - nex <= kHash && !fast_ne) {
      lds_int *tab = lds_i(S.htab);
      int n = 0;
      for (int b0 = 0; b0 < len; b0 += kWave) {
        const int i = b0 + lane;
        const int y = i < len ? base[i] : 0;
        bool keep = i < len;
#pragma unroll
        for (int j = 0; j < 6; j++) keep = keep && y != ex[j];
        const unsigned long long m = __ballot(keep);
        if (keep) tab[n + (int)mbcnt64(m)] = y;
        n += __popcll(m);
      }
      __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
      if (lane == 0) {
        S.tab_kind = 3;
        S.tab_epoch = S.epoch;
        S.tab_lo = n;
      }
    }
  }
  // no condition ties the leaf to its owner: every owner counts all of I
  if (!c_gt && !c_lt && !c_ne) return (unsigned long long)(staged ? S.tab_lo : len - nex) * (unsigned)owners;
  unsigned long long found = 0;
  if (fast_ne && !staged) {
    // NE only: x is in I when it lies in the window and in N(tv)
    const int nI = len - nex;
    if (owner_in && lo < 0 && hi == INT_MAX) return (unsigned long long)(nI - 1) * (unsigned)owners;
    const unsigned *trow = owner_in ? nullptr : hub_row(P, uni(S.M[pl->bn[u][0]]));
    for (int o0 = i0; o0 < i1; o0 += kWave) {
      const int idx = o0 + lane;
      if (idx < i1) {
        const int x = lbase[(long long)(L - 1) * P.dstride + idx];
        const bool inI = x > lo && x < hi && (owner_in || bit_test(trow, x));
        found += (unsigned)(nI - (inI ? 1 : 0));
      }
    }
    return wave_sum(found);
  }
  if (staged || len - nex <= kHash) {
    const lds_int *tab = lds_i(S.htab);
    const int nI = staged ? S.tab_lo : len - nex;
    for (int o0 = i0; o0 < i1; o0 += kWave) {
      const int idx = o0 + lane;
      if (idx < i1) {
        const int x = lbase[(long long)(L - 1) * P.dstride + idx];
        int c = nI;
        if (c_gt || c_lt || (c_ne && x > lo && x < hi)) {
          int p0 = 0, m = nI;  // first element >= x (GT: > x)
          const int key = c_gt ? x + 1 : x;
          while (m > 0) {
            const int h = m >> 1;
            if (tab[p0 + h] < key) {
              p0 += h + 1;
              m -= h + 1;
            } else {
              m = h;
            }
          }
          if (c_gt) c = nI - p0;
          else if (c_lt) c = p0;
          else if (p0 < nI && tab[p0] == x) c--;  // x itself is no candidate (path vertices are never owners)
        }
        found += (unsigned)c;
      }
    }
  } else {
    const int nI = len - nex;
    for (int o0 = i0; o0 < i1; o0 += kWave) {
      const int idx = o0 + lane;
      if (idx < i1) {
        const int x = lbase[(long long)(L - 1) * P.dstride + idx];
        int c = nI;
        if (c_gt || c_lt) {
          const int pos = lower_bound_i(base, len, c_gt ? x + 1 : x);  // elements below the bound
          c = c_gt ? len - pos : pos;
#pragma unroll
          for (int j = 0; j < 6; j++)
            if (ex[j] >= 0 && (c_gt ? ex[j] > x : ex[j] < x)) c--;
        } else if (c_ne && x > lo && x < hi) {
          if (contains(base, len, x)) c--;  // x itself is no candidate (path vertices are never owners)
        }
        found += (unsigned)c;
      }
    }
  }
  return wave_sum(found);
}

// Level-0 filter over directed edges [e0, e1) (src/gpu_match.cu:526-558).
// "exists v in N(c) with deg(v) >= dP(2)" is maxnbr[c] >= dP(2).
template <bool STATS>
__device__ __forceinline__ int filter_edges(const KParams &P, const DevPlan *__restrict__ pl, long long e0, long long e1,
                            int *L0r, int *L0c, unsigned long long &balg) {
  const int lane = lane_id();
  const bool labeled = pl->labeled != 0;
  const int sym01 = pl->sym01;
  const int l0 = pl->label[0], l1 = pl->label[1];
  const int d0 = pl->deg[0], d1 = pl->deg[1], d2 = pl->deg[2];
  unsigned cnt = 0, pre = 0;
  for (long long base = e0; base < e1; base += kWave) {
    const long long e = base + lane;
    bool ok = e < e1;
    int r = 0, c = 0;
    if (ok) {
      r = P.src[e];
      c = P.colidx[e];
      if (!labeled) ok = !(sym01 && r >= c);
      else ok = (int)P.labels[r] == l0 && (int)P.labels[c] == l1;
    }
    if (STATS) pre += (unsigned)__popcll(__ballot(ok));
    if (ok) ok = P.degv[r] >= d0 && P.degv[c] >= d1 && P.maxnbr[c] >= d2;
    const unsigned long long mask = __ballot(ok);
    const unsigned pos = cnt + mbcnt64(mask);
    if (ok) {
      L0r[pos] = r;
      L0c[pos] = c;
    }
    cnt += (unsigned)__popcll(mask);
  }
  if (STATS) balg += 8ull * (unsigned long long)(e1 - e0) + 16ull * pre;
  return (int)cnt;
}

// ---------------------------------------------------------------------------
// Task ring (multi-producer / multi-consumer, bounded, HBM atomics).
// Producers reserve space (q_count), take consecutive positions (q_tail),
// write the payload and publish the slot sequence number pos+1, then make the
// items claimable (q_avail).  Consumers claim an item (q_avail), take a
// position (q_head), wait for its sequence number, read it, and release the
// slot for the next lap (sequence pos + capacity).
// ---------------------------------------------------------------------------

// Reserves n slots; returns the first position or ~0ull when full.  Lane 0
// performs the accounting; the result is wave-uniform.
__device__ __forceinline__ unsigned long long ring_reserve(const KParams &P, int n) {
  unsigned long long pos = ~0ull;
  if (lane_id() == 0) {
    Ctrl *C = P.ctrl;
    const long long cap = (long long)P.q_mask + 1;
    const long long old = fadd(&C->q_count, (long long)n);
    if (old + n > cap) {
      fadd(&C->q_count, -(long long)n);
    } else {
      pos = fadd(&C->q_tail, (unsigned long long)n);
    }
  }
  return uni64(pos);
}

// Lane-parallel publication of one slot (called by the lane that owns it).
__device__ __forceinline__ bool ring_publish(const KParams &P, unsigned long long pos, const TaskRec &t) {
  unsigned long long *seq = P.q_seq + (pos & P.q_mask);
  long long spins = 0;
  while (ld_acq(seq) != pos) {
    if (++spins > kSpinMax) return false;
    __builtin_amdgcn_s_sleep(1);
  }
  P.q_data[pos & P.q_mask] = t;
  st_rel(seq, pos + 1);
  return true;
}

// Claims and reads one task.  Wave-uniform result.
__device__ __forceinline__ bool ring_take(const KParams &P, TaskRec &t) {
  Ctrl *C = P.ctrl;
  int got = 0;
  int err = 0;
  if (lane_id() == 0) {
    if (ld_rlx(&C->q_avail) > 0) {
      const long long old = fadd(&C->q_avail, -1ll);
      if (old <= 0) {
        fadd(&C->q_avail, 1ll);
      } else {
        const unsigned long long pos = fadd(&C->q_head, 1ull);
        unsigned long long *seq = P.q_seq + (pos & P.q_mask);
        long long spins = 0;
        while (ld_acq(seq) != pos + 1) {
          if (++spins > kSpinMax) {
            err = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        if (!err) {
          t = P.q_data[pos & P.q_mask];
          st_rel(seq, pos + P.q_mask + 1);
          fadd(&C->q_count, -1ll);
          got = 1;
        }
      }
    }
    if (err) __hip_atomic_store(&C->error, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  got = uni(got);
  if (got) {
    t.pid = uni(t.pid);
    t.p = uni(t.p);
    t.a = uni(t.a);
    t.b = uni(t.b);
    t.fresh = uni(t.fresh);
    t.mul = uni64(t.mul);
#pragma unroll
    for (int j = 0; j < 8; j++) t.M[j] = uni(t.M[j]);
  }
  return got != 0;
}

// Donates the untaken iterations of the shallowest open level in
// [floor, level], cut into about 2*idle pieces (at least split_chunk
// iterations each, at most 256 pieces).  Ring slots are reserved first, then
// the iterations are taken from the level (a sibling may have stolen part of
// them meanwhile: slots left over carry empty tasks, p = 0).  The stand-in
// level of a free vertex (fmask) is never donated; a task under free levels
// being measured (mmask, levels below the donated one) carries umul times
// 1 + the other candidates of each.  Returns the number of tasks pushed (0
// when the ring is full).  All lanes call it.
template <bool STEAL>
__device__ __forceinline__ int donate(const KParams &P, WaveLds &S, int pid, int floor, int level, const int *L0r,
                      const int *L0c, int idle, int fmask, int mmask, unsigned long long umul) {
  constexpr bool shared = STEAL;
  int ld = -1, rem = 0;
  for (int l = floor; l <= level; l++) {
    if ((fmask >> (l + 1)) & 1) continue;
    const unsigned long long w = ie_read(S, l);
    if (ie_it(w) < ie_end(w)) {
      ld = l;
      rem = ie_end(w) - ie_it(w);
      break;
    }
  }
  if (ld < 0) return 0;
  unsigned long long mul = umul;
  for (int m = floor; m < ld; m++)
    if ((mmask >> m) & 1) mul *= 1ull + (unsigned)uni(S.mrest[m]);
  int piece = 1, ntasks = rem;
  if (ld > 0) {
    const int want = max(1, min(256, 2 * idle));
    piece = max(P.split_chunk, (rem + want - 1) / want);
    ntasks = (rem + piece - 1) / piece;
  } else {
    ntasks = min(rem, 256);
  }
  const unsigned long long pos0 = ring_reserve(P, ntasks);
  if (pos0 == ~0ull) return 0;
  // the tasks are in flight before any of them can be taken
  if (P.pend && lane_id() == 0) fadd_done(&P.pend[(long long)pid * 16], (unsigned long long)ntasks);
  // level > 0: every untaken iteration, in pieces; level 0: the last edge pairs, one task each
  int got = 0;
  const int a0 = level_take_back(S, ld, ld > 0, ntasks, got, shared);
  const int real = ld > 0 ? (got + piece - 1) / piece : got;
  bool okall = true;
  for (int j = lane_id(); j < ntasks; j += kWave) {
    TaskRec t;
    t.pid = pid;
    t.pad0 = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) t.M[q] = 0;
    t.a = t.b = 0;
    t.fresh = 0;
    t.mul = mul;
    if (j >= real) {
      t.p = 0;  // empty task
    } else if (ld == 0) {
      t.p = 2;
      t.M[0] = L0r[a0 + j];
      t.M[1] = L0c[a0 + j];
      t.a = 0;
      t.b = INT_MAX;
      t.fresh = 1;
    } else {
      t.p = ld + 1;
#pragma unroll
      for (int q = 0; q < 8; q++)
        if (q <= ld) t.M[q] = S.M[q];
      t.a = a0 + j * piece;
      t.b = min(t.a + piece, a0 + got);
    }
    okall = ring_publish(P, pos0 + j, t) && okall;
  }
  if (__ballot(!okall) != 0ull) {
    if (lane_id() == 0)
      __hip_atomic_store(&P.ctrl->error, 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (lane_id() == 0) fadd(&P.ctrl->q_avail, (long long)ntasks);
  return real;
}

// After M(0..fhi) are fixed and the n candidates of the next level are
// known: for every u >= umin whose H(u) depends on M(0..f), flo <= f <= fhi,
// builds H(u) when n reaches hoist_min (else marks it unbuilt).  Returns false
// when a built H(u) is empty: then no C(u) below this prefix is non-empty
// (the timed kernel skips the subtree; the statistics kernel walks it).
template <bool STATS, bool MERGE = false>
__device__ __forceinline__ bool prepare_hoists(const KParams &P, const DevPlan *__restrict__ pl, WaveLds &S, int flo,
                                               int fhi, int umin, int n, const int *lbase, int *hbase) {
  const int ri = P.reuse ? 0 : 1;
  bool live = true;
  for (int u = umin; u < pl->k; u++) {
    const int f = pl->hfix[ri][u];
    if (!pl->hoist[ri][u] || f < flo || f > fhi) continue;
    int hv = 0, hn = 0;
    if (P.hoist_min > 0 && n >= P.hoist_min) {
      unsigned long long db = 0, dn = 0;
      hn = (int)extend<STATS, MERGE>(P, pl, S, u, hbase + (long long)(u - 3) * P.dstride, lbase, hbase, kBuildH, false, db,
                              dn);
      hv = 1;
      if (hn == 0) live = false;
    }
    S.hv[u] = hv;
    S.hl[u] = hn;
  }
  return STATS || live;
}

// Marks every H(u) unbuilt (start of a task: the plan or prefix changed).
__device__ __forceinline__ void clear_hoists(WaveLds &S) {
#pragma unroll
  for (int u = 0; u < 8; u++) S.hv[u] = 0;
}

// Completions of the dropped leaf of a factored plan (index k of a k-vertex
// plan) given M(0..la): its candidates, less the later vertices that always
// lie among them and must differ from it.  Vertices after la stand as -1.
template <bool MERGE>
__device__ __forceinline__ long long leaf_factor_count(const KParams &P, const DevPlan *__restrict__ pl, WaveLds &S,
                                                       int la, const int *lbase, int *hbase,
                                                       unsigned long long &balg, unsigned long long &next) {
  if (lane_id() == 0)
    for (int j = la + 1; j < kMaxK; j++) S.M[j] = -1;
  const long long c =
      (long long)extend<false, MERGE>(P, pl, S, pl->k, nullptr, lbase, hbase, kFull, false, balg, next);
  return c - pl->lf_excl;
}

// The completions of a path by the plan's tail, which starts at u0: the
// product over the tail's labels (one group when unlabeled) of the ordered
// tuples (labeled) or subsets (unlabeled) of distinct candidates of the
// group's first vertex, as many as the group has vertices.  M(0..ufix) are
// fixed; the vertices after ufix (a free vertex between, the tail's own)
// stand as -1, which matches no candidate.  mode0: the extension mode of u0
// (H(v) of a later tail vertex is this path's only when built on vertices
// before ufix; one built on M(ufix) is a previous sibling's).  Returns false,
// leaving f undefined, when a product would not fit 64 bits.
// With gx >= 0 (the first tail vertex of one label), f1 receives the same
// product with one candidate of gx's group taken away.
template <bool MERGE>
__device__ __forceinline__ bool tail_product(const KParams &P, const DevPlan *__restrict__ pl, WaveLds &S, int u0,
                                             int ufix, int mode0, const int *lbase, int *hbase,
                                             unsigned long long &balg, unsigned long long &next,
                                             unsigned long long &f, int gx = -1,
                                             unsigned long long *f1 = nullptr) {
  if (lane_id() == 0)
    for (int j = ufix + 1; j < kMaxK; j++) S.M[j] = -1;
  const int k = pl->k;
  const int ri = P.reuse ? 0 : 1;
  f = 1;
  unsigned long long h = 1;
  for (int v = u0; v < k && (f || h); v++) {
    const int g = pl->tail_gsz[v];
    if (!g) continue;
    const int mv = v == u0 ? mode0 : ((S.hv[v] && pl->hfix[ri][v] < ufix) ? kHoisted : kFull);
    const unsigned long long c =
        (unsigned long long)extend<false, MERGE>(P, pl, S, v, nullptr, lbase, hbase, mv, false, balg, next);
    const unsigned long long lim =
        g == 1 ? ~0ull : (g == 2 ? 3000000000ull : (g == 3 ? 2000000ull : (g == 4 ? 50000ull : 6000ull)));
    if (c > lim) return false;  // c^g must fit 63 bits
    unsigned long long fac = 1;
    if (!pl->labeled)
      for (int i = 2; i <= g; i++) fac *= (unsigned)i;
    auto ff = [&](unsigned long long n) {  // n (n-1) ... (n-g+1), or C(n, g) unlabeled
      unsigned long long x = 1;
      for (int i = 0; i < g; i++) x *= n >= (unsigned)i ? n - i : 0ull;
      return x / fac;
    };
    const unsigned long long x = ff(c);
    const unsigned long long x1 = v == gx ? (c ? ff(c - 1) : 0ull) : x;
    if ((x != 0 && f > ~0ull / x) || (x1 != 0 && h > ~0ull / x1)) return false;
    f *= x;
    h *= x1;
  }
  if (f1) *f1 = gx >= 0 ? h : f;
  return true;
}

// Whether candidate x (this lane's: its neighbour-bitmap row, or its
// adjacency list) and y (hub index hy, or -1) are adjacent: a bit of x's row,
// else a bit of y's row, else a binary search of x's list.
__device__ __forceinline__ bool adjacent_pair(const KParams &P, const unsigned *row, const int *lst, int len, int x,
                                              int y, int hy) {
  if (row) return (row[y >> 5] >> (y & 31)) & 1u;
  if (hy >= 0) return (P.hub_bits[(long long)hy * P.hub_words + (x >> 5)] >> (x & 31)) & 1u;
  return contains(lst, len, y);
}

// Triangles of G[S] for a list S of at most 64 vertices: lane i holds S[i]
// and the mask R_i of the indices j > i with S[j] adjacent to it (from the
// bitmap row of either one, else a search of S[i]'s list), and each triangle
// i < j < z is counted once as a bit of R_i & R_j, j in R_i.
__device__ __forceinline__ long long triangles_small(const KParams &P, const int *Sl, int s) {
  const int lane = lane_id();
  const int x = lane < s ? Sl[lane] : 0;
  const int h = (lane < s && P.hub_idx) ? P.hub_idx[x] : -1;
  const unsigned *row = h >= 0 ? P.hub_bits + (long long)h * P.hub_words : nullptr;
  int len = 0;
  const int *lst = (lane < s && h < 0) ? adj_list(P, x, -1, len) : nullptr;
  unsigned long long R = 0;
  for (int j = 1; j < s; j++) {
    const int y = __shfl(x, j);
    const int hy = __shfl(h, j);
    if (lane < j && adjacent_pair(P, row, lst, len, x, y, hy)) R |= 1ull << j;
  }
  unsigned long long acc = 0;
  for (int j = 1; j < s; j++) {
    const unsigned long long Rj = __shfl(R, j);
    if ((R >> j) & 1ull) acc += (unsigned)__popcll(R & Rj);
  }
  return (long long)wave_sum(acc);
}

// Labeled closing triangle a = u, b = u+1, c = u+2 of three distinct labels
// over one set of earlier neighbours: the embeddings are the triples (x, y,
// z) of candidates of labels l(a), l(b), l(c) over a's inputs that are
// mutually adjacent.  The candidates of l(a) are sa[0..na); those of l(b)
// and l(c) are listed into buf in turn; with at most 64 in all, lane i holds
// one, of group g_i, and the mask R_i of the others adjacent to it (as in
// triangles_small), and each triangle is the bit z of R_x & R_y & C for x in
// A and y in B adjacent to x.  Returns -1 (not counted) when the candidates
// exceed one wave.
template <bool MERGE>
__device__ __forceinline__ long long triangles_labeled(const KParams &P, const DevPlan *__restrict__ pl, WaveLds &S,
                                                       int u, const int *sa, int na, int *buf, const int *lbase,
                                                       int *hbase, unsigned long long &balg,
                                                       unsigned long long &next) {
  if (na == 0) return 0;
  const int lane = lane_id();
  int x = lane < na ? sa[lane] : 0;
  int grp = lane < na ? 0 : 3;
  int s = na;
  int b0 = 0, b1 = 0;
  for (int g = 1; g <= 2; g++) {
    const int nb =
        (int)extend<false, MERGE>(P, pl, S, u, buf, lbase, hbase, kFull, false, balg, next, pl->label[u + g]);
    if (nb == 0) return 0;
    if (s + nb > kWave) return -1;
    if (lane >= s && lane < s + nb) {
      x = buf[lane - s];
      grp = g;
    }
    if (g == 1) {
      b0 = s;
      b1 = s + nb;
    }
    s += nb;
  }
  const int h = (lane < s && P.hub_idx) ? P.hub_idx[x] : -1;
  const unsigned *row = h >= 0 ? P.hub_bits + (long long)h * P.hub_words : nullptr;
  int len = 0;
  const int *lst = (lane < s && h < 0) ? adj_list(P, x, -1, len) : nullptr;
  unsigned long long R = 0;
  for (int j = 0; j < s; j++) {
    const int y = __shfl(x, j);
    const int hy = __shfl(h, j);
    if (lane < s && j != lane && adjacent_pair(P, row, lst, len, x, y, hy)) R |= 1ull << j;
  }
  const unsigned long long Cm = __ballot(grp == 2);
  unsigned long long acc = 0;
  for (int j = b0; j < b1; j++) {
    const unsigned long long Rj = __shfl(R, j);
    if (grp == 0 && ((R >> j) & 1ull)) acc += (unsigned)__popcll(R & Rj & Cm);
  }
  return (long long)wave_sum(acc);
}

// Iterative DFS from stack level `floor` (its list, iterator and end already
// set).  Stack level L iterates C(L+1) (L >= 1) or the edge pairs (L = 0);
// taking element x at level L fixes M(L+1) = x and computes C(L+2).
template <bool STATS, bool STEAL, bool MERGE>
__device__ __forceinline__ void dfs(const KParams &P, const DevPlan *__restrict__ pl, WaveLds &S, int pid, int floor,
                    const int *L0r, const int *L0c, int *lbase, int *hbase, unsigned long long &cnt,
                    unsigned long long &balg, unsigned long long &next, unsigned long long &don,
                    long long *prof, long long t_wave, unsigned long long umul) {
  const int k = pl->k;
  constexpr bool shared = STEAL;
  int level = floor;
  long long t0 = wall_clock64();
  long long tprog = t0;
  const bool may_donate = P.timeout_ticks >= 0;
  // Free levels (timed kernel, counting): a free vertex's candidates are
  // counted, not listed; its level holds one stand-in whose subtree count is
  // measured and stands for each candidate.  Work donated from inside a
  // measured subtree carries the factor (the candidate count) for its
  // receiver to apply; the count measured here is that of the work kept.
  // (a free vertex right above the leaf is only ever counted with it)
  const int fmask = (!STATS && !STEAL && !P.emit_n) ? pl->free_mask & (P.tails ? ~0 : ~(1 << (k - 2))) : 0;
  const int la = (!STATS && !STEAL && !P.emit_n) ? pl->lf_a : -1;
  int mmask = 0;  // levels whose first iteration's subtree is being measured
  // Nothing is donated, and no progress reported, before this call has
  // claimed work of its own: a received task could otherwise be passed on
  // whole before anyone works on it, from wave to wave without end (each
  // hop would also keep the idle deadline from seeing a stall).
  bool worked = false;
  for (;;) {
    if ((mmask >> level) & 1) {
      // back at a measuring level: the first iteration's subtree is complete
      const unsigned long long sub = cnt - S.mbase[level];
      cnt += sub * (unsigned long long)(unsigned)uni(S.mrest[level]);
      mmask &= ~(1 << level);
      continue;
    }
    const unsigned long long w = ie_read(S, level);
    if (ie_it(w) >= ie_end(w)) {
      if (level == floor) return;
      level--;
      continue;
    }
    if (P.limit_ticks > 0 && wall_clock64() - t_wave > P.limit_ticks) {
      if (lane_id() == 0) atomicOr(&P.stats[pid * 4 + 3], 1ull);  // this plan's count is incomplete
      return;
    }
    if (level < P.donate_levels) {
      const long long now = wall_clock64();
      if (worked && now - tprog > P.prog_ticks) {  // tell idle waves that this one still advances
        tprog = now;
        if (lane_id() == 0) fadd(&P.ctrl->progress, 1ull);
      }
      if (may_donate && worked && now - t0 > P.timeout_ticks) {
        t0 = now;
        // donate only while idle waves outnumber the tasks already queued
        int idle = 0;
        if (lane_id() == 0) {
          idle = ld_rlx(&P.ctrl->idle);
          const long long queued = ld_rlx(&P.ctrl->q_avail);
          if ((long long)idle <= queued) idle = 0;
        }
        idle = uni(idle);
        if (idle > 0) {
          const long long tp = STATS ? clock64() : 0;
          const int n = donate<STEAL>(P, S, pid, floor, level, L0r, L0c, idle, fmask, mmask, umul);
          if (STATS) prof[5] += clock64() - tp;
          don += (unsigned)n;
          if (n > 0) continue;  // re-read this level: its end has moved
        }
      }
    }
    if (level + 2 == k - 1 && level > 0 && P.leaf_memo && P.leaf_batch && !P.emit_n &&
        pl->leaf_free[P.reuse ? 0 : 1]) {
      int n = 0;
      const int it = level_claim(S, level, 1024, n, shared);
      if (n == 0) continue;
      worked = true;
      const long long tp = STATS ? clock64() : 0;
      cnt += leaf_memo<STATS, MERGE>(P, pl, S, level, it, it + n, lbase, hbase, balg, next);
      if (STATS) prof[2] += clock64() - tp;
      continue;
    }
    if (level + 2 == k - 1 && P.leaf_batch) {
      int n = 0;
      const int it = level_claim(S, level, kWave, n, shared);
      if (n == 0) continue;  // a sibling took the rest
      worked = true;
      int taken = 0;
      const long long tp = STATS ? clock64() : 0;
      cnt += leaf_batch<STATS>(P, pl, S, level, it, it + n, L0r, L0c, lbase, hbase, balg, next, taken);
      count_op<STATS>(S, 0, 1);
      count_op<STATS>(S, 1, (unsigned)taken);
      if (STATS) prof[2] += clock64() - tp;
      if (taken < n) level_unclaim(S, level, it + taken, it + n, shared);
      continue;
    }
    int n1 = 0;
    const int it = level_claim(S, level, 1, n1, shared);
    if (n1 == 0) continue;
    worked = true;
    const bool stand_in = (fmask >> (level + 1)) & 1;
    if (stand_in) {  // the stand-in of a free vertex: its subtree count stands for each candidate
      if (lane_id() == 0) S.mbase[level] = cnt;
      mmask |= 1 << level;
    }
    if (level == 0) {
      S.M[0] = L0r[it];
      S.M[1] = L0c[it];
    } else {
      S.M[level + 1] = stand_in ? -1 : lbase[(long long)(level - 1) * P.dstride + it];
    }
    if (level + 1 == la) {
      // the dropped leaf's completions, fixed by M(0..la): this iteration's
      // subtree count stands for each of them
      const long long lf = leaf_factor_count<MERGE>(P, pl, S, la, lbase, hbase, balg, next);
      if (lf <= 0) continue;
      if (lane_id() == 0) {
        S.mbase[level] = cnt;
        S.mrest[level] = (int)(lf - 1);
      }
      mmask |= 1 << level;
    }
    const int u = level + 2;
    const int mode = S.hv[u] ? kHoisted : kFull;
    const long long tp = STATS ? clock64() : 0;
    if (u == k - 1) {
      cnt += extend<STATS, MERGE>(P, pl, S, u, nullptr, lbase, hbase, mode, true, balg, next);
      if (STATS) prof[2] += clock64() - tp;
      continue;
    }
    if (!wait_readers(S, shared)) {
      if (lane_id() == 0) __hip_atomic_store(&P.ctrl->error, 5, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return;
    }
    int *out = lbase + (long long)level * P.dstride;  // list of level+1
    const bool tails = !STATS && !P.emit_n && P.tails && pl->tail_t;
    if (tails && u == k - pl->tail_t) {
      // the tail's completions, from the candidate counts of its first
      // vertices of each label (timed kernel, counting)
      unsigned long long f = 1;
      if (tail_product<MERGE>(P, pl, S, u, u - 1, mode, lbase, hbase, balg, next, f)) {
        cnt += f;
        continue;
      }
    }
    if ((fmask >> u) & 1) {
      // a free vertex: its candidates are only counted.  Right above the
      // tail, the path's completions are that count times the tail's (taken
      // first: it is often zero); else one stand-in (id -1: no later vertex
      // is adjacent to it or shares its label) is enumerated, whose subtree
      // count stands for each candidate.
      if (!STATS && !P.emit_n && P.tails && u + 1 == k - 1) {
        // right above the leaf: the leaf's count times this vertex's
        if (lane_id() == 0)
          for (int j = u; j < kMaxK; j++) S.M[j] = -1;
        const int ml = (S.hv[k - 1] && pl->hfix[P.reuse ? 0 : 1][k - 1] < u - 1) ? kHoisted : kFull;
        const unsigned long long f =
            extend<false, MERGE>(P, pl, S, k - 1, nullptr, lbase, hbase, ml, false, balg, next);
        if (f) cnt += f * extend<false, MERGE>(P, pl, S, u, nullptr, lbase, hbase, mode, false, balg, next);
        continue;
      }
      if (tails && u + 1 == k - pl->tail_t) {
        unsigned long long f = 1;
        const int m1 = (S.hv[u + 1] && pl->hfix[P.reuse ? 0 : 1][u + 1] < u - 1) ? kHoisted : kFull;
        if (tail_product<MERGE>(P, pl, S, u + 1, u - 1, m1, lbase, hbase, balg, next, f)) {
          if (f) cnt += f * extend<false, MERGE>(P, pl, S, u, nullptr, lbase, hbase, mode, false, balg, next);
          continue;
        }
      }
      const unsigned nf = extend<false, MERGE>(P, pl, S, u, nullptr, lbase, hbase, mode, false, balg, next);
      if (nf > 0) {
        if (lane_id() == 0) S.mrest[level + 1] = (int)nf - 1;
        S.len[level + 1] = 1;
        if (prepare_hoists<STATS, MERGE>(P, pl, S, u - 1, u - 1, u + 1, (int)nf, lbase, hbase)) {
          level_set(S, level + 1, 0, 1, shared);
          level++;
        }
      }
      continue;
    }
    const int n = (int)extend<STATS, MERGE>(P, pl, S, u, out, lbase, hbase, mode, true, balg, next);
    if (tails && u == pl->tx_w && n > 0) {
      // the tail right below u depends on M(u) only through one group's
      // injectivity: over the n candidates x of u, the completions are
      // P0 for x outside that group's candidates and P1 (one fewer) for
      // the m inside (timed kernel, counting)
      unsigned long long p0 = 0, p1 = 0;
      S.len[level + 1] = n;  // C(u), for a tail vertex that starts from it (result reuse)
      if (tail_product<MERGE>(P, pl, S, u + 1, u - 1, kFull, lbase, hbase, balg, next, p0, pl->tx_lead, &p1)) {
        if (p0 | p1) {
          const unsigned long long m = extend<false, MERGE>(P, pl, S, pl->tx_lead, nullptr, lbase, hbase, kFull, false,
                                                            balg, next, -1, out, n);
          cnt += ((unsigned long long)n - m) * p0 + m * p1;
        }
        continue;
      }
    }
    if (!STATS && !P.emit_n && P.tails && pl->tri_tail && u == k - 3 && n <= kWave) {
      // the path's completions are the triangles of G[C(u)] (unlabeled) or
      // of the candidates of the three labels (labeled; listed into the next
      // level's free list) (timed kernel, counting)
      const long long tri = pl->tri_tail == 2
                                ? triangles_labeled<MERGE>(P, pl, S, u, out, n, lbase + (long long)(level + 1) * P.dstride,
                                                           lbase, hbase, balg, next)
                                : (n >= 3 ? triangles_small(P, out, n) : 0);  // -1: not counted here
      if (tri >= 0) {
        cnt += (unsigned long long)tri;
        continue;
      }
    }
    if (n > 0) {
      S.len[level + 1] = n;
      if (prepare_hoists<STATS, MERGE>(P, pl, S, u - 1, u - 1, u + 1, n, lbase, hbase)) {
        level_set(S, level + 1, 0, n, shared);
        level++;
      }
    }
    if (STATS) prof[1] += clock64() - tp;
  }
}

// Serving: one level-0 chunk or ring task of plan pid has ended, or a chunk
// grab found the plan's cursor spent.  The call that brings the plan's
// in-flight count to zero while its cursor is spent publishes the plan's
// count and flags to host-mapped memory (every chunk handed out and every
// task donated has then ended; a second publication writes the same values).
// Lane 0 only; the caller has waited for its own count update to complete.
__device__ __forceinline__ void unit_done(const KParams &P, int pid) {
  const unsigned long long old = fadd(&P.pend[(long long)pid * 16], ~0ull);
  if (old != 1ull) return;
  const DevPlan *pl = P.plans + pid;
  if (ld_rlx(&P.cursors[(long long)pid * 16]) < (unsigned long long)pl->e_span) return;
  const unsigned long long c = ld_rlx(&P.counts[pid]);
  const unsigned long long f = ld_rlx(&P.stats[(long long)pid * 4 + 3]) & 0xffull;
  __hip_atomic_store(&P.done_host[pid * 2 + 0], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&P.done_host[pid * 2 + 1], 1ull | (f << 8), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Intra-workgroup stealing.  An idle wave looks for the shallowest stack
// level of a sibling wave of its workgroup with at least two untaken
// iterations and takes their back half by one LDS compare-and-swap on the
// sibling's packed range.  The path M(0..l) and the list lengths it needs
// are read before the swap, which succeeds only if that range (and so the
// path below it) did not change; the sibling's lists of stack levels 1..l
// (or, at level 0, the edge pairs taken) are copied after it, while the
// sibling's reader count keeps it from overwriting them.  The thief then
// continues level l from its own copies; nothing is recomputed, so the
// statistics equal those of the sibling finishing the work itself.
// Returns the level taken (the DFS floor) or -1; pid gets the sibling's plan.
__device__ __forceinline__ int steal_from_sibling(const KParams &P, WaveLds *lds, int wib, WaveLds &S, int *L0r,
                                                  int *L0c, int *lbase, int &pid) {
  const int lane = lane_id();
  for (int d = 1; d < kWavesPerBlock; d++) {
    const int vib = (wib + d) % kWavesPerBlock;
    WaveLds &V = lds[vib];
    for (int l = 0; l < kMaxK - 2; l++) {
      const unsigned long long w = ie_read(V, l);
      const int it = ie_it(w), end = ie_end(w);
      if (end - it < 2) continue;
      const int vpid = uni(__hip_atomic_load(&V.pid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
      if (lane == 0) {
        __hip_atomic_fetch_add(&V.readers, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (P.pend) fadd_done(&P.pend[(long long)vpid * 16], 1ull);  // in flight before the sibling can end its unit
      }
      int mv = 0, lv = 0;
      if (lane < 8) {
        mv = V.M[lane];
        lv = V.len[lane];
      }
      const int mid = it + (end - it + 1) / 2;
      int ok = 0;
      if (lane == 0) {
        unsigned long long exp = w;
        ok = __hip_atomic_compare_exchange_strong(&V.ie[l], &exp, ie_make(ie_gen(w), it, mid), __ATOMIC_ACQ_REL,
                                                  __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP)
                 ? 1
                 : 0;
      }
      ok = uni(ok);
      if (!ok) {
        if (lane == 0) {
          if (P.pend) fadd(&P.pend[(long long)vpid * 16], ~0ull);  // the sibling's own unit is still in flight
          __hip_atomic_fetch_add(&V.readers, -1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        continue;
      }
      // copy the sibling's lists (agent-scope loads read L2, past this CU's L1)
      const long long vgw = (long long)blockIdx.x * kWavesPerBlock + vib;
      int *vwb = P.wbuf + vgw * P.wstride;
      if (l == 0) {
        for (int i = mid + lane; i < end; i += kWave) {
          L0r[i] = __hip_atomic_load(vwb + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          L0c[i] = __hip_atomic_load(vwb + kChunkMax + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      for (int lev = 1; lev <= l; lev++) {
        const int n = __shfl(lv, lev);
        int *src = vwb + 2 * kChunkMax + (long long)(lev - 1) * P.dstride;
        int *dst = lbase + (long long)(lev - 1) * P.dstride;
        for (int i = lane; i < n; i += kWave) dst[i] = __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
      if (lane == 0) __hip_atomic_fetch_add(&V.readers, -1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      // continue level l from the copies; the plan is known before the level
      // opens (a third wave may steal from it right away)
      if (lane < 8) {
        S.M[lane] = lane <= l ? mv : 0;
        S.len[lane] = lane <= l ? lv : 0;
      }
      S.pid = vpid;
      for (int j = 0; j < l; j++) level_set(S, j, 0, 0, true);
      level_set(S, l, mid, end, true);
      pid = vpid;
      return l;
    }
  }
  return -1;
}

// The persistent match loop of one wave (body of the match kernels below).
// STEAL: idle waves also take work from sibling waves (steal_from_sibling).
// MERGE: long two-input extensions may intersect by merge path (LDS tiles).
template <bool STATS, bool STEAL, bool MERGE = false>
__device__ __forceinline__ void match_body(const KParams &P) {
  __shared__ WaveLds lds[kWavesPerBlock];
  __shared__ int tiles[MERGE ? kWavesPerBlock * 2 * (kTile + 4) : 1];
  __shared__ int htabs[kWavesPerBlock * kHash];
  const int lane = lane_id();
  const int wib = uni(threadIdx.x / kWave);  // wave-uniform: the workspace pointers stay scalar
  const long long gw = (long long)blockIdx.x * kWavesPerBlock + wib;
  int *wb = P.wbuf + gw * P.wstride;
  int *L0r = wb;
  int *L0c = wb + kChunkMax;
  int *lbase = wb + 2 * kChunkMax;  // C(2..k-2): stack levels 1..k-3
  int *hbase = wb + P.hoff;         // H(3..k-1)
  WaveLds &S = lds[wib];
  S.tile = MERGE ? tiles + wib * 2 * (kTile + 4) : nullptr;
  S.htab = htabs + wib * kHash;
  S.epoch = 0;
  S.tab_epoch = ~0u;
  S.tab_kind = 0;
  S.tab_list = nullptr;
  clear_hoists(S);
  if (lane < 8) S.ie[lane] = 0;
  S.readers = 0;
  S.pid = 0;
  constexpr bool shared = STEAL;
  __syncthreads();  // every wave's ranges are valid before a sibling may read them
  Ctrl *C = P.ctrl;
  unsigned long long balg = 0, next = 0, don = 0;
  int stat_pid = -1;
  bool is_idle = false;
  long long prof[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (lane < 8) S.ctr[lane] = 0;
  const long long t_start = STATS ? clock64() : 0;
  const long long t_wave = wall_clock64();
  int cur_plan = (int)(gw % P.nplans);
  int plans_left = P.nplans;
  // interleaved cursors: this workgroup's group (its XCD when workgroups are
  // dealt to the XCDs in turn), and how many groups it has not found spent
  int xg = (int)(blockIdx.x % (unsigned)P.xcd_groups);
  int xleft = P.xcd_groups;

  for (;;) {
    // ---- acquire work: ring first, then the edge cursors, then wait ----
    TaskRec t;
    int kind = 0;  // 1 level-0 chunk, 2 ring task, 3 stolen from a sibling
    int sfloor = -1;
    long long e0 = 0, e1 = 0;
    int pid = 0;
    if (lane == 0) fadd(&C->busy, 1);
    if (ring_take(P, t)) {
      kind = 2;
    } else {
      unsigned long long pos = ~0ull;
      while (P.interleave && plans_left > 0) {
        // cursor g hands out the edge chunks ci = g (mod groups), each to
        // every plan in turn: all plans sweep the same region of the graph
        // together, and with one group per XCD a chunk's plans run on one
        // XCD and share its L2 (the lists and bitmaps of the chunk's
        // vertices); a spent group sends its waves on to the next one
        if (lane == 0) pos = fadd(&P.cursors[(long long)xg * 16], (unsigned long long)P.chunk);
        pos = uni64(pos);
        const unsigned long long c = pos / (unsigned long long)P.chunk;
        const unsigned long long ci = c / (unsigned long long)P.nplans * (unsigned)P.xcd_groups + (unsigned)xg;
        if (ci * (unsigned long long)P.chunk < P.span) {
          cur_plan = (int)(c % (unsigned long long)P.nplans);
          pos = ci * (unsigned long long)P.chunk;
          break;
        }
        if (--xleft == 0) {
          plans_left = 0;
        } else {
          xg = xg + 1 == P.xcd_groups ? 0 : xg + 1;
        }
      }
      // per-plan cursors: this wave stays on its plan until that plan's
      // edges are handed out, then moves to the next plan with edges left
      while (!P.interleave && plans_left > 0) {
        if (lane == 0) {
          // a chunk may be in flight: counted before the grab can be seen
          if (P.pend) fadd_done(&P.pend[(long long)cur_plan * 16], 1ull);
          pos = fadd(&P.cursors[(long long)cur_plan * 16], (unsigned long long)P.chunk);
        }
        pos = uni64(pos);
        if (pos < (unsigned long long)P.plans[cur_plan].e_span) break;
        if (P.pend && lane == 0) unit_done(P, cur_plan);
        cur_plan = cur_plan + 1 == P.nplans ? 0 : cur_plan + 1;
        plans_left--;
      }
      if (plans_left > 0) {
        pid = cur_plan;
        if (P.interleave) {
          e0 = P.edge_lo + (long long)pos;
          e1 = min(e0 + (long long)P.chunk, P.edge_hi);
        } else {
          const DevPlan *cp = P.plans + pid;
          e0 = cp->e_lo + (long long)pos;
          e1 = min(e0 + (long long)P.chunk, cp->e_lo + cp->e_span);
        }
        kind = 1;
      } else {
        if (lane == 0) fadd(&C->busy, -1);
        if (!is_idle) {
          if (lane == 0) fadd(&C->idle, 1);
          is_idle = true;
        }
        long long t_idle = wall_clock64();
        unsigned long long seen = lane == 0 ? ld_rlx(&C->progress) : 0ull;
        const long long tpi = STATS ? clock64() : 0;
        int backoff = 1;
        for (;;) {
          int stop = 0;
          if (lane == 0) {
            if (ld_rlx(&C->error)) {
              stop = 1;
            } else if (ld_rlx(&C->q_avail) <= 0 && ld_rlx(&C->busy) == 0 && ld_rlx(&C->q_avail) <= 0) {
              stop = 1;
            } else if (P.deadline_ticks > 0) {
              // give up only when no busy wave has reported progress for the
              // whole deadline: a lost handshake, never a long subtree
              const long long now = wall_clock64();
              const unsigned long long pr = ld_rlx(&C->progress);
              if (pr != seen) {
                seen = pr;
                t_idle = now;
              } else if (now - t_idle > P.deadline_ticks) {
                __hip_atomic_store(&C->error, 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                stop = 1;
              }
            }
          }
          if (uni(stop)) break;
          // back off: poll every ~2 us at first, up to ~16 us
          for (int s = 0; s < backoff; s++) __builtin_amdgcn_s_sleep(64);
          if (backoff < 8) backoff <<= 1;
          int avail = 0;
          if (lane == 0) avail = ld_rlx(&C->q_avail) > 0;
          if (uni(avail)) {
            if (lane == 0) fadd(&C->busy, 1);
            if (ring_take(P, t)) {
              kind = 2;
              break;
            }
            if (lane == 0) fadd(&C->busy, -1);
          }
          if (shared) {
            // a sibling may still be copying this wave's lists: wait before stealing into them
            if (!wait_readers(S, true)) {
              if (lane == 0) __hip_atomic_store(&C->error, 5, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              continue;
            }
            // Attempts do not count as busy (thousands of idle waves polling
            // would keep the termination test from ever seeing zero); a
            // sibling with work is busy itself, and the thief counts itself
            // busy once it holds the stolen work.
            sfloor = steal_from_sibling(P, lds, wib, S, L0r, L0c, lbase, pid);
            if (sfloor >= 0) {
              kind = 3;
              if (lane == 0) {
                fadd(&C->busy, 1);
                atomicAdd(&P.stats[pid * 4 + 3], 1ull << 32);  // steals, counted per plan
              }
              break;
            }
          }
        }
        if (STATS) prof[3] += clock64() - tpi;
        if (kind == 0) break;
      }
    }
    if (is_idle) {
      if (lane == 0) fadd(&C->idle, -1);
      is_idle = false;
    }
    if (kind == 2) pid = t.pid;
    S.pid = pid;  // before any level of this unit becomes visible to siblings
    const DevPlan *pl = P.plans + pid;
    if (STATS && stat_pid != pid) {
      if (stat_pid >= 0 && lane == 0) {
        atomicAdd(&P.stats[stat_pid * 4 + 0], balg);
        atomicAdd(&P.stats[stat_pid * 4 + 1], next);
      }
      balg = next = 0;
      stat_pid = pid;
    }
    unsigned long long cnt = 0;
    const bool expired = P.limit_ticks > 0 && wall_clock64() - t_wave > P.limit_ticks;
    if (expired) {
      if (lane == 0) atomicOr(&P.stats[pid * 4 + 3], 1ull);  // dropped: count incomplete
    } else {
      int floor = -1;  // stack level the DFS starts from (one call site for both kinds of work)
      const bool lists_free = kind == 3 || wait_readers(S, shared);  // before this unit overwrites them
      if (!lists_free) {
        if (lane == 0) __hip_atomic_store(&C->error, 5, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      } else if (kind == 1) {
        const long long tp = STATS ? clock64() : 0;
        const int n0 = filter_edges<STATS>(P, pl, e0, e1, L0r, L0c, balg);
        if (STATS) prof[0] += clock64() - tp;
        if (n0 > 0) {
          clear_hoists(S);
          if (STATS) next += (unsigned)n0;
          S.len[0] = n0;
          level_set(S, 0, 0, n0, shared);
          floor = 0;
        }
      } else if (kind == 3) {
        // stolen: path and lists copied, level sfloor open (hoisted lists rebuilt here)
        clear_hoists(S);
        if (sfloor == 0 ||
            prepare_hoists<STATS, MERGE>(P, pl, S, 1, sfloor, sfloor + 2, S.len[sfloor], lbase, hbase))
          floor = sfloor;
      } else if (t.p >= 2) {  // p = 0: an empty task
        const int p = t.p;
#pragma unroll
        for (int j = 0; j < 8; j++)
          if (j < p) S.M[j] = t.M[j];
        const int k = pl->k;
        const bool fresh = t.fresh != 0;
        const long long tp = STATS ? clock64() : 0;
        // rebuild C(2..p-1) (the donor counted them), then C(p), counted in
        // the statistics only when the donor never performed it (fresh)
        for (int u = 2; u <= p; u++) {
          const bool last = u == p;
          const bool leaf = last && p == k - 1;
          int *o = leaf ? nullptr : lbase + (long long)(u - 2) * P.dstride;
          const unsigned n = extend<STATS, MERGE>(P, pl, S, u, o, lbase, hbase, kFull, last && fresh, balg, next);
          if (leaf) cnt += n;
          else S.len[u - 1] = (int)n;
        }
        if (STATS) prof[4] += clock64() - tp;
        if (p < k - 1) {
          const int n = S.len[p - 1];
          clear_hoists(S);
          if (n > 0 && prepare_hoists<STATS, MERGE>(P, pl, S, 1, p - 1, p + 1, n, lbase, hbase)) {
            // a free vertex's level (only ever donated whole, with its edge)
            // holds the one stand-in for its n candidates, as in the DFS
            const bool stand_in = !STATS && !STEAL && !P.emit_n && p <= k - 3 && ((pl->free_mask >> p) & 1);
            if (stand_in && lane == 0) S.mrest[p - 1] = n - 1;
            level_set(S, p - 1, stand_in ? 0 : min(t.a, n), stand_in ? 1 : min(t.b, n), shared);
            floor = p - 1;
          }
        }
      }
      unsigned long long umul = kind == 2 ? t.mul : 1ull;
      if (floor == 1 && kind == 2 && t.fresh && !STATS && !STEAL && !P.emit_n && pl->lf_a == 1) {
        // a donated edge (level 0, p = 2) of a factored plan: its leaf factor
        // (the donor never claimed it; tasks from deeper levels carry theirs)
        const long long lf = leaf_factor_count<MERGE>(P, pl, S, 1, lbase, hbase, balg, next);
        if (lf <= 0) floor = -1;
        else umul *= (unsigned long long)lf;
      }
      if (floor >= 0)
        dfs<STATS, STEAL, MERGE>(P, pl, S, pid, floor, L0r, L0c, lbase, hbase, cnt, balg, next, don, prof, t_wave,
                                 umul);
      cnt *= umul;  // a task's count stands for umul such counts
      // a DFS cut short (time limit) leaves levels open: close them before the next unit's plan is published
      if (STEAL)
        for (int j = 0; j < kMaxK - 2; j++) level_set(S, j, 0, 0, true);
    }
    if (lane == 0) {
      if (cnt) {
        const unsigned long long old = fadd(&P.counts[pid], cnt);
        if (old + cnt < old) atomicOr(&P.stats[pid * 4 + 3], 2ull);  // the raw count wrapped past 2^64
      }
      if (don) {
        atomicAdd(&P.stats[pid * 4 + 2], don);
        don = 0;
      }
      if (P.pend) {
        // this unit's count is in place before the unit stops counting as in flight
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        unit_done(P, pid);
      }
      fadd(&C->busy, -1);
    }
    don = 0;
  }
  if (STATS && stat_pid >= 0 && lane == 0) {
    atomicAdd(&P.stats[stat_pid * 4 + 0], balg);
    atomicAdd(&P.stats[stat_pid * 4 + 1], next);
  }
  if (STATS && lane == 0) {
    prof[6] = clock64() - t_start;
#pragma unroll
    for (int i = 0; i < 7; i++) atomicAdd(&P.prof[i], (unsigned long long)prof[i]);
    atomicAdd(&P.prof[7], 1ull);
    for (int i = 0; i < 8; i++) atomicAdd(&P.prof[8 + i], S.ctr[i]);
  }
}

// Timed / statistics kernel: registers as the compiler allocates them, the
// timed one bounded to 3 waves per SIMD (168 VGPRs).
template <bool STATS>
__global__ void __launch_bounds__(kBlock, STATS ? 1 : 3) tdfs_match_kernel(KParams P) {
  match_body<STATS, false>(P);
}

// The same loop compiled for 4, 5, 6 or 8 waves per SIMD: more resident DFS
// stacks to hide memory latency, at the cost of register spills.
__global__ void __launch_bounds__(kBlock, 4) tdfs_match_kernel_w4(KParams P) { match_body<false, false>(P); }
__global__ void __launch_bounds__(kBlock, 5) tdfs_match_kernel_w5(KParams P) { match_body<false, false>(P); }
__global__ void __launch_bounds__(kBlock, 6) tdfs_match_kernel_w6(KParams P) { match_body<false, false>(P); }
__global__ void __launch_bounds__(kBlock, 8) tdfs_match_kernel_w8(KParams P) { match_body<false, false>(P); }

// The same kernels with intra-workgroup stealing compiled in (its code
// raises register use, so it is a variant of its own).
template <bool STATS>
__global__ void __launch_bounds__(kBlock) tdfs_match_kernel_steal(KParams P) {
  match_body<STATS, true>(P);
}
__global__ void __launch_bounds__(kBlock, 4) tdfs_match_kernel_w4_steal(KParams P) { match_body<false, true>(P); }
__global__ void __launch_bounds__(kBlock, 5) tdfs_match_kernel_w5_steal(KParams P) { match_body<false, true>(P); }
__global__ void __launch_bounds__(kBlock, 6) tdfs_match_kernel_w6_steal(KParams P) { match_body<false, true>(P); }
__global__ void __launch_bounds__(kBlock, 8) tdfs_match_kernel_w8_steal(KParams P) { match_body<false, true>(P); }

// And with the merge-path intersection compiled in (tdfs_tuning.merge_min).
template <bool STATS>
__global__ void __launch_bounds__(kBlock) tdfs_match_kernel_merge(KParams P) {
  match_body<STATS, false, true>(P);
}
__global__ void __launch_bounds__(kBlock, 5) tdfs_match_kernel_w5_merge(KParams P) { match_body<false, false, true>(P); }

// Sets bit y of row h for every neighbour y of hub vertex hubs[h].
__global__ void tdfs_hub_build_kernel(const long long *rowptr, const int *colidx, const int *hubs, int nhubs,
                                      unsigned *bits, long long words) {
  for (int h = blockIdx.x; h < nhubs; h += gridDim.x) {
    const int v = hubs[h];
    unsigned *row = bits + (long long)h * words;
    for (long long e = rowptr[v] + threadIdx.x; e < rowptr[v + 1]; e += blockDim.x) {
      const int y = colidx[e];
      atomicOr(row + (y >> 5), 1u << (y & 31));
    }
  }
}

// Resets the control block, results and ring sequence numbers (a kernel, so
// that captured launch sequences hold no memset nodes).
__global__ void tdfs_init_kernel(Ctrl *ctrl, unsigned long long *counts, unsigned long long *stats, int nplans,
                                 unsigned long long *seq, unsigned long long cap, unsigned long long *prof,
                                 unsigned long long *cursors, int ncursors) {
  const unsigned long long tid = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long nth = (unsigned long long)gridDim.x * blockDim.x;
  int *cw = reinterpret_cast<int *>(ctrl);
  for (unsigned long long i = tid; i < sizeof(Ctrl) / sizeof(int); i += nth) cw[i] = 0;
  for (unsigned long long i = tid; i < (unsigned long long)nplans; i += nth) counts[i] = 0;
  for (unsigned long long i = tid; i < (unsigned long long)nplans * 4; i += nth) stats[i] = 0;
  for (unsigned long long i = tid; i < cap; i += nth) seq[i] = i;
  for (unsigned long long i = tid; i < (unsigned long long)kProf; i += nth) prof[i] = 0;
  for (unsigned long long i = tid; i < (unsigned long long)ncursors * 16; i += nth) cursors[i] = 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI (device side)
// ---------------------------------------------------------------------------
// One device copy of the graph under one vertex numbering.
struct GraphView {
  long long *rowptr = nullptr;
  int *colidx = nullptr;
  int *degv = nullptr;
  int *src = nullptr;
  int *maxnbr = nullptr;
  unsigned short *labels = nullptr;  // dense label ids
  int *hub_idx = nullptr;
  unsigned *hub_bits = nullptr;
  long long hub_words = 0;
  int nhubs = 0;
  int *colidx_lab = nullptr;  // lists re-sorted by (label, id)
  int *loff = nullptr;        // n * (nlab + 1) segment offsets
  int nlab = 0;
};

struct tdfs_graph {
  int device;
  int32_t n;
  int64_t nnz;
  int32_t dmax;
  int labeled;
  // distinct label values of the caller, ascending: dense id i stands for
  // label_values[i] on the device
  std::vector<int32_t> label_values;
  // view[0]: the numbering as given.  view[1] (when `ordered`): vertices
  // renumbered by ascending degree (ties by id), so that the symmetry-breaking
  // "smaller id first" conditions start the DFS at low-degree vertices.
  GraphView view[2];
  int ordered;
  std::vector<int32_t> perm;  // new id of each vertex in view[1]
  // host copies of both numberings (work estimates, serving admission)
  std::vector<int64_t> h_rowptr[2];
  std::vector<int32_t> h_colidx[2];
  std::vector<int32_t> h_labels[2];
  std::vector<int32_t> h_deg[2], h_maxnbr[2];
};

// Everything one launch needs, owned: stream, events, per-wave workspace,
// task ring, control block, plan table, results, and pinned host copies of
// the results.  A launch can be captured once into a hipGraph and replayed.
struct Launch {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int *wbuf = nullptr;
  size_t wbuf_ints = 0;
  bool wbuf_external = false;  // wbuf is a region of memory the launch does not own (serving arena)
  TaskRec *q_data = nullptr;
  unsigned long long *q_seq = nullptr;
  unsigned long long q_cap = 0;
  Ctrl *ctrl = nullptr;
  unsigned long long *counts = nullptr, *stats = nullptr, *prof = nullptr;
  unsigned long long *cursors = nullptr;  // 16 words per plan
  DevPlan *plans = nullptr;
  int plan_cap = 0;
  // pinned host mirrors, written by the launch's D2H copies
  unsigned long long *h_counts = nullptr, *h_stats = nullptr, *h_prof = nullptr;
  Ctrl *h_ctrl = nullptr;
  DevPlan *h_plans = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int grid = 0;
};

static int scratch_setup(int device);
constexpr int kMaxDevices = 64;

struct tdfs_ctx {
  int device;
  int cus;
  int wallclock_khz;
  int blocks_per_cu;        // occupancy limit of the timed kernel
  int blocks_per_cu_stats;  // and of the statistics kernel
  int blocks_per_cu_w[9];   // and of the timed kernel compiled for w = 4, 5, 6, 8 waves per SIMD
  int steal_blocks_per_cu, steal_blocks_per_cu_stats, steal_blocks_per_cu_w[9];  // the same with stealing
  int merge_blocks_per_cu, merge_blocks_per_cu_stats, merge_blocks_per_cu_w5;     // and with merge path
  tdfs_tuning tune;
  int stats;
  Launch main;              // used by tdfs_count / tdfs_count_batch
  unsigned long long last_prof[kProf];
};

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) return fail(TDFS_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

extern "C" void tdfs_tuning_default(tdfs_tuning *t) {
  std::memset(t, 0, sizeof(*t));
  t->waves_per_block = kWavesPerBlock;
  t->blocks_per_cu = 0;
  t->chunk_edges = 64;
  t->timeout_us = 200;
  t->split_chunk = 1;
  t->queue_log2 = 18;
  t->donate_levels = 3;
  t->reuse = 1;
  t->leaf_batch = 1;
  t->hub_min_degree = 256;
  t->hub_budget_mb = 4096;
  t->degree_order = 1;
  t->cursor_order = 1;
  t->waves_per_simd = 0;
  t->leaf_pair_cap = 4096;
  t->hoist_min = 2;
  t->idle_deadline_ms = 60000;
  t->steal = 2;
  t->leaf_memo = 1;
  t->leaf_hash = 1;
  t->ext_hash = 1;
  t->tails = 1;
  t->free_order = 1;
}

static void fill_defaults(tdfs_tuning *t) {
  tdfs_tuning d;
  tdfs_tuning_default(&d);
  if (t->waves_per_block <= 0) t->waves_per_block = d.waves_per_block;
  if (t->chunk_edges <= 0) t->chunk_edges = d.chunk_edges;
  if (t->chunk_edges > kChunkMax) t->chunk_edges = kChunkMax;
  if (t->timeout_us == 0) t->timeout_us = d.timeout_us;
  if (t->split_chunk <= 0) t->split_chunk = d.split_chunk;
  if (t->queue_log2 <= 0) t->queue_log2 = d.queue_log2;
  if (t->queue_log2 < 8) t->queue_log2 = 8;
  if (t->queue_log2 > 26) t->queue_log2 = 26;
  if (t->donate_levels <= 0) t->donate_levels = d.donate_levels;
  if (t->reuse <= 0) t->reuse = d.reuse;
  if (t->leaf_batch <= 0) t->leaf_batch = d.leaf_batch;
  if (t->hub_min_degree <= 0) t->hub_min_degree = d.hub_min_degree;
  if (t->hub_budget_mb == 0) t->hub_budget_mb = d.hub_budget_mb;
  if (t->degree_order <= 0) t->degree_order = d.degree_order;
  if (t->cursor_order <= 0) t->cursor_order = d.cursor_order;
  if (t->waves_per_simd != 4 && t->waves_per_simd != 5 && t->waves_per_simd != 6 && t->waves_per_simd != 8)
    t->waves_per_simd = 0;
  if (t->leaf_pair_cap <= 0) t->leaf_pair_cap = d.leaf_pair_cap;
  if (t->leaf_pair_cap < 64) t->leaf_pair_cap = 64;
  if (t->hoist_min == 0) t->hoist_min = d.hoist_min;
  if (t->idle_deadline_ms == 0) t->idle_deadline_ms = d.idle_deadline_ms;
  if (t->steal != 1) t->steal = 2;
  if (t->leaf_memo != 2) t->leaf_memo = 1;
  if (t->leaf_hash != 2 && t->leaf_hash != 3) t->leaf_hash = 1;
  if (t->ext_hash != 2) t->ext_hash = 1;
  if (t->tails != 2) t->tails = 1;
  if (t->free_order != 2) t->free_order = 1;
}

static int launch_init(Launch &L, int device) {
  L.device = device;
  HIP_TRY(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&L.ev0));
  HIP_TRY(hipEventCreate(&L.ev1));
  HIP_TRY(hipMalloc(&L.ctrl, sizeof(Ctrl)));
  HIP_TRY(hipMalloc(&L.prof, kProf * sizeof(unsigned long long)));
  HIP_TRY(hipHostMalloc(&L.h_ctrl, sizeof(Ctrl), hipHostMallocDefault));
  HIP_TRY(hipHostMalloc(&L.h_prof, kProf * sizeof(unsigned long long), hipHostMallocDefault));
  return TDFS_OK;
}

static void launch_free(Launch &L) {
  if (L.stream) (void)hipStreamSynchronize(L.stream);
  if (L.exec) (void)hipGraphExecDestroy(L.exec);
  if (L.graph) (void)hipGraphDestroy(L.graph);
  if (L.wbuf && !L.wbuf_external) (void)hipFree(L.wbuf);
  if (L.q_data) (void)hipFree(L.q_data);
  if (L.q_seq) (void)hipFree(L.q_seq);
  if (L.ctrl) (void)hipFree(L.ctrl);
  if (L.counts) (void)hipFree(L.counts);
  if (L.stats) (void)hipFree(L.stats);
  if (L.prof) (void)hipFree(L.prof);
  if (L.plans) (void)hipFree(L.plans);
  if (L.cursors) (void)hipFree(L.cursors);
  if (L.h_counts) (void)hipHostFree(L.h_counts);
  if (L.h_stats) (void)hipHostFree(L.h_stats);
  if (L.h_prof) (void)hipHostFree(L.h_prof);
  if (L.h_ctrl) (void)hipHostFree(L.h_ctrl);
  if (L.h_plans) (void)hipHostFree(L.h_plans);
  if (L.ev0) (void)hipEventDestroy(L.ev0);
  if (L.ev1) (void)hipEventDestroy(L.ev1);
  if (L.stream) (void)hipStreamDestroy(L.stream);
  L = Launch();
}

extern "C" int tdfs_ctx_create(int32_t device, tdfs_ctx **out) {
  if (!out) return fail(TDFS_E_ARG, "null out");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(TDFS_E_ARG, "device %d out of range (%d devices)", device, ndev);
  HIP_TRY(hipSetDevice(device));
  tdfs_ctx *c = new tdfs_ctx();
  c->device = device;
  c->stats = 0;
  std::memset(c->last_prof, 0, sizeof(c->last_prof));
  tdfs_tuning_default(&c->tune);
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device));
  c->cus = prop.multiProcessorCount;
  int khz = 0;
  HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
  c->wallclock_khz = khz > 0 ? khz : 100000;
  int nb = 0, nb2 = 0, nbw[9] = {}, sb = 0, sb2 = 0, sbw[9] = {};
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbw[4], tdfs_match_kernel_w4, kBlock, 0));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbw[5], tdfs_match_kernel_w5, kBlock, 0));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbw[6], tdfs_match_kernel_w6, kBlock, 0));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbw[8], tdfs_match_kernel_w8, kBlock, 0));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, tdfs_match_kernel<false>, kBlock, 0));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb2, tdfs_match_kernel<true>, kBlock, 0));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&sbw[4], tdfs_match_kernel_w4_steal, kBlock, 0));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&sbw[5], tdfs_match_kernel_w5_steal, kBlock, 0));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&sbw[6], tdfs_match_kernel_w6_steal, kBlock, 0));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&sbw[8], tdfs_match_kernel_w8_steal, kBlock, 0));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&sb, tdfs_match_kernel_steal<false>, kBlock, 0));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&sb2, tdfs_match_kernel_steal<true>, kBlock, 0));
  c->blocks_per_cu = std::max(1, nb);
  c->blocks_per_cu_stats = std::max(1, nb2);
  for (int w = 0; w < 9; w++) c->blocks_per_cu_w[w] = std::max(1, nbw[w]);
  int mb = 0, mb2 = 0, mbw5 = 0;
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&mb, tdfs_match_kernel_merge<false>, kBlock, 0));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&mb2, tdfs_match_kernel_merge<true>, kBlock, 0));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&mbw5, tdfs_match_kernel_w5_merge, kBlock, 0));
  c->merge_blocks_per_cu = std::max(1, mb);
  c->merge_blocks_per_cu_stats = std::max(1, mb2);
  c->merge_blocks_per_cu_w5 = std::max(1, mbw5);
  c->steal_blocks_per_cu = std::max(1, sb);
  c->steal_blocks_per_cu_stats = std::max(1, sb2);
  for (int w = 0; w < 9; w++) c->steal_blocks_per_cu_w[w] = std::max(1, sbw[w]);
  int rc = launch_init(c->main, device);
  if (!rc) rc = scratch_setup(device);
  if (rc) {
    launch_free(c->main);
    delete c;
    return rc;
  }
  *out = c;
  return TDFS_OK;
}

extern "C" int tdfs_ctx_set_tuning(tdfs_ctx *ctx, const tdfs_tuning *t) {
  if (!ctx || !t) return fail(TDFS_E_ARG, "null argument");
  ctx->tune = *t;
  fill_defaults(&ctx->tune);
  return TDFS_OK;
}

extern "C" int tdfs_ctx_get_tuning(const tdfs_ctx *ctx, tdfs_tuning *t) {
  if (!ctx || !t) return fail(TDFS_E_ARG, "null argument");
  *t = ctx->tune;
  fill_defaults(t);
  return TDFS_OK;
}

extern "C" int tdfs_ctx_set_stats(tdfs_ctx *ctx, int32_t on) {
  if (!ctx) return fail(TDFS_E_ARG, "null ctx");
  ctx->stats = on ? 1 : 0;
  return TDFS_OK;
}

extern "C" int tdfs_ctx_info(const tdfs_ctx *ctx, int64_t *out4) {
  if (!ctx || !out4) return fail(TDFS_E_ARG, "null argument");
  out4[0] = ctx->cus;
  out4[1] = ctx->wallclock_khz;
  out4[2] = (int64_t)ctx->cus * ctx->blocks_per_cu * kWavesPerBlock;
  out4[3] = ctx->blocks_per_cu;
  return TDFS_OK;
}

extern "C" void tdfs_ctx_destroy(tdfs_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  launch_free(ctx->main);
  delete ctx;
}

static void free_view(GraphView &v) {
  if (v.rowptr) (void)hipFree(v.rowptr);
  if (v.colidx) (void)hipFree(v.colidx);
  if (v.degv) (void)hipFree(v.degv);
  if (v.src) (void)hipFree(v.src);
  if (v.maxnbr) (void)hipFree(v.maxnbr);
  if (v.labels) (void)hipFree(v.labels);
  if (v.hub_idx) (void)hipFree(v.hub_idx);
  if (v.hub_bits) (void)hipFree(v.hub_bits);
  if (v.colidx_lab) (void)hipFree(v.colidx_lab);
  if (v.loff) (void)hipFree(v.loff);
  v = GraphView();
}

// Copies one numbering of a validated CSR to the device with its derived
// arrays (degree, row of each edge, max neighbour degree, dense 16-bit label
// ids) and the dense neighbour bitmaps of its highest-degree vertices.
// labels: dense ids in [0, nlabels) or null.
static int upload_view(const tdfs_tuning &t, int32_t n, int64_t nnz, const int64_t *rowptr, const int32_t *colidx,
                       const int32_t *labels, int nlabels, GraphView &g) {
  std::vector<int> degv(n > 0 ? n : 1), src(nnz > 0 ? nnz : 1), maxnbr(n > 0 ? n : 1);
  for (int32_t v = 0; v < n; v++) degv[v] = (int)(rowptr[v + 1] - rowptr[v]);
  for (int32_t v = 0; v < n; v++) {
    int m = -1;
    for (int64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
      src[e] = v;
      m = std::max(m, degv[colidx[e]]);
    }
    maxnbr[v] = m;
  }
  auto bail = [&](const char *what, hipError_t e) {
    free_view(g);
    return fail(TDFS_E_OOM, "%s: %s", what, hipGetErrorString(e));
  };
  hipError_t e;
  if ((e = hipMalloc(&g.rowptr, sizeof(long long) * (size_t)(n + 1)))) return bail("rowptr", e);
  if ((e = hipMalloc(&g.colidx, sizeof(int) * (size_t)std::max<int64_t>(nnz, 1)))) return bail("colidx", e);
  if ((e = hipMalloc(&g.degv, sizeof(int) * (size_t)std::max(n, 1)))) return bail("degv", e);
  if ((e = hipMalloc(&g.src, sizeof(int) * (size_t)std::max<int64_t>(nnz, 1)))) return bail("src", e);
  if ((e = hipMalloc(&g.maxnbr, sizeof(int) * (size_t)std::max(n, 1)))) return bail("maxnbr", e);
  if ((e = hipMalloc(&g.labels, sizeof(unsigned short) * (size_t)std::max(n, 1)))) return bail("labels", e);
  std::vector<unsigned short> lab16(n > 0 ? n : 1, 0);
  if (labels)
    for (int32_t v = 0; v < n; v++) lab16[v] = (unsigned short)labels[v];
  if ((e = hipMemcpy(g.rowptr, rowptr, sizeof(long long) * (size_t)(n + 1), hipMemcpyHostToDevice)))
    return bail("copy rowptr", e);
  if (nnz > 0) {
    if ((e = hipMemcpy(g.colidx, colidx, sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice)))
      return bail("copy colidx", e);
    if ((e = hipMemcpy(g.src, src.data(), sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice))) return bail("copy src", e);
  }
  if (n > 0) {
    if ((e = hipMemcpy(g.degv, degv.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice))) return bail("copy degv", e);
    if ((e = hipMemcpy(g.maxnbr, maxnbr.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice)))
      return bail("copy maxnbr", e);
    if ((e = hipMemcpy(g.labels, lab16.data(), sizeof(unsigned short) * (size_t)n, hipMemcpyHostToDevice)))
      return bail("copy labels", e);
  }
  // Label segments: every list re-sorted by (label, id) with per-vertex
  // offsets, so a labeled extension reads only the vertices of its label.
  if (labels && n > 0) {
    const int nlab = nlabels;
    if (nlab <= kMaxSegLabels) {
      const int W = nlab + 1;
      std::vector<int> off((size_t)n * W, 0), cl(nnz > 0 ? nnz : 1);
      for (int32_t v = 0; v < n; v++) {
        int *o = off.data() + (size_t)v * W;
        for (int64_t e = rowptr[v]; e < rowptr[v + 1]; e++) o[lab16[colidx[e]] + 1]++;
        for (int l = 0; l < nlab; l++) o[l + 1] += o[l];
        std::vector<int> fill(o, o + nlab);
        for (int64_t e = rowptr[v]; e < rowptr[v + 1]; e++) cl[rowptr[v] + fill[lab16[colidx[e]]]++] = colidx[e];
      }
      if ((e = hipMalloc(&g.colidx_lab, sizeof(int) * (size_t)std::max<int64_t>(nnz, 1)))) return bail("colidx_lab", e);
      if ((e = hipMalloc(&g.loff, sizeof(int) * off.size()))) return bail("loff", e);
      if (nnz > 0 && (e = hipMemcpy(g.colidx_lab, cl.data(), sizeof(int) * (size_t)nnz, hipMemcpyHostToDevice)))
        return bail("copy colidx_lab", e);
      if ((e = hipMemcpy(g.loff, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice))) return bail("copy loff", e);
      g.nlab = nlab;
    }
  }
  // Dense neighbour bitmaps for the highest-degree vertices, within a byte
  // budget: membership in a hub's list becomes one load.
  const long long words = ((long long)n + 31) / 32;
  std::vector<int> hubs;
  if (t.hub_budget_mb > 0 && n > 0) {
    for (int32_t v = 0; v < n; v++)
      if (degv[v] >= t.hub_min_degree) hubs.push_back(v);
    std::stable_sort(hubs.begin(), hubs.end(), [&](int a, int b) { return degv[a] > degv[b]; });
    const long long cap = ((long long)t.hub_budget_mb << 20) / (words * 4);
    if ((long long)hubs.size() > cap) hubs.resize((size_t)cap);
  }
  std::vector<int> hidx(n > 0 ? n : 1, -1);
  for (size_t h = 0; h < hubs.size(); h++) hidx[hubs[h]] = (int)h;
  g.nhubs = (int)hubs.size();
  g.hub_words = words;
  if ((e = hipMalloc(&g.hub_idx, sizeof(int) * (size_t)std::max(n, 1)))) return bail("hub_idx", e);
  if ((e = hipMemcpy(g.hub_idx, hidx.data(), sizeof(int) * (size_t)std::max(n, 1), hipMemcpyHostToDevice)))
    return bail("copy hub_idx", e);
  if (!hubs.empty()) {
    const size_t bytes = (size_t)hubs.size() * (size_t)words * 4;
    if ((e = hipMalloc(&g.hub_bits, bytes))) return bail("hub bitmaps", e);
    if ((e = hipMemset(g.hub_bits, 0, bytes))) return bail("clear hub bitmaps", e);
    int *dh = nullptr;
    if ((e = hipMalloc(&dh, sizeof(int) * hubs.size()))) return bail("hub list", e);
    if ((e = hipMemcpy(dh, hubs.data(), sizeof(int) * hubs.size(), hipMemcpyHostToDevice))) {
      (void)hipFree(dh);
      return bail("copy hub list", e);
    }
    hipLaunchKernelGGL(tdfs_hub_build_kernel, dim3((unsigned)std::min<size_t>(hubs.size(), 65535)), dim3(256), 0, 0,
                       g.rowptr, g.colidx, dh, (int)hubs.size(), g.hub_bits, words);
    e = hipDeviceSynchronize();
    (void)hipFree(dh);
    if (e) return bail("hub bitmap build", e);
  }
  return TDFS_OK;
}

extern "C" int tdfs_graph_upload(tdfs_ctx *ctx, int32_t n, int64_t nnz, const int64_t *rowptr,
                                 const int32_t *colidx, const int32_t *labels, tdfs_graph **out) {
  if (!ctx || !rowptr || (!colidx && nnz > 0) || !out || n < 0 || nnz < 0)
    return fail(TDFS_E_ARG, "bad graph arguments");
  int rc = tdfs_graph_validate(n, nnz, rowptr, colidx);
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  tdfs_tuning t = ctx->tune;
  fill_defaults(&t);
  tdfs_graph *g = new tdfs_graph();
  g->device = ctx->device;
  g->n = n;
  g->nnz = nnz;
  g->dmax = 0;
  for (int32_t v = 0; v < n; v++) g->dmax = std::max<int32_t>(g->dmax, (int32_t)(rowptr[v + 1] - rowptr[v]));
  g->labeled = labels != nullptr;
  g->ordered = 0;
  // dense label ids: the rank of each distinct value
  std::vector<int32_t> dense;
  if (labels) {
    g->label_values.assign(labels, labels + n);
    std::sort(g->label_values.begin(), g->label_values.end());
    g->label_values.erase(std::unique(g->label_values.begin(), g->label_values.end()), g->label_values.end());
    if (g->label_values.size() > 65535) {
      const size_t nl = g->label_values.size();
      delete g;
      return fail(TDFS_E_ARG, "%zu distinct labels (at most 65535)", nl);
    }
    dense.resize(n);
    for (int32_t v = 0; v < n; v++)
      dense[v] = (int32_t)(std::lower_bound(g->label_values.begin(), g->label_values.end(), labels[v]) -
                           g->label_values.begin());
  }
  const int nlabels = (int)g->label_values.size();
  rc = upload_view(t, n, nnz, rowptr, colidx, labels ? dense.data() : nullptr, nlabels, g->view[0]);
  if (rc) {
    delete g;
    return rc;
  }
  g->h_rowptr[0].assign(rowptr, rowptr + n + 1);
  g->h_colidx[0].assign(colidx, colidx + nnz);
  if (labels) g->h_labels[0].assign(labels, labels + n);
  auto derive = [&](int v) {
    const std::vector<int64_t> &rp = g->h_rowptr[v];
    const std::vector<int32_t> &ci = g->h_colidx[v];
    g->h_deg[v].resize(n);
    g->h_maxnbr[v].assign(n, -1);
    for (int32_t x = 0; x < n; x++) g->h_deg[v][x] = (int32_t)(rp[x + 1] - rp[x]);
    for (int32_t x = 0; x < n; x++)
      for (int64_t e = rp[x]; e < rp[x + 1]; e++) g->h_maxnbr[v][x] = std::max(g->h_maxnbr[v][x], g->h_deg[v][ci[e]]);
  };
  derive(0);
  if (t.degree_order == 1 && n > 0) {
    // new id = rank by (degree, id)
    std::vector<int32_t> byrank(n);
    for (int32_t v = 0; v < n; v++) byrank[v] = v;
    std::stable_sort(byrank.begin(), byrank.end(), [&](int32_t a, int32_t b) {
      return rowptr[a + 1] - rowptr[a] < rowptr[b + 1] - rowptr[b];
    });
    g->perm.assign(n, 0);
    for (int32_t r = 0; r < n; r++) g->perm[byrank[r]] = r;
    std::vector<int64_t> rp2(n + 1, 0);
    for (int32_t r = 0; r < n; r++) rp2[r + 1] = rp2[r] + (rowptr[byrank[r] + 1] - rowptr[byrank[r]]);
    std::vector<int32_t> ci2(nnz > 0 ? nnz : 1), lab2;
    for (int32_t r = 0; r < n; r++) {
      const int32_t v = byrank[r];
      int32_t *dst = ci2.data() + rp2[r];
      for (int64_t e = rowptr[v]; e < rowptr[v + 1]; e++) *dst++ = g->perm[colidx[e]];
      std::sort(ci2.data() + rp2[r], ci2.data() + rp2[r + 1]);
    }
    std::vector<int32_t> dense2;
    if (labels) {
      lab2.resize(n);
      dense2.resize(n);
      for (int32_t v = 0; v < n; v++) {
        lab2[g->perm[v]] = labels[v];
        dense2[g->perm[v]] = dense[v];
      }
    }
    rc = upload_view(t, n, nnz, rp2.data(), ci2.data(), labels ? dense2.data() : nullptr, nlabels, g->view[1]);
    if (rc) {
      free_view(g->view[0]);
      delete g;
      return rc;
    }
    g->ordered = 1;
    g->h_rowptr[1] = std::move(rp2);
    ci2.resize((size_t)nnz);
    g->h_colidx[1] = std::move(ci2);
    g->h_labels[1] = std::move(lab2);
    derive(1);
  }
  *out = g;
  return TDFS_OK;
}

extern "C" void tdfs_graph_free(tdfs_graph *g) {
  if (!g) return;
  (void)hipSetDevice(g->device);
  free_view(g->view[0]);
  free_view(g->view[1]);
  delete g;
}

extern "C" int tdfs_graph_info(const tdfs_graph *g, int64_t *out4) {
  if (!g || !out4) return fail(TDFS_E_ARG, "null argument");
  out4[0] = g->n;
  out4[1] = g->nnz;
  out4[2] = g->ordered;
  out4[3] = g->ordered ? g->view[1].nhubs : g->view[0].nhubs;
  return TDFS_OK;
}

extern "C" int tdfs_graph_order(const tdfs_graph *g, int32_t *perm_out) {
  if (!g || !perm_out) return fail(TDFS_E_ARG, "null argument");
  if (!g->ordered) return fail(TDFS_E_STATE, "graph was uploaded without the degree-ordered view");
  std::memcpy(perm_out, g->perm.data(), sizeof(int32_t) * (size_t)g->n);
  return TDFS_OK;
}

// The level-0 prune "some neighbour of M(1) has degree >= dP(2)"
// (src/gpu_match.cu:540-548) never drops an embedding when vertex 2 is
// adjacent to vertex 1 or dP(2) <= dP(0); only then is the raw count
// independent of the vertex numbering.
static bool prune_is_sound(const tdfs_plan &p) { return p.adj[2][1] != 0 || p.deg[2] <= p.deg[0]; }

extern "C" int tdfs_graph_estimate(const tdfs_graph *g, const tdfs_plan *plan, int32_t nsample, uint64_t seed,
                                   double cap, double *out3) {
  if (!g || !plan) return fail(TDFS_E_ARG, "null argument");
  const int v = (g->ordered && prune_is_sound(*plan)) ? 1 : 0;
  return tdfs::estimate(g->n, g->nnz, g->h_rowptr[v].data(), g->h_colidx[v].data(),
                        g->labeled ? g->h_labels[v].data() : nullptr, plan, nsample, seed, cap, out3,
                        g->h_deg[v].data(), g->h_maxnbr[v].data());
}

// Length t >= 2 of the plan's tail (0: none): the last t vertices have one
// backward-neighbour set and are pairwise non-adjacent, and, unlabeled, are
// bound to earlier vertices only by injectivity and totally ordered among
// themselves (transitive closure of LT / GT, no cycle).  The longest such
// suffix that leaves vertices 0 and 1 out.
static int tail_length(const tdfs_plan &p) {
  int best = 0;
  for (int t = 2; t <= p.k - 2; t++) {  // the tail starts at u0 = k - t >= 2
    const int u0 = p.k - t;
    bool ok = true;
    for (int v = u0; v < p.k && ok; v++) {
      ok = p.num_bn[v] == p.num_bn[u0];
      for (int i = 0; ok && i < p.num_bn[v]; i++) ok = p.bn[v][i] == p.bn[u0][i];  // ascending lists
      for (int w = u0; ok && w < p.k; w++) ok = w == v || p.adj[v][w] == 0;
      for (int j = 0; ok && !p.labeled && j < u0; j++) ok = p.cond[v][j] == TDFS_NE;
    }
    if (ok && !p.labeled) {
      // every pair of tail vertices ordered (transitive closure of LT / GT)
      bool less[kMaxK][kMaxK] = {};
      for (int v = u0; v < p.k; v++)
        for (int j = u0; j < v; j++) {
          if (p.cond[v][j] == TDFS_LT) less[v][j] = true;
          if (p.cond[v][j] == TDFS_GT) less[j][v] = true;
        }
      for (int m = u0; m < p.k; m++)
        for (int a = u0; a < p.k; a++)
          for (int b = u0; b < p.k; b++)
            if (less[a][m] && less[m][b]) less[a][b] = true;
      for (int a = u0; a < p.k && ok; a++) {
        ok = !less[a][a];  // a cycle of LT / GT admits no match
        for (int b = a + 1; b < p.k && ok; b++) ok = less[a][b] || less[b][a];
      }
    }
    if (ok) best = t;
  }
  return best;
}

// label_values: the graph's distinct labels (ascending); a labeled plan's
// labels become their dense ids, and a label no vertex carries becomes the
// id one past the last, which no vertex carries (the plan still runs, so its
// statistics are those of the reference's DFS).
static void to_dev_plan(const tdfs_plan &p, const std::vector<int32_t> &label_values, DevPlan &d) {
  std::memset(&d, 0, sizeof(d));
  d.k = p.k;
  d.labeled = p.labeled;
  d.sym01 = p.sym01;
  for (int u = 0; u < p.k; u++) {
    d.label[u] = p.label[u];
    if (p.labeled) {
      auto it = std::lower_bound(label_values.begin(), label_values.end(), p.label[u]);
      if (it == label_values.end() || *it != p.label[u]) {
        d.label[u] = (int)label_values.size();  // carried by no vertex
      } else {
        d.label[u] = (int)(it - label_values.begin());
      }
    }
    d.deg[u] = p.deg[u];
    d.num_bn[u] = p.num_bn[u];
    for (int i = 0; i < p.num_bn[u]; i++) d.bn[u][i] = p.bn[u][i];
    d.gt_mask[u] = d.lt_mask[u] = 0;
    for (int j = 0; j < u; j++) {
      if (p.cond[u][j] == TDFS_GT) d.gt_mask[u] |= 1 << j;
      if (p.cond[u][j] == TDFS_LT) d.lt_mask[u] |= 1 << j;
    }
    d.same_label_mask[u] = 0;
    for (int j = 0; j < u; j++)
      if (p.label[j] == p.label[u]) d.same_label_mask[u] |= 1 << j;
    d.dep[u] = p.shared_lvl[u] >= 2 ? p.shared_lvl[u] : -1;
    d.num_bn_sh[u] = d.dep[u] >= 0 ? p.num_bn_sh[u] : 0;
    for (int i = 0; i < d.num_bn_sh[u]; i++) d.bn_sh[u][i] = p.bn_sh[u][i];
  }
  // leaf_free: no input of C(k-1) involves M(k-2)
  for (int ri = 0; ri < 2; ri++) {
    const int u = p.k - 1;
    d.leaf_free[ri] = 0;
    if (p.k < 4) continue;
    const int dep = ri == 0 ? d.dep[u] : -1;
    const int n_in = dep >= 0 ? d.num_bn_sh[u] : d.num_bn[u];
    const int *in = dep >= 0 ? d.bn_sh[u] : d.bn[u];
    bool free_ = true;
    for (int i = 0; i < n_in; i++) free_ = free_ && in[i] != u - 1;
    d.leaf_free[ri] = free_ ? 1 : 0;
  }
  d.free_mask = 0;
  d.leaf_in = 0;
  d.lf_a = -1;
  d.lf_excl = 0;
  d.tx_w = -1;
  d.tx_lead = -1;
  d.tail_t = 0;
  d.tri_tail = 0;
  if (!p.labeled && p.k >= 5) {  // the triangle's first vertex a = k-3 >= 2
    const int a = p.k - 3, b = p.k - 2, c = p.k - 1;
    auto bnset = [&](int v) {
      int m = 0;
      for (int i = 0; i < p.num_bn[v]; i++) m |= 1 << p.bn[v][i];
      return m;
    };
    bool ok = bnset(b) == (bnset(a) | 1 << a) && bnset(c) == (bnset(a) | 1 << a | 1 << b);
    // less[x][y]: M(x) < M(y) in every match (transitive closure of LT / GT)
    bool less[kMaxK][kMaxK] = {};
    for (int v = 1; v < p.k; v++)
      for (int j = 0; j < v; j++) {
        if (p.cond[v][j] == TDFS_LT) less[v][j] = true;
        if (p.cond[v][j] == TDFS_GT) less[j][v] = true;
      }
    for (int m = 0; m < p.k; m++)
      for (int x = 0; x < p.k; x++)
        for (int y = 0; y < p.k; y++)
          if (less[x][m] && less[m][y]) less[x][y] = true;
    for (int x = a; ok && x <= c; x++) {
      ok = !less[x][x];  // a cycle of LT / GT admits no match
      for (int y = x + 1; ok && y <= c; y++) ok = less[x][y] || less[y][x];  // a, b, c totally ordered
    }
    // every one of a, b, c stands to each earlier vertex j exactly as a's own
    // condition on j says: the candidates of a then hold every match's
    // M(b) and M(c) too, and every element of them meets b's and c's
    // conditions on j
    for (int j = 0; ok && j < a; j++) {
      const int o = p.cond[a][j];
      for (int v = a; ok && v <= c; v++)
        ok = less[v][j] == (o == TDFS_LT) && less[j][v] == (o == TDFS_GT);
    }
    d.tri_tail = ok ? 1 : 0;
  }
  if (p.labeled && p.k >= 5) {
    // labeled: a triangle over one set of earlier neighbours, three labels
    const int a = p.k - 3, b = p.k - 2, c = p.k - 1;
    auto bnset = [&](int v) {
      int m = 0;
      for (int i = 0; i < p.num_bn[v]; i++) m |= 1 << p.bn[v][i];
      return m;
    };
    const bool ok = bnset(b) == (bnset(a) | 1 << a) && bnset(c) == (bnset(a) | 1 << a | 1 << b) &&
                    p.label[a] != p.label[b] && p.label[a] != p.label[c] && p.label[b] != p.label[c];
    d.tri_tail = ok ? 2 : 0;
  }
  d.tail_t = tail_length(p);
  if (d.tail_t) {
    const int u0 = p.k - d.tail_t;
    for (int v = u0; v < p.k; v++) {
      bool lead = true;
      for (int w = u0; w < v; w++) lead = lead && (p.labeled ? p.label[w] != p.label[v] : false);
      int g = 0;
      for (int w = v; lead && w < p.k; w++) g += !p.labeled || p.label[w] == p.label[v];
      d.tail_gsz[v] = g;
    }
    const int w = u0 - 1;
    int lead = -1;
    bool ok = w >= 2;
    for (int v = u0; ok && v < p.k; v++) {
      ok = p.adj[v][w] == 0;
      if (!p.labeled) ok = ok && p.cond[v][w] == TDFS_NE;
      else if (lead < 0 && p.label[v] == p.label[w]) lead = v;
    }
    if (!p.labeled) lead = u0;
    if (ok && lead >= 0) {
      d.tx_w = w;
      d.tx_lead = lead;
    }
  }
  for (int j = 0; j + 1 < p.k; j++) {
    const int u = p.k - 1;
    bool in = p.num_bn[u] > 0;
    for (int i = 0; i < p.num_bn[u]; i++) in = in && p.adj[j][p.bn[u][i]] != 0;  // never for j in BN(u)
    if (in) d.leaf_in |= 1 << j;
  }
  for (int w = 2; p.labeled && w <= p.k - 2; w++) {
    bool free_ = true;
    for (int v = w + 1; v < p.k; v++) {
      if (d.label[v] == d.label[w]) free_ = false;
      for (int i = 0; i < p.num_bn[v]; i++) free_ = free_ && p.bn[v][i] != w;
    }
    if (free_) d.free_mask |= 1 << w;
  }
  // H(u): the inputs of C(u) other than N(M(u-1)); worth a list of its own
  // when at least two of them remain to intersect
  for (int ri = 0; ri < 2; ri++) {
    for (int u = 3; u < p.k; u++) {
      const int dep = ri == 0 ? d.dep[u] : -1;
      const int n_in = dep >= 0 ? d.num_bn_sh[u] : d.num_bn[u];
      const int *in = dep >= 0 ? d.bn_sh[u] : d.bn[u];
      bool has_prev = false;
      int nx = 0;
      for (int i = 0; i < n_in; i++) {
        if (in[i] == u - 1) has_prev = true;
        else d.hx[ri][u][nx++] = in[i];
      }
      d.hx_n[ri][u] = nx;
      d.hoist[ri][u] = (dep >= 0 ? 1 : 0) + nx >= 2 ? (has_prev ? 1 : 2) : 0;
      // the last vertex the hoisted inputs depend on (C(dep) on M(0..dep-1))
      int f = dep >= 0 ? dep - 1 : -1;
      for (int i = 0; i < nx; i++) f = std::max(f, d.hx[ri][u][i]);
      d.hfix[ri][u] = d.hoist[ri][u] ? std::max(f, 1) : -1;
    }
  }
}

static int check_plan(const tdfs_plan &p) {
  if (p.k < 3 || p.k > kMaxK) return fail(TDFS_E_ARG, "plan k=%d outside [3, %d]", p.k, kMaxK);
  for (int u = 1; u < p.k; u++) {
    if (p.num_bn[u] < 1 || p.num_bn[u] > u) return fail(TDFS_E_ARG, "vertex %d has no backward neighbour", u);
    for (int i = 0; i < p.num_bn[u]; i++)
      if (p.bn[u][i] < 0 || p.bn[u][i] >= u) return fail(TDFS_E_ARG, "bad backward neighbour");
    const int dep = p.shared_lvl[u];
    if (dep >= 0) {
      if (dep < 2 || dep >= u) return fail(TDFS_E_ARG, "bad shared_lvl[%d]=%d", u, dep);
      if (p.num_bn_sh[u] < 0 || p.num_bn_sh[u] > 6) return fail(TDFS_E_ARG, "bad num_bn_sh");
      for (int i = 0; i < p.num_bn_sh[u]; i++)
        if (p.bn_sh[u][i] < 0 || p.bn_sh[u][i] >= u) return fail(TDFS_E_ARG, "bad bn_sh");
    }
  }
  return TDFS_OK;
}

static int launch_reserve(Launch &L, int nplans, long long waves, long long wstride, unsigned long long qcap) {
  const size_t need = (size_t)waves * (size_t)wstride;
  if (L.wbuf_external) {
    if (need > L.wbuf_ints)
      return fail(TDFS_E_BUDGET, "workspace of %zu bytes exceeds the region of %zu bytes", need * sizeof(int),
                  L.wbuf_ints * sizeof(int));
  } else if (need > L.wbuf_ints) {
    if (L.wbuf) HIP_TRY(hipFree(L.wbuf));
    L.wbuf = nullptr;
    L.wbuf_ints = 0;
    hipError_t e = hipMalloc(&L.wbuf, need * sizeof(int));
    if (e != hipSuccess) return fail(TDFS_E_OOM, "workspace of %zu bytes: %s", need * sizeof(int), hipGetErrorString(e));
    L.wbuf_ints = need;
  }
  if (qcap != L.q_cap) {
    if (L.q_data) HIP_TRY(hipFree(L.q_data));
    if (L.q_seq) HIP_TRY(hipFree(L.q_seq));
    L.q_data = nullptr;
    L.q_seq = nullptr;
    HIP_TRY(hipMalloc(&L.q_data, sizeof(TaskRec) * qcap));
    HIP_TRY(hipMalloc(&L.q_seq, sizeof(unsigned long long) * qcap));
    L.q_cap = qcap;
  }
  if (nplans > L.plan_cap) {
    if (L.plans) HIP_TRY(hipFree(L.plans));
    if (L.cursors) HIP_TRY(hipFree(L.cursors));
    if (L.counts) HIP_TRY(hipFree(L.counts));
    if (L.stats) HIP_TRY(hipFree(L.stats));
    if (L.h_counts) HIP_TRY(hipHostFree(L.h_counts));
    if (L.h_stats) HIP_TRY(hipHostFree(L.h_stats));
    if (L.h_plans) HIP_TRY(hipHostFree(L.h_plans));
    L.plans = nullptr;
    L.cursors = nullptr;
    L.counts = L.stats = L.h_counts = L.h_stats = nullptr;
    L.h_plans = nullptr;
    L.plan_cap = 0;
    HIP_TRY(hipMalloc(&L.plans, sizeof(DevPlan) * nplans));
    HIP_TRY(hipMalloc(&L.cursors, sizeof(unsigned long long) * 16 * std::max(nplans, kXcds)));
    HIP_TRY(hipMalloc(&L.counts, sizeof(unsigned long long) * nplans));
    HIP_TRY(hipMalloc(&L.stats, sizeof(unsigned long long) * 4 * nplans));
    HIP_TRY(hipHostMalloc(&L.h_counts, sizeof(unsigned long long) * nplans, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(&L.h_stats, sizeof(unsigned long long) * 4 * nplans, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc(&L.h_plans, sizeof(DevPlan) * nplans, hipHostMallocDefault));
    L.plan_cap = nplans;
  }
  return TDFS_OK;
}

// Per-wave workspace (ints) for plans of at most kmax vertices: the level-0
// edge pairs, C(2..kmax-2) and H(3..kmax-1), each list dstride long.
static long long wave_ints(int kmax, int dstride) {
  return 2ll * kChunkMax + 2ll * std::max(0, kmax - 3) * dstride;
}

static int graph_dstride(const tdfs_graph *g) { return std::max(64, (g->dmax + 63) / 64 * 64); }

// Checks the plans, sizes and fills the launch's buffers and parameters.
// blocks_cap > 0 limits the persistent grid to that many workgroups per CU.
// ranges (2 per plan, may be null): each plan's own level-0 edge range
// [ranges[2i], ranges[2i+1]) (end < 0: all), counted with per-plan cursors.
// fixed: every buffer of L is already in place (plan table of plan_cap
// entries, ring of q_cap tasks, workspace of wbuf_ints); nothing is
// allocated, and the grid shrinks to the workgroups whose stacks fit.
// Positions 2..k-1 of a labeled plan that the counting kernel enumerates
// one candidate at a time: all but its free levels (w <= k-2, no later
// vertex adjacent or of the same label: one candidate's subtree stands for
// the others), its tail (counted from candidate counts) and the vertex right
// above the tail when no tail vertex is adjacent to it (summed over).  *fsum: the sum
// of the free levels' positions; *emask: bit w for each enumerated position.
static int enumerated_levels(const tdfs_plan &p, int *fsum, int *emask) {
  const int t = tail_length(p);
  const int tail0 = t ? p.k - t : p.k;
  int n = 0;
  *fsum = 0;
  *emask = 0;
  for (int w = 2; w < tail0; w++) {
    bool f = w <= p.k - 2;
    for (int v = w + 1; v < p.k && f; v++) f = p.adj[w][v] == 0 && p.label[v] != p.label[w];
    // right above the tail and adjacent to none of it: summed over, not enumerated
    bool x = t && w == tail0 - 1;
    for (int v = tail0; v < p.k && x; v++) x = p.adj[w][v] == 0;
    if (f || x) {
      *fsum += w;
    } else {
      n++;
      *emask |= 1 << w;
    }
  }
  return n;
}

// Matching order of a labeled plan for the counting kernel.  A labeled
// count is that of embeddings (injective only among vertices of one label),
// which no connected order changes.  Vertices 0 and 1 (the level-0 edge)
// stay; of the connected orders of the others whose level-0 degree filter
// stays exact (prune_is_sound) and in which every enumerated vertex keeps at
// least the backward neighbours it has in the plan's order (no enumerated
// candidate set grows for want of constraints), the one with the fewest
// enumerated levels, then with its free levels at the shallowest positions
// (fewer prefixes to count them under), then closest to the plan's order, is
// taken when it beats the plan's own on the first two.  Otherwise q = p.
static void counting_order(const tdfs_plan &p, tdfs_plan &q, int *chosen = nullptr) {
  q = p;
  if (chosen)
    for (int i = 0; i < p.k; i++) chosen[i] = i;
  if (!p.labeled || p.k < 4 || !prune_is_sound(p)) return;
  int fs0 = 0, em0 = 0;
  const int base = enumerated_levels(p, &fs0, &em0);
  int best = base, best_fs = fs0, best_disp = 0;
  bool found = false;
  int best_order[kMaxK] = {};
  int rest[kMaxK];
  for (int i = 0; i < p.k - 2; i++) rest[i] = i + 2;
  do {
    int order[kMaxK] = {0, 1};
    bool conn = true;
    for (int i = 2; i < p.k; i++) {
      order[i] = rest[i - 2];
      bool any = false;
      for (int j = 0; j < i && !any; j++) any = p.adj[order[i]][order[j]] != 0;
      conn = conn && any;
    }
    if (!conn) continue;
    // the fields the scores read (labels, adjacency, degrees, backward
    // neighbours); the plan itself is built for the order taken
    tdfs_plan r;
    std::memset(&r, 0, sizeof(r));
    r.k = p.k;
    r.labeled = p.labeled;
    for (int i = 0; i < p.k; i++) {
      r.label[i] = p.label[order[i]];
      r.deg[i] = p.deg[order[i]];
      for (int j = 0; j < p.k; j++) r.adj[i][j] = p.adj[order[i]][order[j]];
      for (int j = 0; j < i; j++)
        if (r.adj[i][j]) r.bn[i][r.num_bn[i]++] = j;
    }
    if (!prune_is_sound(r)) continue;
    int fs = 0, em = 0, disp = 0;
    const int e = enumerated_levels(r, &fs, &em);
    bool kept = true;
    for (int i = 2; i < p.k; i++) {
      disp += std::abs(order[i] - i);
      if ((em >> i) & 1) kept = kept && r.num_bn[i] >= p.num_bn[order[i]];
    }
    if (!kept) continue;
    if (e < best || (e == best && (fs < best_fs || (fs == best_fs && found && disp < best_disp)))) {
      found = true;
      best = e;
      best_fs = fs;
      best_disp = disp;
      std::memcpy(best_order, order, sizeof(order));
    }
  } while (std::next_permutation(rest, rest + p.k - 2));
  tdfs_plan r;
  if (found && tdfs::plan_permuted(p, best_order, &r) == TDFS_OK) {
    q = r;
    if (chosen)
      for (int i = 0; i < p.k; i++) chosen[i] = best_order[i];
  }
}

extern "C" int tdfs_plan_counting_order(const tdfs_plan *plan, tdfs_plan *out, int32_t *order) {
  if (!plan || !out) return fail(TDFS_E_ARG, "null argument");
  int rc = check_plan(*plan);
  if (rc) return rc;
  int chosen[kMaxK];
  tdfs_plan q;
  counting_order(*plan, q, chosen);
  *out = q;
  if (order)
    for (int i = 0; i < plan->k; i++) order[i] = chosen[i];
  return TDFS_OK;
}

// Leaf factor of a plan for the counting kernel.  The last vertex v = k-1
// depends on M(0..a) only, a = max(1, last backward neighbour of v) <= k-4,
// when every later vertex j in (a, v) that v must differ from (all of them
// unlabeled, those of v's label labeled) is adjacent to every backward
// neighbour of v -- M(j) then always lies among v's candidates and is one of
// them to exclude -- and, unlabeled, v has no LT / GT condition (an excluded
// M(j) always meets v's conditions).  v's completions are then the same for
// every completion of vertices a+1..k-2: the count of v's candidates given
// M(0..a), less the excluded ones.  Returns a (and *excl), or -1.
static int leaf_factor(const tdfs_plan &p, int *excl) {
  const int v = p.k - 1;
  if (p.k < 5 || p.num_bn[v] < 1) return -1;
  int f = 0;
  for (int i = 0; i < p.num_bn[v]; i++) f = std::max(f, (int)p.bn[v][i]);
  const int a = std::max(f, 1);
  if (a > p.k - 4) return -1;
  for (int j = 0; !p.labeled && j < v; j++)
    if (p.cond[v][j] != TDFS_NE) return -1;
  int n = 0;
  for (int j = a + 1; j < v; j++) {
    if (p.labeled && p.label[j] != p.label[v]) continue;
    for (int i = 0; i < p.num_bn[v]; i++)
      if (!p.adj[j][p.bn[v][i]]) return -1;
    n++;
  }
  *excl = n;
  return a;
}

// The plan the counting kernel runs: the pattern's first k-1 vertices, with
// the dropped leaf kept at index k-1 for the per-path factor.  Derived
// fields are the first k-1 vertices' own, except that a free level must also
// be free of the dropped leaf.
static void to_dev_plan_factored(const tdfs_plan &p, int a, int excl, const std::vector<int32_t> &label_values,
                                 DevPlan &d) {
  DevPlan full;
  to_dev_plan(p, label_values, full);
  tdfs_plan r = p;
  r.k = p.k - 1;
  to_dev_plan(r, label_values, d);
  const int v = p.k - 1;
  d.label[v] = full.label[v];
  d.deg[v] = full.deg[v];
  d.num_bn[v] = full.num_bn[v];
  for (int i = 0; i < 8; i++) d.bn[v][i] = full.bn[v][i];
  d.gt_mask[v] = full.gt_mask[v];
  d.lt_mask[v] = full.lt_mask[v];
  d.same_label_mask[v] = full.same_label_mask[v];
  d.dep[v] = -1;
  d.num_bn_sh[v] = 0;
  for (int ri = 0; ri < 2; ri++) {
    d.hoist[ri][v] = 0;
    d.hx_n[ri][v] = 0;
    d.hfix[ri][v] = -1;
  }
  d.free_mask &= full.free_mask;
  // the factor is taken when M(a) is fixed: a shortcut that counts vertex a
  // itself (a tail or closing triangle starting at or before a) would skip it
  if (d.tail_t && p.k - 1 - d.tail_t <= a) d.tail_t = 0;
  if (!d.tail_t || d.tx_w <= a) d.tx_w = -1;
  if (d.tri_tail && p.k - 1 - 3 <= a) d.tri_tail = 0;
  d.lf_a = a;
  d.lf_excl = excl;
}

static int launch_prepare(const tdfs_ctx *ctx, Launch &L, const tdfs_graph *g, const tdfs_plan *plans, int nplans,
                          int64_t edge_begin, int64_t edge_end, bool stats, int blocks_cap, KParams &P,
                          const int64_t *ranges = nullptr, bool fixed = false, bool emit = false) {
  if (!ctx || !g || !plans || nplans < 1) return fail(TDFS_E_ARG, "bad arguments");
  if (g->device != ctx->device) return fail(TDFS_E_ARG, "graph and context on different devices");
  for (int i = 0; i < nplans; i++) {
    int rc = check_plan(plans[i]);
    if (rc) return rc;
    if (plans[i].labeled && !g->labeled) return fail(TDFS_E_ARG, "labeled plan on a graph uploaded without labels");
  }
  HIP_TRY(hipSetDevice(ctx->device));
  tdfs_tuning t = ctx->tune;
  fill_defaults(&t);
  if (edge_begin < 0) edge_begin = 0;
  if (edge_end < 0 || edge_end > g->nnz) edge_end = g->nnz;
  if (edge_begin > edge_end) edge_begin = edge_end;
  const int chunk = t.chunk_edges;
  const long long span = edge_end - edge_begin;
  const int dstride = graph_dstride(g);
  if (g->dmax > kIeMax) return fail(TDFS_E_ARG, "maximum degree %d beyond %d", g->dmax, kIeMax);
  int kmax = 3;
  for (int i = 0; i < nplans; i++) kmax = std::max(kmax, (int)plans[i].k);
  const long long wstride = wave_ints(kmax, dstride);
  const int wv = stats ? 0 : t.waves_per_simd;
  const bool st = t.steal == 1 && t.merge_min <= 0;  // merge path and stealing are separate variants
  const bool mg = t.merge_min > 0;
  const int occ = mg ? (stats ? ctx->merge_blocks_per_cu_stats
                              : (wv == 5 ? ctx->merge_blocks_per_cu_w5 : ctx->merge_blocks_per_cu))
                : st ? (stats ? ctx->steal_blocks_per_cu_stats
                              : (wv > 0 ? ctx->steal_blocks_per_cu_w[wv] : ctx->steal_blocks_per_cu))
                     : (stats ? ctx->blocks_per_cu_stats
                              : (wv > 0 ? ctx->blocks_per_cu_w[wv] : ctx->blocks_per_cu));
  int bpc = t.blocks_per_cu > 0 ? std::min(t.blocks_per_cu, occ) : occ;
  if (blocks_cap > 0) bpc = std::min(bpc, blocks_cap);
  int grid = ctx->cus * bpc;
  if (fixed) {
    const long long fit = (long long)(L.wbuf_ints / (size_t)wstride) / kWavesPerBlock;
    if (fit < 1)
      return fail(TDFS_E_BUDGET, "one workgroup's DFS stacks need %lld bytes (%d-vertex plans, max degree %d); "
                  "the buffer holds %zu", (long long)wstride * kWavesPerBlock * (long long)sizeof(int), kmax,
                  g->dmax, L.wbuf_ints * sizeof(int));
    if (nplans > L.plan_cap) return fail(TDFS_E_ARG, "%d plans, room for %d", nplans, L.plan_cap);
    grid = (int)std::min<long long>(grid, fit);
  } else {
    const long long waves = (long long)grid * kWavesPerBlock;
    int rc = launch_reserve(L, nplans, waves, wstride, 1ull << t.queue_log2);
    if (rc) return rc;
  }
  L.grid = grid;
  bool all_sound = true;
  for (int i = 0; i < nplans; i++) all_sound = all_sound && prune_is_sound(plans[i]);
  const GraphView &V = (g->ordered && all_sound) ? g->view[1] : g->view[0];
  for (int i = 0; i < nplans; i++) {
    // the counting kernel may take another order (the statistics follow the
    // plan's; enumeration writes vertices in the plan's order)
    tdfs_plan q;
    if (!stats && !emit && t.free_order == 1) counting_order(plans[i], q);
    else q = plans[i];
    int excl = 0;
    const int la = (!stats && !emit && t.tails == 1 && !st) ? leaf_factor(q, &excl) : -1;
    if (la >= 1) to_dev_plan_factored(q, la, excl, g->label_values, L.h_plans[i]);
    else to_dev_plan(q, g->label_values, L.h_plans[i]);
    long long lo = edge_begin, hi = edge_end;
    if (ranges) {
      lo = std::max<int64_t>(0, ranges[2 * i]);
      hi = ranges[2 * i + 1] < 0 || ranges[2 * i + 1] > g->nnz ? g->nnz : ranges[2 * i + 1];
      if (lo > hi) lo = hi;
    }
    L.h_plans[i].e_lo = lo;
    L.h_plans[i].e_span = hi - lo;
  }
  HIP_TRY(hipMemcpyAsync(L.plans, L.h_plans, sizeof(DevPlan) * nplans, hipMemcpyHostToDevice, L.stream));

  std::memset(&P, 0, sizeof(P));
  P.rowptr = V.rowptr;
  P.colidx = V.colidx;
  P.degv = V.degv;
  P.src = V.src;
  P.maxnbr = V.maxnbr;
  P.labels = V.labels;
  P.hub_idx = V.hub_idx;
  P.hub_bits = V.hub_bits;
  P.hub_words = V.hub_words;
  P.colidx_lab = V.colidx_lab;
  P.loff = V.loff;
  P.nlab = V.nlab;
  P.plans = L.plans;
  P.nplans = nplans;
  P.chunk = chunk;
  P.edge_lo = edge_begin;
  P.edge_hi = edge_end;
  P.span = (unsigned long long)span;
  P.cursors = L.cursors;
  P.interleave = ((t.cursor_order == 1 || t.cursor_order == 3) && !ranges) ? 1 : 0;
  P.xcd_groups = P.interleave && t.cursor_order == 1 && nplans > 1 ? kXcds : 1;
  P.wbuf = L.wbuf;
  P.wstride = wstride;
  P.hoff = 2ll * kChunkMax + (long long)std::max(0, kmax - 3) * dstride;
  P.dstride = dstride;
  P.q_data = L.q_data;
  P.q_seq = L.q_seq;
  P.q_mask = L.q_cap - 1;
  P.counts = L.counts;
  P.stats = L.stats;
  P.prof = L.prof;
  P.ctrl = L.ctrl;
  const long long ticks_per_us = std::max(1, ctx->wallclock_khz / 1000);
  P.timeout_ticks = t.timeout_us < 0 ? -1 : (long long)t.timeout_us * ticks_per_us;
  P.deadline_ticks = t.idle_deadline_ms < 0 ? 0 : (long long)t.idle_deadline_ms * 1000 * ticks_per_us;
  P.prog_ticks = 10ll * 1000 * ticks_per_us;  // 10 ms
  P.split_chunk = t.split_chunk;
  P.donate_levels = t.donate_levels;
  P.reuse = t.reuse == 1 ? 1 : 0;
  P.leaf_batch = t.leaf_batch == 1 ? 1 : 0;
  P.leaf_memo = t.leaf_memo == 1 && P.leaf_batch ? 1 : 0;
  P.leaf_hash = t.leaf_hash == 1 ? 3 : (t.leaf_hash == 3 ? 1 : 0);
  P.ext_hash = t.ext_hash == 1 ? 1 : 0;
  P.tails = t.tails == 1 ? 1 : 0;
  P.limit_ticks = t.time_limit_ms > 0 ? (long long)t.time_limit_ms * 1000 * ticks_per_us : 0;
  P.variant = wv;
  P.leaf_pair_cap = t.leaf_pair_cap;
  P.merge_min = t.merge_min > 0 ? t.merge_min : 0;
  P.steal = t.steal == 1 && !P.merge_min ? 1 : 0;
  P.hoist_min = t.hoist_min > 0 ? t.hoist_min : 0;
  return TDFS_OK;
}

// Enqueues reset + match kernel + result copies on the launch's stream.
// No memset and no event nodes, so the sequence can be stream-captured.
static int launch_enqueue(Launch &L, const KParams &P, bool stats) {
  const int init_blocks = (int)std::min<unsigned long long>(1024, (L.q_cap + 255) / 256);
  hipLaunchKernelGGL(tdfs_init_kernel, dim3(init_blocks), dim3(256), 0, L.stream, L.ctrl, L.counts, L.stats,
                     P.nplans, L.q_seq, L.q_cap, L.prof, L.cursors, std::max(P.nplans, P.xcd_groups));
  HIP_TRY(hipGetLastError());
  if (P.merge_min) {
    if (stats)
      hipLaunchKernelGGL(tdfs_match_kernel_merge<true>, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
    else if (P.variant == 5)
      hipLaunchKernelGGL(tdfs_match_kernel_w5_merge, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
    else
      hipLaunchKernelGGL(tdfs_match_kernel_merge<false>, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
  } else if (P.steal) {
    if (stats)
      hipLaunchKernelGGL(tdfs_match_kernel_steal<true>, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
    else if (P.variant == 8)
      hipLaunchKernelGGL(tdfs_match_kernel_w8_steal, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
    else if (P.variant == 6)
      hipLaunchKernelGGL(tdfs_match_kernel_w6_steal, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
    else if (P.variant == 5)
      hipLaunchKernelGGL(tdfs_match_kernel_w5_steal, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
    else if (P.variant == 4)
      hipLaunchKernelGGL(tdfs_match_kernel_w4_steal, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
    else
      hipLaunchKernelGGL(tdfs_match_kernel_steal<false>, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
  } else if (stats)
    hipLaunchKernelGGL(tdfs_match_kernel<true>, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
  else if (P.variant == 8)
    hipLaunchKernelGGL(tdfs_match_kernel_w8, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
  else if (P.variant == 6)
    hipLaunchKernelGGL(tdfs_match_kernel_w6, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
  else if (P.variant == 5)
    hipLaunchKernelGGL(tdfs_match_kernel_w5, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
  else if (P.variant == 4)
    hipLaunchKernelGGL(tdfs_match_kernel_w4, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
  else
    hipLaunchKernelGGL(tdfs_match_kernel<false>, dim3(L.grid), dim3(kBlock), 0, L.stream, P);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(L.h_counts, L.counts, sizeof(unsigned long long) * P.nplans, hipMemcpyDeviceToHost,
                         L.stream));
  HIP_TRY(hipMemcpyAsync(L.h_stats, L.stats, sizeof(unsigned long long) * 4 * P.nplans, hipMemcpyDeviceToHost,
                         L.stream));
  HIP_TRY(hipMemcpyAsync(L.h_ctrl, L.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, L.stream));
  HIP_TRY(hipMemcpyAsync(L.h_prof, L.prof, kProf * sizeof(unsigned long long), hipMemcpyDeviceToHost, L.stream));
  return TDFS_OK;
}

// Reads the pinned result mirrors after the launch's stream has drained.
static int launch_collect(const Launch &L, const tdfs_plan *plans, int nplans, double ms, tdfs_result *out,
                          unsigned long long *prof_out) {
  if (L.h_ctrl->error) {
    const int e = L.h_ctrl->error;
    return fail(TDFS_E_HIP, "device scheduler error %d (%s)", e,
                e == 4 ? "idle deadline: no wave reported progress" : "task ring handshake");
  }
  if (prof_out) std::memcpy(prof_out, L.h_prof, kProf * sizeof(unsigned long long));
  for (int i = 0; i < nplans; i++) {
    tdfs_result &r = out[i];
    std::memset(&r, 0, sizeof(r));
    r.raw = L.h_counts[i];
    r.alg_bytes = L.h_stats[4 * i + 0];
    r.extensions = L.h_stats[4 * i + 1];
    r.donated = L.h_stats[4 * i + 2];
    r.flags = (uint32_t)(L.h_stats[4 * i + 3] & 1ull);
    r.stolen = (uint32_t)(L.h_stats[4 * i + 3] >> 32);
    r.kernel_ms = ms;
    if (L.h_stats[4 * i + 3] & 2ull) return fail(TDFS_E_OVERFLOW, "raw count of plan %d overflows 64 bits", i);
    if (plans[i].labeled) {
      r.count = r.raw;
    } else {
      const unsigned __int128 prod = (unsigned __int128)r.raw * plans[i].multiplicity;
      if (prod >> 64) return fail(TDFS_E_OVERFLOW, "count of plan %d overflows 64 bits", i);
      r.count = (uint64_t)prod;
    }
  }
  return TDFS_OK;
}

static int run_plans(tdfs_ctx *ctx, const tdfs_graph *g, const tdfs_plan *plans, int nplans, int64_t edge_begin,
                     int64_t edge_end, tdfs_result *out) {
  if (!out) return fail(TDFS_E_ARG, "null out");
  Launch &L = ctx->main;
  KParams P;
  int rc = launch_prepare(ctx, L, g, plans, nplans, edge_begin, edge_end, ctx->stats != 0, 0, P);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(L.ev0, L.stream));
  rc = launch_enqueue(L, P, ctx->stats != 0);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(L.ev1, L.stream));
  HIP_TRY(hipStreamSynchronize(L.stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, L.ev0, L.ev1));
  return launch_collect(L, plans, nplans, ms, out, ctx->stats ? ctx->last_prof : nullptr);
}

// ---------------------------------------------------------------------------
// Prepared batches: plans uploaded once, the launch sequence captured into a
// hipGraph on the batch's own stream, replayed by tdfs_batch_run.
// ---------------------------------------------------------------------------
struct tdfs_batch {
  tdfs_ctx *ctx;
  const tdfs_graph *g;
  std::vector<tdfs_plan> plans;
  Launch L;
  bool stats;
};

extern "C" int tdfs_batch_create(tdfs_ctx *ctx, const tdfs_graph *g, const tdfs_plan *plans, int32_t nplans,
                                 int32_t blocks_per_cu, tdfs_batch **out) {
  if (!ctx || !g || !plans || nplans < 1 || !out) return fail(TDFS_E_ARG, "bad arguments");
  HIP_TRY(hipSetDevice(ctx->device));
  tdfs_batch *b = new tdfs_batch();
  b->ctx = ctx;
  b->g = g;
  b->plans.assign(plans, plans + nplans);
  b->stats = ctx->stats != 0;
  int rc = launch_init(b->L, ctx->device);
  KParams P;
  if (!rc) rc = launch_prepare(ctx, b->L, g, plans, nplans, -1, -1, b->stats, blocks_per_cu, P);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(b->L.stream);  // plan upload done before capture
    if (e) rc = fail(TDFS_E_HIP, "sync: %s", hipGetErrorString(e));
  }
  if (!rc) {
    hipError_t e = hipStreamBeginCapture(b->L.stream, hipStreamCaptureModeThreadLocal);
    if (e) rc = fail(TDFS_E_HIP, "begin capture: %s", hipGetErrorString(e));
    if (!rc) {
      const int erc = launch_enqueue(b->L, P, b->stats);
      e = hipStreamEndCapture(b->L.stream, &b->L.graph);
      if (erc) rc = erc;
      else if (e) rc = fail(TDFS_E_HIP, "end capture: %s", hipGetErrorString(e));
    }
  }
  if (!rc) {
    hipError_t e = hipGraphInstantiate(&b->L.exec, b->L.graph, nullptr, nullptr, 0);
    if (e) rc = fail(TDFS_E_HIP, "instantiate: %s", hipGetErrorString(e));
  }
  if (rc) {
    launch_free(b->L);
    delete b;
    return rc;
  }
  *out = b;
  return TDFS_OK;
}

// Starts one replay without waiting (results via tdfs_batch_wait).
extern "C" int tdfs_batch_launch(tdfs_batch *b) {
  if (!b) return fail(TDFS_E_ARG, "null batch");
  HIP_TRY(hipSetDevice(b->ctx->device));
  HIP_TRY(hipEventRecord(b->L.ev0, b->L.stream));
  HIP_TRY(hipGraphLaunch(b->L.exec, b->L.stream));
  HIP_TRY(hipEventRecord(b->L.ev1, b->L.stream));
  return TDFS_OK;
}

// Returns 1 when the last replay has finished, 0 if still running.
extern "C" int tdfs_batch_done(const tdfs_batch *b) {
  if (!b) return fail(TDFS_E_ARG, "null batch");
  const hipError_t e = hipEventQuery(b->L.ev1);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return fail(TDFS_E_HIP, "event query: %s", hipGetErrorString(e));
}

extern "C" int tdfs_batch_wait(tdfs_batch *b, tdfs_result *out) {
  if (!b || !out) return fail(TDFS_E_ARG, "null argument");
  HIP_TRY(hipEventSynchronize(b->L.ev1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, b->L.ev0, b->L.ev1));
  return launch_collect(b->L, b->plans.data(), (int)b->plans.size(), ms, out, nullptr);
}

extern "C" int tdfs_batch_run(tdfs_batch *b, tdfs_result *out) {
  int rc = tdfs_batch_launch(b);
  if (rc) return rc;
  return tdfs_batch_wait(b, out);
}

extern "C" void tdfs_batch_destroy(tdfs_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->ctx->device);
  launch_free(b->L);
  delete b;
}

// ---------------------------------------------------------------------------
// Serving: continuous batching over a device-resident graph.
//
// The server owns `slots` launches (own stream, task ring, control block,
// plan table), a device-memory arena for the DFS stacks of the running
// launches, and host-mapped completion records.  Submitted queries wait in a
// FIFO.  Admission is by DFS-stack memory: a launch over plans of at most
// kmax vertices needs waves x (level-0 pairs + 2 (kmax - 3) lists of dmax
// entries) of stack, carved from the arena; the next waiting queries start on
// an idle slot as soon as a region of that size is free, so light (small-k)
// queries pack more launches into the same memory.  Every plan of a running
// launch reports its own completion (count and flags in host-mapped memory,
// written by the wave that ends the plan's last unit of work), so a query is
// collected as soon as its own work is done, not when its batch-mates are.
// ---------------------------------------------------------------------------
struct ServeQuery {
  uint64_t ticket;
  tdfs_plan plan;
  int64_t lo, hi;  // level-0 edge range (hi < 0: all)
  int view;        // numbering counted on: 1 degree-ordered copy, 0 as given
  double t_submit;
};

struct ServeSlot {
  Launch L;
  std::vector<ServeQuery> running;
  std::vector<char> reported;
  unsigned long long *done_host = nullptr;  // host-mapped: per plan {raw count, 1 | flags << 8}
  unsigned long long *done_dev = nullptr;   // the same memory as the device sees it
  size_t off = 0, bytes = 0;                // stack region of the arena in use
  double t_launch = 0;
  long long launch_no = 0;                  // server-wide number of the launch in this slot (1-based)
  bool busy = false;
};

struct tdfs_server {
  tdfs_ctx *ctx;
  const tdfs_graph *g;
  int blocks_per_cu;
  int max_batch;
  double budget;
  int32_t est_samples;
  long long waves;     // per launch
  int dstride;
  char *arena = nullptr;
  size_t arena_bytes = 0;
  std::vector<ServeSlot> slots;
  std::deque<ServeQuery> waiting;
  uint64_t next_ticket = 1;
  long long launches = 0;  // launches started (or attempted)
  // error-path tests: launch number whose start fails (TDFS_SERVE_FAIL_START)
  // or whose device run is taken as failed (TDFS_SERVE_FAIL_RUN); 0: none
  long long fail_start = 0, fail_run = 0;
};

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static size_t serve_need_bytes(const tdfs_server *sv, int kmax) {
  return (size_t)sv->waves * (size_t)wave_ints(kmax, sv->dstride) * sizeof(int);
}

static void serve_free(tdfs_server *sv) {
  for (auto &sl : sv->slots) {
    launch_free(sl.L);
    if (sl.done_host) (void)hipHostFree(sl.done_host);
    sl.done_host = nullptr;
  }
  if (sv->arena) (void)hipFree(sv->arena);
  delete sv;
}

extern "C" int tdfs_serve_create(tdfs_ctx *ctx, const tdfs_graph *g, int32_t slots, int32_t blocks_per_cu,
                                 int32_t max_batch, int64_t stack_mb, double budget, tdfs_server **out) {
  if (!ctx || !g || !out || slots < 1 || slots > 64 || max_batch < 1 || max_batch > 4096)
    return fail(TDFS_E_ARG, "bad server arguments");
  if (g->device != ctx->device) return fail(TDFS_E_ARG, "graph and context on different devices");
  HIP_TRY(hipSetDevice(ctx->device));
  tdfs_server *sv = new tdfs_server();
  sv->ctx = ctx;
  sv->g = g;
  sv->blocks_per_cu = blocks_per_cu > 0 ? std::min(blocks_per_cu, ctx->blocks_per_cu) : 1;
  sv->max_batch = max_batch;
  sv->budget = budget;
  sv->est_samples = 4000;
  sv->waves = (long long)ctx->cus * sv->blocks_per_cu * kWavesPerBlock;
  sv->dstride = graph_dstride(g);
  // default arena: every slot can run a launch of 7-vertex plans at once
  sv->arena_bytes = stack_mb > 0 ? (size_t)stack_mb << 20 : (size_t)slots * serve_need_bytes(sv, kMaxK);
  if (sv->arena_bytes < serve_need_bytes(sv, 3)) {
    const size_t need = serve_need_bytes(sv, 3);
    delete sv;
    return fail(TDFS_E_BUDGET, "stack memory of %lld MiB holds no launch (3-vertex plans need %zu bytes)",
                (long long)stack_mb, need);
  }
  if (hipMalloc(&sv->arena, sv->arena_bytes) != hipSuccess) {
    const size_t b = sv->arena_bytes;
    delete sv;
    return fail(TDFS_E_OOM, "serving stack arena of %zu bytes", b);
  }
  // host-forced failures for tests of the error paths (the kernels run normally)
  if (const char *s = std::getenv("TDFS_SERVE_FAIL_START")) sv->fail_start = std::atoll(s);
  if (const char *s = std::getenv("TDFS_SERVE_FAIL_RUN")) sv->fail_run = std::atoll(s);
  tdfs_tuning t = ctx->tune;
  fill_defaults(&t);
  sv->slots.resize(slots);
  for (auto &sl : sv->slots) {
    int rc = launch_init(sl.L, ctx->device);
    // the ring, control block and plan table are fixed here; stacks come from the arena
    sl.L.wbuf_external = true;
    if (!rc) rc = launch_reserve(sl.L, max_batch, 0, 0, 1ull << t.queue_log2);
    if (!rc && hipHostMalloc(&sl.done_host, sizeof(unsigned long long) * 2 * max_batch,
                             hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
      rc = fail(TDFS_E_OOM, "completion records");
    if (!rc && hipHostGetDevicePointer((void **)&sl.done_dev, sl.done_host, 0) != hipSuccess)
      rc = fail(TDFS_E_HIP, "completion records: no device address");
    if (rc) {
      serve_free(sv);
      return rc;
    }
  }
  *out = sv;
  return TDFS_OK;
}

extern "C" int tdfs_serve_submit_range(tdfs_server *sv, const tdfs_plan *plan, int64_t edge_begin, int64_t edge_end,
                                       uint64_t *ticket) {
  if (!sv || !plan || !ticket) return fail(TDFS_E_ARG, "null argument");
  int rc = check_plan(*plan);
  if (rc) return rc;
  if (plan->labeled && !sv->g->labeled) return fail(TDFS_E_ARG, "labeled query on an unlabeled graph");
  const size_t need = serve_need_bytes(sv, plan->k);
  if (need > sv->arena_bytes)
    return fail(TDFS_E_BUDGET, "query refused: its DFS stacks need %zu bytes, the server holds %zu", need,
                sv->arena_bytes);
  if (sv->budget > 0) {
    double est[3];
    rc = tdfs_graph_estimate(sv->g, plan, sv->est_samples, 1, sv->budget, est);
    if (rc) return rc;
    if (est[2] != 0.0 || est[0] > sv->budget)
      return fail(TDFS_E_BUDGET, "query refused: estimated work %.3g exceeds the budget %.3g", est[0], sv->budget);
  }
  ServeQuery q;
  q.ticket = sv->next_ticket++;
  q.plan = *plan;
  q.lo = edge_begin;
  q.hi = edge_end;
  q.view = (sv->g->ordered && prune_is_sound(*plan)) ? 1 : 0;
  q.t_submit = now_s();
  sv->waiting.push_back(q);
  *ticket = q.ticket;
  return TDFS_OK;
}

extern "C" int tdfs_serve_submit(tdfs_server *sv, const tdfs_plan *plan, uint64_t *ticket) {
  return tdfs_serve_submit_range(sv, plan, 0, -1, ticket);
}

// First free region of the arena of at least `bytes` (regions of busy slots
// are taken), or SIZE_MAX.
static size_t arena_fit(const tdfs_server *sv, size_t bytes) {
  std::vector<std::pair<size_t, size_t>> used;
  for (const auto &sl : sv->slots)
    if (sl.busy) used.emplace_back(sl.off, sl.bytes);
  std::sort(used.begin(), used.end());
  size_t at = 0;
  for (const auto &u : used) {
    if (u.first >= at + bytes) return at;
    at = std::max(at, u.first + u.second);
  }
  return sv->arena_bytes >= at + bytes ? at : SIZE_MAX;
}

// The result of running query i of a slot from its completion record
// (or, done_rec false, from the launch's final counts).
static void serve_result(const ServeSlot &sl, size_t i, bool done_rec, tdfs_result &r) {
  std::memset(&r, 0, sizeof(r));
  unsigned long long f = 0;
  if (done_rec) {
    r.raw = __atomic_load_n(&sl.done_host[2 * i + 0], __ATOMIC_RELAXED);
    f = __atomic_load_n(&sl.done_host[2 * i + 1], __ATOMIC_RELAXED) >> 8;
  } else {
    r.raw = sl.L.h_counts[i];
    f = sl.L.h_stats[4 * i + 3] & 0xffull;
  }
  r.flags = (uint32_t)(f & 1ull);
  if (f & 2ull) r.flags |= TDFS_FAILED;  // raw count overflowed 64 bits
  r.kernel_ms = (now_s() - sl.t_launch) * 1e3;
  const tdfs_plan &p = sl.running[i].plan;
  if (p.labeled) {
    r.count = r.raw;
  } else {
    const unsigned __int128 prod = (unsigned __int128)r.raw * p.multiplicity;
    if (prod >> 64) r.flags |= TDFS_FAILED;
    r.count = (uint64_t)prod;
  }
}

// Collects finished queries into (tickets, results, latency_s), at most cap,
// and starts idle slots on waiting queries.  *nout = entries written.  A
// launch that fails on the device reports its unfinished queries with
// TDFS_FAILED and frees its slot; a launch that cannot be started puts its
// queries back at the head of the queue and returns the error.
extern "C" int tdfs_serve_poll(tdfs_server *sv, uint64_t *tickets, tdfs_result *results, double *latency_s,
                               int32_t cap, int32_t *nout) {
  if (!sv || !nout || (cap > 0 && (!tickets || !results))) return fail(TDFS_E_ARG, "null argument");
  *nout = 0;
  HIP_TRY(hipSetDevice(sv->ctx->device));
  int n = 0;
  auto report = [&](ServeSlot &sl, size_t i, const tdfs_result &r, double t) {
    tickets[n] = sl.running[i].ticket;
    results[n] = r;
    if (latency_s) latency_s[n] = t - sl.running[i].t_submit;
    sl.reported[i] = 1;
    n++;
  };
  for (auto &sl : sv->slots) {
    if (!sl.busy) continue;
    // a launch forced to fail reports every query at its end, all as failed
    const bool forced = sv->fail_run > 0 && sl.launch_no == sv->fail_run;
    // queries whose own work has ended
    for (size_t i = 0; i < sl.running.size() && n < cap && !forced; i++) {
      if (sl.reported[i]) continue;
      if (!(__atomic_load_n(&sl.done_host[2 * i + 1], __ATOMIC_ACQUIRE) & 1ull)) continue;
      tdfs_result r;
      serve_result(sl, i, true, r);
      report(sl, i, r, now_s());
    }
    const hipError_t e = hipEventQuery(sl.L.ev1);
    if (e == hipErrorNotReady) continue;
    const bool failed = e != hipSuccess || sl.L.h_ctrl->error != 0 || forced;
    size_t left = 0;
    for (size_t i = 0; i < sl.running.size(); i++) left += !sl.reported[i];
    if ((int)left > cap - n) continue;  // no room to report the rest of this slot yet
    const double t = now_s();
    for (size_t i = 0; i < sl.running.size(); i++) {
      if (sl.reported[i]) continue;
      tdfs_result r;
      serve_result(sl, i, false, r);
      if (failed) r.flags |= TDFS_FAILED;
      report(sl, i, r, t);
    }
    if (failed)
      fail(TDFS_E_HIP, "serving launch failed: %s",
           e != hipSuccess ? hipGetErrorString(e)
                           : (sl.L.h_ctrl->error != 0 ? "device scheduler error" : "forced (TDFS_SERVE_FAIL_RUN)"));
    sl.running.clear();
    sl.reported.clear();
    sl.busy = false;
  }
  for (auto &sl : sv->slots) {
    if (sl.busy || sv->waiting.empty()) continue;
    // the longest FIFO prefix (same numbering, at most max_batch) whose stacks fit a free region
    int kmax = 3;
    size_t take = 0, off = SIZE_MAX, bytes = 0;
    const int view = sv->waiting.front().view;
    while (take < sv->waiting.size() && (int)take < sv->max_batch && sv->waiting[take].view == view) {
      const int k2 = std::max(kmax, (int)sv->waiting[take].plan.k);
      const size_t b2 = serve_need_bytes(sv, k2);
      const size_t o2 = arena_fit(sv, b2);
      if (o2 == SIZE_MAX) break;
      kmax = k2;
      off = o2;
      bytes = b2;
      take++;
    }
    if (take == 0) break;  // no stack memory free: wait for a running launch to end
    std::vector<tdfs_plan> plans;
    std::vector<int64_t> ranges;
    for (size_t i = 0; i < take; i++) {
      sl.running.push_back(sv->waiting.front());
      plans.push_back(sv->waiting.front().plan);
      ranges.push_back(sv->waiting.front().lo);
      ranges.push_back(sv->waiting.front().hi);
      sv->waiting.pop_front();
    }
    sl.reported.assign(take, 0);
    sl.L.wbuf = reinterpret_cast<int *>(sv->arena + off);
    sl.L.wbuf_ints = bytes / sizeof(int);
    std::memset(sl.done_host, 0, sizeof(unsigned long long) * 2 * take);
    const long long no = ++sv->launches;
    KParams P;
    int rc = sv->fail_start > 0 && no == sv->fail_start
                 ? fail(TDFS_E_OOM, "launch %lld not started (forced: TDFS_SERVE_FAIL_START)", no)
                 : launch_prepare(sv->ctx, sl.L, sv->g, plans.data(), (int)plans.size(), -1, -1, false,
                                  sv->blocks_per_cu, P, ranges.data(), true);
    if (!rc) {
      P.pend = sl.L.cursors + 8;
      P.done_host = sl.done_dev;
      hipError_t e = hipEventRecord(sl.L.ev0, sl.L.stream);
      if (e != hipSuccess) rc = fail(TDFS_E_HIP, "event record: %s", hipGetErrorString(e));
      if (!rc) rc = launch_enqueue(sl.L, P, false);
      if (!rc && (e = hipEventRecord(sl.L.ev1, sl.L.stream)) != hipSuccess)
        rc = fail(TDFS_E_HIP, "event record: %s", hipGetErrorString(e));
    }
    if (rc) {
      // nothing was started: the queries go back to the head of the queue in order
      for (size_t i = sl.running.size(); i-- > 0;) sv->waiting.push_front(sl.running[i]);
      sl.running.clear();
      sl.reported.clear();
      *nout = n;
      return rc;
    }
    sl.off = off;
    sl.bytes = bytes;
    sl.t_launch = now_s();
    sl.launch_no = no;
    sl.busy = true;
  }
  *nout = n;
  return TDFS_OK;
}

// out4 = {queries waiting, slots busy, slots, queries running (not yet collected)}
extern "C" int tdfs_serve_status(const tdfs_server *sv, int64_t *out4) {
  if (!sv || !out4) return fail(TDFS_E_ARG, "null argument");
  int busy = 0;
  int64_t running = 0;
  for (const auto &sl : sv->slots) {
    busy += sl.busy;
    for (char r : sl.reported) running += !r;
  }
  out4[0] = (int64_t)sv->waiting.size();
  out4[1] = busy;
  out4[2] = (int64_t)sv->slots.size();
  out4[3] = running;
  return TDFS_OK;
}

extern "C" void tdfs_serve_destroy(tdfs_server *sv) {
  if (!sv) return;
  (void)hipSetDevice(sv->ctx->device);
  for (auto &sl : sv->slots)
    if (sl.L.stream) (void)hipStreamSynchronize(sl.L.stream);
  serve_free(sv);
}

extern "C" int tdfs_ctx_last_profile(const tdfs_ctx *ctx, uint64_t *out8) {
  if (!ctx || !out8) return fail(TDFS_E_ARG, "null argument");
  for (int i = 0; i < 8; i++) out8[i] = ctx->last_prof[i];
  return TDFS_OK;
}

extern "C" int tdfs_ctx_last_counters(const tdfs_ctx *ctx, uint64_t *out8) {
  if (!ctx || !out8) return fail(TDFS_E_ARG, "null argument");
  for (int i = 0; i < 8; i++) out8[i] = ctx->last_prof[8 + i];
  return TDFS_OK;
}

extern "C" int tdfs_count(tdfs_ctx *ctx, const tdfs_graph *g, const tdfs_plan *plan, int64_t edge_begin,
                          int64_t edge_end, tdfs_result *out) {
  return run_plans(ctx, g, plan, 1, edge_begin, edge_end, out);
}

extern "C" int tdfs_count_batch(tdfs_ctx *ctx, const tdfs_graph *g, const tdfs_plan *plans, int32_t nplans,
                                tdfs_result *out) {
  return run_plans(ctx, g, plans, nplans, -1, -1, out);
}

extern "C" int tdfs_enumerate(tdfs_ctx *ctx, const tdfs_graph *g, const tdfs_plan *plan, int64_t edge_begin,
                              int64_t edge_end, int32_t *rows, uint64_t cap, uint64_t *written, tdfs_result *out) {
  if (!ctx || !g || !plan || !written || !out || (cap > 0 && !rows)) return fail(TDFS_E_ARG, "bad arguments");
  const int k = plan->k;
  if (cap > (uint64_t)1 << 40) return fail(TDFS_E_ARG, "cap %llu too large", (unsigned long long)cap);
  Launch &L = ctx->main;
  KParams P;
  int rc = launch_prepare(ctx, L, g, plan, 1, edge_begin, edge_end, false, 0, P, nullptr, false, true);
  if (rc) return rc;
  int *d_rows = nullptr;
  unsigned long long *d_n = nullptr;
  auto release = [&]() {
    if (d_rows) (void)hipFree(d_rows);
    if (d_n) (void)hipFree(d_n);
  };
  if (cap > 0 && hipMalloc(&d_rows, sizeof(int) * (size_t)cap * (size_t)k) != hipSuccess) {
    release();
    return fail(TDFS_E_OOM, "enumeration buffer of %llu rows", (unsigned long long)cap);
  }
  if (hipMalloc(&d_n, sizeof(unsigned long long)) != hipSuccess) {
    release();
    return fail(TDFS_E_OOM, "enumeration counter");
  }
  P.emit = d_rows;
  P.emit_cap = cap;
  P.emit_n = d_n;
  P.leaf_batch = 0;  // leaves are walked one parent at a time to write their rows
  P.leaf_memo = 0;
  hipError_t e = hipMemsetAsync(d_n, 0, sizeof(unsigned long long), L.stream);
  if (e == hipSuccess) e = hipEventRecord(L.ev0, L.stream);
  if (e != hipSuccess) {
    release();
    return fail(TDFS_E_HIP, "enumerate: %s", hipGetErrorString(e));
  }
  rc = launch_enqueue(L, P, false);
  if (rc) {
    release();
    return rc;
  }
  unsigned long long produced = 0;
  e = hipEventRecord(L.ev1, L.stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&produced, d_n, sizeof(produced), hipMemcpyDeviceToHost, L.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(L.stream);
  const uint64_t nw = std::min<uint64_t>(produced, cap);
  if (e == hipSuccess && nw > 0)
    e = hipMemcpy(rows, d_rows, sizeof(int) * (size_t)nw * (size_t)k, hipMemcpyDeviceToHost);
  release();
  if (e != hipSuccess) return fail(TDFS_E_HIP, "enumerate: %s", hipGetErrorString(e));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, L.ev0, L.ev1));
  rc = launch_collect(L, plan, 1, ms, out, nullptr);
  if (rc) return rc;
  // rows are in the numbering counted on: map the degree-ordered copy back
  if (g->ordered && prune_is_sound(*plan)) {
    std::vector<int32_t> inv(g->n);
    for (int32_t v = 0; v < g->n; v++) inv[g->perm[v]] = v;
    for (uint64_t i = 0; i < nw * (uint64_t)k; i++) rows[i] = inv[rows[i]];
  }
  *written = nw;
  return TDFS_OK;
}

// ---------------------------------------------------------------------------
// tdfs_count_noalloc: tdfs_count for callers that cannot allocate.  Every
// buffer of the launch (control block, result and statistics words, cursor,
// plan, a task ring of kScratchRing entries, and the per-wave DFS stacks) is
// carved from `scratch`, a file-scope device array of kScratchBytes that
// each GPU holds from module load until the process ends.  Host code never
// uses the array's name as an address: tdfs_ctx_create looks up its device
// address once per device (hipGetSymbolAddress, with that device current)
// and every later use goes through that pointer.  The grid shrinks to the
// workgroups whose DFS stacks fit the array; a plan whose single workgroup
// does not fit is refused with TDFS_E_BUDGET.  All calls on one GPU share
// the array and the context's stream: calls that overlap on one GPU (from
// two contexts, or two threads) race on it.  The launch's init kernel clears
// what must start at zero, so the previous call's contents do not matter.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = 256ull << 20;
constexpr unsigned long long kScratchRing = 1ull << 14;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];  // device address of `scratch` on each device

// Set-up, once per device, with that device current (from tdfs_ctx_create).
static int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(TDFS_E_ARG, "device %d beyond %d", device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return TDFS_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return fail(TDFS_E_HIP, "scratch address: %s", hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return TDFS_OK;
}

extern "C" int tdfs_count_noalloc(tdfs_ctx *ctx, const tdfs_graph *g, const tdfs_plan *plan, int64_t edge_begin,
                                  int64_t edge_end, tdfs_result *out) {
  if (!ctx || !g || !plan || !out) return fail(TDFS_E_ARG, "null argument");
  void *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = g_scratch_ptr[ctx->device];
  }
  if (!base) return fail(TDFS_E_STATE, "scratch of device %d not set up", ctx->device);
  // carve the scratch array: 256-byte aligned regions, the stacks last
  char *at = static_cast<char *>(base);
  auto take = [&](size_t bytes) {
    char *p = at;
    at += (bytes + 255) / 256 * 256;
    return p;
  };
  Launch L;
  L.device = ctx->device;
  L.stream = ctx->main.stream;
  L.ev0 = ctx->main.ev0;
  L.ev1 = ctx->main.ev1;
  L.ctrl = reinterpret_cast<Ctrl *>(take(sizeof(Ctrl)));
  L.counts = reinterpret_cast<unsigned long long *>(take(sizeof(unsigned long long)));
  L.stats = reinterpret_cast<unsigned long long *>(take(4 * sizeof(unsigned long long)));
  L.prof = reinterpret_cast<unsigned long long *>(take(kProf * sizeof(unsigned long long)));
  L.cursors = reinterpret_cast<unsigned long long *>(take(16 * sizeof(unsigned long long)));
  L.plans = reinterpret_cast<DevPlan *>(take(sizeof(DevPlan)));
  L.plan_cap = 1;
  L.q_seq = reinterpret_cast<unsigned long long *>(take(kScratchRing * sizeof(unsigned long long)));
  L.q_data = reinterpret_cast<TaskRec *>(take(kScratchRing * sizeof(TaskRec)));
  L.q_cap = kScratchRing;
  const size_t used = (size_t)(at - static_cast<char *>(base));
  L.wbuf = reinterpret_cast<int *>(at);
  L.wbuf_ints = (kScratchBytes - used) / sizeof(int);
  L.wbuf_external = true;
  // host sides of the plan upload and result copies: this call's own frame
  DevPlan h_plan;
  unsigned long long h_counts[1], h_stats[4], h_prof[kProf];
  Ctrl h_ctrl;
  L.h_plans = &h_plan;
  L.h_counts = h_counts;
  L.h_stats = h_stats;
  L.h_prof = h_prof;
  L.h_ctrl = &h_ctrl;
  const bool stats = ctx->stats != 0;
  KParams P;
  int rc = launch_prepare(ctx, L, g, plan, 1, edge_begin, edge_end, stats, 0, P, nullptr, true);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(L.ev0, L.stream));
  rc = launch_enqueue(L, P, stats);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(L.ev1, L.stream));
  HIP_TRY(hipStreamSynchronize(L.stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, L.ev0, L.ev1));
  return launch_collect(L, plan, 1, ms, out, stats ? ctx->last_prof : nullptr);
}
