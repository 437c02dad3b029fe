// tdfs_host.cpp -- host side of libtdfs: .g parsing, the matching planner,
// the reference binary CSR format, and seeded graph/pattern generators.
//
// Planner semantics follow STMatch::PatternPreprocessor (src/pattern.h):
//   Aut(P)                 : every automorphism of the uncoloured pattern graph
//                            (bliss generators + closure, src/pattern.h:581-663;
//                            PatternMultiplicity by enumeration, :279-310).  Here the
//                            group is found by a degree-pruned backtracking search.
//   symmetry-breaking pairs: GetConditions, src/pattern.h:665-707
//   condition list         : src/pattern.h:416-448
//   backward neighbours    : src/pattern.h:342-358
//   result reuse + veto    : src/pattern.h:370-413, 459-480
// The dead STMatch slot outputs (get_set_ops ... convert2oned, :217-270,
// :493-579) are not produced: the kernel never reads them.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "tdfs.h"
#include "tdfs_internal.h"

namespace tdfs {

static thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

const char *last_error() { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// Pattern parsing (src/pattern.h:110-143).
// ---------------------------------------------------------------------------
struct RawPattern {
  int k = 0;
  int adj[TDFS_MAX_K][TDFS_MAX_K] = {};
  int label[TDFS_MAX_K] = {};
};

static int parse_g(const char *text, int labeled, RawPattern *rp) {
  std::istringstream in(text);
  std::string line;
  int k = 0;
  bool saw_edge = false;
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ls(line);
    char tag = 0;
    ls >> tag;
    if (tag == 'v') {
      if (saw_edge) return fail(TDFS_E_FORMAT, "vertex line after edge lines");
      int id, lab;
      if (!(ls >> id >> lab)) return fail(TDFS_E_FORMAT, "bad vertex line '%s'", line.c_str());
      if (id != k) return fail(TDFS_E_FORMAT, "vertex ids must be 0..k-1 in order (got %d)", id);
      if (k >= TDFS_MAX_K) return fail(TDFS_E_FORMAT, "pattern has more than %d vertices", TDFS_MAX_K);
      rp->label[k++] = labeled ? lab : 1;
    } else if (tag == 'e') {
      saw_edge = true;
      int a, b, lab;
      if (!(ls >> a >> b >> lab)) return fail(TDFS_E_FORMAT, "bad edge line '%s'", line.c_str());
      if (a < 0 || b < 0 || a >= k || b >= k || a == b)
        return fail(TDFS_E_FORMAT, "bad edge endpoints %d %d", a, b);
      if (lab != 1) return fail(TDFS_E_FORMAT, "edge label must be 1 (got %d)", lab);
      rp->adj[a][b] = rp->adj[b][a] = 1;
    } else {
      return fail(TDFS_E_FORMAT, "unrecognised line '%s'", line.c_str());
    }
  }
  rp->k = k;
  if (k < 3) return fail(TDFS_E_FORMAT, "pattern needs at least 3 vertices (got %d)", k);
  return TDFS_OK;
}

// ---------------------------------------------------------------------------
// Aut(P) by backtracking: image of vertex i is chosen among vertices of equal
// degree, consistent with the adjacency to every already-mapped vertex.
// ---------------------------------------------------------------------------
using Perm = std::vector<int>;

static void aut_search(const RawPattern &p, const int *deg, int i, Perm &img, std::vector<bool> &used,
                       std::vector<Perm> &out) {
  const int k = p.k;
  if (i == k) {
    out.push_back(img);
    return;
  }
  for (int c = 0; c < k; c++) {
    if (used[c] || deg[c] != deg[i]) continue;
    bool ok = true;
    for (int j = 0; j < i && ok; j++) ok = (p.adj[i][j] == p.adj[c][img[j]]);
    if (!ok) continue;
    used[c] = true;
    img[i] = c;
    aut_search(p, deg, i + 1, img, used, out);
    used[c] = false;
  }
}

static std::vector<Perm> automorphisms(const RawPattern &p, const int *deg) {
  std::vector<Perm> out;
  Perm img(p.k, -1);
  std::vector<bool> used(p.k, false);
  aut_search(p, deg, 0, img, used, out);
  return out;
}

// GetConditions (src/pattern.h:665-707).
static std::vector<std::pair<int, int>> conditions(int k, std::vector<Perm> aut) {
  std::vector<std::pair<int, int>> result;
  auto classes = [&](const std::vector<Perm> &g) {
    std::map<int, std::set<int>> ec;
    for (int i = 0; i < k; i++) {
      std::set<int> orbit;
      for (const Perm &q : g) orbit.insert(q[i]);
      if (orbit.empty()) orbit.insert(i);  // trivial group: {id}
      ec[*orbit.begin()].insert(orbit.begin(), orbit.end());
    }
    return ec;
  };
  for (;;) {
    auto ec = classes(aut);
    auto it = std::find_if(ec.begin(), ec.end(), [](const auto &e) { return e.second.size() > 1; });
    if (it == ec.end()) break;
    const std::set<int> &cls = it->second;
    const int n0 = *cls.begin();
    for (const Perm &q : aut) {
      int best = -1;
      for (int x : cls) {
        if (x == n0) continue;
        if (best < 0 || q[x] < q[best]) best = x;
      }
      result.emplace_back(n0, best);
    }
    aut.erase(std::remove_if(aut.begin(), aut.end(), [n0](const Perm &q) { return q[n0] != n0; }),
              aut.end());
  }
  std::sort(result.begin(), result.end());
  result.erase(std::unique(result.begin(), result.end()), result.end());
  return result;
}

static int build_plan(const RawPattern &rp, int labeled, tdfs_plan *pl) {
  std::memset(pl, 0, sizeof(*pl));
  const int k = rp.k;
  pl->k = k;
  pl->labeled = labeled ? 1 : 0;
  for (int i = 0; i < k; i++) {
    pl->label[i] = rp.label[i];
    for (int j = 0; j < k; j++) pl->adj[i][j] = rp.adj[i][j];
  }
  // pattern degrees (src/pattern.h:189-195)
  for (int i = 0; i < k; i++) {
    int d = 0;
    for (int j = 0; j < k; j++) d += rp.adj[i][j] > 0;
    pl->deg[i] = d;
  }
  std::vector<Perm> aut = automorphisms(rp, pl->deg);
  pl->multiplicity = aut.size();
  // sym01: vertex 1 lies in the orbit of vertex 0 (src/pattern.h:312-340)
  pl->sym01 = 0;
  for (const Perm &q : aut)
    if (q[0] == 1) pl->sym01 = 1;
  auto pairs = conditions(k, aut);
  if (pairs.size() > 21) return fail(TDFS_E_FORMAT, "too many symmetry pairs");
  pl->npairs = (int)pairs.size();
  for (size_t i = 0; i < pairs.size(); i++) {
    pl->pairs[i][0] = pairs[i].first;
    pl->pairs[i][1] = pairs[i].second;
  }
  // backward neighbours (src/pattern.h:342-358)
  for (int u = 1; u < k; u++)
    for (int j = 0; j < u; j++)
      if (rp.adj[u][j]) pl->bn[u][pl->num_bn[u]++] = j;
  // conditions (src/pattern.h:416-448): SetConditions puts (LT, b) on a and
  // (GT, a) on b for each pair (a, b); every other j < u is NE.
  for (int u = 0; u < k; u++)
    for (int j = 0; j < u; j++) {
      int op = TDFS_NE;
      for (auto &pr : pairs) {
        if (pr.first == u && pr.second == j) { op = TDFS_LT; break; }
        if (pr.second == u && pr.first == j) { op = TDFS_GT; break; }
      }
      pl->cond[u][j] = op;
    }
  // result reuse (src/pattern.h:370-413)
  unsigned nb[TDFS_MAX_K] = {};
  for (int u = 1; u < k; u++)
    for (int j = 0; j < u; j++)
      if (rp.adj[j][u]) nb[u] |= 1u << j;
  for (int u = 0; u < k; u++) pl->shared_lvl[u] = -1;
  for (int u = 2; u < k; u++) {
    for (int j = u - 1; j >= 1; j--) {
      if ((nb[u] & nb[j]) == nb[j] && pl->num_bn[u] >= 2 && pl->num_bn[j] >= 2 &&
          rp.label[u] == rp.label[j]) {
        pl->shared_lvl[u] = j;
        break;
      }
    }
    const int dep = pl->shared_lvl[u];
    if (dep < 0) continue;
    for (int j = u - 1; j >= 0; j--)
      if ((nb[u] >> j & 1u) && !(nb[dep] >> j & 1u)) pl->bn_sh[u][pl->num_bn_sh[u]++] = j;
  }
  // unlabeled veto (src/pattern.h:459-480): every condition of dep must
  // appear unchanged among the conditions of u.
  if (!labeled)
    for (int u = 0; u < k; u++) {
      const int dep = pl->shared_lvl[u];
      if (dep < 0) continue;
      for (int j = 0; j < dep; j++)
        if (pl->cond[dep][j] != pl->cond[u][j]) {
          pl->shared_lvl[u] = -1;
          break;
        }
    }
  for (int u = 0; u < k; u++)
    if (pl->shared_lvl[u] < 0) pl->num_bn_sh[u] = 0;
  // every vertex u >= 1 needs a backward neighbour (device assert, src/gpu_match.cu:638-639)
  for (int u = 1; u < k; u++)
    if (pl->num_bn[u] == 0)
      return fail(TDFS_E_FORMAT, "vertex %d has no backward neighbour (order is not connected)", u);
  return TDFS_OK;
}

// ---------------------------------------------------------------------------
// PRNG: splitmix64 seeding + xoshiro256**.
// ---------------------------------------------------------------------------
struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto &w : s) w = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
  uint64_t next() {
    const uint64_t res = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return res;
  }
  // unbiased integer in [0, n) by rejection on the top bits
  uint64_t below(uint64_t n) {
    if (n <= 1) return 0;
    const uint64_t lim = (~0ull) - ((~0ull) % n);
    for (;;) {
      const uint64_t x = next();
      if (x < lim) return x % n;
    }
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Open-addressing set of undirected edges keyed (min << 32 | max).
struct EdgeSet {
  std::vector<uint64_t> slots;
  uint64_t mask;
  explicit EdgeSet(uint64_t expect) {
    uint64_t cap = 16;
    while (cap < expect * 2) cap <<= 1;
    slots.assign(cap, ~0ull);
    mask = cap - 1;
  }
  bool insert(uint64_t key) {
    uint64_t h = key * 0x9E3779B97F4A7C15ull;
    h ^= h >> 29;
    for (uint64_t i = h & mask;; i = (i + 1) & mask) {
      if (slots[i] == key) return false;
      if (slots[i] == ~0ull) {
        slots[i] = key;
        return true;
      }
    }
  }
};

static void edges_to_csr(int32_t n, const std::vector<uint64_t> &edges, int64_t *rowptr, int32_t *colidx) {
  std::vector<int64_t> deg(n + 1, 0);
  for (uint64_t e : edges) {
    deg[(int32_t)(e >> 32)]++;
    deg[(int32_t)(e & 0xffffffffu)]++;
  }
  rowptr[0] = 0;
  for (int32_t v = 0; v < n; v++) rowptr[v + 1] = rowptr[v] + deg[v];
  std::vector<int64_t> fill(rowptr, rowptr + n);
  for (uint64_t e : edges) {
    const int32_t a = (int32_t)(e >> 32), b = (int32_t)(e & 0xffffffffu);
    colidx[fill[a]++] = b;
    colidx[fill[b]++] = a;
  }
  for (int32_t v = 0; v < n; v++) std::sort(colidx + rowptr[v], colidx + rowptr[v + 1]);
}

static inline uint64_t ekey(uint32_t a, uint32_t b) {
  if (a > b) std::swap(a, b);
  return ((uint64_t)a << 32) | b;
}

// Walker alias table over weights w.
struct Alias {
  std::vector<double> prob;
  std::vector<int32_t> alias;
  explicit Alias(const std::vector<double> &w) {
    const size_t n = w.size();
    prob.resize(n);
    alias.resize(n);
    double sum = 0;
    for (double x : w) sum += x;
    std::vector<double> sc(n);
    std::vector<int32_t> small, large;
    for (size_t i = 0; i < n; i++) {
      sc[i] = w[i] * (double)n / sum;
      (sc[i] < 1.0 ? small : large).push_back((int32_t)i);
    }
    while (!small.empty() && !large.empty()) {
      int32_t s = small.back(), l = large.back();
      small.pop_back();
      prob[s] = sc[s];
      alias[s] = l;
      sc[l] = (sc[l] + sc[s]) - 1.0;
      if (sc[l] < 1.0) {
        large.pop_back();
        small.push_back(l);
      }
    }
    for (int32_t l : large) prob[l] = 1.0, alias[l] = l;
    for (int32_t s : small) prob[s] = 1.0, alias[s] = s;
  }
  int32_t draw(Rng &r) const {
    const uint64_t i = r.below(prob.size());
    return r.uniform() < prob[i] ? (int32_t)i : alias[i];
  }
};

}  // namespace tdfs

// ---------------------------------------------------------------------------
// Work estimate: the DFS of the counting path (same level-0 filter, lists,
// conditions and reuse) expanded from nsample directed edges drawn uniformly
// within 16 equal edge-index strata (stratified estimator),
// charging each extension the probe elements it tests times the lists it
// tests them against.  Once the scaled estimate passes `cap` the expansion
// stops and the estimate is marked capped (a lower bound: the pattern is then
// treated as infeasible), so one call costs about cap * nsample / nnz work.
// ---------------------------------------------------------------------------
namespace {
struct Estimator {
  const int64_t *rowptr;
  const int32_t *colidx;
  const int32_t *labels;
  const tdfs_plan *p;
  int32_t M[TDFS_MAX_K];
  std::vector<int32_t> cand[TDFS_MAX_K];
  double cost = 0, raw = 0, cap = 0;
  double base = 0, scale = 1;  // scaled estimate so far = base + cost * scale
  bool capped = false;

  bool pass(int u, int32_t x) const {
    if (!p->labeled) {
      for (int j = 0; j < u; j++) {
        const int op = p->cond[u][j];
        if (op == TDFS_LT ? !(x < M[j]) : op == TDFS_GT ? !(x > M[j]) : x == M[j]) return false;
      }
      return true;
    }
    if (labels[x] != p->label[u]) return false;
    for (int j = 0; j < u; j++)
      if (M[j] == x) return false;
    return true;
  }

  void dfs(int u) {
    if (capped) return;
    const int32_t *lp[TDFS_MAX_K] = {};
    int64_t ln[TDFS_MAX_K] = {};  // every vertex u >= 1 has an input (checked plans)
    int nl = 0;
    const int dep = p->shared_lvl[u];
    if (dep >= 2) {
      lp[nl] = cand[dep].data();
      ln[nl++] = (int64_t)cand[dep].size();
      for (int i = 0; i < p->num_bn_sh[u]; i++) {
        const int32_t t = M[p->bn_sh[u][i]];
        lp[nl] = colidx + rowptr[t];
        ln[nl++] = rowptr[t + 1] - rowptr[t];
      }
    } else {
      for (int i = 0; i < p->num_bn[u]; i++) {
        const int32_t t = M[p->bn[u][i]];
        lp[nl] = colidx + rowptr[t];
        ln[nl++] = rowptr[t + 1] - rowptr[t];
      }
    }
    int pi = 0;
    for (int i = 1; i < nl; i++)
      if (ln[i] < ln[pi]) pi = i;
    cost += (double)ln[pi] * std::max(1, nl - 1) + nl;
    std::vector<int32_t> &out = cand[u];
    out.clear();
    for (int64_t e = 0; e < ln[pi]; e++) {
      const int32_t x = lp[pi][e];
      if (!pass(u, x)) continue;
      bool ok = true;
      for (int i = 0; i < nl && ok; i++)
        if (i != pi) ok = std::binary_search(lp[i], lp[i] + ln[i], x);
      if (ok) out.push_back(x);
    }
    if (u == p->k - 1) {
      raw += (double)out.size();
      return;
    }
    if (base + cost * scale > cap) {
      capped = true;
      return;
    }
    for (size_t i = 0; i < out.size() && !capped; i++) {
      M[u] = out[i];
      dfs(u + 1);
    }
  }
};
}  // namespace

namespace tdfs {
int estimate(int32_t n, int64_t nnz, const int64_t *rowptr, const int32_t *colidx, const int32_t *labels,
             const tdfs_plan *plan, int32_t nsample, uint64_t seed, double cap, double *out3,
             const int32_t *deg_in, const int32_t *maxnbr_in) {
  if (!rowptr || !plan || !out3 || nsample < 1) return fail(TDFS_E_ARG, "bad estimate arguments");
  if (plan->labeled && !labels) return fail(TDFS_E_ARG, "labeled plan needs labels");
  out3[0] = out3[1] = out3[2] = 0;
  if (nnz == 0) return TDFS_OK;
  auto degree = [&](int32_t v) { return deg_in ? deg_in[v] : (int32_t)(rowptr[v + 1] - rowptr[v]); };
  std::vector<int32_t> maxnbr_local;
  if (!maxnbr_in) maxnbr_local.assign(n, -1);
  auto max_nbr_deg = [&](int32_t c) {
    if (maxnbr_in) return maxnbr_in[c];
    if (maxnbr_local[c] < 0) {
      int m = 0;
      for (int64_t q = rowptr[c]; q < rowptr[c + 1]; q++) m = std::max(m, degree(colidx[q]));
      maxnbr_local[c] = m;
    }
    return maxnbr_local[c];
  };
  Rng r(seed);
  Estimator E;
  E.rowptr = rowptr;
  E.colidx = colidx;
  E.labels = labels;
  E.p = plan;
  E.cap = cap;
  // Stratified by edge index: rows sorted by degree put the hub rows in the
  // last strata, so every estimate samples them.
  const int32_t nstrata = std::max<int32_t>(1, std::min<int32_t>(16, nsample));
  const int32_t per = std::max<int32_t>(1, nsample / nstrata);
  double raw_scaled = 0;
  for (int32_t st = 0; st < nstrata && !E.capped; st++) {
    const int64_t lo = nnz * st / nstrata, hi = nnz * (st + 1) / nstrata;
    if (hi <= lo) continue;
    E.scale = (double)(hi - lo) / (double)per;
    E.cost = 0;
    E.raw = 0;
  for (int32_t s = 0; s < per && !E.capped; s++) {
    const int64_t e = lo + (int64_t)r.below((uint64_t)(hi - lo));
    const int32_t c = colidx[e];
    const int32_t rr = (int32_t)(std::upper_bound(rowptr, rowptr + n + 1, e) - rowptr - 1);
    E.cost += 2;
    if (!plan->labeled) {
      if (plan->sym01 && !(rr < c)) continue;
    } else if (labels[rr] != plan->label[0] || labels[c] != plan->label[1]) {
      continue;
    }
    if (degree(rr) < plan->deg[0] || degree(c) < plan->deg[1]) continue;
    if (max_nbr_deg(c) < plan->deg[2]) continue;
    E.M[0] = rr;
    E.M[1] = c;
    E.dfs(2);
    if (E.base + E.cost * E.scale > E.cap) E.capped = true;
  }
    E.base += E.cost * E.scale;
    raw_scaled += E.raw * E.scale;
  }
  out3[0] = E.base;
  out3[1] = raw_scaled;
  out3[2] = E.capped ? 1.0 : 0.0;
  return TDFS_OK;
}

int plan_permuted(const tdfs_plan &p, const int *order, tdfs_plan *out) {
  RawPattern rp;
  rp.k = p.k;
  for (int i = 0; i < p.k; i++) {
    if (order[i] < 0 || order[i] >= p.k) return fail(TDFS_E_ARG, "bad vertex order");
    rp.label[i] = p.label[order[i]];
    for (int j = 0; j < p.k; j++) rp.adj[i][j] = p.adj[order[i]][order[j]];
  }
  return build_plan(rp, p.labeled, out);
}
}  // namespace tdfs

using namespace tdfs;

extern "C" {

const char *tdfs_last_error(void) { return tdfs::last_error(); }
const char *tdfs_version(void) { return "tdfs-mi355x 0.1"; }

int tdfs_pattern_parse(const char *text, int32_t labeled, tdfs_plan *out) {
  if (!text || !out) return fail(TDFS_E_ARG, "null argument");
  RawPattern rp;
  int rc = parse_g(text, labeled, &rp);
  if (rc) return rc;
  return build_plan(rp, labeled, out);
}

int tdfs_pattern_load(const char *path, int32_t labeled, tdfs_plan *out) {
  if (!path) return fail(TDFS_E_ARG, "null path");
  std::ifstream f(path);
  if (!f.good()) return fail(TDFS_E_IO, "cannot open pattern file %s", path);
  std::stringstream ss;
  ss << f.rdbuf();
  return tdfs_pattern_parse(ss.str().c_str(), labeled, out);
}

static int write_g(int k, const int (*adj)[TDFS_MAX_K], const int *label, char *text_out, size_t cap) {
  std::string s = "# t 0\n";
  for (int i = 0; i < k; i++) s += "v " + std::to_string(i) + " " + std::to_string(label[i]) + "\n";
  for (int i = 0; i < k; i++)
    for (int j = i + 1; j < k; j++)
      if (adj[i][j]) s += "e " + std::to_string(i) + " " + std::to_string(j) + " 1\n";
  if (!text_out || s.size() + 1 > cap) return fail(TDFS_E_ARG, "text buffer too small (%zu)", s.size() + 1);
  std::memcpy(text_out, s.c_str(), s.size() + 1);
  return TDFS_OK;
}

int tdfs_pattern_reorder(const char *text, int32_t *order_out, char *text_out, size_t cap) {
  if (!text) return fail(TDFS_E_ARG, "null argument");
  RawPattern rp;
  int rc = parse_g(text, 1, &rp);
  if (rc) return rc;
  const int k = rp.k;
  int deg[TDFS_MAX_K];
  for (int i = 0; i < k; i++) {
    deg[i] = 0;
    for (int j = 0; j < k; j++) deg[i] += rp.adj[i][j];
  }
  // Start from the edge (a, b) maximising (deg a + deg b); then greedily add
  // the vertex with the most already-placed neighbours (ties: higher degree,
  // then lower id).  Position 2 is forced adjacent to position 1, which makes
  // the level-0 prune sound for every connected pattern.
  int best = -1, oa = -1, ob = -1;
  for (int a = 0; a < k; a++)
    for (int b = 0; b < k; b++) {
      if (a == b || !rp.adj[a][b]) continue;
      int score = (deg[a] + deg[b]) * 16 + deg[a];
      bool has2 = false;
      for (int c = 0; c < k; c++) has2 |= (c != a && c != b && rp.adj[b][c]);
      if (!has2) continue;
      if (score > best) best = score, oa = a, ob = b;
    }
  if (oa < 0) return fail(TDFS_E_FORMAT, "pattern is not connected or has no 3-vertex path");
  std::vector<int> order = {oa, ob};
  std::vector<bool> placed(k, false);
  placed[oa] = placed[ob] = true;
  while ((int)order.size() < k) {
    int pick = -1, pb = -1, pd = -1;
    const bool need_adj_b = order.size() == 2;
    for (int c = 0; c < k; c++) {
      if (placed[c]) continue;
      if (need_adj_b && !rp.adj[c][ob]) continue;
      int nb = 0;
      for (int x : order) nb += rp.adj[c][x];
      if (nb == 0) continue;
      if (nb > pb || (nb == pb && deg[c] > pd)) pick = c, pb = nb, pd = deg[c];
    }
    if (pick < 0) return fail(TDFS_E_FORMAT, "pattern is not connected");
    order.push_back(pick);
    placed[pick] = true;
  }
  int nadj[TDFS_MAX_K][TDFS_MAX_K] = {}, nlab[TDFS_MAX_K];
  for (int i = 0; i < k; i++) {
    nlab[i] = rp.label[order[i]];
    for (int j = 0; j < k; j++) nadj[i][j] = rp.adj[order[i]][order[j]];
    if (order_out) order_out[i] = order[i];
  }
  if (!text_out) return TDFS_OK;
  return write_g(k, nadj, nlab, text_out, cap);
}

int tdfs_graph_read_meta(const char *prefix, int64_t *meta4) {
  if (!prefix || !meta4) return fail(TDFS_E_ARG, "null argument");
  std::ifstream f(std::string(prefix) + ".meta.txt");
  if (!f.good()) return fail(TDFS_E_IO, "cannot open %s.meta.txt", prefix);
  long long a, b, c, d;
  if (!(f >> a >> b >> c >> d)) return fail(TDFS_E_FORMAT, "meta.txt needs 4 numbers");
  if (c != 4) return fail(TDFS_E_FORMAT, "vid_size must be 4 (src/graph.h:145), got %lld", c);
  if (a < 0 || b < 0 || a > INT32_MAX) return fail(TDFS_E_FORMAT, "bad sizes n=%lld nnz=%lld", a, b);
  meta4[0] = a, meta4[1] = b, meta4[2] = c, meta4[3] = d;
  return TDFS_OK;
}

static int read_exact(const std::string &path, void *dst, size_t bytes) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return fail(TDFS_E_IO, "cannot open %s", path.c_str());
  size_t got = bytes ? fread(dst, 1, bytes, f) : 0;
  fclose(f);
  if (got != bytes) return fail(TDFS_E_IO, "%s: short read (%zu of %zu bytes)", path.c_str(), got, bytes);
  return TDFS_OK;
}

int tdfs_graph_read(const char *prefix, int64_t n, int64_t nnz, int64_t *rowptr, int32_t *colidx,
                    int32_t *labels) {
  if (!prefix || !rowptr || (!colidx && nnz > 0)) return fail(TDFS_E_ARG, "null argument");
  std::string p(prefix);
  int rc = read_exact(p + ".vertex.bin", rowptr, sizeof(int64_t) * (size_t)(n + 1));
  if (rc) return rc;
  rc = read_exact(p + ".edge.bin", colidx, sizeof(int32_t) * (size_t)nnz);
  if (rc) return rc;
  if (labels) {
    rc = read_exact(p + ".label.bin", labels, sizeof(int32_t) * (size_t)n);
    if (rc) return rc;
  }
  if (rowptr[0] != 0 || rowptr[n] != nnz) return fail(TDFS_E_FORMAT, "rowptr does not span [0, nnz]");
  return TDFS_OK;
}

static int write_exact(const std::string &path, const void *src, size_t bytes) {
  FILE *f = fopen(path.c_str(), "wb");
  if (!f) return fail(TDFS_E_IO, "cannot create %s", path.c_str());
  size_t put = bytes ? fwrite(src, 1, bytes, f) : 0;
  fclose(f);
  if (put != bytes) return fail(TDFS_E_IO, "%s: short write", path.c_str());
  return TDFS_OK;
}

int tdfs_graph_write(const char *prefix, int64_t n, int64_t nnz, const int64_t *rowptr,
                     const int32_t *colidx, const int32_t *labels) {
  if (!prefix || !rowptr) return fail(TDFS_E_ARG, "null argument");
  std::string p(prefix);
  int64_t dmax = 0;
  for (int64_t v = 0; v < n; v++) dmax = std::max(dmax, rowptr[v + 1] - rowptr[v]);
  {
    FILE *f = fopen((p + ".meta.txt").c_str(), "w");
    if (!f) return fail(TDFS_E_IO, "cannot create %s.meta.txt", prefix);
    fprintf(f, "%lld %lld 4 %lld\n", (long long)n, (long long)nnz, (long long)dmax);
    fclose(f);
  }
  int rc = write_exact(p + ".vertex.bin", rowptr, sizeof(int64_t) * (size_t)(n + 1));
  if (rc) return rc;
  rc = write_exact(p + ".edge.bin", colidx, sizeof(int32_t) * (size_t)nnz);
  if (rc) return rc;
  if (labels) return write_exact(p + ".label.bin", labels, sizeof(int32_t) * (size_t)n);
  return TDFS_OK;
}

int tdfs_graph_validate(int64_t n, int64_t nnz, const int64_t *rowptr, const int32_t *colidx) {
  if (n < 0 || nnz < 0 || !rowptr) return fail(TDFS_E_ARG, "bad arguments");
  if (rowptr[0] != 0 || rowptr[n] != nnz) return fail(TDFS_E_FORMAT, "rowptr does not span [0, nnz]");
  for (int64_t v = 0; v < n; v++) {
    if (rowptr[v + 1] < rowptr[v]) return fail(TDFS_E_FORMAT, "rowptr decreases at %lld", (long long)v);
    for (int64_t e = rowptr[v]; e < rowptr[v + 1]; e++) {
      const int32_t w = colidx[e];
      if (w < 0 || w >= n) return fail(TDFS_E_FORMAT, "neighbour id %d out of range", w);
      if (w == v) return fail(TDFS_E_FORMAT, "self-loop at vertex %lld", (long long)v);
      if (e > rowptr[v] && colidx[e - 1] >= w)
        return fail(TDFS_E_FORMAT, "adjacency of %lld not strictly ascending", (long long)v);
      // symmetry: v must appear in N(w)
      const int32_t *b = colidx + rowptr[w], *en = colidx + rowptr[w + 1];
      if (!std::binary_search(b, en, (int32_t)v))
        return fail(TDFS_E_FORMAT, "edge %lld-%d has no reverse", (long long)v, w);
    }
  }
  return TDFS_OK;
}

int tdfs_gen_er(int32_t n, int64_t m, uint64_t seed, int64_t *rowptr, int32_t *colidx) {
  if (n < 2 || m < 0 || (double)m > (double)n * (n - 1) / 2) return fail(TDFS_E_ARG, "bad ER sizes");
  Rng r(seed);
  EdgeSet set(m);
  std::vector<uint64_t> edges;
  edges.reserve(m);
  while ((int64_t)edges.size() < m) {
    const uint32_t a = (uint32_t)r.below(n), b = (uint32_t)r.below(n);
    if (a == b) continue;
    const uint64_t key = ekey(a, b);
    if (set.insert(key)) edges.push_back(key);
  }
  edges_to_csr(n, edges, rowptr, colidx);
  return TDFS_OK;
}

int tdfs_gen_chung_lu(int32_t n, int64_t m, double gamma, double dmax, uint64_t seed, int64_t *rowptr,
                      int32_t *colidx) {
  if (n < 2 || m < 0 || gamma <= 1.0 || dmax <= 0) return fail(TDFS_E_ARG, "bad Chung-Lu parameters");
  const double alpha = 1.0 / (gamma - 1.0);
  const double total = 2.0 * (double)m;
  // choose i0 so that the largest expected degree total * w0 / sum(w) is dmax
  auto top = [&](double i0) {
    double s = 0;
    for (int32_t i = 0; i < n; i++) s += std::pow(i + i0, -alpha);
    return total * std::pow(i0, -alpha) / s;
  };
  double lo = 1e-3, hi = (double)n;
  for (int it = 0; it < 60; it++) {
    const double mid = std::sqrt(lo * hi);
    if (top(mid) > dmax) lo = mid; else hi = mid;
  }
  const double i0 = std::sqrt(lo * hi);
  std::vector<double> w(n);
  for (int32_t i = 0; i < n; i++) w[i] = std::pow(i + i0, -alpha);
  Alias alias(w);
  Rng r(seed);
  EdgeSet set(m);
  std::vector<uint64_t> edges;
  edges.reserve(m);
  int64_t attempts = 0;
  while ((int64_t)edges.size() < m) {
    if (++attempts > 50 * m + 1000) return fail(TDFS_E_ARG, "Chung-Lu: too many rejections");
    const int32_t a = alias.draw(r), b = alias.draw(r);
    if (a == b) continue;
    const uint64_t key = ekey((uint32_t)a, (uint32_t)b);
    if (set.insert(key)) edges.push_back(key);
  }
  edges_to_csr(n, edges, rowptr, colidx);
  return TDFS_OK;
}

int tdfs_gen_labels(int32_t n, int32_t nlabels, uint64_t seed, int32_t *labels) {
  if (n < 0 || nlabels < 1 || !labels) return fail(TDFS_E_ARG, "bad label arguments");
  Rng r(seed);
  for (int32_t v = 0; v < n; v++) labels[v] = (int32_t)r.below((uint64_t)nlabels);
  return TDFS_OK;
}

int tdfs_gen_pattern(int32_t k, int32_t m, int32_t nlabels, uint64_t seed, char *text_out, size_t cap) {
  if (k < 3 || k > TDFS_MAX_K || m < k - 1 || m > k * (k - 1) / 2)
    return fail(TDFS_E_ARG, "bad pattern sizes k=%d m=%d", k, m);
  Rng r(seed);
  int adj[TDFS_MAX_K][TDFS_MAX_K] = {};
  // random spanning tree: vertex i attaches to a uniform earlier vertex
  std::vector<int> perm(k);
  for (int i = 0; i < k; i++) perm[i] = i;
  for (int i = k - 1; i > 0; i--) std::swap(perm[i], perm[r.below(i + 1)]);
  int edges = 0;
  for (int i = 1; i < k; i++) {
    const int a = perm[i], b = perm[r.below(i)];
    adj[a][b] = adj[b][a] = 1;
    edges++;
  }
  std::vector<std::pair<int, int>> rest;
  for (int a = 0; a < k; a++)
    for (int b = a + 1; b < k; b++)
      if (!adj[a][b]) rest.emplace_back(a, b);
  for (size_t i = rest.size(); i > 1; i--) std::swap(rest[i - 1], rest[r.below(i)]);
  for (size_t i = 0; edges < m && i < rest.size(); i++, edges++)
    adj[rest[i].first][rest[i].second] = adj[rest[i].second][rest[i].first] = 1;
  int label[TDFS_MAX_K];
  for (int i = 0; i < k; i++) label[i] = nlabels > 0 ? (int)r.below((uint64_t)nlabels) : 1;
  char raw[1024];
  int rc = write_g(k, adj, label, raw, sizeof(raw));
  if (rc) return rc;
  return tdfs_pattern_reorder(raw, nullptr, text_out, cap);
}

}  // extern "C"

extern "C" int tdfs_estimate_work(int32_t n, int64_t nnz, const int64_t *rowptr, const int32_t *colidx,
                                  const int32_t *labels, const tdfs_plan *plan, int32_t nsample, uint64_t seed,
                                  double cap, double *out3) {
  return tdfs::estimate(n, nnz, rowptr, colidx, labels, plan, nsample, seed, cap, out3);
}
