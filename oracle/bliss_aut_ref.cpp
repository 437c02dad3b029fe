// bliss_aut_ref.cpp -- TEST INFRASTRUCTURE ONLY (part of the oracle).
//
// Links bliss 0.73 compiled from the reference's vendored sources
// (bliss/graph.hh:215-219 find_automorphisms, :652 Graph(N), :725 add_edge)
// and returns Aut(P) the way the reference obtains it: the generators bliss
// reports (src/pattern.h:637-649), closed under composition until the set stops
// growing, at most 100 rounds (src/pattern.h:603-635, 651-660).  A pattern
// with a trivial group yields no generators and therefore an empty set, as in
// the reference.
#include <set>
#include <vector>

#include "graph.hh"

namespace {
void hook(void *param, const unsigned int n, const unsigned int *aut) {
  auto *out = static_cast<std::vector<std::vector<int>> *>(param);
  out->emplace_back(aut, aut + n);
}
}  // namespace

extern "C" int bliss_ref_aut(int k, const int *adj, int *out_perms, int cap) {
  bliss::Graph g(k);
  for (int i = 0; i < k; i++)
    for (int j = 0; j < k; j++)
      if (adj[i * k + j] > 0) g.add_edge(i, j);
  std::vector<std::vector<int>> gens;
  bliss::Stats stats;
  g.find_automorphisms(stats, hook, &gens);
  std::set<std::vector<int>> seen(gens.begin(), gens.end());
  std::vector<std::vector<int>> group(seen.begin(), seen.end());
  size_t last = 0;
  for (int round = 0; round <= 100 && group.size() != last; round++) {
    last = group.size();
    std::vector<std::vector<int>> add;
    for (const auto &p1 : group)
      for (const auto &p2 : group) {
        std::vector<int> prod(k);
        for (int i = 0; i < k; i++) prod[i] = p1[p2[i]];
        if (seen.insert(prod).second) add.push_back(prod);
      }
    group.insert(group.end(), add.begin(), add.end());
  }
  int n = 0;
  for (const auto &p : group) {
    if (n >= cap) return -1;
    for (int i = 0; i < k; i++) out_perms[n * k + i] = p[i];
    n++;
  }
  return n;
}
