// tdfs_cli.cpp -- the ulb / lb command-line drivers (cu_test.cu:10-110).
//
//   ulb <graph_prefix> <pattern.g>      unlabeled (built with TDFS_CLI_LABELED=0)
//   lb  <graph_prefix> <pattern.g>      labeled   (built with TDFS_CLI_LABELED=1)
//
// Diagnostics go to stdout first, as in the reference (src/graph.h:166-183,
// src/pattern.h:359-483, cu_test.cu:19-26); the last line keeps the
// reference's format "<pattern>\t<kernel ms>\t<count>\n" (cu_test.cu:107).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "tdfs.h"

#ifndef TDFS_CLI_LABELED
#define TDFS_CLI_LABELED 0
#endif

static int die(const char *what, int rc) {
  std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, tdfs_last_error());
  return 1;
}

int main(int argc, char **argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <graph_prefix> <pattern.g>\n", argv[0]);
    return 2;
  }
  const int labeled = TDFS_CLI_LABELED;
  int64_t meta[4];
  int rc = tdfs_graph_read_meta(argv[1], meta);
  if (rc) return die("read meta", rc);
  const int64_t n = meta[0], nnz = meta[1];
  std::vector<int64_t> rowptr(n + 1);
  std::vector<int32_t> colidx(nnz > 0 ? nnz : 1), labels;
  if (labeled) labels.resize(n > 0 ? n : 1);
  rc = tdfs_graph_read(argv[1], n, nnz, rowptr.data(), colidx.data(), labeled ? labels.data() : nullptr);
  if (rc) return die("read graph", rc);
  for (int t : {2, 3, 4}) {
    int64_t c = 0;
    for (int64_t v = 0; v < n; v++) c += (rowptr[v + 1] - rowptr[v]) <= t;
    std::printf("there're %lld vertices degree <= %d, percentage: %f%%\n", (long long)c, t,
                n ? 100.0 * (double)c / (double)n : 0.0);
  }
  tdfs_plan plan;
  rc = tdfs_pattern_load(argv[2], labeled, &plan);
  if (rc) return die("pattern", rc);
  std::printf("# BNs:\n");
  for (int u = 0; u < plan.k; u++) {
    for (int i = 0; i < plan.num_bn[u]; i++) std::printf("%d ", plan.bn[u][i]);
    std::printf("\n");
  }
  std::printf("# shared levels:");
  for (int u = 0; u < plan.k; u++) std::printf(" %d", plan.shared_lvl[u]);
  std::printf("\n# Conditions:\n");
  for (int u = 0; u < plan.k; u++) {
    for (int j = 0; j < u; j++) std::printf("%d %d | ", plan.cond[u][j], j);
    std::printf("\n");
  }
  tdfs_ctx *ctx = nullptr;
  rc = tdfs_ctx_create(0, &ctx);
  if (rc) return die("context", rc);
  tdfs_graph *g = nullptr;
  rc = tdfs_graph_upload(ctx, (int32_t)n, nnz, rowptr.data(), colidx.data(), labeled ? labels.data() : nullptr, &g);
  if (rc) return die("graph upload", rc);
  tdfs_result res;
  rc = tdfs_count(ctx, g, &plan, 0, -1, &res);
  if (rc) return die("count", rc);
  std::printf("%s\t%f\t%llu\n", argv[2], res.kernel_ms, (unsigned long long)res.count);
  tdfs_graph_free(g);
  tdfs_ctx_destroy(ctx);
  return 0;
}
