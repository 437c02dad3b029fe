// tdfs_internal.h -- shared between the host (.cpp) and device (.hip) halves
// of libtdfs.  Not part of the public ABI.
#ifndef TDFS_INTERNAL_H_
#define TDFS_INTERNAL_H_

#include <cstdint>

struct tdfs_plan;

namespace tdfs {
// Records a formatted message for tdfs_last_error() and returns code.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
const char *last_error();

// Sampled work estimate of one plan on a host CSR (see tdfs_estimate_work).
// out3 = {estimated cost in list elements touched, estimated raw count,
// 1 if some sampled subtree exceeded cap}.
// deg / maxnbr (per vertex: degree, max neighbour degree) may be null; they
// are then derived on the fly.
int estimate(int32_t n, int64_t nnz, const int64_t *rowptr, const int32_t *colidx, const int32_t *labels,
             const struct tdfs_plan *plan, int32_t nsample, uint64_t seed, double cap, double *out3,
             const int32_t *deg = nullptr, const int32_t *maxnbr = nullptr);

// The plan of the same pattern with its vertices taken in the order
// order[0..k-1] (vertex i of *out is vertex order[i] of p), built as
// tdfs_pattern_parse builds one.
int plan_permuted(const struct tdfs_plan &p, const int *order, struct tdfs_plan *out);
}  // namespace tdfs

#endif
