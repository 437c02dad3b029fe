// tdfs_internal.h -- shared between the host (.cpp) and device (.hip) halves
// of libtdfs.  Not part of the public ABI.
#ifndef TDFS_INTERNAL_H_
#define TDFS_INTERNAL_H_

namespace tdfs {
// Records a formatted message for tdfs_last_error() and returns code.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
const char *last_error();
}  // namespace tdfs

#endif
