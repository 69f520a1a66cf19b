// Batched local-MP2 pair domains (filled in below).
#include <hip/hip_runtime.h>

#include "fh_host.h"

namespace fh {
struct GraphCache {};
void release_graphs(State &s) {
    delete s.graphs;
    s.graphs = nullptr;
}
}  // namespace fh

using namespace fh;

extern "C" fh_status fh_rimp2_batched_dev(const fh_domain *, size_t, const double *, size_t,
                                          const double *, size_t, fh_precision, int, void *,
                                          double *, double *) {
    return fail(FH_UNSUPPORTED, "batched domains not built yet");
}
