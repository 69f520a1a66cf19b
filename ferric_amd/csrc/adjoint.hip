// Adjoint of the RI-MP2 pair energy (filled in below).
#include <hip/hip_runtime.h>

#include "fh_host.h"

using namespace fh;

extern "C" fh_status fh_rimp2_energy_adjoint(const double *, size_t, size_t, size_t,
                                             const double *, size_t, size_t, size_t, double,
                                             double, fh_precision, fh_spin_components *, double *,
                                             double *) {
    return fail(FH_UNSUPPORTED, "adjoint not built yet");
}

extern "C" fh_status fh_rimp2_energy_adjoint_dev(const void *, fh_dtype, size_t, size_t, size_t,
                                                 const double *, size_t, size_t, size_t, double,
                                                 double, fh_precision, void *,
                                                 fh_spin_components *, double *, double *) {
    return fail(FH_UNSUPPORTED, "adjoint not built yet");
}
