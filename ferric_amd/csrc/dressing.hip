// Dressing / MO transform upstream of the pair stage (rimp2.rs:648-756) on the
// device: B = M (C_left^T (Q|mu nu) C_right), the RI factor the pair kernels
// consume, built where it is used. Three plain dense GEMMs, so they go to
// rocBLAS (column-major; every row-major operand is passed as its transpose):
//   1. X[Q, nu, l] = sum_mu (Q|nu mu) C_left[mu, l]     one GEMM, (naux*nao x nao)(nao x nleft),
//      using (Q|mu nu) = (Q|nu mu) so the occupied index is contracted first
//      (the reference's order, rimp2.rs:714-729);
//   2. Y[Q, l, r] = sum_nu X[Q, nu, l] C_right[nu, r]   strided-batched over Q;
//   3. B[P, lr] = sum_Q M[P, Q] Y[Q, lr]                 one GEMM over the band.
// The reference streams Q in chunks of 256 with beta = 1 accumulation; here Y
// is held whole (naux * nleft * nright doubles) and dressed in one GEMM.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <mutex>

#include "fh_host.h"

namespace fh {
namespace {

fh_status blas_check(rocblas_status st, const char *what) {
    if (st == rocblas_status_success) return FH_OK;
    return fail(FH_KERNEL, "%s: rocBLAS status %d", what, (int)st);
}

}  // namespace

fh_status blas_handle(State &s, hipStream_t stream, void **handle) {
    if (!s.blas) {
        rocblas_handle h = nullptr;
        fh_status st = blas_check(rocblas_create_handle(&h), "rocblas_create_handle");
        if (st != FH_OK) return st;
        st = blas_check(rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed),
                        "rocblas_set_atomics_mode");
        if (st != FH_OK) return st;
        s.blas = h;
    }
    *handle = s.blas;
    return blas_check(rocblas_set_stream((rocblas_handle)s.blas, stream), "rocblas_set_stream");
}

namespace {

fh_status dress_core(State &s, const double *j3, size_t naux, size_t nao, const double *cl,
                     size_t nleft, const double *cr, size_t nright, const double *m, size_t nband,
                     hipStream_t stream, double *out) {
    void *hv = nullptr;
    fh_status st = blas_handle(s, stream, &hv);
    if (st != FH_OK) return st;
    rocblas_handle h = (rocblas_handle)hv;
    const size_t width = nleft * nright;
    st = reserve(s, {{kSlotDressX, sizeof(double) * naux * nao * nleft},
                     {kSlotDressY, sizeof(double) * naux * width}});
    if (st != FH_OK) return st;
    double *X = (double *)s.slots[kSlotDressX].ptr;
    double *Y = (double *)s.slots[kSlotDressY].ptr;
    const double one = 1.0, zero = 0.0;
    mark(s, 0, stream);
    // 1. X^T (nleft x naux*nao) = C_left^T (nleft x nao) . J^T (nao x naux*nao)
    st = blas_check(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (int)nleft,
                                  (int)(naux * nao), (int)nao, &one, cl, (int)nleft, j3, (int)nao,
                                  &zero, X, (int)nleft),
                    "dress GEMM 1");
    if (st != FH_OK) return st;
    mark(s, 1, stream);
    // 2. per Q: Y_Q^T (nright x nleft) = C_right^T (nright x nao) . X_Q (nao x nleft)
    st = blas_check(rocblas_dgemm_strided_batched(
                        h, rocblas_operation_none, rocblas_operation_transpose, (int)nright,
                        (int)nleft, (int)nao, &one, cr, (int)nright, 0, X, (int)nleft,
                        (rocblas_stride)(nao * nleft), &zero, Y, (int)nright,
                        (rocblas_stride)width, (int)naux),
                    "dress GEMM 2");
    if (st != FH_OK) return st;
    mark(s, 2, stream);
    // 3. B^T (width x nband) = Y^T (width x naux) . M^T (naux x nband)
    st = blas_check(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (int)width,
                                  (int)nband, (int)naux, &one, Y, (int)width, m, (int)naux, &zero,
                                  out, (int)width),
                    "dress GEMM 3");
    if (st != FH_OK) return st;
    mark(s, 3, stream);
    st = hip_check(hipStreamSynchronize(stream), "stream sync");
    if (st != FH_OK) return st;
    s.pending.kernel_launches += 3;
    finish_timing(s, 0, 0);
    return FH_OK;
}

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" fh_status fh_dress_b_ov_dev(const double *j3, size_t naux, size_t nao,
                                       const double *c_left, size_t nleft, const double *c_right,
                                       size_t nright, const double *m, size_t nband,
                                       void *stream, double *b_out) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    if (naux == 0 || nao == 0 || nleft == 0 || nright == 0 || nband == 0) {
        if (nband > naux) return fail(FH_LAYOUT, "band of %zu rows > naux %zu", nband, naux);
        return FH_OK;   // nothing to compute (an empty output)
    }
    if (!j3 || !c_left || !c_right || !m || !b_out) return fail(FH_LAYOUT, "null input pointer");
    if (nband > naux) return fail(FH_LAYOUT, "band of %zu rows > naux %zu", nband, naux);
    const size_t lim = (size_t)0x7fffffff;
    if (naux * nao > lim || nleft * nright > lim || nao > lim)
        return fail(FH_UNSUPPORTED, "dimensions beyond 32-bit GEMM sizes");
    fh_status st = ensure_device(s);
    if (st != FH_OK) return st;
    st = dress_core(s, j3, naux, nao, c_left, nleft, c_right, nright, m, nband,
                    (hipStream_t)stream, b_out);
    if (st != FH_OK) return st;
    commit(s, Prec::F64, false, false);
    return FH_OK;
}
