// Per-pair MP2 energies on the device: the estimates behind the pair-energy
// screen of local (DLPNO) MP2, estimate_pair_energies_from_b_ov
// (crates/ferric-mp2/src/dlpno_mp2.rs:527-594),
//   e_ij = fac sum_ab (2 (ia|jb) - (ib|ja)) (ia|jb) / D,   fac = 1 (i == j) or 2,
// split here into its opposite-spin part fac sum g^2 / D and same-spin part
// fac sum g (g - g^T) / D (the reference's e_ij is their sum). One pass of the
// fused pair-energy kernel (any precision) leaves one compensated partial per
// work item, already weighted; the items of pair p are p*ntp .. p*ntp + ntp-1,
// and one workgroup per pair folds them in item order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "fh_device.h"
#include "fh_host.h"
#include "fh_kernels.h"

namespace fh {
namespace {

__global__ __launch_bounds__(256) void reduce_pairs_kernel(const double *__restrict__ ip, int ntp,
                                                           double *__restrict__ out) {
    const int64_t p = blockIdx.x;
    KSum os, ss;
    for (int t = threadIdx.x; t < ntp; t += 256) {
        const double *q = ip + 4 * (p * ntp + t);
        os.merge(q[0], q[1]);
        ss.merge(q[2], q[3]);
    }
    wave_merge(os);
    wave_merge(ss);
    __shared__ double red[4][4];
    const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (lane == 0) {
        red[wid][0] = os.s;
        red[wid][1] = os.c;
        red[wid][2] = ss.s;
        red[wid][3] = ss.c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        KSum o, s;
        for (int w = 0; w < 4; ++w) {
            o.merge(red[w][0], red[w][1]);
            s.merge(red[w][2], red[w][3]);
        }
        out[2 * p] = o.value();
        out[2 * p + 1] = s.value();
    }
}

fh_status pair_energies_core(State &s, const void *b_dev, fh_dtype b_dtype, size_t naux,
                             size_t nocc, size_t nvir, const double *eps, size_t first_occ,
                             size_t nocc_total, Prec prec, hipStream_t stream, double *e_os,
                             double *e_ss) {
    // The reference refuses a non-negative denominator (a zero or inverted
    // gap is not a valid MP2 starting point): the largest D is
    // 2 max eps_occ - 2 min eps_vir.
    double emax = -INFINITY, vmin = INFINITY;
    for (size_t i = 0; i < nocc; ++i) emax = std::max(emax, eps[first_occ + i]);
    for (size_t a = 0; a < nvir; ++a) vmin = std::min(vmin, eps[nocc_total + a]);
    if (!(2.0 * emax - 2.0 * vmin < 0.0))
        return fail(FH_LAYOUT,
                    "non-negative MP2 denominator %g (max eps_occ %g, min eps_vir %g): the "
                    "reference has a zero or inverted gap",
                    2.0 * emax - 2.0 * vmin, emax, vmin);
    const EnergyPlan pl = plan_energy(nocc, nvir, naux, prec);
    const size_t npairs = nocc * (nocc + 1) / 2;
    fh_status st = reserve(s, {{kSlotPairs, sizeof(double) * 2 * npairs}});
    if (st != FH_OK) return st;
    fh_spin_components tot;
    st = energy_core(s, b_dev, b_dtype, naux, nocc, nvir, eps, first_occ, nocc_total, NAN, prec,
                     stream, &tot, nullptr, &pl);
    if (st != FH_OK) return st;
    double *dev = (double *)s.slots[kSlotPairs].ptr;
    hipLaunchKernelGGL(reduce_pairs_kernel, dim3((unsigned)npairs), dim3(256), 0, stream,
                       (const double *)s.slots[kSlotItems].ptr, pl.ntp, dev);
    st = hip_check(hipGetLastError(), "pair reduction");
    if (st != FH_OK) return st;
    std::vector<double> h(2 * npairs);
    st = hip_check(hipMemcpyAsync(h.data(), dev, sizeof(double) * h.size(),
                                  hipMemcpyDeviceToHost, stream),
                   "D2H pair energies");
    if (st != FH_OK) return st;
    st = hip_check(hipStreamSynchronize(stream), "stream sync");
    if (st != FH_OK) return st;
    s.pending.d2h_bytes += sizeof(double) * h.size();
    s.pending.kernel_launches += 1;
    size_t p = 0;
    for (size_t i = 0; i < nocc; ++i)
        for (size_t j = i; j < nocc; ++j, ++p) {
            e_os[i * nocc + j] = e_os[j * nocc + i] = h[2 * p];
            e_ss[i * nocc + j] = e_ss[j * nocc + i] = h[2 * p + 1];
        }
    return FH_OK;
}

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" fh_status fh_rimp2_pair_energies_dev(const void *b_dev, fh_dtype b_dtype, size_t naux,
                                                size_t nocc, size_t nvir, const double *eps,
                                                size_t n_eps, size_t first_occ,
                                                size_t nocc_total, fh_precision precision,
                                                void *stream, double *e_os, double *e_ss) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    if (!e_os || !e_ss) return fail(FH_LAYOUT, "null output pointer");
    Prec prec;
    if (!to_prec(precision, &prec)) return fail(FH_LAYOUT, "unknown precision %d", (int)precision);
    if (b_dtype != FH_DT_F64 && b_dtype != FH_DT_F32)
        return fail(FH_LAYOUT, "unknown b_dtype %d", (int)b_dtype);
    bool empty = false;
    fh_status st = check_shapes(naux, nocc, nvir, n_eps, first_occ, nocc_total, &empty);
    if (st != FH_OK) return st;
    if (empty) {
        for (size_t k = 0; k < nocc * nocc; ++k) e_os[k] = e_ss[k] = 0.0;
        commit(s, prec, false, false);
        return FH_OK;
    }
    if (!b_dev || !eps) return fail(FH_LAYOUT, "null input pointer");
    st = ensure_device(s);
    if (st != FH_OK) return st;
    st = pair_energies_core(s, b_dev, b_dtype, naux, nocc, nvir, eps, first_occ, nocc_total, prec,
                            (hipStream_t)stream, e_os, e_ss);
    if (st != FH_OK) return st;
    commit(s, prec, false, false);
    return FH_OK;
}

extern "C" fh_status fh_rimp2_pair_energies(const double *b_ov, size_t naux, size_t nocc,
                                            size_t nvir, const double *eps, size_t n_eps,
                                            size_t first_occ, size_t nocc_total,
                                            fh_precision precision, double *e_os, double *e_ss) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    if (!e_os || !e_ss) return fail(FH_LAYOUT, "null output pointer");
    Prec prec;
    if (!to_prec(precision, &prec)) return fail(FH_LAYOUT, "unknown precision %d", (int)precision);
    bool empty = false;
    fh_status st = check_shapes(naux, nocc, nvir, n_eps, first_occ, nocc_total, &empty);
    if (st != FH_OK) return st;
    if (empty) {
        for (size_t k = 0; k < nocc * nocc; ++k) e_os[k] = e_ss[k] = 0.0;
        commit(s, prec, false, false);
        return FH_OK;
    }
    if (!b_ov || !eps) return fail(FH_LAYOUT, "null input pointer");
    st = ensure_device(s);
    if (st != FH_OK) return st;
    const size_t bbytes = sizeof(double) * naux * nocc * nvir;
    st = reserve(s, {{kSlotStage, bbytes}});
    if (st != FH_OK) return st;
    st = hip_check(hipMemcpyAsync(s.slots[kSlotStage].ptr, b_ov, bbytes, hipMemcpyHostToDevice,
                                  nullptr),
                   "H2D b_ov");
    if (st != FH_OK) return st;
    s.pending.h2d_bytes += bbytes;
    st = pair_energies_core(s, s.slots[kSlotStage].ptr, FH_DT_F64, naux, nocc, nvir, eps,
                            first_occ, nocc_total, prec, nullptr, e_os, e_ss);
    if (st != FH_OK) return st;
    commit(s, prec, false, false);
    return FH_OK;
}
