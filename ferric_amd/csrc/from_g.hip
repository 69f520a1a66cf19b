// MP2 energies from explicit (ia|jb) blocks on the device, for integrals that
// are not a Gram product B^T B (the robust-fit three-term G, domain-local fits):
//
//   spin_components_from_g (crates/ferric-mp2/src/rimp2.rs:984-1018)
//     g is the dense (nocc nvir) x (nocc nvir) matrix; pair (i, j) reads
//     (ia|jb) = g[i nvir + a][j nvir + b] and (ib|ja) = g[i nvir + b][j nvir + a];
//   per-pair blocks (the layout of fh_domain_fit_pair_blocks)
//     block k of pair (i_k, j_k) is nvir x nvir with (ia|jb) at [a][b] and
//     (ib|ja) at [b][a].
//
// Both are one pass over the blocks: HBM-bound, every element read once. A work
// item is one (pair, 64-wide row tile A); it walks the tile pairs (A, B >= A),
// reading the tiles (A, B) and (B, A) of the pair's block with coalesced row
// reads (the next pair of tiles loading into registers meanwhile): g_ab stays
// in the registers of the thread that consumes it, its partner
// g_ba = (B, A)[b][a] comes from tile (B, A) staged in LDS.
// Terms as in the fused kernel: for A < B each (a, b) of tile (A, B) adds
// (g_ab^2 + g_ba^2) / D to e_os and (g_ab - g_ba)^2 / D to e_ss (covering both
// orders of the pair of virtuals), for A = B it adds half of that, all times
// f_ij = 1 (i == j) or 2. One compensated partial per item, folded per pair in
// item order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "fh_device.h"
#include "fh_host.h"
#include "fh_kernels.h"
#include "pair_body.h"

namespace fh {
namespace {

constexpr int kT = 64;         // virtual tile
constexpr int kLd = kT + 1;    // LDS row stride (doubles): the transposed reads avoid bank conflicts
constexpr int kThreads = 512;

struct FromGArgs {
    const double *g;        // dense g, or the per-pair blocks
    int64_t ld;             // row stride of a block: nocc nvir (dense) or nvir (blocks)
    int dense;              // 1: block of pair (i, j) at g + i nvir ld + j nvir; 0: at g + k nvir^2
    const int32_t *pairs;   // (i, j) per pair
    const double *eps_occ;  // active occupied energies
    const double *eps_vir;
    int nvir, nt;           // nt row tiles of 64 virtuals: the items of a pair
    int64_t nitems;
    double *item_partials;  // 4 per item: os (s, c), ss (s, c)
};

// Loads this thread's share of tiles (A, B) and (B, A), and of the virtual
// energies of tile B, into registers; rows and columns beyond nvir read as zero.
template <int kPer>
__device__ __forceinline__ void load_tiles(const FromGArgs &a, const double *blk, int A0, int B0,
                                           double (&va)[kPer], double (&vb)[kPer], double &veb) {
    veb = (threadIdx.x < kT && B0 + (int)threadIdx.x < a.nvir) ? a.eps_vir[B0 + threadIdx.x] : 0.0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const int e = threadIdx.x + k * kThreads;
        const int r = e / kT, c = e % kT;
        const int ra = A0 + r, cb = B0 + c, rb = B0 + r, ca = A0 + c;
        va[k] = (ra < a.nvir && cb < a.nvir) ? blk[(int64_t)ra * a.ld + cb] : 0.0;
        vb[k] = (rb < a.nvir && ca < a.nvir) ? blk[(int64_t)rb * a.ld + ca] : 0.0;
    }
}

// 1/d for a normal, nonzero d: the hardware reciprocal estimate refined by two
// Newton-Raphson steps r <- r + r (1 - d r), each of which squares the relative
// error, so the result is within an ulp or two of the rounded quotient.
__device__ __forceinline__ double recip(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(r, fma(-d, r, 1.0), r);
}

// One item = (pair, row tile A): the tile pairs (A, B) for B = A .. nt - 1 in
// turn, the next pair of tiles loading into registers while the current one is
// consumed from LDS. The loop reads nothing else from global memory, so the
// wait for the next tiles comes only at their LDS write.
// kExact: 1/D by division, for denominator ranges that may hold zero or values
// outside [2^-1000, 2^1000], where the refined reciprocal is not accurate.
template <bool kExact>
__global__ __launch_bounds__(kThreads) void energy_from_g_kernel(FromGArgs a) {
    // Only (B, A) goes through LDS (its transpose is read); each thread keeps
    // its own elements of (A, B) in registers, in the order it consumes them.
    __shared__ double tba[kT * kLd];
    __shared__ double eps_a[kT], eps_b[kT];
    constexpr int kPer = kT * kT / kThreads;
    const int64_t item = blockIdx.x;
    const int64_t p = item / a.nt;
    const int ta = (int)(item % a.nt);
    const int i = a.pairs[2 * p], j = a.pairs[2 * p + 1];
    const double *blk = a.dense ? a.g + (int64_t)i * a.nvir * a.ld + (int64_t)j * a.nvir
                                : a.g + p * (int64_t)a.nvir * a.nvir;
    const int A0 = ta * kT;
    const double eij = a.eps_occ[i] + a.eps_occ[j];
    double va[kPer], vb[kPer], veb, cur[kPer];
    load_tiles<kPer>(a, blk, A0, A0, va, vb, veb);
    if (threadIdx.x < kT) eps_a[threadIdx.x] = veb;  // tile B = A first
    KSum os, ss, osd, ssd;  // osd, ssd: the diagonal tile pair (A, A), weight 1/2
    for (int tb = ta; tb < a.nt; ++tb) {
        const int B0 = tb * kT;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int e = threadIdx.x + k * kThreads;
            tba[(e / kT) * kLd + e % kT] = vb[k];
            cur[k] = va[k];
        }
        if (threadIdx.x < kT) eps_b[threadIdx.x] = veb;
        __syncthreads();
        if (tb + 1 < a.nt) load_tiles<kPer>(a, blk, A0, B0 + kT, va, vb, veb);
        // The kPer terms of a thread are all of one sign (D < 0), so their plain
        // sum is within kPer ulps of exact; it enters the compensated sums once
        // per tile pair.
        double to = 0.0, tq = 0.0;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            const int e = threadIdx.x + k * kThreads;
            const int r = e / kT, c = e % kT;
            if (A0 + r >= a.nvir || B0 + c >= a.nvir) continue;
            const double g1 = cur[k];            // (ia|jb)
            const double g2 = tba[c * kLd + r];  // (ib|ja)
            const double d = eij - eps_a[r] - eps_b[c];
            const double inv = kExact ? 1.0 / d : recip(d);
            const double dd = g1 - g2;
            to = fma(fma(g1, g1, g2 * g2), inv, to);
            tq = fma(dd * dd, inv, tq);
        }
        (tb == ta ? osd : os).add(to);
        (tb == ta ? ssd : ss).add(tq);
        __syncthreads();
    }
    os.merge(0.5 * osd.s, 0.5 * osd.c);
    ss.merge(0.5 * ssd.s, 0.5 * ssd.c);
    pair::block_store(os, ss, i == j ? 1.0 : 2.0, a.item_partials + 4 * item);
}

// Per-pair fold of the item partials, in item order: out[2 p], out[2 p + 1].
__global__ __launch_bounds__(256) void fold_pair_items_kernel(const double *__restrict__ ip,
                                                              int per, double *__restrict__ out) {
    const int64_t p = blockIdx.x;
    KSum os, ss;
    for (int t = threadIdx.x; t < per; t += 256) {
        const double *q = ip + 4 * (p * per + t);
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

// Runs the pass and copies the per-pair (e_os, e_ss) to host (2 npairs).
fh_status from_g_core(State &s, const double *g, bool dense, const int32_t *pairs, size_t npairs,
                      size_t nocc, size_t nvir, const double *eps, size_t first_occ,
                      size_t nocc_total, hipStream_t stream, std::vector<double> &pair_out) {
    const int nt = (int)((nvir + kT - 1) / kT);
    const int64_t nitems = (int64_t)npairs * nt;
    if (nitems > (int64_t)INT32_MAX) return fail(FH_LAYOUT, "too many work items (%lld)",
                                                 (long long)nitems);
    fh_status st = reserve(s, {{kSlotItems, sizeof(double) * 4 * (size_t)nitems},
                               {kSlotPairs, sizeof(double) * 2 * npairs},
                               {kSlotEps, sizeof(double) * (nocc + nvir)},
                               {kSlotGPairs, sizeof(int32_t) * 2 * npairs}});
    if (st != FH_OK) return st;
    std::vector<double> he(nocc + nvir);
    for (size_t k = 0; k < nocc; ++k) he[k] = eps[first_occ + k];
    for (size_t v = 0; v < nvir; ++v) he[nocc + v] = eps[nocc_total + v];
    double *deps = (double *)s.slots[kSlotEps].ptr;
    int32_t *dpairs = (int32_t *)s.slots[kSlotGPairs].ptr;
    st = hip_check(hipMemcpyAsync(deps, he.data(), sizeof(double) * he.size(),
                                  hipMemcpyHostToDevice, stream),
                   "H2D eps");
    if (st != FH_OK) return st;
    st = hip_check(hipMemcpyAsync(dpairs, pairs, sizeof(int32_t) * 2 * npairs,
                                  hipMemcpyHostToDevice, stream),
                   "H2D pairs");
    if (st != FH_OK) return st;
    s.pending.h2d_bytes += sizeof(double) * he.size() + sizeof(int32_t) * 2 * npairs;
    FromGArgs a;
    a.g = g;
    a.ld = dense ? (int64_t)(nocc * nvir) : (int64_t)nvir;
    a.dense = dense ? 1 : 0;
    a.pairs = dpairs;
    a.eps_occ = deps;
    a.eps_vir = deps + nocc;
    a.nvir = (int)nvir;
    a.nt = nt;
    a.nitems = nitems;
    a.item_partials = (double *)s.slots[kSlotItems].ptr;
    mark(s, 0, stream);
    mark(s, 1, stream);
    // Every denominator lies in [2 min e_occ - 2 max e_vir, 2 max e_occ - 2 min e_vir]
    // (over the active orbitals); the reciprocal kernel needs that range negative and
    // of normal magnitude with margin.
    const auto [omin, omax] = std::minmax_element(he.begin(), he.begin() + nocc);
    const auto [vmin, vmax] = std::minmax_element(he.begin() + nocc, he.end());
    const double dlo = 2.0 * *omin - 2.0 * *vmax, dhi = 2.0 * *omax - 2.0 * *vmin;
    const bool exact = !(dhi < 0.0 && -dhi >= 0x1p-1000 && -dlo <= 0x1p1000);
    if (exact)
        hipLaunchKernelGGL(energy_from_g_kernel<true>, dim3((unsigned)nitems), dim3(kThreads), 0,
                           stream, a);
    else
        hipLaunchKernelGGL(energy_from_g_kernel<false>, dim3((unsigned)nitems), dim3(kThreads), 0,
                           stream, a);
    mark(s, 2, stream);
    double *dout = (double *)s.slots[kSlotPairs].ptr;
    hipLaunchKernelGGL(fold_pair_items_kernel, dim3((unsigned)npairs), dim3(256), 0, stream,
                       (const double *)a.item_partials, nt, dout);
    st = hip_check(hipGetLastError(), "energy from g kernels");
    if (st != FH_OK) return st;
    s.pending.kernel_launches += 2;
    pair_out.resize(2 * npairs);
    st = hip_check(hipMemcpyAsync(pair_out.data(), dout, sizeof(double) * 2 * npairs,
                                  hipMemcpyDeviceToHost, stream),
                   "D2H pair energies");
    if (st != FH_OK) return st;
    st = hip_check(hipStreamSynchronize(stream), "stream sync");
    if (st != FH_OK) return st;
    mark(s, 3, stream);
    s.pending.d2h_bytes += sizeof(double) * 2 * npairs;
    finish_timing(s, (uint64_t)nitems, kT);
    return FH_OK;
}

fh_status check_pairs(const int32_t *pairs, size_t npairs, size_t nocc) {
    if (npairs > 0 && !pairs) return fail(FH_LAYOUT, "null pair list");
    for (size_t k = 0; k < npairs; ++k) {
        const int32_t i = pairs[2 * k], j = pairs[2 * k + 1];
        if (i < 0 || i > j || (size_t)j >= nocc)
            return fail(FH_LAYOUT, "pair %zu is (%d, %d): need 0 <= i <= j < nocc = %zu", k, i,
                        j, nocc);
    }
    return FH_OK;
}

fh_status dense_energy(State &s, const double *g_dev, size_t nocc, size_t nvir, const double *eps,
                       size_t first_occ, size_t nocc_total, hipStream_t stream,
                       fh_spin_components *out) {
    std::vector<int32_t> pairs;
    pairs.reserve(nocc * (nocc + 1));
    for (size_t i = 0; i < nocc; ++i)
        for (size_t j = i; j < nocc; ++j) {
            pairs.push_back((int32_t)i);
            pairs.push_back((int32_t)j);
        }
    std::vector<double> po;
    fh_status st = from_g_core(s, g_dev, true, pairs.data(), pairs.size() / 2, nocc, nvir, eps,
                               first_occ, nocc_total, stream, po);
    if (st != FH_OK) return st;
    // Ascending pair order, compensated (the reference folds i ascending, j >= i).
    double os = 0.0, osc = 0.0, ss = 0.0, ssc = 0.0;
    for (size_t p = 0; p < pairs.size() / 2; ++p) {
        double t = os + po[2 * p];
        osc += std::fabs(os) >= std::fabs(po[2 * p]) ? (os - t) + po[2 * p] : (po[2 * p] - t) + os;
        os = t;
        t = ss + po[2 * p + 1];
        ssc += std::fabs(ss) >= std::fabs(po[2 * p + 1]) ? (ss - t) + po[2 * p + 1]
                                                         : (po[2 * p + 1] - t) + ss;
        ss = t;
    }
    out->e_os = os + osc;
    out->e_ss = ss + ssc;
    out->e_total = out->e_os + out->e_ss;
    return FH_OK;
}

void begin(State &s) {
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
}

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" fh_status fh_rimp2_energy_from_g_dev(const double *g_dev, size_t nocc, size_t nvir,
                                                const double *eps, size_t n_eps,
                                                size_t first_occ, size_t nocc_total,
                                                void *stream, fh_spin_components *out) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    begin(s);
    if (!out) return fail(FH_LAYOUT, "null output pointer");
    bool empty = false;
    fh_status st = check_shapes(1, nocc, nvir, n_eps, first_occ, nocc_total, &empty);
    if (st != FH_OK) return st;
    *out = fh_spin_components{0.0, 0.0, 0.0};
    if (empty) {
        commit(s, Prec::F64, false, false);
        return FH_OK;
    }
    if (!g_dev || !eps) return fail(FH_LAYOUT, "null input pointer");
    st = ensure_device(s);
    if (st != FH_OK) return st;
    st = dense_energy(s, g_dev, nocc, nvir, eps, first_occ, nocc_total, (hipStream_t)stream, out);
    if (st != FH_OK) return st;
    commit(s, Prec::F64, false, false);
    return FH_OK;
}

extern "C" fh_status fh_rimp2_energy_from_g(const double *g, size_t nocc, size_t nvir,
                                            const double *eps, size_t n_eps, size_t first_occ,
                                            size_t nocc_total, fh_spin_components *out) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    begin(s);
    if (!out) return fail(FH_LAYOUT, "null output pointer");
    bool empty = false;
    fh_status st = check_shapes(1, nocc, nvir, n_eps, first_occ, nocc_total, &empty);
    if (st != FH_OK) return st;
    *out = fh_spin_components{0.0, 0.0, 0.0};
    if (empty) {
        commit(s, Prec::F64, false, false);
        return FH_OK;
    }
    if (!g || !eps) return fail(FH_LAYOUT, "null input pointer");
    st = ensure_device(s);
    if (st != FH_OK) return st;
    const size_t nov = nocc * nvir;
    const size_t bytes = sizeof(double) * nov * nov;
    st = reserve(s, {{kSlotStage, bytes}});
    if (st != FH_OK) return st;
    st = hip_check(hipMemcpyAsync(s.slots[kSlotStage].ptr, g, bytes, hipMemcpyHostToDevice,
                                  nullptr),
                   "H2D g");
    if (st != FH_OK) return st;
    s.pending.h2d_bytes += bytes;
    st = dense_energy(s, (const double *)s.slots[kSlotStage].ptr, nocc, nvir, eps, first_occ,
                      nocc_total, nullptr, out);
    if (st != FH_OK) return st;
    commit(s, Prec::F64, false, false);
    return FH_OK;
}

extern "C" fh_status fh_pair_energies_from_blocks_dev(const double *blocks_dev,
                                                      const int32_t *pairs, size_t npairs,
                                                      size_t nocc, size_t nvir, const double *eps,
                                                      size_t n_eps, size_t first_occ,
                                                      size_t nocc_total, void *stream,
                                                      double *e_os, double *e_ss) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    begin(s);
    if (!e_os || !e_ss) return fail(FH_LAYOUT, "null output pointer");
    bool empty = false;
    fh_status st = check_shapes(1, nocc, nvir, n_eps, first_occ, nocc_total, &empty);
    if (st != FH_OK) return st;
    st = check_pairs(pairs, npairs, nocc);
    if (st != FH_OK) return st;
    if (empty || npairs == 0) {
        for (size_t k = 0; k < npairs; ++k) e_os[k] = e_ss[k] = 0.0;
        commit(s, Prec::F64, false, false);
        return FH_OK;
    }
    if (!blocks_dev || !eps) return fail(FH_LAYOUT, "null input pointer");
    st = ensure_device(s);
    if (st != FH_OK) return st;
    std::vector<double> po;
    st = from_g_core(s, blocks_dev, false, pairs, npairs, nocc, nvir, eps, first_occ, nocc_total,
                     (hipStream_t)stream, po);
    if (st != FH_OK) return st;
    for (size_t k = 0; k < npairs; ++k) {
        e_os[k] = po[2 * k];
        e_ss[k] = po[2 * k + 1];
    }
    commit(s, Prec::F64, false, false);
    return FH_OK;
}

extern "C" fh_status fh_pair_energies_from_g_dev(const double *g_dev, size_t nocc, size_t nvir,
                                                 const double *eps, size_t n_eps,
                                                 size_t first_occ, size_t nocc_total,
                                                 void *stream, double *e_os, double *e_ss) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    begin(s);
    if (!e_os || !e_ss) return fail(FH_LAYOUT, "null output pointer");
    bool empty = false;
    fh_status st = check_shapes(1, nocc, nvir, n_eps, first_occ, nocc_total, &empty);
    if (st != FH_OK) return st;
    if (empty) {
        for (size_t k = 0; k < nocc * nocc; ++k) e_os[k] = e_ss[k] = 0.0;
        commit(s, Prec::F64, false, false);
        return FH_OK;
    }
    if (!g_dev || !eps) return fail(FH_LAYOUT, "null input pointer");
    // estimate_pair_energies refuses a non-negative denominator
    // (pair_energy_screen.rs:160-167): the largest is 2 max e_occ - 2 min e_vir.
    double emax = -INFINITY, vmin = INFINITY;
    for (size_t i = 0; i < nocc; ++i) emax = std::max(emax, eps[first_occ + i]);
    for (size_t a = 0; a < nvir; ++a) vmin = std::min(vmin, eps[nocc_total + a]);
    if (!(2.0 * emax - 2.0 * vmin < 0.0))
        return fail(FH_LAYOUT,
                    "non-negative MP2 denominator %g (max eps_occ %g, min eps_vir %g): the "
                    "reference has a zero or inverted gap",
                    2.0 * emax - 2.0 * vmin, emax, vmin);
    st = ensure_device(s);
    if (st != FH_OK) return st;
    std::vector<int32_t> pairs;
    pairs.reserve(nocc * (nocc + 1));
    for (size_t i = 0; i < nocc; ++i)
        for (size_t j = i; j < nocc; ++j) {
            pairs.push_back((int32_t)i);
            pairs.push_back((int32_t)j);
        }
    std::vector<double> po;
    st = from_g_core(s, g_dev, true, pairs.data(), pairs.size() / 2, nocc, nvir, eps, first_occ,
                     nocc_total, (hipStream_t)stream, po);
    if (st != FH_OK) return st;
    size_t p = 0;
    for (size_t i = 0; i < nocc; ++i)
        for (size_t j = i; j < nocc; ++j, ++p) {
            e_os[i * nocc + j] = e_os[j * nocc + i] = po[2 * p];
            e_ss[i * nocc + j] = e_ss[j * nocc + i] = po[2 * p + 1];
        }
    commit(s, Prec::F64, false, false);
    return FH_OK;
}
