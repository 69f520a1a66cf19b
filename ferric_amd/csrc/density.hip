// MP2 one-particle densities P_oo, P_vv on the device
// (build_mp2_density, crates/ferric-mp2/src/oo_rimp2.rs:1192-1300, with the
// spin scaling of apply_spin_scaling / opposite_spin_density,
// crates/ferric-mp2/src/rimp2.rs:1917-2052):
//   P_ij = - sum_k sum_ab t_ik^ab T~_jk^ab
//   P_ab =   sum_ij sum_c t_ij^ac T~_ij^bc,   T~ = (c_os + c_ss) t - c_ss t^T.
// (c_os, c_ss) = (1, 1) gives the MP2 density (T~ = 2t - t^T); otherwise
// c_ss P_MP2 + (c_os - c_ss) P_OS.
//
// The fused FP64 energy kernel runs once with an epilogue that stores
// t = g / D of every ordered occupied pair (both orientations of each unique
// pair), the layout the reference's t2 has per (i, j) block. Both densities
// are then one split-K FP64 MFMA product each over those blocks, every
// operand row K-contiguous:
//   P_vv: K = (pair ij, c); A row a = t_ij[a][.], B row b = cd t_ij[b][.] - ce t_ji[b][.]
//   P_oo: K = (k, ab);      A row i = t_ik,       B row j = cd t_jk - ce t_kj
// (t_ij^ba = t_ji^ab makes the transposed reads of T~ plain reads of the
// mirrored block). Each work item is one 64 x 64 output tile over one K slice;
// a second kernel sums the slices of every tile in slice order, so the result
// does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "../../include/ferric_hip.h"
#include "fh_device.h"
#include "fh_host.h"
#include "fh_kernels.h"

namespace fh {
namespace {

constexpr int kDT = 64;   // output tile side
constexpr int kDKP = 16;  // K panel depth
constexpr int kDLdr = kDKP + 2;  // LDS row stride (doubles): 16 B of padding

// Row pointers of the three operands at panel kp (K = kp * 16 .. + 16).
struct DensRows {
    const double *a, *b1, *b2;
};

__device__ __forceinline__ DensRows dens_rows(const DensArgs &d, int64_t kp, int m, int n) {
    const int64_t vp = d.nvir_pad;
    const int64_t k0 = kp * kDKP;
    DensRows r;
    if (d.mode == DensMode::VV) {
        const int64_t q = k0 / vp, c = k0 % vp;
        const int i = (int)(q / d.nocc), j = (int)(q % d.nocc);
        const int64_t qt = (int64_t)j * d.nocc + i;
        r.a = d.tamp + (q * vp + m) * vp + c;
        r.b1 = d.tamp + (q * vp + n) * vp + c;
        r.b2 = d.tamp + (qt * vp + n) * vp + c;
    } else {
        const int64_t blk = vp * vp;
        const int64_t kk = k0 / blk, f = k0 % blk;
        r.a = d.tamp + ((int64_t)m * d.nocc + kk) * blk + f;
        r.b1 = d.tamp + ((int64_t)n * d.nocc + kk) * blk + f;
        r.b2 = d.tamp + (kk * d.nocc + n) * blk + f;
    }
    return r;
}

// One item = (output tile tm, tn; K slice s). 256 threads: four waves in a
// 2 x 2 grid, each a 32 x 32 quadrant of 16 x 16 v_mfma_f64_16x16x4 blocks.
// Per panel each thread stages two 16-byte chunks of A and two of B (each B
// chunk formed as cd B1 - ce B2 while in registers); double-buffered LDS, one
// barrier per panel.
__global__ __launch_bounds__(256) void density_gemm_kernel(DensArgs d) {
    __shared__ __attribute__((aligned(16))) double lds[2][2][kDT][kDLdr];
    const int64_t nt2 = (int64_t)d.ntile * d.ntile;
    const int64_t item = blockIdx.x;
    const int s = (int)(item / nt2);
    const int t = (int)(item % nt2);
    const int tm = t / d.ntile, tn = t % d.ntile;
    const int64_t per = (d.kpanels + d.splits - 1) / d.splits;
    const int64_t p0 = (int64_t)s * per, p1 = min(d.kpanels, p0 + per);
    const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
    const int wr = wid / 2, wc = wid % 2, r = lane % 16, g = lane / 16;

    // Staging: chunk c in [0, 512) of an operand = (row c / 8, 16-byte piece c % 8).
    v4d sa[2], sb[2];
    auto gload = [&](int64_t kp) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int c = tid + 256 * u;
            const int row = c / 8, piece = c % 8;
            const int m = tm * kDT + row, n = tn * kDT + row;
            v4d va = {0, 0, 0, 0}, vb = {0, 0, 0, 0};
            if (kp < p1) {
                const DensRows rr = dens_rows(d, kp, m < d.mdim ? m : 0, n < d.mdim ? n : 0);
                if (m < d.mdim) {
                    const double2 x = *(const double2 *)(rr.a + 2 * piece);
                    va[0] = x.x;
                    va[1] = x.y;
                }
                if (n < d.mdim) {
                    const double2 y1 = *(const double2 *)(rr.b1 + 2 * piece);
                    const double2 y2 = *(const double2 *)(rr.b2 + 2 * piece);
                    vb[0] = d.cd * y1.x - d.ce * y2.x;
                    vb[1] = d.cd * y1.y - d.ce * y2.y;
                }
            }
            sa[u] = va;
            sb[u] = vb;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int c = tid + 256 * u;
            const int row = c / 8, piece = c % 8;
            double *pa = &lds[buf][0][row][2 * piece];
            double *pb = &lds[buf][1][row][2 * piece];
            pa[0] = sa[u][0];
            pa[1] = sa[u][1];
            pb[0] = sb[u][0];
            pb[1] = sb[u][1];
        }
    };

    v4d acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = (v4d){0, 0, 0, 0};

    if (p0 < p1) {
        gload(p0);
        lstore(0);
        __syncthreads();
        for (int64_t kp = p0; kp < p1; ++kp) {
            const int cur = (int)((kp - p0) & 1);
            if (kp + 1 < p1) gload(kp + 1);
            // lane group g owns k = 4 g .. 4 g + 3 of the panel for A and B alike
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                double fa[2], fb[2];
#pragma unroll
                for (int x = 0; x < 2; ++x) fa[x] = lds[cur][0][wr * 32 + x * 16 + r][4 * g + kk];
#pragma unroll
                for (int y = 0; y < 2; ++y) fb[y] = lds[cur][1][wc * 32 + y * 16 + r][4 * g + kk];
#pragma unroll
                for (int x = 0; x < 2; ++x)
#pragma unroll
                    for (int y = 0; y < 2; ++y)
                        acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[x], fb[y], acc[x][y],
                                                                         0, 0, 0);
            }
            if (kp + 1 < p1) lstore(cur ^ 1);
            __syncthreads();
        }
    }
    // partial tile, row-major 64 x 64
    double *out = d.partial + item * (int64_t)(kDT * kDT);
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = wr * 32 + x * 16 + out_row<double>(g, v);
                const int col = wc * 32 + y * 16 + r;
                out[row * kDT + col] = acc[x][y][v];
            }
}

// out[m][n] = scale * sum over slices s = 0, 1, ... of the tile's partials.
__global__ __launch_bounds__(256) void density_reduce_kernel(DensArgs d) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t total = (int64_t)d.mdim * d.mdim;
    if (e >= total) return;
    const int m = (int)(e / d.mdim), n = (int)(e % d.mdim);
    const int tm = m / kDT, tn = n / kDT;
    const int64_t nt2 = (int64_t)d.ntile * d.ntile;
    const int64_t t = (int64_t)tm * d.ntile + tn;
    const int off = (m % kDT) * kDT + (n % kDT);
    double acc = 0.0;
    for (int s = 0; s < d.splits; ++s) acc += d.partial[((int64_t)s * nt2 + t) * kDT * kDT + off];
    d.out[e] = d.scale * acc;
}

}  // namespace

hipError_t launch_density(const DensArgs &d, hipStream_t st) {
    const int64_t nitems = (int64_t)d.ntile * d.ntile * d.splits;
    if (nitems <= 0 || nitems > 0x7fffffffLL || d.mdim <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(density_gemm_kernel, dim3((unsigned)nitems), dim3(256), 0, st, d);
    const int64_t total = (int64_t)d.mdim * d.mdim;
    hipLaunchKernelGGL(density_reduce_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       st, d);
    return hipGetLastError();
}

namespace {

// Slices per output tile: enough items to fill the chip (about 2048), each
// slice at least 8 panels deep.
int density_splits(int ntile, int64_t kpanels) {
    const int64_t tiles = (int64_t)ntile * ntile;
    int64_t s = std::max<int64_t>(1, 2048 / tiles);
    s = std::min<int64_t>(s, std::max<int64_t>(1, kpanels / 8));
    return (int)s;
}

fh_status density_core(State &s, const void *b_dev, fh_dtype b_dtype, size_t naux, size_t nocc,
                       size_t nvir, const double *eps, size_t first_occ, size_t nocc_total,
                       double c_os, double c_ss, hipStream_t stream, fh_spin_components *out,
                       double *p_oo, double *p_vv) {
    const EnergyPlan pl = plan_energy(nocc, nvir, naux, Prec::F64);
    const size_t vp = (size_t)pl.nvir_pad;
    const size_t amp_bytes = sizeof(double) * nocc * nocc * vp * vp;
    DensArgs vv, oo;
    vv.mode = DensMode::VV;
    vv.nocc = (int)nocc;
    vv.nvir_pad = (int)vp;
    vv.mdim = (int)vp;
    vv.ntile = (int)((vp + kDT - 1) / kDT);
    vv.kpanels = (int64_t)(nocc * nocc * vp / kDKP);
    vv.splits = density_splits(vv.ntile, vv.kpanels);
    vv.cd = c_os + c_ss;
    vv.ce = c_ss;
    vv.scale = 1.0;
    oo = vv;
    oo.mode = DensMode::OO;
    oo.mdim = (int)nocc;
    oo.ntile = (int)((nocc + kDT - 1) / kDT);
    oo.kpanels = (int64_t)(nocc * vp * vp / kDKP);
    oo.splits = density_splits(oo.ntile, oo.kpanels);
    oo.scale = -1.0;
    const size_t part_elems =
        std::max((size_t)vv.ntile * vv.ntile * vv.splits, (size_t)oo.ntile * oo.ntile * oo.splits) *
        kDT * kDT;
    std::vector<std::pair<int, size_t>> req = {
        {kSlotDensAmp, amp_bytes},
        {kSlotDensPart, sizeof(double) * part_elems},
        {kSlotDensOut, sizeof(double) * (vp * vp + nocc * nocc)},
    };
    fh_status st = reserve(s, req);
    if (st != FH_OK) return st;
    double *amp = (double *)s.slots[kSlotDensAmp].ptr;
    st = energy_core(s, b_dev, b_dtype, naux, nocc, nvir, eps, first_occ, nocc_total, NAN,
                     Prec::F64, stream, out, nullptr, &pl, nullptr, 0, amp);
    if (st != FH_OK) return st;
    double *dout = (double *)s.slots[kSlotDensOut].ptr;
    vv.tamp = oo.tamp = amp;
    vv.partial = oo.partial = (double *)s.slots[kSlotDensPart].ptr;
    vv.out = dout;
    oo.out = dout + vp * vp;
    hipError_t he = launch_density(vv, stream);
    if (he == hipSuccess) he = launch_density(oo, stream);
    if (he != hipSuccess) {
        (void)hipGetLastError();
        return fail(FH_KERNEL, "density kernels: %s", hipGetErrorString(he));
    }
    s.pending.kernel_launches += 4;
    std::vector<double> h(vp * vp + nocc * nocc);
    st = hip_check(hipMemcpyAsync(h.data(), dout, sizeof(double) * h.size(),
                                  hipMemcpyDeviceToHost, stream),
                   "D2H densities");
    if (st != FH_OK) return st;
    st = hip_check(hipStreamSynchronize(stream), "stream sync");
    if (st != FH_OK) return st;
    s.pending.d2h_bytes += sizeof(double) * h.size();
    for (size_t a = 0; a < nvir; ++a)
        for (size_t b = 0; b < nvir; ++b) p_vv[a * nvir + b] = h[a * vp + b];
    for (size_t k = 0; k < nocc * nocc; ++k) p_oo[k] = h[vp * vp + k];
    return FH_OK;
}

fh_status density_entry(const void *b, bool host, fh_dtype b_dtype, size_t naux, size_t nocc,
                        size_t nvir, const double *eps, size_t n_eps, size_t first_occ,
                        size_t nocc_total, double c_os, double c_ss, void *stream,
                        fh_spin_components *out, double *p_oo, double *p_vv) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    if (!out || (nocc && !p_oo) || (nvir && !p_vv)) return fail(FH_LAYOUT, "null output pointer");
    if (!std::isfinite(c_os) || !std::isfinite(c_ss))
        return fail(FH_LAYOUT, "spin scaling coefficients must be finite");
    if (b_dtype != FH_DT_F64 && b_dtype != FH_DT_F32)
        return fail(FH_LAYOUT, "unknown b_dtype %d", (int)b_dtype);
    bool empty = false;
    fh_status st = check_shapes(naux, nocc, nvir, n_eps, first_occ, nocc_total, &empty);
    if (st != FH_OK) return st;
    if (empty) {
        *out = fh_spin_components{0.0, 0.0, 0.0};
        for (size_t k = 0; k < nocc * nocc; ++k) p_oo[k] = 0.0;
        for (size_t k = 0; k < nvir * nvir; ++k) p_vv[k] = 0.0;
        commit(s, Prec::F64, true, false);
        return FH_OK;
    }
    if (!b || !eps) return fail(FH_LAYOUT, "null input pointer");
    st = ensure_device(s);
    if (st != FH_OK) return st;
    const void *b_dev = b;
    hipStream_t strm = (hipStream_t)stream;
    if (host) {
        const size_t bbytes = sizeof(double) * naux * nocc * nvir;
        st = reserve(s, {{kSlotStage, bbytes}});
        if (st != FH_OK) return st;
        st = hip_check(hipMemcpyAsync(s.slots[kSlotStage].ptr, b, bbytes, hipMemcpyHostToDevice,
                                      strm),
                       "H2D b_ov");
        if (st != FH_OK) return st;
        s.pending.h2d_bytes += bbytes;
        b_dev = s.slots[kSlotStage].ptr;
    }
    st = density_core(s, b_dev, b_dtype, naux, nocc, nvir, eps, first_occ, nocc_total, c_os, c_ss,
                      strm, out, p_oo, p_vv);
    if (st != FH_OK) return st;
    commit(s, Prec::F64, true, false);
    return FH_OK;
}

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" {

fh_status fh_rimp2_mp2_density(const double *b_ov, size_t naux, size_t nocc, size_t nvir,
                               const double *eps, size_t n_eps, size_t first_occ,
                               size_t nocc_total, double c_os, double c_ss,
                               fh_spin_components *out, double *p_oo, double *p_vv) {
    return density_entry(b_ov, true, FH_DT_F64, naux, nocc, nvir, eps, n_eps, first_occ,
                         nocc_total, c_os, c_ss, nullptr, out, p_oo, p_vv);
}

fh_status fh_rimp2_mp2_density_dev(const void *b_dev, fh_dtype b_dtype, size_t naux, size_t nocc,
                                   size_t nvir, const double *eps, size_t n_eps, size_t first_occ,
                                   size_t nocc_total, double c_os, double c_ss, void *stream,
                                   fh_spin_components *out, double *p_oo, double *p_vv) {
    return density_entry(b_dev, false, b_dtype, naux, nocc, nvir, eps, n_eps, first_occ,
                         nocc_total, c_os, c_ss, stream, out, p_oo, p_vv);
}

}  // extern "C"
