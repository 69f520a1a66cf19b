// Fused RI-MP2 pair kernels for gfx950 (CDNA4 MFMA): the energy, and phase A
// of the adjoint.
//
// Replaces the per-i tail GEMM G_i = B_i^T B_tail plus the host pair loop of
// the reference (crates/ferric-mp2/src/rimp2.rs:1399-1544; device twin
// crates/ferric-mp2/src/rimp2_gpu.rs:198-271) with one launch in which
// K_ij = B_i^T B_j never leaves the chip.
//
// Work item = (unique occupied pair i <= j, tile pair tA <= tB over the
// virtuals). A workgroup (4 waves, 2x2) computes, over the whole Q range,
//   C1(a, b) = sum_Q B_i[Q, a] B_j[Q, b] = (ia|jb)       a in A, b in B
//   C2(a, b) = sum_Q B_j[Q, a] B_i[Q, b] = (ib|ja)
// so g_ab and g_ba of every (a, b) sit in the same lane and register. The
// pair loop's terms for (a, b) and for the mirrored (b, a) are
//   OS: (g_ab^2 + g_ba^2)/D        SS: (g_ab - g_ba)^2/D
// (the reference's g_ab^2/D and g_ab(g_ab - g_ba)/D summed over both
// orientations; D is symmetric in a, b). A diagonal tile (tA == tB) covers
// each unordered {a, b} twice and takes weight 1/2; the occupied pair takes
// the reference's fac = 1 (i == j) or 2 (i < j). Every term is FP64 with one
// divide; lane sums are compensated and every reduction runs in a fixed
// order, so the result does not depend on scheduling.
//
// Adjoint epilogue (same contraction): with t = g/D and
// T~ = (c_os + c_ss) t - c_ss t^T (rimp2.rs:1987-2011), it writes the T~ block
// of the pair for the phase-B/C GEMMs (adjoint.hip) and the per-item sums of
// t*T~ by virtual row, virtual column and in total, from which dE/deps
// follows (dE/dD = -t T~).
//
// Operands come from the packed K-major layout of pack.hip:
// row (i*nvir_pad + a) holds B^Q_ia for Q = 0..naux_pad-1.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "fh_device.h"
#include "fh_kernels.h"
#include "pair_body.h"

namespace fh {

namespace {

__device__ __forceinline__ pair::Item decode_item(const EnergyArgs &a) {
    pair::Item it;
    it.id = xcd_remap((int64_t)blockIdx.x, a.nitems);
    const int64_t p = it.id / a.ntp;
    const int tp = (int)(it.id % a.ntp);
    if (a.i_fixed >= 0) {
        it.i = a.i_fixed;
        it.j = a.i_fixed + (int)p;
    } else {
        decode_pair(p, a.nocc, it.i, it.j);
    }
    decode_tile_pair(tp, a.ntile, it.ta, it.tb);
    return it;
}

template <typename T, int TILE, int KP, bool ADJ, bool SIGNED = false, bool FLUSH = false,
          bool AMP = false>
__global__ __launch_bounds__(256) void pair_k4_kernel(EnergyArgs a) {
    __shared__ __attribute__((aligned(16))) pair::K4Lds<T, TILE, KP> lds;
    pair::pair_k4_body<T, TILE, KP, ADJ, false, SIGNED, FLUSH, AMP>(a, decode_item(a), lds);
}

// Opposite-spin rectangle: item (i, J, tA, tB), J and tB on the beta side;
// the items of alpha row i are contiguous.
template <typename T, int TILE, int KP, bool FLUSH = false>
__global__ __launch_bounds__(256) void pair_rect_kernel(EnergyArgs a) {
    __shared__ __attribute__((aligned(16))) pair::K4Lds<T, TILE, KP> lds;
    pair::Item it;
    it.id = xcd_remap((int64_t)blockIdx.x, a.nitems);
    const int64_t ntp = (int64_t)a.ntile * a.ntile_b;
    const int64_t p = it.id / ntp;
    const int t = (int)(it.id % ntp);
    it.i = (int)(p / a.nocc_b);
    it.j = (int)(p % a.nocc_b);
    it.ta = t / a.ntile_b;
    it.tb = t % a.ntile_b;
    pair::pair_k4_body<T, TILE, KP, false, true, false, FLUSH>(a, it, lds);
}


// BF16x3 energy on eight waves from the fragment-order operand
// (pair_bf16x3_w8f_body), TILE = 128 or 64 (at 64: 106 VGPRs, two workgroups
// per CU).
template <int TILE>
__global__ __launch_bounds__(512) void pair_bf16x3_w8f_kernel(EnergyArgs a) {
    __shared__ __attribute__((aligned(16))) pair::Bf16Lds<TILE> lds;
    pair::pair_bf16x3_w8f_body<TILE>(a, decode_item(a), lds);
}

template <int TILE, bool ADJ>
__global__ __launch_bounds__(256) void pair_bf16x3_kernel(EnergyArgs a) {
    __shared__ __attribute__((aligned(16))) pair::Bf16Lds<TILE> lds;
    pair::pair_bf16x3_body<TILE, ADJ>(a, decode_item(a), lds);
}

// ---------------------------------------------------------------------------
// Deterministic reductions: per occupied row i (items of row i are contiguous:
// pairs (i, j >= i) start at p0 = i*nocc - i(i-1)/2), then an ascending fold
// over i (the reference's i-ordered serial sum, rimp2.rs:1533-1538).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void reduce_rows_kernel(const double *__restrict__ ip, int nocc,
                                                          int ntp, double *__restrict__ rows) {
    const int i = blockIdx.x;
    const int64_t p0 = (int64_t)i * nocc - (int64_t)i * (i - 1) / 2;
    const int64_t w0 = p0 * ntp, w1 = (p0 + (nocc - i)) * ntp;
    KSum os, ss;
    for (int64_t w = w0 + threadIdx.x; w < w1; w += 256) {
        os.merge(ip[4 * w], ip[4 * w + 1]);
        ss.merge(ip[4 * w + 2], ip[4 * w + 3]);
    }
    pair::block_store(os, ss, 1.0, rows + 4 * i);
}

// Rectangle rows: items [i*per_row, (i+1)*per_row) belong to alpha row i.
__global__ __launch_bounds__(256) void reduce_rect_rows_kernel(const double *__restrict__ ip,
                                                               int64_t per_row,
                                                               double *__restrict__ rows) {
    const int i = blockIdx.x;
    KSum os, ss;
    for (int64_t w = (int64_t)i * per_row + threadIdx.x; w < (int64_t)(i + 1) * per_row;
         w += 256) {
        os.merge(ip[4 * w], ip[4 * w + 1]);
        ss.merge(ip[4 * w + 2], ip[4 * w + 3]);
    }
    pair::block_store(os, ss, 1.0, rows + 4 * i);
}

__global__ void fold_rows_kernel(const double *__restrict__ rows, int nocc,
                                 double *__restrict__ result) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    KSum os, ss;
    for (int i = 0; i < nocc; ++i) {
        os.merge(rows[4 * i], rows[4 * i + 1]);
        ss.merge(rows[4 * i + 2], rows[4 * i + 3]);
        result[2 + 2 * i] = rows[4 * i] + rows[4 * i + 1];
        result[3 + 2 * i] = rows[4 * i + 2] + rows[4 * i + 3];
    }
    result[0] = os.value();
    result[1] = ss.value();
}

template <typename K>
hipError_t launch_items(K kern, const EnergyArgs &a, hipStream_t st, int threads = 256) {
    hipLaunchKernelGGL(kern, dim3((unsigned)a.nitems), dim3(threads), 0, st, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_energy(const EnergyArgs &a, hipStream_t st) {
    if (a.nitems <= 0 || a.nitems > 0x7fffffffLL) return hipErrorInvalidValue;
    hipError_t e = hipErrorInvalidValue;
    if (a.tamp) {  // MP2 densities: FP64 energy that also stores the amplitudes
        if (a.prec != Prec::F64 || a.q_neg < a.naux_pad) return hipErrorInvalidValue;
        if (a.tile == 128) e = launch_items(pair_k4_kernel<double, 128, 16, false, false, false, true>, a, st);
        else if (a.tile == 64) e = launch_items(pair_k4_kernel<double, 64, 16, false, false, false, true>, a, st);
        else if (a.tile == 32) e = launch_items(pair_k4_kernel<double, 32, 16, false, false, false, true>, a, st);
        return e;
    }
    if (a.q_neg < a.naux_pad) {  // signed factor: FP64 / FP32 only
        if (a.prec == Prec::F64) {
            if (a.tile == 128) e = launch_items(pair_k4_kernel<double, 128, 16, false, true>, a, st);
            else if (a.tile == 64) e = launch_items(pair_k4_kernel<double, 64, 16, false, true>, a, st);
            else if (a.tile == 32) e = launch_items(pair_k4_kernel<double, 32, 16, false, true>, a, st);
        } else if (a.prec == Prec::F32) {
            if (a.tile == 128) e = launch_items(pair_k4_kernel<float, 128, 16, false, true>, a, st);
            else if (a.tile == 64) e = launch_items(pair_k4_kernel<float, 64, 16, false, true>, a, st);
            else if (a.tile == 32) e = launch_items(pair_k4_kernel<float, 32, 16, false, true>, a, st);
        }
        return e;
    }
    switch (a.prec) {
    case Prec::F64:
        if (a.tile == 128) e = launch_items(pair_k4_kernel<double, 128, 16, false>, a, st);
        else if (a.tile == 64) e = launch_items(pair_k4_kernel<double, 64, 16, false>, a, st);
        else if (a.tile == 32) e = launch_items(pair_k4_kernel<double, 32, 16, false>, a, st);
        break;
    case Prec::F32:
        if (a.tile == 128) e = launch_items(pair_k4_kernel<float, 128, 16, false>, a, st);
        else if (a.tile == 64) e = launch_items(pair_k4_kernel<float, 64, 16, false>, a, st);
        else if (a.tile == 32) e = launch_items(pair_k4_kernel<float, 32, 16, false>, a, st);
        break;
    case Prec::MIXED:
        if (a.tile == 128) e = launch_items(pair_k4_kernel<float, 128, 16, false, false, true>, a, st);
        else if (a.tile == 64) e = launch_items(pair_k4_kernel<float, 64, 16, false, false, true>, a, st);
        else if (a.tile == 32) e = launch_items(pair_k4_kernel<float, 32, 16, false, false, true>, a, st);
        break;
    case Prec::BF16X3:
        if (a.tile == 128) {
            if (a.opf) e = launch_items(pair_bf16x3_w8f_kernel<128>, a, st, 512);
        }
        else if (a.tile == 64 && a.opf) e = launch_items(pair_bf16x3_w8f_kernel<64>, a, st, 512);
        else if (a.tile == 64) e = launch_items(pair_bf16x3_kernel<64, false>, a, st);
        else if (a.tile == 32) e = launch_items(pair_bf16x3_kernel<32, false>, a, st);
        break;
    case Prec::FP8:
        e = launch_energy_fp8(a, st);
        break;
    case Prec::I8X4:
        e = launch_energy_i8x4(a, st);
        break;
    }
    return e;
}

hipError_t launch_adjoint_pairs(const EnergyArgs &a, hipStream_t st) {
    if (a.nitems <= 0 || a.nitems > 0x7fffffffLL) return hipErrorInvalidValue;
    if (a.prec == Prec::I8X4) return launch_adjoint_i8x4(a, st);
    if (a.i_fixed < 0) return hipErrorInvalidValue;
    hipError_t e = hipErrorInvalidValue;
    switch (a.prec) {
    case Prec::F64:
        // 32 or 64 wide (the host caps the FP64 adjoint tile at 64)
        if (a.tile == 64) e = launch_items(pair_k4_kernel<double, 64, 16, true>, a, st);
        else if (a.tile == 32) e = launch_items(pair_k4_kernel<double, 32, 8, true>, a, st);
        break;
    case Prec::BF16X3:
        if (a.tile == 128) e = launch_items(pair_bf16x3_kernel<128, true>, a, st);
        else if (a.tile == 64) e = launch_items(pair_bf16x3_kernel<64, true>, a, st);
        else if (a.tile == 32) e = launch_items(pair_bf16x3_kernel<32, true>, a, st);
        break;
    case Prec::I8X4:
        e = launch_adjoint_i8x4(a, st);
        break;
    default:
        break;
    }
    return e;
}

hipError_t launch_energy_rect(const EnergyArgs &a, hipStream_t st) {
    if (a.nitems <= 0 || a.nitems > 0x7fffffffLL || a.ntile_b <= 0 || a.nocc_b <= 0)
        return hipErrorInvalidValue;
    hipError_t e = hipErrorInvalidValue;
    switch (a.prec) {
    case Prec::F64:
        if (a.tile == 128) e = launch_items(pair_rect_kernel<double, 128, 16>, a, st);
        else if (a.tile == 64) e = launch_items(pair_rect_kernel<double, 64, 16>, a, st);
        else if (a.tile == 32) e = launch_items(pair_rect_kernel<double, 32, 16>, a, st);
        break;
    case Prec::F32:
        if (a.tile == 128) e = launch_items(pair_rect_kernel<float, 128, 16>, a, st);
        else if (a.tile == 64) e = launch_items(pair_rect_kernel<float, 64, 16>, a, st);
        else if (a.tile == 32) e = launch_items(pair_rect_kernel<float, 32, 16>, a, st);
        break;
    case Prec::MIXED:
        if (a.tile == 128) e = launch_items(pair_rect_kernel<float, 128, 16, true>, a, st);
        else if (a.tile == 64) e = launch_items(pair_rect_kernel<float, 64, 16, true>, a, st);
        else if (a.tile == 32) e = launch_items(pair_rect_kernel<float, 32, 16, true>, a, st);
        break;
    case Prec::I8X4:
        e = launch_energy_i8x4_rect(a, st);
        break;
    default:
        break;
    }
    return e;
}

hipError_t launch_rect_reduce(const EnergyArgs &a, hipStream_t st) {
    const int64_t per_row = (int64_t)a.nocc_b * a.ntile * a.ntile_b;
    hipLaunchKernelGGL(reduce_rect_rows_kernel, dim3((unsigned)a.nocc), dim3(256), 0, st,
                       a.item_partials, per_row, a.row_partials);
    hipLaunchKernelGGL(fold_rows_kernel, dim3(1), dim3(64), 0, st, a.row_partials, a.nocc,
                       a.result);
    return hipGetLastError();
}

hipError_t launch_energy_reduce(const EnergyArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(reduce_rows_kernel, dim3((unsigned)a.nocc), dim3(256), 0, st,
                       a.item_partials, a.nocc, a.ntp, a.row_partials);
    hipLaunchKernelGGL(fold_rows_kernel, dim3(1), dim3(64), 0, st, a.row_partials, a.nocc,
                       a.result);
    return hipGetLastError();
}

}  // namespace fh
