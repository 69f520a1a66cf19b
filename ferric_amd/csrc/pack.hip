// Operand packing for the RI-MP2 contraction kernels (gfx950).
//
// The caller's tensor is b_ov[Q][i*nvir + a] (row-major (naux, nocc*nvir),
// the reference layout, crates/ferric-mp2/src/rimp2.rs:1445-1499). The
// contraction runs over Q, so every kernel reads "K-major" rows: one row per
// (i, a) holding all Q contiguously. pack_* writes
//
//   bp[(i*nvir_pad + a)*naux_pad + Q]
//
// with nvir_pad = round_up(nvir, 16) and naux_pad = round_up(naux, KP) of the
// consuming kernel; padding rows and columns are written as zeros, so no
// memset is needed and the kernels' K loops need no bounds checks. A signed
// factor (B+ and B-, rimp2.rs:1559-1607) fills two column segments of the same
// rows: the FP64 / FP32 packers then write with a row stride ld wider than the
// segment, starting at the segment's first column.
//
// Formats: FP64, FP32, BF16 hi/lo split (hi = RN(x), lo = RN(x - hi)),
// OCP FP8 E4M3 with one power-of-two (E8M0) scale per (occupied i, 64-row Q
// block), and INT8 digit slices: per packed row one exponent f with
// max_Q |x| <= (1 - 2^-7) 2^(f + 1), and per element the four signed byte
// digits of Y = RN(x 2^(30 - f)), stored in MFMA fragment order: one 4 KB
// block per (16 packed rows, 64-deep Q panel), digit by digit (energy_i8.hip).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "fh_device.h"
#include "fh_kernels.h"

namespace fh {

namespace {

constexpr int PT = 64;  // transpose tile: 64 Q x 64 output rows

template <typename Tin>
__device__ __forceinline__ float load_as_float(const Tin *p) {
    return (float)(*p);
}

// One 64x64 tile: rows r in [r0, r0+64) of the packed layout, Q in [q0, q0+64).
// Reads are coalesced along ia (input rows), writes along Q (output rows).
template <typename Tin, typename Tout, typename Conv>
__device__ __forceinline__ void transpose_tile(const Tin *__restrict__ b, int64_t nov, int naux,
                                               int nvir, int nvir_pad, int64_t nrows,
                                               int naux_pad, int64_t ld, Conv conv,
                                               Tout *__restrict__ out0, Tout *__restrict__ out1) {
    __shared__ float tile[PT][PT + 1];
    const int tx = threadIdx.x % PT, ty = threadIdx.x / PT;  // 64 x 4
    const int64_t r0 = (int64_t)blockIdx.x * PT;
    const int q0 = blockIdx.y * PT;
    // Load: tile[q][r] = b[q0+q][col(r0+r)].
    const int64_t r = r0 + tx;
    int64_t col = -1;
    if (r < nrows) {
        const int64_t i = r / nvir_pad;
        const int a = (int)(r % nvir_pad);
        if (a < nvir) col = i * nvir + a;
    }
    for (int q = ty; q < PT; q += 4) {
        const int qq = q0 + q;
        float v = 0.0f;
        if (col >= 0 && qq < naux) v = load_as_float(b + (int64_t)qq * nov + col);
        tile[q][tx] = v;
    }
    __syncthreads();
    // Store: out[(r0+rr)*ld + q0 + tx] = tile[tx][rr].
    for (int rr = ty; rr < PT; rr += 4) {
        const int64_t orow = r0 + rr;
        if (orow < nrows && q0 + tx < naux_pad) {
            const int64_t o = orow * ld + q0 + tx;
            conv(tile[tx][rr], out0 + o, out1 ? out1 + o : nullptr);
        }
    }
}

// FP64 is transposed in double (no float round trip).
template <typename Tin>
__global__ __launch_bounds__(256) void pack_f64_kernel(const Tin *__restrict__ b, int64_t nov,
                                                       int naux, int nvir, int nvir_pad,
                                                       int64_t nrows, int naux_pad,
                                                       double *__restrict__ out, int64_t ld) {
    __shared__ double tile[PT][PT + 1];
    const int tx = threadIdx.x % PT, ty = threadIdx.x / PT;
    const int64_t r0 = (int64_t)blockIdx.x * PT;
    const int q0 = blockIdx.y * PT;
    const int64_t r = r0 + tx;
    int64_t col = -1;
    if (r < nrows) {
        const int64_t i = r / nvir_pad;
        const int a = (int)(r % nvir_pad);
        if (a < nvir) col = i * nvir + a;
    }
    for (int q = ty; q < PT; q += 4) {
        const int qq = q0 + q;
        double v = 0.0;
        if (col >= 0 && qq < naux) v = (double)b[(int64_t)qq * nov + col];
        tile[q][tx] = v;
    }
    __syncthreads();
    for (int rr = ty; rr < PT; rr += 4) {
        const int64_t orow = r0 + rr;
        if (orow < nrows && q0 + tx < naux_pad) out[orow * ld + q0 + tx] = tile[tx][rr];
    }
}

struct ConvF32 {
    __device__ void operator()(float x, float *o, float *) const { *o = x; }
};

struct ConvBf16Split {
    __device__ void operator()(float x, uint16_t *hi, uint16_t *lo) const {
        __bf16 h = (__bf16)x;  // round to nearest even
        float hf = (float)h;
        __bf16 l = (__bf16)(x - hf);  // x - hf is exact (Sterbenz)
        *hi = __builtin_bit_cast(uint16_t, h);
        *lo = __builtin_bit_cast(uint16_t, l);
    }
};

template <typename Tin>
__global__ __launch_bounds__(256) void pack_f32_kernel(const Tin *__restrict__ b, int64_t nov,
                                                       int naux, int nvir, int nvir_pad,
                                                       int64_t nrows, int naux_pad,
                                                       float *__restrict__ out, int64_t ld) {
    transpose_tile<Tin, float>(b, nov, naux, nvir, nvir_pad, nrows, naux_pad, ld, ConvF32{}, out,
                               (float *)nullptr);
}

template <typename Tin>
__global__ __launch_bounds__(256) void pack_bf16x3_kernel(const Tin *__restrict__ b, int64_t nov,
                                                          int naux, int nvir, int nvir_pad,
                                                          int64_t nrows, int naux_pad,
                                                          uint16_t *__restrict__ hi,
                                                          uint16_t *__restrict__ lo) {
    transpose_tile<Tin, uint16_t>(b, nov, naux, nvir, nvir_pad, nrows, naux_pad, naux_pad,
                                  ConvBf16Split{}, hi, lo);
}

// Per (i, 64-row Q block) maximum |B| over finite entries -> E8M0 scale
// exponent byte e + 127 with 2^e >= amax / 448 (so every scaled value fits
// E4M3's finite range). Grid (nqb, nocc), 256 threads.
template <typename Tin>
__global__ __launch_bounds__(256) void fp8_scale_kernel(const Tin *__restrict__ b, int64_t nov,
                                                        int naux, int nvir, int nqb,
                                                        uint8_t *__restrict__ scale) {
    const int qb = blockIdx.x, i = blockIdx.y;
    float m = 0.0f;
    for (int q = qb * 64; q < qb * 64 + 64 && q < naux; ++q) {
        const Tin *row = b + (int64_t)q * nov + (int64_t)i * nvir;
        for (int a = threadIdx.x; a < nvir; a += 256) {
            float v = fabsf((float)row[a]);
            if (v <= 3.0e38f && v > m) m = v;  // finite only
        }
    }
    __shared__ float red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        scale[(int64_t)i * nqb + qb] = (uint8_t)(fp8_scale_exponent(red[0]) + 127);
    }
}

template <typename Tin>
__global__ __launch_bounds__(256) void pack_fp8_kernel(const Tin *__restrict__ b, int64_t nov,
                                                       int naux, int nvir, int nvir_pad,
                                                       int64_t nrows, int naux_pad, int nqb,
                                                       const uint8_t *__restrict__ scale,
                                                       uint8_t *__restrict__ out) {
    __shared__ float tile[PT][PT + 1];
    const int tx = threadIdx.x % PT, ty = threadIdx.x / PT;
    const int64_t r0 = (int64_t)blockIdx.x * PT;
    const int q0 = blockIdx.y * PT;  // one 64-row Q block per tile
    const int64_t r = r0 + tx;
    int64_t col = -1;
    if (r < nrows) {
        const int64_t i = r / nvir_pad;
        const int a = (int)(r % nvir_pad);
        if (a < nvir) col = i * nvir + a;
    }
    for (int q = ty; q < PT; q += 4) {
        const int qq = q0 + q;
        float v = 0.0f;
        if (col >= 0 && qq < naux) v = (float)b[(int64_t)qq * nov + col];
        tile[q][tx] = v;
    }
    __syncthreads();
    for (int rr = ty; rr < PT; rr += 4) {
        const int64_t orow = r0 + rr;
        if (orow < nrows && q0 + tx < naux_pad) {
            const int64_t i = orow / nvir_pad;
            const int e = (int)scale[i * nqb + blockIdx.y] - 127;
            out[orow * naux_pad + q0 + tx] = f32_to_e4m3(ldexpf(tile[tx][rr], -e));
        }
    }
}

// The E4M3 bytes of pack_fp8_kernel in MFMA fragment order (PackFmt::FP8F):
// per (16-row group, 128-deep Q panel) one 2 KB block, lane l = 16 g + r at
// byte 32 l holding row r's Q = 16 g .. 16 g + 15 and then 64 + 16 g .. 64 + 16 g
// + 15 of the panel (the two 16-byte chunks g and g + 4 of its 128-byte row).
// One 128 Q x 64 row tile per block; thread t then writes lane t % 64 of row
// group t / 64, so every wave writes two whole kilobytes.
template <typename Tin>
__global__ __launch_bounds__(256) void pack_fp8f_kernel(const Tin *__restrict__ b, int64_t nov,
                                                        int naux, int nvir, int nvir_pad,
                                                        int64_t nrows, int naux_pad, int nqb,
                                                        const uint8_t *__restrict__ scale,
                                                        char *__restrict__ out) {
    __shared__ float tile[2 * PT][PT + 1];
    const int tx = threadIdx.x % PT, ty = threadIdx.x / PT;
    const int64_t r0 = (int64_t)blockIdx.x * PT;
    const int P = blockIdx.y, q0 = P * 2 * PT;
    const int64_t r = r0 + tx;
    int64_t col = -1;
    if (r < nrows) {
        const int64_t i = r / nvir_pad;
        const int a = (int)(r % nvir_pad);
        if (a < nvir) col = i * nvir + a;
    }
    for (int q = ty; q < 2 * PT; q += 4) {
        const int qq = q0 + q;
        float v = 0.0f;
        if (col >= 0 && qq < naux) v = (float)b[(int64_t)qq * nov + col];
        tile[q][tx] = v;
    }
    __syncthreads();
    const int l = threadIdx.x % 64, g = l / 16, rr = (threadIdx.x / 64) * 16 + l % 16;
    const int64_t orow = r0 + rr;
    if (orow >= nrows) return;
    const int64_t i = orow / nvir_pad;
    uint32_t w[8];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int e = (int)scale[i * nqb + 2 * P + h] - 127;
#pragma unroll
        for (int k = 0; k < 16; k += 4) {
            uint32_t word = 0;
#pragma unroll
            for (int u = 0; u < 4; ++u)
                word |= (uint32_t)f32_to_e4m3(ldexpf(tile[64 * h + 16 * g + k + u][rr], -e))
                        << (8 * u);
            w[4 * h + k / 4] = word;
        }
    }
    char *o = out + ((orow / 16) * (int64_t)(naux_pad / (2 * PT)) + P) * 2048 + l * 32;
    *(uint4 *)o = make_uint4(w[0], w[1], w[2], w[3]);
    *(uint4 *)(o + 16) = make_uint4(w[4], w[5], w[6], w[7]);
}

// Row sums of lo^2 over Q (FP64; each lo^2 is exact in FP32): the lo*lo term
// the three-pass BF16 product omits. On the diagonal elements
// K_ii[a, a] = sum_Q B^Q_ia^2 every omitted term is positive, so the omission
// is a systematic bias there (elsewhere it has either sign and averages out);
// the energy epilogue adds these sums back on exactly those elements.
// One wave per packed row, 4 rows per workgroup.
__global__ __launch_bounds__(256) void lo_sq_kernel(const uint16_t *__restrict__ lo,
                                                    int64_t nrows, int naux_pad,
                                                    double *__restrict__ out) {
    const int64_t row = (int64_t)blockIdx.x * 4 + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (row >= nrows) return;
    const uint16_t *p = lo + row * naux_pad;
    double s = 0.0;
    for (int q = lane; q < naux_pad; q += kWave) {
        const float v = (float)__builtin_bit_cast(__bf16, p[q]);
        s += (double)(v * v);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, kWave);
    if (lane == 0) out[row] = s;
}

// max over finite |x| (for the reduced-precision range check).
template <typename Tin>
__global__ __launch_bounds__(256) void absmax_kernel(const Tin *__restrict__ b, int64_t n,
                                                     unsigned int *__restrict__ out_bits) {
    float m = 0.0f;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256) {
        double v = fabs((double)b[k]);
        if (v <= 1.7e308 && (float)v > m) m = (float)fmin(v, 3.0e38);
    }
    __shared__ float red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    // Non-negative floats order like their bit patterns: an integer max is exact.
    if (threadIdx.x == 0) atomicMax(out_bits, __float_as_uint(red[0]));
}

__global__ void zero_u32_kernel(unsigned int *p) { *p = 0u; }

// INT8 slices, step 1: the row exponent f of every packed row (i, a). With E
// the frexp exponent of m = max_Q |B^Q_ia| (m < 2^E), f = E - 1 when
// m <= (1 - 2^-7) 2^E and f = E otherwise, so that |Y| = |RN(x 2^(30 - f))| stays
// at most 2^31 - 2^24 (top digit in [-127, 127]) and uses 31 bits for most rows;
// 0 for a zero or padding row, kI8NonFinite for a row holding a NaN or an
// infinity.
// Block = 64 columns x 4 Q lanes (reads coalesced along a), grid
// (ceil(nvir_pad / 64), nocc).
template <typename Tin>
__global__ __launch_bounds__(256) void i8_rowexp_kernel(const Tin *__restrict__ b, int64_t nov,
                                                        int naux, int nvir, int nvir_pad,
                                                        int *__restrict__ rexp) {
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    const int a = blockIdx.x * 64 + tx, i = blockIdx.y;
    double m = 0.0;
    int bad = 0;
    if (a < nvir) {
        const Tin *p = b + (int64_t)i * nvir + a;
        for (int q = ty; q < naux; q += 4) {
            const double v = fabs((double)p[(int64_t)q * nov]);
            if (v <= 1.7976931348623157e308)
                m = fmax(m, v);
            else
                bad = 1;  // NaN or infinity
        }
    }
    __shared__ double red[4][64];
    __shared__ int flag[4][64];
    red[ty][tx] = m;
    flag[ty][tx] = bad;
    __syncthreads();
    if (ty == 0 && a < nvir_pad) {
        const double mm = fmax(fmax(red[0][tx], red[1][tx]), fmax(red[2][tx], red[3][tx]));
        int e = 0;
        if (flag[0][tx] | flag[1][tx] | flag[2][tx] | flag[3][tx])
            e = kI8NonFinite;
        else if (mm > 0.0) {
            (void)frexp(mm, &e);
            if (mm <= ldexp(1.0 - 0x1p-7, e)) e -= 1;
        }
        rexp[(int64_t)i * nvir_pad + a] = e;
    }
}

// INT8 slices, step 2: one 64x64 tile (64 Q = one panel, 64 packed rows).
// Y = RN(x 2^(30 - f)) is exact arithmetic on the FP64 value
// (|Y| <= 2^31 - 2^24); its digits are peeled from the bottom as signed bytes,
// so d4, d3, d2 lie in [-128, 127] and the remaining top digit d1 in
// [-127, 127].
template <typename Tin>
__global__ __launch_bounds__(256) void pack_i8x4_kernel(const Tin *__restrict__ b, int64_t nov,
                                                        int naux, int nvir, int nvir_pad,
                                                        int64_t nrows, int naux_pad,
                                                        const int *__restrict__ rexp,
                                                        int8_t *__restrict__ out) {
    __shared__ double tile[PT][PT + 1];
    const int tx = threadIdx.x % PT, ty = threadIdx.x / PT;
    const int64_t r0 = (int64_t)blockIdx.x * PT;
    const int q0 = blockIdx.y * PT;
    const int64_t r = r0 + tx;
    int64_t col = -1;
    if (r < nrows) {
        const int64_t i = r / nvir_pad;
        const int a = (int)(r % nvir_pad);
        if (a < nvir) col = i * nvir + a;
    }
    for (int q = ty; q < PT; q += 4) {
        const int qq = q0 + q;
        double v = 0.0;
        if (col >= 0 && qq < naux) v = (double)b[(int64_t)qq * nov + col];
        tile[q][tx] = v;
    }
    __syncthreads();
    // Digit phase: thread t takes packed row r0 + rr, rr = 16 (t / 64) + t % 16,
    // and the 16 Q values 16 g .. 16 g + 15 of the panel, g = (t / 16) % 4:
    // the 16 bytes of lane 16 g + rr % 16 in each digit's kilobyte of the MFMA
    // fragment order (block (row group orow / 16, panel q0 / 64) of 4 KB, digit
    // by digit), so every wave writes four whole kilobytes.
    const int t = threadIdx.x;
    const int g = (t / 16) % 4, rr = (t / 64) * 16 + t % 16;
    const int64_t orow = r0 + rr;
    if (orow < nrows) {
        const int e = rexp[orow];
        uint32_t w[4][4];  // [digit d1..d4][word]: byte k % 4 of word k / 4 is Q = 16 g + k
#pragma unroll
        for (int d = 0; d < 4; ++d)
#pragma unroll
            for (int k = 0; k < 4; ++k) w[d][k] = 0u;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            int y = 0;
            if (e != kI8NonFinite) y = (int)rint(ldexp(tile[16 * g + k][rr], 30 - e));
            const int d4 = (int)(int8_t)(y & 0xff);
            y = (y - d4) >> 8;
            const int d3 = (int)(int8_t)(y & 0xff);
            y = (y - d3) >> 8;
            const int d2 = (int)(int8_t)(y & 0xff);
            const int d1 = (y - d2) >> 8;
            const int sh = 8 * (k % 4);
            w[0][k / 4] |= (uint32_t)(uint8_t)d1 << sh;
            w[1][k / 4] |= (uint32_t)(uint8_t)d2 << sh;
            w[2][k / 4] |= (uint32_t)(uint8_t)d3 << sh;
            w[3][k / 4] |= (uint32_t)(uint8_t)d4 << sh;
        }
        int8_t *o = out + ((orow / 16) * (int64_t)(naux_pad / 64) + q0 / 64) * 4096 +
                    (g * 16 + (int)(orow % 16)) * 16;
#pragma unroll
        for (int d = 0; d < 4; ++d)
            *(uint4 *)(o + 1024 * d) = make_uint4(w[d][0], w[d][1], w[d][2], w[d][3]);
    }
}

// BF16 hi / lo split (ConvBf16Split) written in MFMA fragment order
// (PackFmt::BF16F): per (16-row group, 32-deep Q panel) one 2 KB block, the hi
// fragment then the lo fragment, lane l = 16 g + r at byte 16 l holding row r,
// Q = 8 g .. 8 g + 7 of the panel. One 64 x 64 tile per block (as
// transpose_tile); thread t then writes lane t % 64 of row group t / 64 for
// both panels of the tile, so every wave writes whole kilobytes.
template <typename Tin>
__global__ __launch_bounds__(256) void pack_bf16f_kernel(const Tin *__restrict__ b, int64_t nov,
                                                         int naux, int nvir, int nvir_pad,
                                                         int64_t nrows, int naux_pad,
                                                         char *__restrict__ out) {
    __shared__ float tile[PT][PT + 1];
    const int tx = threadIdx.x % PT, ty = threadIdx.x / PT;
    const int64_t r0 = (int64_t)blockIdx.x * PT;
    const int q0 = blockIdx.y * PT;
    const int64_t r = r0 + tx;
    int64_t col = -1;
    if (r < nrows) {
        const int64_t i = r / nvir_pad;
        const int a = (int)(r % nvir_pad);
        if (a < nvir) col = i * nvir + a;
    }
    for (int q = ty; q < PT; q += 4) {
        const int qq = q0 + q;
        float v = 0.0f;
        if (col >= 0 && qq < naux) v = load_as_float(b + (int64_t)qq * nov + col);
        tile[q][tx] = v;
    }
    __syncthreads();
    const int l = threadIdx.x % 64, g = l / 16, rr = (threadIdx.x / 64) * 16 + l % 16;
    const int64_t orow = r0 + rr;
    if (orow >= nrows) return;
    const int np = naux_pad / 32;
#pragma unroll
    for (int pp = 0; pp < 2; ++pp) {
        const int P = q0 / 32 + pp;
        if (P >= np) break;
        uint32_t hw[4], lw[4];
#pragma unroll
        for (int k = 0; k < 8; k += 2) {
            uint16_t h0, l0, h1, l1;
            ConvBf16Split{}(tile[32 * pp + 8 * g + k][rr], &h0, &l0);
            ConvBf16Split{}(tile[32 * pp + 8 * g + k + 1][rr], &h1, &l1);
            hw[k / 2] = (uint32_t)h0 | ((uint32_t)h1 << 16);
            lw[k / 2] = (uint32_t)l0 | ((uint32_t)l1 << 16);
        }
        char *o = out + ((orow / 16) * (int64_t)np + P) * 2048 + l * 16;
        *(uint4 *)o = make_uint4(hw[0], hw[1], hw[2], hw[3]);
        *(uint4 *)(o + 1024) = make_uint4(lw[0], lw[1], lw[2], lw[3]);
    }
}

// Sum over Q of lo^2 per packed row from the fragment-order layout, one wave
// per group of 16 rows: lane 16 g + r reads its 16-byte lo chunk (row r, Q =
// 32 P + 8 g .. + 7) of every panel P, one contiguous kilobyte per wave load,
// and sums its squares in Q order; the four lanes of a row then combine by an
// xor pair (g, g ^ 1) and (g, g ^ 2).
__global__ __launch_bounds__(256) void lo_sq_frag_kernel(const char *__restrict__ frag,
                                                         int64_t nrows, int naux_pad,
                                                         double *__restrict__ out) {
    const int64_t grp = (int64_t)blockIdx.x * 4 + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (grp * 16 >= nrows) return;
    const int np = naux_pad / 32;
    const char *p = frag + grp * (int64_t)np * 2048 + 1024 + lane * 16;
    double s = 0.0;
    for (int P = 0; P < np; ++P) {
        const uint4 w = *(const uint4 *)(p + (int64_t)P * 2048);
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float v =
                (float)__builtin_bit_cast(__bf16, (uint16_t)(ws[e / 2] >> (16 * (e % 2))));
            s += (double)(v * v);
        }
    }
    s += __shfl_xor(s, 16, kWave);
    s += __shfl_xor(s, 32, kWave);
    if (lane < 16) out[grp * 16 + lane] = s;
}

}  // namespace

static inline dim3 pack_grid(int64_t nrows, int naux_pad) {
    return dim3((unsigned)((nrows + PT - 1) / PT), (unsigned)((naux_pad + PT - 1) / PT));
}

hipError_t launch_pack(const PackArgs &a, hipStream_t st) {
    const int64_t nrows = (int64_t)a.nocc * a.nvir_pad;
    const int64_t nov = (int64_t)a.nocc * a.nvir;
    dim3 grid = pack_grid(nrows, a.naux_pad);
    const bool f32in = a.in_f32;
    const int64_t ld = a.ld > 0 ? a.ld : a.naux_pad;
    if (ld != a.naux_pad && a.fmt != PackFmt::F64 && a.fmt != PackFmt::F32)
        return hipErrorInvalidValue;
#define FH_LAUNCH(K, ...)                                                                      \
    do {                                                                                       \
        if (f32in)                                                                             \
            hipLaunchKernelGGL(K<float>, grid, dim3(256), 0, st, (const float *)a.b, nov,      \
                               a.naux, a.nvir, a.nvir_pad, nrows, a.naux_pad, __VA_ARGS__);    \
        else                                                                                   \
            hipLaunchKernelGGL(K<double>, grid, dim3(256), 0, st, (const double *)a.b, nov,    \
                               a.naux, a.nvir, a.nvir_pad, nrows, a.naux_pad, __VA_ARGS__);    \
    } while (0)
    switch (a.fmt) {
    case PackFmt::F64:
        FH_LAUNCH(pack_f64_kernel, (double *)a.out0, ld);
        break;
    case PackFmt::F32:
        FH_LAUNCH(pack_f32_kernel, (float *)a.out0, ld);
        break;
    case PackFmt::BF16X2:
        FH_LAUNCH(pack_bf16x3_kernel, (uint16_t *)a.out0, (uint16_t *)a.out1);
        if (a.lo_sq)
            hipLaunchKernelGGL(lo_sq_kernel, dim3((unsigned)((nrows + 3) / 4)), dim3(256), 0, st,
                               (const uint16_t *)a.out1, nrows, a.naux_pad, a.lo_sq);
        break;
    case PackFmt::BF16F:
        if (a.naux_pad % 32 != 0 || nrows % 16 != 0) return hipErrorInvalidValue;
        FH_LAUNCH(pack_bf16f_kernel, (char *)a.out0);
        if (a.lo_sq)
            hipLaunchKernelGGL(lo_sq_frag_kernel, dim3((unsigned)((nrows / 16 + 3) / 4)), dim3(256),
                               0, st, (const char *)a.out0, nrows, a.naux_pad, a.lo_sq);
        break;
    case PackFmt::FP8: {
        const int nqb = a.naux_pad / 64;
        dim3 sg((unsigned)nqb, (unsigned)a.nocc);
        if (f32in)
            hipLaunchKernelGGL(fp8_scale_kernel<float>, sg, dim3(256), 0, st, (const float *)a.b,
                               nov, a.naux, a.nvir, nqb, a.scale);
        else
            hipLaunchKernelGGL(fp8_scale_kernel<double>, sg, dim3(256), 0, st,
                               (const double *)a.b, nov, a.naux, a.nvir, nqb, a.scale);
        FH_LAUNCH(pack_fp8_kernel, nqb, (const uint8_t *)a.scale, (uint8_t *)a.out0);
        break;
    }
    case PackFmt::FP8F: {
        if (a.naux_pad % (2 * PT) != 0 || nrows % 16 != 0) return hipErrorInvalidValue;
        const int nqb = a.naux_pad / 64;
        dim3 sg((unsigned)nqb, (unsigned)a.nocc);
        if (f32in)
            hipLaunchKernelGGL(fp8_scale_kernel<float>, sg, dim3(256), 0, st, (const float *)a.b,
                               nov, a.naux, a.nvir, nqb, a.scale);
        else
            hipLaunchKernelGGL(fp8_scale_kernel<double>, sg, dim3(256), 0, st,
                               (const double *)a.b, nov, a.naux, a.nvir, nqb, a.scale);
        grid.y = (unsigned)(a.naux_pad / (2 * PT));
        FH_LAUNCH(pack_fp8f_kernel, nqb, (const uint8_t *)a.scale, (char *)a.out0);
        break;
    }
    case PackFmt::I8X4: {
        if (!a.rexp || a.naux_pad % PT != 0) return hipErrorInvalidValue;
        dim3 eg((unsigned)((a.nvir_pad + 63) / 64), (unsigned)a.nocc);
        if (f32in)
            hipLaunchKernelGGL(i8_rowexp_kernel<float>, eg, dim3(256), 0, st, (const float *)a.b,
                               nov, a.naux, a.nvir, a.nvir_pad, a.rexp);
        else
            hipLaunchKernelGGL(i8_rowexp_kernel<double>, eg, dim3(256), 0, st,
                               (const double *)a.b, nov, a.naux, a.nvir, a.nvir_pad, a.rexp);
        FH_LAUNCH(pack_i8x4_kernel, (const int *)a.rexp, (int8_t *)a.out0);
        break;
    }
    }
#undef FH_LAUNCH
    return hipGetLastError();
}

hipError_t launch_absmax(const void *b, bool in_f32, int64_t n, unsigned int *out_bits,
                         hipStream_t st) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, st, out_bits);
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    if (in_f32)
        hipLaunchKernelGGL(absmax_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, st,
                           (const float *)b, n, out_bits);
    else
        hipLaunchKernelGGL(absmax_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, st,
                           (const double *)b, n, out_bits);
    return hipGetLastError();
}

}  // namespace fh
