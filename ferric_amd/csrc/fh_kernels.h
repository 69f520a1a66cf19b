// Launch interfaces of the RI-MP2 kernels (host side) and the scalar format
// helpers shared by host and device code.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace fh {

// ---- OCP FP8 E4M3 (bias 7, max 448, no infinities, NaN = S.1111.111) ------

// Round-to-nearest-even conversion with saturation to +-448.
__host__ __device__ inline uint8_t f32_to_e4m3(float x) {
    uint32_t bits;
    __builtin_memcpy(&bits, &x, 4);
    const uint8_t sign = (uint8_t)((bits >> 24) & 0x80u);
    const float a = fabsf(x);
    if (!(a == a)) return 0x7F;           // NaN
    if (a >= 448.0f) return sign | 0x7E;  // saturate
    if (a < 0.015625f) {                  // below 2^-6: subnormal step 2^-9
        const float k = rintf(a * 512.0f);  // 0..8; 8 is the smallest normal, code 0x08
        return sign | (uint8_t)k;
    }
    int e2;
    (void)frexpf(a, &e2);
    int ex = e2 - 1;                          // a in [2^ex, 2^(ex+1))
    float q = rintf(ldexpf(a, 3 - ex));      // 8..16
    if (q >= 16.0f) {
        q = 8.0f;
        ex += 1;
    }
    uint8_t code = (uint8_t)(((ex + 7) << 3) | ((int)q - 8));
    if (code >= 0x7F) code = 0x7E;
    return sign | code;
}

__host__ __device__ inline float e4m3_to_f32(uint8_t c) {
    const float s = (c & 0x80) ? -1.0f : 1.0f;
    const int e = (c >> 3) & 0xF, m = c & 7;
    if (e == 15 && m == 7) return NAN;
    if (e == 0) return s * ldexpf((float)m, -9);
    return s * ldexpf((float)(8 + m), e - 10);
}

// Smallest e with 448 * 2^e >= amax (0 for amax == 0), clamped to E8M0's range.
__host__ __device__ inline int fp8_scale_exponent(float amax) {
    if (!(amax > 0.0f)) return 0;
    int e2;
    (void)frexpf(amax / 448.0f, &e2);
    int e = e2;
    while (e > -127 && ldexpf(448.0f, e - 1) >= amax) --e;
    while (e < 127 && ldexpf(448.0f, e) < amax) ++e;
    return e;
}

// ---- INT8 digit slices -----------------------------------------------------

// Row exponent of a packed row holding a non-finite value (its digits are 0
// and its energy terms come out NaN). The epilogues see the SUM of two row
// exponents; finite rows have |e| < 1100, so a sum holding the marker is at
// least 2^28 - 1100 and one without it stays below 2200: any sum at or above
// kI8NonFiniteSum carries a non-finite row, whichever operand it came from.
constexpr int kI8NonFinite = 1 << 28;
constexpr int kI8NonFiniteSum = 1 << 27;
// Largest naux_pad whose level sums fit int32: with |d1| <= 127 and the
// other digits in [-128, 127], one level sums at most
// 2 * 127 * 128 + 2 * 128 * 128 = 65280 per Q, and 65280 * 32832 < 2^31.
constexpr int kI8MaxNauxPad = 32832;

// ---- Packing ---------------------------------------------------------------

// BF16F: the BF16X2 hi / lo split in MFMA fragment order (per 16-row group
// and 32-deep Q panel a 2 KB block: the hi then the lo fragment, lane 16 g + r
// at byte 16 (16 g + r) holding row r, Q = 8 g .. 8 g + 7), read by the
// eight-wave BF16x3 energy kernel straight into registers.
// FP8F: the FP8 bytes in MFMA fragment order (per 16-row group and 128-deep
// Q panel a 2 KB block, lane 16 g + r at byte 32 (16 g + r) holding chunks g
// and g + 4 of row r's 128-byte panel row), read by the eight-wave FP8 kernel.
enum class PackFmt { F64, F32, BF16X2, FP8, I8X4, BF16F, FP8F };

struct PackArgs {
    const void *b;    // caller layout, row-major (naux, nocc*nvir)
    bool in_f32;      // element type of b
    int naux, nocc, nvir, nvir_pad, naux_pad;
    PackFmt fmt;
    void *out0;       // F64/F32/FP8 rows, or BF16 hi
    void *out1;       // BF16 lo
    uint8_t *scale;   // FP8: (nocc, naux_pad/64) E8M0 bytes
    double *lo_sq;    // BF16: per packed row, sum over Q of lo^2 (may be null)
    int *rexp = nullptr;  // I8X4: per packed row, the exponent e of its scale 2^(e - 30)
    // F64/F32 only: row stride of the packed rows when this call fills one
    // segment of them (the two factors of a signed kernel); 0 = naux_pad. The
    // segment is naux_pad columns wide and out0 points at its first column.
    int64_t ld = 0;
};

hipError_t launch_pack(const PackArgs &a, hipStream_t st);
hipError_t launch_absmax(const void *b, bool in_f32, int64_t n, unsigned int *out_bits,
                         hipStream_t st);

// ---- Energy ----------------------------------------------------------------

// MIXED: FP32 operands and FP32 MFMA accumulation over Q panels of at most
// kMixedPanel rows, each panel's sums added into FP64 accumulators (the
// reference's Precision::Mixed, crates/ferric-core/src/gpu/mixed.rs:70,
// 122-180).
enum class Prec { F64 = 0, F32 = 1, BF16X3 = 2, FP8 = 3, I8X4 = 4, MIXED = 5 };

// Q rows of one FP32 partial sum in MIXED (MIXED_K_PANEL_DEFAULT).
constexpr int kMixedPanel = 64;

// Panel depth along Q of each precision's kernel (naux_pad is a multiple).
inline int panel_depth(Prec p) {
    switch (p) {
    case Prec::F64: return 16;
    case Prec::F32: return 16;
    case Prec::MIXED: return 16;
    case Prec::BF16X3: return 32;
    case Prec::FP8: return 128;
    case Prec::I8X4: return 64;
    }
    return 128;
}

struct EnergyArgs {
    Prec prec;
    int tile;                 // 32, 64 or 128
    const void *op0;          // packed rows (F64/F32/FP8) or BF16 hi
    const void *op1;          // BF16 lo
    const uint8_t *scale;     // FP8 scales (nocc, nqb)
    const double *diag_corr;  // BF16: sum_Q lo^2 per packed row, added to K_ii[a, a]
    const int *rexp = nullptr;  // I8X4: per packed row, the exponent of its digit scale
    const void *opf = nullptr;  // BF16X3 / FP8, tile 128: the operand in MFMA fragment order
    // I8X4: > 0 starts every item's Q panels at (device wall clock / q_rot_ticks)
    // mod the panel count, so items running at the same moment read the same Q
    // panel of their shared occupied rows; the int32 level sums are exact (modular),
    // so the panel order changes no bit of the result. 0: panels from 0.
    int q_rot_ticks = 0;
    int nocc, nvir, nvir_pad, naux_pad, nqb;
    const double *eps_occ;    // device, active occupied (nocc)
    const double *eps_vir;    // device, virtual (nvir)
    double kappa;             // NaN = none
    int ntile, ntp;           // tiles over nvir, tile pairs (tA <= tB)
    int64_t nitems;           // npairs * ntp
    double *item_partials;    // 4 doubles per item
    double *row_partials;     // 4 doubles per i
    double *result;           // e_os, e_ss, then 2*nocc row values
    // Adjoint (phase A of one occupied row) only:
    int i_fixed = -1;         // >= 0: items are the pairs (i_fixed, j >= i_fixed)
    double c_os = 1.0, c_ss = 1.0;
    void *tbuf = nullptr;     // T~_ij^ab at tbuf[a*ld_t + (j-i)*nvir_pad + b]
    void *ubuf = nullptr;     // T~_ij^ab at ubuf[((j-i-1)*nvir_pad + b)*nvir_pad + a], j > i
    int64_t ld_t = 0;
    bool t_f32 = false;       // T stored FP32 (reduced-precision adjoint) or FP64
    double *item_vec = nullptr;  // per item: row sums, column sums, total of t*T~
    int vec_stride = 0;
    // Row-group adjoint (INT8X4): phase A of the occupied rows [g0, g1) writes
    // T~ (symmetric over the padded (i, a) index space, column pad vpg) as BF16
    // hi / lo into two slabs:
    //   S [(i - g0) vpg + a][j vpg + b]   = T~_ij^ab, i in the group, every j >= g0
    //   St[j vpg + b][(i - g0) vpg + a]   = T~_ij^ab, i in the group, j >= g1
    // (row strides ld_s = nocc vpg and ld_st = (g1 - g0) vpg).
    uint16_t *s_hi = nullptr, *s_lo = nullptr, *st_hi = nullptr, *st_lo = nullptr;
    int64_t ld_s = 0, ld_st = 0;
    int g0 = 0, g1 = 0, vpg = 0;
    int64_t item0 = 0;        // first item of the launch (row groups)
    // Signed factor (K = B+^T B+ - B-^T B-): packed Q columns from q_neg on
    // hold B- and enter with -1.
    int q_neg = 1 << 30;
    // Opposite-spin rectangle: op0 holds the alpha rows (nocc, nvir, nvir_pad,
    // eps_occ, eps_vir), op1 the beta rows; items are (i, J, tA, tB) over every
    // alpha i, beta J and tile of each side.
    int nocc_b = 0, nvir_b = 0, nvir_pad_b = 0, ntile_b = 0;
    const double *eps_occ_b = nullptr, *eps_vir_b = nullptr;
    const int *rexp_b = nullptr;  // I8X4: row exponents of the beta rows
    // MP2 densities (FP64): the energy kernel also stores t = g / D of every
    // ordered occupied pair, tamp[((i*nocc + j)*nvir_pad + a)*nvir_pad + b].
    double *tamp = nullptr;
};

hipError_t launch_energy(const EnergyArgs &a, hipStream_t st);
// Phase A of the adjoint: the pair kernel with the T~ / dE/deps epilogue.
hipError_t launch_adjoint_pairs(const EnergyArgs &a, hipStream_t st);

// Grouped NT GEMM of the adjoint's phases B and C for one occupied row i:
//   gamma[Q, i*nvir + a]      += alpha sum_k B2[Q, i*nvir_pad + k] tbuf[a, k], k < ntail
//   gamma[Q, (i+jj)*nvir + b] += alpha sum_a B2[Q, i*nvir_pad + a] ubuf_jj[b, a], jj >= 1
struct GemmArgs {
    Prec prec;                // F64 (double operands) or BF16X3 (float operands, split)
    const void *b2;           // (m_pad, nov_pad) Q-major copy of B
    int64_t ld_b2;            // nov_pad
    const void *tbuf;         // (nvir_pad, ld_t)
    int64_t ld_t;
    const void *ubuf;         // (nocc-1-i) blocks of (nvir_pad, nvir_pad)
    int i, nocc, nvir, nvir_pad, naux, m_pad;
    int64_t nov;              // gamma row stride (nocc*nvir)
    double alpha;
    double *gamma;            // FP64 (naux, nov), accumulated into
    int tile;                 // 64 or 128
    int splits;               // phase-B split-K factor
    int kchunk;               // phase-B k per split (multiple of the panel depth)
    double *partial;          // (splits, m_pad, nvir_pad) phase-B partial tiles
};
hipError_t launch_adjoint_gemm(const GemmArgs &g, hipStream_t st);
hipError_t launch_energy_fp8(const EnergyArgs &a, hipStream_t st);
hipError_t launch_energy_i8x4(const EnergyArgs &a, hipStream_t st);
hipError_t launch_energy_i8x4_rect(const EnergyArgs &a, hipStream_t st);
// Phase A of the adjoint on the INT8 digit-slice kernel over the items of the
// occupied rows [g0, g1) (i_fixed < 0), T~ into the group's slabs.
hipError_t launch_adjoint_i8x4(const EnergyArgs &a, hipStream_t st);

// C[m][n] += alpha sum_k A[m][k] B[n][k] for m < M, n < N, with A and B given
// as BF16 hi / lo pairs (k contiguous, K a multiple of 32) and C FP32: the
// three products hi*hi + hi*lo + lo*hi, FP32 accumulation, 256 x 256 tiles.
struct Bf16x3Gemm {
    const uint16_t *ahi, *alo;
    int64_t lda;
    const uint16_t *bhi, *blo;
    int64_t ldb;
    float *c;
    int64_t ldc;
    int M, N, K;
    float alpha;
};
hipError_t launch_bf16x3_nt(const Bf16x3Gemm &g, hipStream_t st);
// Opposite-spin rectangle (F64 / F32) and its row reduction + fold.
hipError_t launch_energy_rect(const EnergyArgs &a, hipStream_t st);
hipError_t launch_rect_reduce(const EnergyArgs &a, hipStream_t st);
// Row reductions and the ascending fold into a.result.
hipError_t launch_energy_reduce(const EnergyArgs &a, hipStream_t st);

// Split-K FP64 MFMA product of the MP2 densities from stored amplitudes:
//   out[m][n] = scale * sum_k A[m][k] (cd B1[n][k] - ce B2[n][k])
// with the rows of A, B1, B2 addressed per mode:
//   DENS_VV (m, n virtual; k = (q = i*nocc + j, c)):  A = t_ij[m][c],
//            B1 = t_ij[n][c], B2 = t_ji[n][c]          (P_ab, oo_rimp2.rs:1192-1300)
//   DENS_OO (m, n occupied; k = (kk, f = a*nvir_pad + b)): A = t_m,kk[f],
//            B1 = t_n,kk[f], B2 = t_kk,n[f]            (P_ij, scale -1)
enum class DensMode { VV = 0, OO = 1 };
struct DensArgs {
    DensMode mode;
    const double *tamp;   // ordered-pair amplitude blocks
    int nocc, nvir_pad;
    int mdim;             // M = N: nvir_pad (VV) or nocc (OO)
    int ntile;            // 64-wide output tiles per side
    int64_t kpanels;      // 16-deep K panels in total
    int splits;           // K slices per output tile
    double cd, ce, scale;
    double *partial;      // (ntile^2 * splits) tiles of 64 x 64
    double *out;          // (mdim, mdim) row-major
};
hipError_t launch_density(const DensArgs &d, hipStream_t st);

}  // namespace fh
