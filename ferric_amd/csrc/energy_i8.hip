// INT8 digit-sliced variant of the fused RI-MP2 pair-energy kernel (gfx950):
// the FP64-accurate reduced-precision mode.
//
// Operand format (pack.hip, PackFmt::I8X4). Every packed row r = (i, a) has
// one exponent e_r with max_Q |B^Q_ia| <= (1 - 2^-7) 2^(e_r + 1), and every
// element is the integer Y = RN(B 2^(30 - e_r)), |Y| <= 2^31 - 2^24, written as
// four signed byte digits Y = d1 2^24 + d2 2^16 + d3 2^8 + d4 (d2, d3, d4 in
// [-128, 127], |d1| <= 127). A row's panel of 64 Q values is 256 bytes: its
// 64 d1, then the 64 d2, the 64 d3 and the 64 d4.
//
// Contraction. sum_Q Y_ia Y_jb = sum_L 2^(8 (8 - L)) S_L with the level sums
// S_L = sum_Q sum_{p + q = L} d_p d'_q. v_mfma_i32_16x16x64_i8 accumulates
// every S_L exactly in int32 (|S_L| <= 65280 naux_pad < 2^31). Levels 2..6
// are kept: 13 digit products per 64-deep panel, against 16 FP64 MFMAs of
// four times the cycles each for the same panel. Levels 7 and 8 are dropped;
// their weight is 2^-40 of level 2's and they have either sign. The epilogue
// forms g = 2^(e_ia + e_jb - 60) sum_L 2^(8 (8 - L)) S_L in FP64 and runs the
// same FP64 energy epilogue as the other kernels, so the only departures from
// the FP64 contraction are the rounding of Y (half a unit, 2^-32 of the row
// maximum's power of two) and the dropped levels.
//
// Work decomposition, single-product items and the tile-pair order are those
// of the other pair kernels (pair_body.h). The int32 MFMA result layout is
// that of the FP32-accumulating 16x16 instructions (row 4 g + v).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "fh_device.h"
#include "fh_kernels.h"
#include "pair_body.h"

namespace fh {

namespace {

constexpr int kLevels = 5;  // digit-product levels L = 2..6, index L - 2

// sum_L 2^(8 (8 - L)) S_L * 2^(e - 60), summed from the smallest level up.
__device__ __forceinline__ double combine_levels(int s2, int s3, int s4, int s5, int s6, int e) {
    if (e >= kI8NonFinite) return NAN;
    double t = (double)s6 * 0x1p-32;
    t += (double)s5 * 0x1p-24;
    t += (double)s4 * 0x1p-16;
    t += (double)s3 * 0x1p-8;
    t += (double)s2;
    return ldexp(t, e - 12);
}

template <int TILE>
__global__ __launch_bounds__(256) void energy_i8x4_kernel(EnergyArgs a) {
    constexpr int KP = 64;             // Q values per panel (one MFMA K)
    constexpr int TW = TILE / 2;
    constexpr int NB = TW / 16;
    constexpr int CPR = 16;            // 16-byte chunks per row and panel (4 digits x 64)
    constexpr int OPCH = TILE * CPR;
    constexpr int NCH = 4 * OPCH;
    constexpr int CPT = (NCH + 255) / 256;

    // Chunk c of row r is stored at slot c ^ (r & 15), so that the fragment
    // reads of 16 consecutive rows hit 16 distinct slots.
    __shared__ __attribute__((aligned(16))) v4i lds[2][4][TILE][CPR];

    pair::Item it;
    it.id = xcd_remap((int64_t)blockIdx.x, a.nitems);
    decode_pair(it.id / a.ntp, a.nocc, it.i, it.j);
    decode_tile_pair((int)(it.id % a.ntp), a.ntile, it.ta, it.tb);

    const int8_t *__restrict__ bp = (const int8_t *)a.op0;
    const int A0 = it.ta * TILE, B0 = it.tb * TILE;
    const pair::SplitQ sq(it, a.naux_pad, KP);
    // Operand rows as in pair_body.h: 0 = (i, A), 1 = (j, B), 2 = (j, A),
    // 3 = (i, B), or 0 and 1 one panel ahead for a single-product item.
    const int64_t base0 = (int64_t)it.i * a.nvir_pad + A0;
    const int64_t base1 = (int64_t)it.j * a.nvir_pad + B0;
    const int64_t base2 = sq.single ? base0 : (int64_t)it.j * a.nvir_pad + A0;
    const int64_t base3 = sq.single ? base1 : (int64_t)it.i * a.nvir_pad + B0;
    const int limA = a.nvir_pad - A0, limB = a.nvir_pad - B0;
    const int64_t ldr = (int64_t)a.naux_pad * 4;  // bytes per packed row
    const int tid = threadIdx.x;

    v4i stg[CPT];
    auto gload = [&](int q0) {
        const int q23 = q0 + sq.off23;
#pragma unroll
        for (int t = 0; t < CPT; ++t) {
            const int c = tid + 256 * t;
            const int op = c / OPCH, rem = c % OPCH;
            const int row = rem / CPR, ch = rem % CPR;
            const int64_t base = op == 0 ? base0 : op == 1 ? base1 : op == 2 ? base2 : base3;
            const int lim = (op & 1) ? limB : limA;
            const int q = op < 2 ? q0 : q23;
            v4i v = {0, 0, 0, 0};
            if (c < NCH && row < lim && q < a.naux_pad)
                v = *(const v4i *)(bp + (base + row) * ldr + (int64_t)q * 4 + ch * 16);
            stg[t] = v;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int t = 0; t < CPT; ++t) {
            const int c = tid + 256 * t;
            if (c < NCH) {
                const int op = c / OPCH, rem = c % OPCH;
                const int row = rem / CPR, ch = rem % CPR;
                lds[buf][op][row][ch ^ (row & 15)] = stg[t];
            }
        }
    };

    const int lane = tid % kWave, wid = tid / kWave;
    const pair::Lane L{A0, B0, wid / 2, wid % 2, lane % 16, lane / 16};
    bool rv[NB], cv[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        rv[k] = A0 + L.wr * TW + k * 16 < a.nvir_pad;
        cv[k] = B0 + L.wc * TW + k * 16 < a.nvir_pad;
    }
    v4i acc1[kLevels][NB][NB], acc2[kLevels][NB][NB];
#pragma unroll
    for (int l = 0; l < kLevels; ++l)
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y) {
                acc1[l][x][y] = (v4i){0, 0, 0, 0};
                acc2[l][x][y] = (v4i){0, 0, 0, 0};
            }

    // Lane (r, g) supplies the 16 bytes [16 g, 16 g + 16) of digit d's 64 for
    // A rows and B columns alike, so every MFMA contracts matching Q.
    auto frag = [&](int buf, int op, int row, int d) -> v4i {
        return lds[buf][op][row][(d * 4 + L.g) ^ (row & 15)];
    };
    auto compute = [&](int buf) {
        v4i f0[4][NB], f2[4][NB], f1[4][NB], f3[4][NB];
#pragma unroll
        for (int d = 0; d < 4; ++d)
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int ra = L.wr * TW + k * 16 + L.r, rb = L.wc * TW + k * 16 + L.r;
                f0[d][k] = frag(buf, 0, ra, d);
                f2[d][k] = frag(buf, 2, ra, d);
                f1[d][k] = frag(buf, 1, rb, d);
                f3[d][k] = frag(buf, 3, rb, d);
            }
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
                if (rv[x] && cv[y]) {
#pragma unroll
                    for (int p = 0; p < 4; ++p)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (p + q < kLevels) {
                                acc1[p + q][x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                    f0[p][x], f1[q][y], acc1[p + q][x][y], 0, 0, 0);
                                acc2[p + q][x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                    f2[p][x], f3[q][y], acc2[p + q][x][y], 0, 0, 0);
                            }
                }
    };

    gload(0);
    lstore(0);
    __syncthreads();
    for (int p = 0; p < sq.iters; ++p) {
        const int cur = p & 1;
        if (p + 1 < sq.iters) gload((p + 1) * sq.step);
        compute(cur);
        if (p + 1 < sq.iters) lstore(cur ^ 1);
        __syncthreads();
    }

    // A single-product item's two accumulator sets both hold C1 (even and odd
    // panels): merge them exactly in int32 before the conversion.
    if (sq.single) {
#pragma unroll
        for (int l = 0; l < kLevels; ++l)
#pragma unroll
            for (int x = 0; x < NB; ++x)
#pragma unroll
                for (int y = 0; y < NB; ++y) {
                    acc1[l][x][y] += acc2[l][x][y];
                    acc2[l][x][y] = (v4i){0, 0, 0, 0};
                }
    }

    // Level sums -> FP64 g with the row scales: C1(a, b) pairs rows (i, a)
    // and (j, b), C2(a, b) rows (j, a) and (i, b).
    const int64_t ri = (int64_t)it.i * a.nvir_pad, rj = (int64_t)it.j * a.nvir_pad;
    v4d g1[NB][NB], g2[NB][NB];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            const int bb = L.B0 + L.wc * TW + y * 16 + L.r;
            const int ej_b = cv[y] ? a.rexp[rj + bb] : 0;
            const int ei_b = cv[y] ? a.rexp[ri + bb] : 0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int aa = L.A0 + L.wr * TW + x * 16 + out_row<float>(L.g, v);
                const int ei_a = rv[x] ? a.rexp[ri + aa] : 0;
                const int ej_a = rv[x] ? a.rexp[rj + aa] : 0;
                g1[x][y][v] = combine_levels(acc1[0][x][y][v], acc1[1][x][y][v], acc1[2][x][y][v],
                                             acc1[3][x][y][v], acc1[4][x][y][v],
                                             ei_a + ej_b);
                g2[x][y][v] = sq.single
                                  ? 0.0
                                  : combine_levels(acc2[0][x][y][v], acc2[1][x][y][v],
                                                   acc2[2][x][y][v], acc2[3][x][y][v],
                                                   acc2[4][x][y][v], ej_a + ei_b);
            }
        }
    if (sq.single) {
        static_assert(sizeof(lds) >= sizeof(double) * TW * (TILE + 1), "LDS holds a C1 half-tile");
        pair::derive_c2<TILE, double>(it, L, g1, g2, pair::RowOf<float>{},
                                      (double *)&lds[0][0][0][0]);
    }
    pair::energy_epilogue<TILE>(a, it, L, g1, g2, rv, cv, pair::RowOf<float>{}, pair::NoCorr{});
}

}  // namespace

hipError_t launch_energy_i8x4(const EnergyArgs &a, hipStream_t st) {
    if (!a.rexp || a.naux_pad % 64 != 0 || a.naux_pad > kI8MaxNauxPad) return hipErrorInvalidValue;
    if (a.tile == 64)
        hipLaunchKernelGGL(energy_i8x4_kernel<64>, dim3((unsigned)a.nitems), dim3(256), 0, st, a);
    else if (a.tile == 32)
        hipLaunchKernelGGL(energy_i8x4_kernel<32>, dim3((unsigned)a.nitems), dim3(256), 0, st, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace fh
