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

#include <type_traits>

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
    const int64_t ldr = (int64_t)a.naux_pad * 4;  // bytes per packed row
    const int tid = threadIdx.x;

    // Global -> LDS staging. Thread t owns chunk ch = t % 16 of rows
    // r0 + 16 k (r0 = t / 16) of every operand, so its four row pointers are
    // loop-invariant and every load is pointer + uniform offset. Rows past
    // nvir_pad (a last tile narrower than TILE) are read from the next occupied
    // index or from the buffer's TILE spare rows: they feed only the 16-row
    // blocks that rv / cv switch off, so no row test is needed.
    constexpr int RPS = 256 / CPR;  // rows per pass of the workgroup (16)
    constexpr int SPO = TILE / RPS; // passes per operand
    static_assert(CPT == 4 * SPO && NCH == 256 * CPT, "one chunk per thread and pass");
    const int ch = tid % CPR, r0 = tid / CPR;
    const int8_t *src[4];
    src[0] = bp + (base0 + r0) * ldr + ch * 16;
    src[1] = bp + (base1 + r0) * ldr + ch * 16;
    src[2] = bp + (base2 + r0) * ldr + ch * 16;
    src[3] = bp + (base3 + r0) * ldr + ch * 16;
    v4i stg[CPT];
    auto gload = [&](int q0) {
        const int q23 = q0 + sq.off23;
        // uniform: a single-product item's extra panel past the end is zeros
        const bool hi_ok = q23 < a.naux_pad;
#pragma unroll
        for (int op = 0; op < 4; ++op)
#pragma unroll
            for (int k = 0; k < SPO; ++k) {
                const int64_t off = (int64_t)k * RPS * ldr + (int64_t)(op < 2 ? q0 : q23) * 4;
                v4i v = {0, 0, 0, 0};
                if (op < 2 || hi_ok) v = *(const v4i *)(src[op] + off);
                stg[op * SPO + k] = v;
            }
    };
    // (r0 + 16 k) & 15 == r0: one swizzled slot per thread.
    auto lstore = [&](int buf) {
#pragma unroll
        for (int op = 0; op < 4; ++op)
#pragma unroll
            for (int k = 0; k < SPO; ++k) lds[buf][op][r0 + RPS * k][ch ^ r0] = stg[op * SPO + k];
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
    // FULL: every 16x16 block of the wave lies inside nvir_pad (the common
    // case), so the MFMA sequence needs no per-block test.
    auto compute = [&](int buf, auto full) {
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
                if (decltype(full)::value || (rv[x] && cv[y])) {
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

    bool full = true;
#pragma unroll
    for (int k = 0; k < NB; ++k) full = full && rv[k] && cv[k];

    gload(0);
    lstore(0);
    __syncthreads();
    for (int p = 0; p < sq.iters; ++p) {
        const int cur = p & 1;
        if (p + 1 < sq.iters) gload((p + 1) * sq.step);
        if (full)
            compute(cur, std::true_type{});
        else
            compute(cur, std::false_type{});
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

// 64-wide tile on 8 waves: waves 0-3 accumulate C1 (operands 0 and 1), waves
// 4-7 accumulate C2 (operands 2 and 3), each wave one 32x32 quadrant. Half the
// accumulators per wave of the 4-wave kernel let two waves share each SIMD, so
// one wave's LDS fragment reads overlap the other's MFMAs, and the registers
// left over hold a second staged panel: global loads run two panels ahead of
// the MFMAs. For a single-product item waves 4-7 take the odd panels of C1.
// The halves meet through LDS after the loop (exact int32 merge of the level
// sums for a single-product item; FP64 C2 for a two-product item), and waves
// 0-3 run the energy epilogue.
__global__ __launch_bounds__(512) void energy_i8x4_wide_kernel(EnergyArgs a) {
    constexpr int TILE = 64, KP = 64;
    constexpr int TW = 32, NB = 2;
    constexpr int CPR = 16;             // 16-byte chunks per row and panel
    constexpr int RPS = 512 / CPR;      // rows per pass of the workgroup (32)
    constexpr int SPO = TILE / RPS;     // passes per operand (2)
    constexpr int CPT = 4 * SPO;        // staged chunks per thread and panel (8)

    __shared__ __attribute__((aligned(16))) v4i lds[2][4][TILE][CPR];

    pair::Item it;
    it.id = xcd_remap((int64_t)blockIdx.x, a.nitems);
    decode_pair(it.id / a.ntp, a.nocc, it.i, it.j);
    decode_tile_pair((int)(it.id % a.ntp), a.ntile, it.ta, it.tb);

    const int8_t *__restrict__ bp = (const int8_t *)a.op0;
    const int A0 = it.ta * TILE, B0 = it.tb * TILE;
    const pair::SplitQ sq(it, a.naux_pad, KP);
    const int64_t base0 = (int64_t)it.i * a.nvir_pad + A0;
    const int64_t base1 = (int64_t)it.j * a.nvir_pad + B0;
    const int64_t base2 = sq.single ? base0 : (int64_t)it.j * a.nvir_pad + A0;
    const int64_t base3 = sq.single ? base1 : (int64_t)it.i * a.nvir_pad + B0;
    const int64_t ldr = (int64_t)a.naux_pad * 4;
    const int tid = threadIdx.x;

    // Staging as in the 4-wave kernel (rows past nvir_pad only feed switched-off
    // blocks; the buffer has TILE spare rows): thread t owns chunk t % 16 of
    // rows t / 16 + 32 k of every operand.
    const int ch = tid % CPR, r0 = tid / CPR;
    const int8_t *src[4];
    src[0] = bp + (base0 + r0) * ldr + ch * 16;
    src[1] = bp + (base1 + r0) * ldr + ch * 16;
    src[2] = bp + (base2 + r0) * ldr + ch * 16;
    src[3] = bp + (base3 + r0) * ldr + ch * 16;
    v4i stg[2][CPT];
    auto gload = [&](int p, v4i(&s)[CPT]) {
        const int q0 = p * sq.step, q23 = q0 + sq.off23;
        const bool hi_ok = q23 < a.naux_pad;  // uniform
#pragma unroll
        for (int op = 0; op < 4; ++op)
#pragma unroll
            for (int k = 0; k < SPO; ++k) {
                const int64_t off = (int64_t)k * RPS * ldr + (int64_t)(op < 2 ? q0 : q23) * 4;
                v4i v = {0, 0, 0, 0};
                if (op < 2 || hi_ok) v = *(const v4i *)(src[op] + off);
                s[op * SPO + k] = v;
            }
    };
    auto lstore = [&](int buf, const v4i(&s)[CPT]) {
#pragma unroll
        for (int op = 0; op < 4; ++op)
#pragma unroll
            for (int k = 0; k < SPO; ++k)
                lds[buf][op][r0 + RPS * k][ch ^ (r0 & 15)] = s[op * SPO + k];
    };

    const int lane = tid % kWave, wid = tid / kWave;
    const int prod = wid >> 2, quad = wid & 3;
    const pair::Lane L{A0, B0, quad >> 1, quad & 1, lane % 16, lane / 16};
    const int opA = 2 * prod, opB = 2 * prod + 1;
    bool rv[NB], cv[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        rv[k] = A0 + L.wr * TW + k * 16 < a.nvir_pad;
        cv[k] = B0 + L.wc * TW + k * 16 < a.nvir_pad;
    }
    v4i acc[kLevels][NB][NB];
#pragma unroll
    for (int l = 0; l < kLevels; ++l)
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y) acc[l][x][y] = (v4i){0, 0, 0, 0};

    auto compute = [&](int buf, auto full) {
        v4i fa[4][NB], fb[4][NB];
#pragma unroll
        for (int d = 0; d < 4; ++d)
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                const int ra = L.wr * TW + k * 16 + L.r, rb = L.wc * TW + k * 16 + L.r;
                fa[d][k] = lds[buf][opA][ra][(d * 4 + L.g) ^ L.r];
                fb[d][k] = lds[buf][opB][rb][(d * 4 + L.g) ^ L.r];
            }
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
                if (decltype(full)::value || (rv[x] && cv[y])) {
#pragma unroll
                    for (int p = 0; p < 4; ++p)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (p + q < kLevels)
                                acc[p + q][x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                    fa[p][x], fb[q][y], acc[p + q][x][y], 0, 0, 0);
                }
    };
    bool full = true;
#pragma unroll
    for (int k = 0; k < NB; ++k) full = full && rv[k] && cv[k];
    auto panel = [&](int buf) {
        if (full)
            compute(buf, std::true_type{});
        else
            compute(buf, std::false_type{});
    };

    // Two panels in flight: at iteration p, LDS buffer p & 1 holds panel p,
    // stg[(p + 1) & 1] panel p + 1 (landing), and panel p + 2 is fetched into
    // stg[p & 1]. The loop is unrolled by two so the staging index is static.
    const int iters = sq.iters;
    gload(0, stg[0]);
    if (iters > 1) gload(1, stg[1]);
    lstore(0, stg[0]);
    __syncthreads();
    for (int p = 0; p < iters; p += 2) {
        if (p + 2 < iters) gload(p + 2, stg[0]);
        panel(0);
        if (p + 1 < iters) lstore(1, stg[1]);
        __syncthreads();
        if (p + 1 >= iters) break;
        if (p + 3 < iters) gload(p + 3, stg[1]);
        panel(1);
        if (p + 2 < iters) lstore(0, stg[0]);
        __syncthreads();
    }

    // The two halves meet through LDS (the loop's last barrier has passed).
    const int64_t ri = (int64_t)it.i * a.nvir_pad, rj = (int64_t)it.j * a.nvir_pad;
    constexpr int LDG = TILE + 1;
    double *gx = (double *)&lds[0][0][0][0];
    int *lx = (int *)&lds[0][0][0][0];
    if (sq.single) {
        // exact merge: waves 4-7 hand their int32 level sums to waves 0-3
        if (prod == 1) {
#pragma unroll
            for (int l = 0; l < kLevels; ++l)
#pragma unroll
                for (int x = 0; x < NB; ++x)
#pragma unroll
                    for (int y = 0; y < NB; ++y)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int ra = L.wr * TW + x * 16 + out_row<float>(L.g, v);
                            const int cb = L.wc * TW + y * 16 + L.r;
                            lx[(l * TILE + ra) * TILE + cb] = acc[l][x][y][v];
                        }
        }
        __syncthreads();
        if (prod == 0) {
#pragma unroll
            for (int l = 0; l < kLevels; ++l)
#pragma unroll
                for (int x = 0; x < NB; ++x)
#pragma unroll
                    for (int y = 0; y < NB; ++y)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int ra = L.wr * TW + x * 16 + out_row<float>(L.g, v);
                            const int cb = L.wc * TW + y * 16 + L.r;
                            acc[l][x][y][v] += lx[(l * TILE + ra) * TILE + cb];
                        }
        }
        __syncthreads();
    }
    // Level sums -> FP64 with the row scales of this wave's operand rows:
    // C1 pairs (i, a) with (j, b), C2 pairs (j, a) with (i, b); a single-product
    // item's merged C1 is in waves 0-3.
    const int64_t rA = prod == 0 ? ri : rj, rB = prod == 0 ? rj : ri;
    v4d g1[NB][NB], g2[NB][NB];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            const int bb = B0 + L.wc * TW + y * 16 + L.r;
            const int eb = cv[y] ? a.rexp[rB + bb] : 0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int aa = A0 + L.wr * TW + x * 16 + out_row<float>(L.g, v);
                const int ea = rv[x] ? a.rexp[rA + aa] : 0;
                g1[x][y][v] = combine_levels(acc[0][x][y][v], acc[1][x][y][v], acc[2][x][y][v],
                                             acc[3][x][y][v], acc[4][x][y][v], ea + eb);
                g2[x][y][v] = 0.0;
            }
        }
    // C2 into waves 0-3: from waves 4-7 (two-product item), or C1 itself
    // (i == j), or C1 transposed (diagonal tile pair).
    const bool copy = sq.single && it.i == it.j;
    if (!copy) {
        if (prod == (sq.single ? 0 : 1)) {
#pragma unroll
            for (int x = 0; x < NB; ++x)
#pragma unroll
                for (int y = 0; y < NB; ++y)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int ra = L.wr * TW + x * 16 + out_row<float>(L.g, v);
                        const int cb = L.wc * TW + y * 16 + L.r;
                        gx[ra * LDG + cb] = g1[x][y][v];
                    }
        }
        __syncthreads();
        if (prod == 0) {
#pragma unroll
            for (int x = 0; x < NB; ++x)
#pragma unroll
                for (int y = 0; y < NB; ++y)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int ra = L.wr * TW + x * 16 + out_row<float>(L.g, v);
                        const int cb = L.wc * TW + y * 16 + L.r;
                        g2[x][y][v] = sq.single ? gx[cb * LDG + ra] : gx[ra * LDG + cb];
                    }
        }
    } else if (prod == 0) {
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y) g2[x][y] = g1[x][y];
    }
    // Waves 4-7 take part in the epilogue's block reduction with nothing to add.
    bool rve[NB], cve[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        rve[k] = prod == 0 && rv[k];
        cve[k] = prod == 0 && cv[k];
    }
    pair::energy_epilogue<TILE>(a, it, L, g1, g2, rve, cve, pair::RowOf<float>{},
                                pair::NoCorr{});
}

}  // namespace

hipError_t launch_energy_i8x4(const EnergyArgs &a, hipStream_t st) {
    if (!a.rexp || a.naux_pad % 64 != 0 || a.naux_pad > kI8MaxNauxPad) return hipErrorInvalidValue;
    if (a.tile == 64)
        hipLaunchKernelGGL(energy_i8x4_wide_kernel, dim3((unsigned)a.nitems), dim3(512), 0, st, a);
    else if (a.tile == 32)
        hipLaunchKernelGGL(energy_i8x4_kernel<32>, dim3((unsigned)a.nitems), dim3(256), 0, st, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace fh
