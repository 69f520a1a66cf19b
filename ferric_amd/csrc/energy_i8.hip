// INT8 digit-sliced variant of the fused RI-MP2 pair-energy kernel (gfx950):
// the FP64-accurate reduced-precision mode.
//
// Operand format (pack.hip, PackFmt::I8X4). Every packed row r = (i, a) has
// one exponent e_r with max_Q |B^Q_ia| <= (1 - 2^-7) 2^(e_r + 1), and every
// element is the integer Y = RN(B 2^(30 - e_r)), |Y| <= 2^31 - 2^24, written as
// four signed byte digits Y = d1 2^24 + d2 2^16 + d3 2^8 + d4 (d2, d3, d4 in
// [-128, 127], |d1| <= 127). The digits are stored in MFMA fragment order:
// for each group of 16 packed rows and each 64-deep Q panel one 4 KB block,
// digit by digit, and within a digit's kilobyte the 16 bytes of lane
// l = 16 g + r (row r of the group, Q = 64 P + 16 g .. 16 g + 15) at l * 16.
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
    if (e >= kI8NonFiniteSum) return NAN;  // either row non-finite
    double t = (double)s6 * 0x1p-32;
    t += (double)s5 * 0x1p-24;
    t += (double)s4 * 0x1p-16;
    t += (double)s3 * 0x1p-8;
    t += (double)s2;
    return ldexp(t, e - 12);
}

// One work item on 8 waves (512 threads): waves 0-3 accumulate C1 (operand
// rows (i, A) and (j, B)), waves 4-7 C2 ((j, A) and (i, B)), each wave one
// TILE/2-square quadrant; for a single-product item waves 4-7 take the odd
// panels of C1 instead. Every fragment (one digit of 16 rows x 64 Q) is one
// coalesced 1 KB wave load straight into registers, so the main loop has no
// LDS traffic and no barrier: each wave streams its own fragments one panel
// ahead (double-buffered registers) while two waves share each SIMD. Blocks
// of a last tile narrower than TILE read the next rows or the buffer's TILE
// spare rows; they feed only the 16-row blocks that rv / cv switch off. The
// halves meet through LDS after the loop (an exact int32 merge of the level
// sums for a single-product item; FP64 C2 for a two-product item), and waves
// 0-3 run the energy epilogue from LDS-staged orbital energies and row
// exponents. With a.q_rot_ticks > 0 an item starts at the Q panel the device
// wall clock names (one panel per q_rot_ticks ticks, wrapping), so the items an
// XCD runs at one moment read the same panel of the occupied rows they share
// and the XCD's L2 holds it once; the int32 level sums are exact, so this
// changes no bit. ADJ: phase A of the adjoint over every item at once, ending in
// the T~ epilogue of pair_body.h (writing the whole-T~ matrix) instead of the
// energy terms.
template <int TILE, bool ADJ>
__global__ __launch_bounds__(512) void energy_i8x4_kernel(EnergyArgs a) {
    constexpr int TW = TILE / 2, NB = TW / 16;
    constexpr int FB = 4096;  // one (16-row group, 64-deep panel) block: 4 digits x 1 KB

    __shared__ __attribute__((aligned(16))) int xbuf[kLevels * TILE * TILE];
    // Epilogue inputs, fetched at the start: the virtual energies of tiles A
    // and B and the row exponents of the four operand row sets.
    __shared__ double eps_t[2][TILE];
    __shared__ int rexp_t[4][TILE];
    __shared__ int qrot;  // first Q panel of this item (one value for all waves)

    pair::Item it;
    it.id = a.item0 + pair_interleave((int64_t)blockIdx.x, a.ntp, a.nitems / a.ntp);
    decode_pair(it.id / a.ntp, a.nocc, it.i, it.j);
    if (ADJ)  // the adjoint's row reduction reads each item's sums in decode_tile_pair's order
        decode_tile_pair((int)(it.id % a.ntp), a.ntile, it.ta, it.tb);
    else
        decode_tile_pair_banded((int)(it.id % a.ntp), a.ntile, it.ta, it.tb);

    const int8_t *__restrict__ bp = (const int8_t *)a.op0;
    const int A0 = it.ta * TILE, B0 = it.tb * TILE;
    const pair::SplitQ sq(it, a.naux_pad, 64);
    const int np = a.naux_pad / 64, ng = a.nvir_pad / 16;
    const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
    const int prod = wid >> 2, quad = wid & 3;
    const pair::Lane L{A0, B0, quad >> 1, quad & 1, lane % 16, lane / 16};

    // Operand rows of this wave: C1 (i, A) x (j, B), C2 (j, A) x (i, B); a
    // single-product item's waves 4-7 read C1's rows at the odd panels.
    const bool c2 = prod == 1 && !sq.single;
    const int occA = c2 ? it.j : it.i, occB = c2 ? it.i : it.j;
    const int pstep = sq.single ? 2 : 1, poff = (prod == 1 && sq.single) ? 1 : 0;
    const int8_t *pa[NB], *pb[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        pa[k] = bp + (int64_t)(occA * ng + (A0 + L.wr * TW) / 16 + k) * np * FB + lane * 16;
        pb[k] = bp + (int64_t)(occB * ng + (B0 + L.wc * TW) / 16 + k) * np * FB + lane * 16;
    }

    if (tid < 2 * TILE) {
        const int c = tid % TILE, v = (tid < TILE ? A0 : B0) + c;
        eps_t[tid / TILE][c] = v < a.nvir ? a.eps_vir[v] : 0.0;
    } else if (tid < 6 * TILE) {
        const int t = tid - 2 * TILE, op = t / TILE, c = t % TILE;
        // operand row sets 0 (i, A), 1 (j, B), 2 (j, A), 3 (i, B); a
        // single-product item's 2 and 3 are 0 and 1
        const bool swap = op >= 2 && !sq.single;
        const int occ = ((op & 1) ^ swap) ? it.j : it.i;
        const int v = ((op & 1) ? B0 : A0) + c;
        rexp_t[op][c] = v < a.nvir_pad ? a.rexp[(int64_t)occ * a.nvir_pad + v] : 0;
    }

    bool rv[NB], cv[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        rv[k] = A0 + L.wr * TW + k * 16 < a.nvir_pad;
        cv[k] = B0 + L.wc * TW + k * 16 < a.nvir_pad;
    }
    bool full = true, anyr = false, anyc = false;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        full = full && rv[k] && cv[k];
        anyr = anyr || rv[k];
        anyc = anyc || cv[k];
    }
    v4i acc[kLevels][NB][NB];
#pragma unroll
    for (int l = 0; l < kLevels; ++l)
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y) acc[l][x][y] = (v4i){0, 0, 0, 0};

    // The waves of a single-product item split the panels by parity of their
    // position after the rotation, so all of them must rotate by the same amount.
    if (tid == 0)
        qrot = a.q_rot_ticks > 0 ? (int)((wall_clock64() / (uint64_t)a.q_rot_ticks) % np) : 0;
    __syncthreads();
    const int rot = qrot;

    v4i fa[2][4][NB], fb[2][4][NB];
    // FULL: every block is loaded. Otherwise only the blocks inside nvir_pad,
    // the only ones the MFMAs below read (a 16-row block is used when both its
    // row block and its column block are inside).
    auto fetch = [&](int P, v4i(&ra)[4][NB], v4i(&rb)[4][NB], auto fullc) {
        const int R = P + rot;  // P < np and rot < np
        const int64_t off = (int64_t)(R >= np ? R - np : R) * FB;
#pragma unroll
        for (int d = 0; d < 4; ++d)
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                if (decltype(fullc)::value || rv[k])
                    ra[d][k] = *(const v4i *)(pa[k] + off + d * 1024);
                if (decltype(fullc)::value || cv[k])
                    rb[d][k] = *(const v4i *)(pb[k] + off + d * 1024);
            }
    };
    // FULL: every 16x16 block of the wave lies inside nvir_pad (the common
    // case), so the MFMA sequence needs no per-block test.
    auto mma = [&](const v4i(&ra)[4][NB], const v4i(&rb)[4][NB], auto fullc) {
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
                if (decltype(fullc)::value || (rv[x] && cv[y])) {
#pragma unroll
                    for (int p = 0; p < 4; ++p)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (p + q < kLevels)
                                acc[p + q][x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                    ra[p][x], rb[q][y], acc[p + q][x][y], 0, 0, 0);
                }
    };
    // Iteration p covers panel p * pstep + poff (a single-product item's odd
    // panel past the end is skipped by its waves); unrolled by two so the
    // register buffer of each step is static. The guarded fetches make the
    // compiler wait for every outstanding load before a panel's MFMAs, the
    // next panel's included, so the two waves of a SIMD take turns; fetching
    // unconditionally (the next panel's loads in flight during the MFMAs) was
    // measured 5 % slower at 100/900/3600.
    const int iters = sq.iters;
    auto valid = [&](int p) { return p < iters && p * pstep + poff < np; };
    auto run = [&](auto fullc) {
        if (valid(0)) fetch(poff, fa[0], fb[0], fullc);
        for (int p = 0; p < iters; p += 2) {
            if (valid(p + 1)) fetch((p + 1) * pstep + poff, fa[1], fb[1], fullc);
            if (valid(p)) mma(fa[0], fb[0], fullc);
            if (valid(p + 2)) fetch((p + 2) * pstep + poff, fa[0], fb[0], fullc);
            if (valid(p + 1)) mma(fa[1], fb[1], fullc);
        }
    };
    // A wave of a last tile whose quadrant lies wholly past nvir_pad (no row
    // block or no column block inside) has nothing to contract: its level sums
    // stay zero.
    if (full)
        run(std::true_type{});
    else if (anyr && anyc)
        run(std::false_type{});
    __syncthreads();  // eps_t / rexp_t written; every wave done with its loop

    constexpr int LDG = TILE + 1;
    double *gx = (double *)xbuf;
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
                            xbuf[(l * TILE + ra) * TILE + cb] = acc[l][x][y][v];
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
                            acc[l][x][y][v] += xbuf[(l * TILE + ra) * TILE + cb];
                        }
        }
        __syncthreads();
    }
    // Level sums -> FP64 with the row scales of this wave's operand rows.
    const int eA = c2 ? 2 : 0, eB = c2 ? 3 : 1;
    v4d g1[NB][NB], g2[NB][NB];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            const int eb = rexp_t[eB][L.wc * TW + y * 16 + L.r];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int ea = rexp_t[eA][L.wr * TW + x * 16 + out_row<float>(L.g, v)];
                g1[x][y][v] = combine_levels(acc[0][x][y][v], acc[1][x][y][v], acc[2][x][y][v],
                                             acc[3][x][y][v], acc[4][x][y][v], ea + eb);
                g2[x][y][v] = 0.0;
            }
        }
    // C2 into waves 0-3: from waves 4-7 (two-product item), C1 itself
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
    if constexpr (ADJ) {
        // xbuf is free again: every read of it above precedes this point for
        // all waves (the single-product merge ends in a barrier; the C2
        // handover reads gx before the barrier inside the epilogue, after
        // which tlds is read).
        static_assert(sizeof(xbuf) >= 2 * TILE * (TILE + 1) * sizeof(uint32_t),
                      "xbuf holds the two transposed T~ tiles");
        __syncthreads();
        pair::adjoint_epilogue<TILE, v4d, pair::RowOf<float>, pair::NoCorr, true>(
            a, it, L, g1, g2, rv, cv, pair::RowOf<float>{}, pair::NoCorr{}, prod == 0,
            (uint32_t *)xbuf);
        return;
    }
    // The energy terms (pair::energy_epilogue's, with the virtual energies
    // from LDS); waves 4-7 join the block reduction with nothing to add.
    const double eij = a.eps_occ[it.i] + a.eps_occ[it.j];
    const bool kap = !isnan(a.kappa);
    KSum os, ss;
    if (prod == 0) {
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y) {
                if (!(rv[x] && cv[y])) continue;  // wave-uniform
                const int cb = L.wc * TW + y * 16 + L.r;
                const bool okb = B0 + cb < a.nvir;
                const double eb = eps_t[1][cb];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    // Branch-free: a virtual past nvir adds an exact 0.0, which
                    // leaves both compensated sums bit for bit as they were
                    // (eps_t holds 0.0 there, so d stays eij: finite, nonzero).
                    const int ra = L.wr * TW + x * 16 + out_row<float>(L.g, v);
                    const bool ok = okb && A0 + ra < a.nvir;
                    const double d = eij - eps_t[0][ra] - eb;
                    const double g1v = g1[x][y][v], g2v = g2[x][y][v];
                    const double inv = 1.0 / d;
                    const double dd = g1v - g2v;
                    const double dmp = pair::damping(kap, a.kappa, d);
                    const double to = dmp * ((g1v * g1v + g2v * g2v) * inv);
                    const double ts = dmp * ((dd * dd) * inv);
                    os.add(ok ? to : 0.0);
                    ss.add(ok ? ts : 0.0);
                }
            }
    }
    const double w = (it.i == it.j ? 1.0 : 2.0) * (it.ta == it.tb ? 0.5 : 1.0);
    pair::block_store(os, ss, w, a.item_partials + 4 * it.id);
}

// Opposite-spin rectangle (u_rimp2.rs:731-760) on the INT8 digit slices: item
// (i, J, tA, tB) over every alpha i, beta J and tile of each side, one product
// C = B_alpha(i, A)^T B_beta(J, B). Waves 0-3 take the even Q panels, waves 4-7
// the odd ones; their int32 level sums merge exactly through LDS, and waves 0-3
// add sum (ia|JB)^2 / D over the tile with weight 1. Operand format, fetch,
// wall-clock Q-panel rotation and level combination as energy_i8x4_kernel;
// op0 / rexp hold the alpha rows (nvir_pad), op1 / rexp_b the beta rows
// (nvir_pad_b), each buffer with TILE spare rows after its last.
template <int TILE>
__global__ __launch_bounds__(512) void energy_i8x4_rect_kernel(EnergyArgs a) {
    constexpr int TW = TILE / 2, NB = TW / 16;
    constexpr int FB = 4096;
    __shared__ __attribute__((aligned(16))) int xbuf[kLevels * TILE * TILE];
    __shared__ int rexp_t[2][TILE];
    __shared__ int qrot;  // first Q panel (the wall-clock rotation of energy_i8x4_kernel)

    pair::Item it;
    it.id = xcd_remap((int64_t)blockIdx.x, a.nitems);
    const int64_t ntp = (int64_t)a.ntile * a.ntile_b;
    const int64_t pij = it.id / ntp;
    const int t = (int)(it.id % ntp);
    it.i = (int)(pij / a.nocc_b);
    it.j = (int)(pij % a.nocc_b);
    it.ta = t / a.ntile_b;
    it.tb = t % a.ntile_b;

    const int A0 = it.ta * TILE, B0 = it.tb * TILE;
    const int np = a.naux_pad / 64, nga = a.nvir_pad / 16, ngb = a.nvir_pad_b / 16;
    const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
    const int prod = wid >> 2, quad = wid & 3;
    const pair::Lane L{A0, B0, quad >> 1, quad & 1, lane % 16, lane / 16};
    const int8_t *pa[NB], *pb[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        pa[k] = (const int8_t *)a.op0 +
                (int64_t)(it.i * nga + (A0 + L.wr * TW) / 16 + k) * np * FB + lane * 16;
        pb[k] = (const int8_t *)a.op1 +
                (int64_t)(it.j * ngb + (B0 + L.wc * TW) / 16 + k) * np * FB + lane * 16;
    }
    if (tid < TILE) {
        const int v = A0 + tid;
        rexp_t[0][tid] = v < a.nvir_pad ? a.rexp[(int64_t)it.i * a.nvir_pad + v] : 0;
    } else if (tid < 2 * TILE) {
        const int c = tid - TILE, v = B0 + c;
        rexp_t[1][c] = v < a.nvir_pad_b ? a.rexp_b[(int64_t)it.j * a.nvir_pad_b + v] : 0;
    }

    bool rv[NB], cv[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        rv[k] = A0 + L.wr * TW + k * 16 < a.nvir_pad;
        cv[k] = B0 + L.wc * TW + k * 16 < a.nvir_pad_b;
    }
    bool full = true;
#pragma unroll
    for (int k = 0; k < NB; ++k) full = full && rv[k] && cv[k];
    v4i acc[kLevels][NB][NB];
#pragma unroll
    for (int l = 0; l < kLevels; ++l)
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y) acc[l][x][y] = (v4i){0, 0, 0, 0};

    // The even / odd panel split of the two halves needs one rotation for all waves.
    if (tid == 0)
        qrot = a.q_rot_ticks > 0 ? (int)((wall_clock64() / (uint64_t)a.q_rot_ticks) % np) : 0;
    __syncthreads();
    const int rot = qrot;

    v4i fa[2][4][NB], fb[2][4][NB];
    auto fetch = [&](int P, v4i(&ra)[4][NB], v4i(&rb)[4][NB]) {
        const int R = P + rot;  // P < np and rot < np
        const int64_t off = (int64_t)(R >= np ? R - np : R) * FB;
#pragma unroll
        for (int d = 0; d < 4; ++d)
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                ra[d][k] = *(const v4i *)(pa[k] + off + d * 1024);
                rb[d][k] = *(const v4i *)(pb[k] + off + d * 1024);
            }
    };
    auto mma = [&](const v4i(&ra)[4][NB], const v4i(&rb)[4][NB], auto fullc) {
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
                if (decltype(fullc)::value || (rv[x] && cv[y])) {
#pragma unroll
                    for (int p = 0; p < 4; ++p)
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (p + q < kLevels)
                                acc[p + q][x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(
                                    ra[p][x], rb[q][y], acc[p + q][x][y], 0, 0, 0);
                }
    };
    auto compute = [&](const v4i(&ra)[4][NB], const v4i(&rb)[4][NB]) {
        if (full)
            mma(ra, rb, std::true_type{});
        else
            mma(ra, rb, std::false_type{});
    };
    // Iteration p covers panel 2 p + prod.
    const int iters = (np + 1) / 2;
    auto valid = [&](int p) { return p < iters && 2 * p + prod < np; };
    if (valid(0)) fetch(prod, fa[0], fb[0]);
    for (int p = 0; p < iters; p += 2) {
        if (valid(p + 1)) fetch(2 * (p + 1) + prod, fa[1], fb[1]);
        if (valid(p)) compute(fa[0], fb[0]);
        if (valid(p + 2)) fetch(2 * (p + 2) + prod, fa[0], fb[0]);
        if (valid(p + 1)) compute(fa[1], fb[1]);
    }
    __syncthreads();  // rexp_t written; every wave done with its loop
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
                        xbuf[(l * TILE + ra) * TILE + cb] = acc[l][x][y][v];
                    }
    }
    __syncthreads();
    const double eij = a.eps_occ[it.i] + a.eps_occ_b[it.j];
    KSum os, ss;
    if (prod == 0) {
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y) {
                if (!(rv[x] && cv[y])) continue;
                const int cb = L.wc * TW + y * 16 + L.r;
                if (B0 + cb >= a.nvir_b) continue;
                const double eb = a.eps_vir_b[B0 + cb];
                const int xb = rexp_t[1][cb];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int ra = L.wr * TW + x * 16 + out_row<float>(L.g, v);
                    if (A0 + ra >= a.nvir) continue;
                    int sl[kLevels];
#pragma unroll
                    for (int l = 0; l < kLevels; ++l)
                        sl[l] = acc[l][x][y][v] + xbuf[(l * TILE + ra) * TILE + cb];
                    const double g = combine_levels(sl[0], sl[1], sl[2], sl[3], sl[4],
                                                    rexp_t[0][ra] + xb);
                    os.add((g * g) / (eij - a.eps_vir[A0 + ra] - eb));
                }
            }
    }
    pair::block_store(os, ss, 1.0, a.item_partials + 4 * it.id);
}

}  // namespace

hipError_t launch_energy_i8x4_rect(const EnergyArgs &a, hipStream_t st) {
    if (!a.rexp || !a.rexp_b || a.naux_pad % 64 != 0 || a.naux_pad > kI8MaxNauxPad ||
        a.nvir_pad % 16 != 0 || a.nvir_pad_b % 16 != 0)
        return hipErrorInvalidValue;
    if (a.tile == 64)
        hipLaunchKernelGGL((energy_i8x4_rect_kernel<64>), dim3((unsigned)a.nitems), dim3(512), 0,
                           st, a);
    else if (a.tile == 32)
        hipLaunchKernelGGL((energy_i8x4_rect_kernel<32>), dim3((unsigned)a.nitems), dim3(512), 0,
                           st, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_energy_i8x4(const EnergyArgs &a, hipStream_t st) {
    if (!a.rexp || a.naux_pad % 64 != 0 || a.naux_pad > kI8MaxNauxPad || a.nvir_pad % 16 != 0)
        return hipErrorInvalidValue;
    if (a.tile == 64)
        hipLaunchKernelGGL((energy_i8x4_kernel<64, false>), dim3((unsigned)a.nitems), dim3(512), 0,
                           st, a);
    else if (a.tile == 32)
        hipLaunchKernelGGL((energy_i8x4_kernel<32, false>), dim3((unsigned)a.nitems), dim3(512), 0,
                           st, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_adjoint_i8x4(const EnergyArgs &a, hipStream_t st) {
    if (!a.rexp || a.naux_pad % 64 != 0 || a.naux_pad > kI8MaxNauxPad || a.nvir_pad % 16 != 0 ||
        a.i_fixed >= 0 || !a.s_hi || !a.s_lo || a.vpg < a.nvir_pad || a.g0 < 0 ||
        a.g1 <= a.g0 || a.g1 > a.nocc || a.ld_s != (int64_t)a.nocc * a.vpg ||
        a.ld_st != (int64_t)(a.g1 - a.g0) * a.vpg || (a.g1 < a.nocc && (!a.st_hi || !a.st_lo)) ||
        a.nitems % a.ntp != 0)
        return hipErrorInvalidValue;
    if (a.tile == 64)
        hipLaunchKernelGGL((energy_i8x4_kernel<64, true>), dim3((unsigned)a.nitems), dim3(512), 0,
                           st, a);
    else if (a.tile == 32)
        hipLaunchKernelGGL((energy_i8x4_kernel<32, true>), dim3((unsigned)a.nitems), dim3(512), 0,
                           st, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace fh
