// Device bodies of the fused RI-MP2 pair kernels, shared by the one-problem
// launches (energy.hip) and the persistent batched launch (batched.hip).
//
// One call computes one work item = (occupied pair i <= j, virtual tile pair
// tA <= tB): over the whole Q range it accumulates
//   C1(a, b) = sum_Q B_i[Q, a] B_j[Q, b] = (ia|jb)       a in A, b in B
//   C2(a, b) = sum_Q B_j[Q, a] B_i[Q, b] = (ib|ja)
// so g_ab and g_ba of every (a, b) sit in the same lane and register, then
// runs the FP64 epilogue: the energy terms (both orientations of the
// reference's pair loop, crates/ferric-mp2/src/rimp2.rs:1399-1441) or, for the
// adjoint, T~ and the t*T~ sums. Every argument the body reads comes from the
// EnergyArgs it is given (operand pointers, shapes, eps, output slots).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <type_traits>

#include "fh_device.h"
#include "fh_kernels.h"

namespace fh {
namespace pair {

template <typename T>
struct Mfma4;
template <>
struct Mfma4<double> {
    using acc = v4d;
    __device__ __forceinline__ static acc mma(double x, double y, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
    }
};
template <>
struct Mfma4<float> {
    using acc = v4f;
    __device__ __forceinline__ static acc mma(float x, float y, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
    }
};

struct Item {
    int i, j, ta, tb;
    int64_t id;
};

// Weighted, compensated workgroup sum of (os, ss) -> 4 doubles at out4.
__device__ __forceinline__ void block_store(KSum os, KSum ss, double w, double *out4) {
    os.s *= w;
    os.c *= w;
    ss.s *= w;
    ss.c *= w;
    wave_merge(os);
    wave_merge(ss);
    __shared__ double red[16][4];
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
        const int nw = blockDim.x / kWave;
        for (int k = 0; k < nw; ++k) {
            o.merge(red[k][0], red[k][1]);
            s.merge(red[k][2], red[k][3]);
        }
        out4[0] = o.s;
        out4[1] = o.c;
        out4[2] = s.s;
        out4[3] = s.c;
    }
}

__device__ __forceinline__ double damping(bool kap, double kappa, double d) {
    if (!kap) return 1.0;
    const double d1 = 1.0 - exp(kappa * d);
    return d1 * d1;
}

struct NoCorr {
    __device__ double operator()(int, int) const { return 0.0; }
};

// The lane's coordinates inside the tile pair.
struct Lane {
    int A0, B0, wr, wc, r, g;
};

// Energy epilogue: acc1/acc2 hold g_ab / g_ba; row_of maps (g, v) to the row
// in a 16x16 block; corr(i, a) is added to K_ii[a, a] (0 except for BF16x3).
// The lane's virtual energies are read ahead of the arithmetic that uses
// them, a row block at a time (indices clamped into range; elements past nvir
// are skipped below), so their latencies overlap instead of following one
// another through the element loop.
template <int TILE, typename Acc, typename RowOfT, typename Corr>
__device__ __forceinline__ void energy_epilogue(const EnergyArgs &a, const Item &it, const Lane &L,
                                                Acc (&acc1)[TILE / 32][TILE / 32],
                                                Acc (&acc2)[TILE / 32][TILE / 32],
                                                const bool *rv, const bool *cv, RowOfT row_of,
                                                Corr corr) {
    constexpr int TW = TILE / 2, NB = TW / 16;
    constexpr bool kCorr = !std::is_same<Corr, NoCorr>::value;
    const double eij = a.eps_occ[it.i] + a.eps_occ[it.j];
    const bool kap = !isnan(a.kappa);
    const int last = a.nvir - 1;
    double eb[NB];
#pragma unroll
    for (int y = 0; y < NB; ++y) eb[y] = a.eps_vir[min(L.B0 + L.wc * TW + y * 16 + L.r, last)];
    KSum os, ss;
#pragma unroll
    for (int x = 0; x < NB; ++x) {
        double ea[4];
#pragma unroll
        for (int v = 0; v < 4; ++v)
            ea[v] = a.eps_vir[min(L.A0 + L.wr * TW + x * 16 + row_of(L.g, v), last)];
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            if (!(rv[x] && cv[y])) continue;  // wave-uniform
            const int bb = L.B0 + L.wc * TW + y * 16 + L.r;
            const bool okb = bb < a.nvir;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                // Branch-free: a virtual past nvir adds an exact 0.0, which
                // leaves both compensated sums bit for bit as they were (its
                // clamped eps keeps d finite and nonzero).
                const int aa = L.A0 + L.wr * TW + x * 16 + row_of(L.g, v);
                const bool ok = okb && aa < a.nvir;
                const double d = eij - ea[v] - eb[y];
                double g1 = acc1[x][y][v], g2 = acc2[x][y][v];
                if constexpr (kCorr) {
                    if (ok && it.i == it.j && aa == bb) {
                        const double c = corr(it.i, aa);
                        g1 += c;
                        g2 += c;
                    }
                }
                const double inv = 1.0 / d;
                const double dd = g1 - g2;
                const double dmp = damping(kap, a.kappa, d);
                const double to = dmp * ((g1 * g1 + g2 * g2) * inv);
                const double ts = dmp * ((dd * dd) * inv);
                os.add(ok ? to : 0.0);
                ss.add(ok ? ts : 0.0);
            }
        }
    }
    const double w = (it.i == it.j ? 1.0 : 2.0) * (it.ta == it.tb ? 0.5 : 1.0);
    block_store(os, ss, w, a.item_partials + 4 * it.id);
}

// BF16 hi = RN(x), lo = RN(x - hi) of x = RN_f32(v) (the split of pack.hip),
// packed as hi | lo << 16.
__device__ __forceinline__ uint32_t bf16_pair(double v) {
    const float f = (float)v;
    const __bf16 h = (__bf16)f;
    const __bf16 l = (__bf16)(f - (float)h);
    return (uint32_t)__builtin_bit_cast(uint16_t, h) |
           ((uint32_t)__builtin_bit_cast(uint16_t, l) << 16);
}

// The row group's T~ slabs (EnergyArgs::s_hi, st_hi): slab_put writes T~_ij^xy
// (i in [g0, g1), j >= i) at S row (i - g0, x), column (j, y); slab_put_mirror
// writes its mirror T~_ji^yx at S row (j - g0, y), column (i, x) when j is in
// the group, else at St row (j, y), column (i - g0, x) (nothing for j == i:
// the diagonal pair's elements each have their own writer).
__device__ __forceinline__ void slab_put(const EnergyArgs &a, int i, int x, int j, int y,
                                         uint32_t hl) {
    const int64_t vp = a.vpg;
    const int64_t k = ((int64_t)(i - a.g0) * vp + x) * a.ld_s + (int64_t)j * vp + y;
    a.s_hi[k] = (uint16_t)hl;
    a.s_lo[k] = (uint16_t)(hl >> 16);
}
__device__ __forceinline__ void slab_put_mirror(const EnergyArgs &a, int i, int x, int j, int y,
                                                uint32_t hl) {
    if (j == i) return;
    const int64_t vp = a.vpg;
    if (j < a.g1) {
        const int64_t m = ((int64_t)(j - a.g0) * vp + y) * a.ld_s + (int64_t)i * vp + x;
        a.s_hi[m] = (uint16_t)hl;
        a.s_lo[m] = (uint16_t)(hl >> 16);
    } else {
        const int64_t m = ((int64_t)j * vp + y) * a.ld_st + (int64_t)(i - a.g0) * vp + x;
        a.st_hi[m] = (uint16_t)hl;
        a.st_lo[m] = (uint16_t)(hl >> 16);
    }
}

__device__ __forceinline__ void tstore(void *buf, bool f32, int64_t idx, double v) {
    if (f32)
        ((float *)buf)[idx] = (float)v;
    else
        ((double *)buf)[idx] = v;
}

// Adjoint epilogue (phase A): energy partials as above, T~ written to the
// row's scratch blocks, and per-item sums of t*T~ (by row a, by column b,
// total) for dE/deps. Padding elements (nvir <= a, b < nvir_pad) of T~ are
// written as zeros: the phase-B/C GEMMs contract over them. The 2x2 waves
// holding the tile pass active = true; further waves of a larger workgroup
// (the INT8 kernel's) pass false and only join the barriers and the block sum.
// FULLT: T~ goes to the row group's slabs at (i a, j b) and (i b, j a), with
// their mirrors, instead of the row blocks tbuf / ubuf. The tile's T~_ij^ab
// and T~_ij^ba go through LDS (tlds, two TILE x (TILE + 1) words) and all
// four writes are made after a barrier by all threads of the workgroup, along
// b for the slab rows indexed by a and along a for those indexed by b: every
// store instruction then writes whole 64-byte runs (a lane's own elements run
// along b in 16-lane pieces, which left half-filled runs).
template <int TILE, typename Acc, typename RowOfT, typename Corr, bool FULLT = false>
__device__ __forceinline__ void adjoint_epilogue(const EnergyArgs &a, const Item &it,
                                                 const Lane &L,
                                                 Acc (&acc1)[TILE / 32][TILE / 32],
                                                 Acc (&acc2)[TILE / 32][TILE / 32],
                                                 const bool *rv, const bool *cv, RowOfT row_of,
                                                 Corr corr, bool active = true,
                                                 uint32_t *tlds = nullptr) {
    constexpr int TW = TILE / 2, NB = TW / 16;
    const double eij = a.eps_occ[it.i] + a.eps_occ[it.j];
    const double cd = a.c_os + a.c_ss, ce = a.c_ss;
    const bool diag_tile = it.ta == it.tb;
    const double wt = diag_tile ? 0.5 : 1.0;
    const int jj = it.j - it.i;
    const int64_t ub = (int64_t)(jj - 1) * a.nvir_pad * a.nvir_pad;
    KSum os, ss;
    double rsum[NB][4], csum[NB];
#pragma unroll
    for (int x = 0; x < NB; ++x) {
        csum[x] = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) rsum[x][v] = 0.0;
    }
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            if (!(active && rv[x] && cv[y])) continue;
            const int bb = L.B0 + L.wc * TW + y * 16 + L.r;
            const bool bok = bb < a.nvir;
            const double eb = bok ? a.eps_vir[bb] : 0.0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int aa = L.A0 + L.wr * TW + x * 16 + row_of(L.g, v);
                double T1 = 0.0, T2 = 0.0;
                if (bok && aa < a.nvir) {
                    const double d = eij - a.eps_vir[aa] - eb;
                    double g1 = acc1[x][y][v], g2 = acc2[x][y][v];
                    if (it.i == it.j && aa == bb) {
                        const double c = corr(it.i, aa);
                        g1 += c;
                        g2 += c;
                    }
                    const double inv = 1.0 / d;
                    const double dd = g1 - g2;
                    os.add((g1 * g1 + g2 * g2) * inv);
                    ss.add((dd * dd) * inv);
                    const double t1 = g1 * inv, t2 = g2 * inv;
                    T1 = cd * t1 - ce * t2;
                    T2 = cd * t2 - ce * t1;
                    const double val = wt * (t1 * T1 + t2 * T2);
                    rsum[x][v] += val;
                    csum[y] += val;
                }
                if constexpr (FULLT) {
                    constexpr int LT = TILE + 1;
                    const int la = L.wr * TW + x * 16 + row_of(L.g, v);
                    const int lb = L.wc * TW + y * 16 + L.r;
                    tlds[la * LT + lb] = bf16_pair(T1);
                    if (!diag_tile) tlds[TILE * LT + la * LT + lb] = bf16_pair(T2);
                } else {
                    // T~_ij^ab and T~_ij^ba into the row-i blocks.
                    tstore(a.tbuf, a.t_f32,
                           (int64_t)aa * a.ld_t + (int64_t)jj * a.nvir_pad + bb, T1);
                    if (!diag_tile)
                        tstore(a.tbuf, a.t_f32,
                               (int64_t)bb * a.ld_t + (int64_t)jj * a.nvir_pad + aa, T2);
                    if (jj > 0) {
                        tstore(a.ubuf, a.t_f32, ub + (int64_t)bb * a.nvir_pad + aa, T1);
                        if (!diag_tile)
                            tstore(a.ubuf, a.t_f32, ub + (int64_t)aa * a.nvir_pad + bb, T2);
                    }
                }
            }
        }
    const double w = (it.i == it.j ? 1.0 : 2.0) * wt;
    block_store(os, ss, w, a.item_partials + 4 * it.id);

    // Row sums: lanes of one group g share rows; reduce over r (xor 1..8),
    // then over the two column waves through LDS. Column sums: reduce over g
    // (xor 16, 32), then over the two row waves.
    __shared__ double rows_lds[2][TILE];
    __shared__ double cols_lds[2][TILE];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double s = rsum[x][v];
            s += __shfl_xor(s, 1, kWave);
            s += __shfl_xor(s, 2, kWave);
            s += __shfl_xor(s, 4, kWave);
            s += __shfl_xor(s, 8, kWave);
            if (active && L.r == 0) rows_lds[L.wc][L.wr * TW + x * 16 + row_of(L.g, v)] = s;
        }
#pragma unroll
    for (int y = 0; y < NB; ++y) {
        double s = csum[y];
        s += __shfl_xor(s, 16, kWave);
        s += __shfl_xor(s, 32, kWave);
        if (active && L.g == 0) cols_lds[L.wr][L.wc * TW + y * 16 + L.r] = s;
    }
    // Blocks skipped above (beyond nvir_pad) leave their LDS slots unset:
    // clear those slots first in a separate pass is unnecessary because the
    // reduction below reads only rows/cols < nvir_pad - A0 / - B0.
    __syncthreads();
    double *vec = a.item_vec + it.id * (int64_t)a.vec_stride;
    const int rmax = min(TILE, a.nvir_pad - L.A0), cmax = min(TILE, a.nvir_pad - L.B0);
    for (int k = threadIdx.x; k < TILE; k += blockDim.x) {
        vec[k] = k < rmax ? rows_lds[0][k] + rows_lds[1][k] : 0.0;
        vec[TILE + k] = k < cmax ? cols_lds[0][k] + cols_lds[1][k] : 0.0;
    }
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < rmax; ++k) s += rows_lds[0][k] + rows_lds[1][k];
        vec[2 * TILE] = s;
    }
    if constexpr (FULLT) {
        // The slab writes, from tlds (filled before the barrier above) by all
        // threads of the workgroup: first the slab rows indexed by a (threads
        // run along b), then those indexed by b (threads run along a), so every
        // store instruction writes whole 64-byte runs of a slab row.
        constexpr int LT = TILE + 1;
        for (int e = threadIdx.x; e < TILE * TILE; e += blockDim.x) {
            const int la = e / TILE, lb = e % TILE;
            const int aa = L.A0 + la, bb = L.B0 + lb;
            if (aa >= a.nvir_pad || bb >= a.nvir_pad) continue;
            slab_put(a, it.i, aa, it.j, bb, tlds[la * LT + lb]);
            if (!diag_tile) slab_put_mirror(a, it.i, bb, it.j, aa, tlds[TILE * LT + la * LT + lb]);
        }
        for (int e = threadIdx.x; e < TILE * TILE; e += blockDim.x) {
            const int la = e % TILE, lb = e / TILE;
            const int aa = L.A0 + la, bb = L.B0 + lb;
            if (aa >= a.nvir_pad || bb >= a.nvir_pad) continue;
            slab_put_mirror(a, it.i, aa, it.j, bb, tlds[la * LT + lb]);
            if (!diag_tile) slab_put(a, it.i, bb, it.j, aa, tlds[TILE * LT + la * LT + lb]);
        }
    }
}

// Amplitude store (the MP2 densities): t = g / D of the item's elements into
// the ordered-pair amplitude blocks, a.tamp[((p*nocc + q)*nvir_pad + x)*nvir_pad + y]
// = t_pq^xy for every ordered occupied pair (p, q). C1 gives t_ij^ab, stored as
// (i, j, a, b) and as its mirror (j, i, b, a); C2 gives t_ij^ba, stored as
// (i, j, b, a) and (j, i, a, b) (on a diagonal tile pair C2 = C1^T, whose
// elements other lanes store). Elements past nvir are stored as zeros, so the
// padded blocks read back whole. Every element has exactly one writer.
template <int TILE, typename Acc, typename RowOfT>
__device__ __forceinline__ void amp_store(const EnergyArgs &a, const Item &it, const Lane &L,
                                          Acc (&acc1)[TILE / 32][TILE / 32],
                                          Acc (&acc2)[TILE / 32][TILE / 32], const bool *rv,
                                          const bool *cv, RowOfT row_of) {
    constexpr int TW = TILE / 2, NB = TW / 16;
    const double eij = a.eps_occ[it.i] + a.eps_occ[it.j];
    const int64_t vp = a.nvir_pad;
    double *bij = a.tamp + ((int64_t)it.i * a.nocc + it.j) * vp * vp;
    double *bji = a.tamp + ((int64_t)it.j * a.nocc + it.i) * vp * vp;
    const bool diag_tile = it.ta == it.tb;
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            if (!(rv[x] && cv[y])) continue;
            const int bb = L.B0 + L.wc * TW + y * 16 + L.r;
            const bool bok = bb < a.nvir;
            const double eb = bok ? a.eps_vir[bb] : 0.0;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int aa = L.A0 + L.wr * TW + x * 16 + row_of(L.g, v);
                double t1 = 0.0, t2 = 0.0;
                if (bok && aa < a.nvir) {
                    const double inv = 1.0 / (eij - a.eps_vir[aa] - eb);
                    t1 = acc1[x][y][v] * inv;
                    t2 = acc2[x][y][v] * inv;
                }
                bij[aa * vp + bb] = t1;
                if (!diag_tile) bij[bb * vp + aa] = t2;
                // For i == j the mirror block is the block itself: each of its
                // elements has exactly one writer above (on a diagonal tile
                // pair the lane of (b, a) writes t^ba, which need not equal
                // this lane's t^ab to the last bit).
                if (it.i != it.j) {
                    bji[bb * vp + aa] = t1;
                    if (!diag_tile) bji[aa * vp + bb] = t2;
                }
            }
        }
}

struct LoSqCorr {
    const double *p;
    int nvir_pad;
    __device__ double operator()(int i, int aa) const { return p[(int64_t)i * nvir_pad + aa]; }
};
template <typename T>
struct RowOf {
    __device__ int operator()(int g, int v) const { return out_row<T>(g, v); }
};

// Single-product items: for i == j the operands of C2 are those of C1
// (C2 = C1), and on a diagonal tile pair (A == B) C2[a][b] = g_ba = C1[b][a].
// Such an item computes C1 alone, with both accumulator sets: operands 2 and 3
// are operands 0 and 1 again, one Q panel ahead, so acc1 sums the even panels
// and acc2 the odd ones and the loop runs half as many iterations (a panel
// past naux_pad reads as zeros). derive_c2 then forms C1 = acc1 + acc2 and
// C2 from it.
struct SplitQ {
    bool single;
    int off23, step, iters;
    __device__ SplitQ(const Item &it, int64_t naux_pad, int kp, bool always = false) {
        const int np = (int)(naux_pad / kp);
        single = always || it.i == it.j || it.ta == it.tb;
        off23 = single ? kp : 0;
        step = single ? 2 * kp : kp;
        iters = single ? (np + 1) / 2 : np;
    }
};

// After the main loop of a single-product item: acc1 += acc2 gives C1, then
// acc2 = C1 (i == j) or C1^T. The transpose goes through LDS (cb, at least
// TW x (TILE + 1) elements, free once the loop's last barrier has passed) in
// two rounds of TW rows: in round h the waves with wr == h write their rows of
// C1 and the waves with wc == h read them transposed. Blocks the loop skipped
// (beyond nvir_pad) hold zeros and stay zeros.

template <int TILE, typename T, typename Acc, typename RowOfT>
__device__ __forceinline__ void derive_c2(const Item &it, const Lane &L,
                                          Acc (&acc1)[TILE / 32][TILE / 32],
                                          Acc (&acc2)[TILE / 32][TILE / 32], RowOfT row_of,
                                          T *cb) {
    constexpr int TW = TILE / 2, NB = TW / 16, LDC = TILE + 1;
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            acc1[x][y] += acc2[x][y];
            acc2[x][y] = acc1[x][y];
        }
    if (it.i == it.j) return;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (L.wr == h) {
#pragma unroll
            for (int x = 0; x < NB; ++x)
#pragma unroll
                for (int y = 0; y < NB; ++y)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        cb[(x * 16 + row_of(L.g, v)) * LDC + L.wc * TW + y * 16 + L.r] =
                            acc1[x][y][v];
        }
        __syncthreads();
        if (L.wc == h) {
#pragma unroll
            for (int x = 0; x < NB; ++x)
#pragma unroll
                for (int y = 0; y < NB; ++y)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        acc2[x][y][v] =
                            cb[(y * 16 + L.r) * LDC + L.wr * TW + x * 16 + row_of(L.g, v)];
        }
        __syncthreads();
    }
}

// Opposite-spin rectangle epilogue (u_rimp2.rs:731-760): acc holds
// (ia|JB) for a in the alpha tile A and B in the beta tile; the item adds
// sum (ia|JB)^2 / D with D = e_i + e_J - e_a - e_B, weight 1 (every (i, J) and
// every tile pair is its own item), into the os slot of its partial.
template <int TILE, typename Acc, typename RowOfT>
__device__ __forceinline__ void rect_epilogue(const EnergyArgs &a, const Item &it, const Lane &L,
                                              Acc (&acc)[TILE / 32][TILE / 32], const bool *rv,
                                              const bool *cv, RowOfT row_of) {
    constexpr int TW = TILE / 2, NB = TW / 16;
    const double eij = a.eps_occ[it.i] + a.eps_occ_b[it.j];
    KSum os, ss;
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            if (!(rv[x] && cv[y])) continue;
            const int bb = L.B0 + L.wc * TW + y * 16 + L.r;
            if (bb >= a.nvir_b) continue;
            const double eb = a.eps_vir_b[bb];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int aa = L.A0 + L.wr * TW + x * 16 + row_of(L.g, v);
                if (aa >= a.nvir) continue;
                const double g = acc[x][y][v];
                os.add((g * g) / (eij - a.eps_vir[aa] - eb));
            }
        }
    block_store(os, ss, 1.0, a.item_partials + 4 * it.id);
}

// ---------------------------------------------------------------------------
// FP64 / FP32: v_mfma_{f64,f32}_16x16x4. Lane (r = l%16, g = l/16) supplies
// A[row r][k] and B[k][col r] for its k values; the output element v of lane
// l is C[out_row(g, v)][r]. Within a panel of KP Q values lane group g owns
// the contiguous k range [g*KPL, g*KPL + KPL), read with one 16-byte LDS load;
// A and B use the same k assignment, so each MFMA contracts matching Q.
// ---------------------------------------------------------------------------
template <typename T, int TILE, int KP>
using K4Lds = T[2][4][TILE][KP + 16 / sizeof(T)];

// RECT: the opposite-spin rectangle (op0 alpha rows, op1 beta rows, it.j the
// beta occupied index); every item is single-product and ends in
// rect_epilogue. SIGNED: panels whose first Q column lies at or past a.q_neg
// belong to the negative factor of a signed kernel; their A-side fragments are
// negated, so the MFMA chain accumulates B+^T B+ - B-^T B- (a compile-time
// switch, so the other kernels carry no extra registers for it).
// FLUSH (FP32 only, the MIXED mode): the FP32 accumulators restart every
// kMixedPanel Q rows and each panel's sums are added into FP64 accumulators,
// which the epilogue reads; a two-product item's accumulators each cover a
// whole 64-row panel, a single-product item's (even / odd 16-row panels) 32
// rows of it.
template <typename T, int TILE, int KP, bool ADJ, bool RECT = false, bool SIGNED = false,
          bool FLUSH = false, bool AMP = false>
__device__ __forceinline__ void pair_k4_body(const EnergyArgs &a, const Item &it,
                                             K4Lds<T, TILE, KP> &lds) {
    static_assert(!FLUSH || (std::is_same<T, float>::value && !ADJ && !SIGNED),
                  "the FP64 flush is an FP32 energy mode");
    static_assert(!AMP || (std::is_same<T, double>::value && !ADJ && !RECT && !SIGNED && !FLUSH),
                  "the amplitude store is an FP64 energy mode");
    static_assert(kMixedPanel % (2 * KP) == 0, "flushes fall on whole loop iterations");
    using M = Mfma4<T>;
    constexpr int TW = TILE / 2;          // wave sub-tile side (2x2 waves)
    constexpr int NB = TW / 16;           // 16x16 blocks per wave side
    constexpr int EPC = 16 / sizeof(T);   // elements per 16-byte chunk
    constexpr int CPR = KP / EPC;         // chunks per row per panel
    constexpr int OPCH = TILE * CPR;      // chunks per operand panel
    constexpr int NCH = 4 * OPCH;         // chunks per panel (4 operands)
    constexpr int CPT = (NCH + 255) / 256;
    constexpr int LDR = KP + EPC;         // LDS row stride: one chunk of padding
    static_assert(sizeof(lds[0][0][0]) == LDR * sizeof(T), "K4Lds rows carry one spare chunk");
    constexpr int KPL = KP / 4;           // k values per lane per panel
    static_assert((KPL * sizeof(T)) % 16 == 0, "fragments are whole 16-byte chunks");
    typedef T frag_t __attribute__((ext_vector_type(KPL)));
    typedef T chunk_t __attribute__((ext_vector_type(EPC)));

    const T *__restrict__ bpa = (const T *)a.op0;
    const T *__restrict__ bpb = RECT ? (const T *)a.op1 : bpa;
    const int vpa = a.nvir_pad, vpb = RECT ? a.nvir_pad_b : a.nvir_pad;
    const int A0 = it.ta * TILE, B0 = it.tb * TILE;
    const SplitQ sq(it, a.naux_pad, KP, RECT);
    // Operand rows: 0 = (i, A), 1 = (j, B), 2 = (j, A), 3 = (i, B); for a
    // single-product item 2 and 3 are 0 and 1 again, one Q panel further on.
    // Operands 0 and 2 come from bpa, 1 and 3 from bpb.
    const int64_t base0 = (int64_t)it.i * vpa + A0;
    const int64_t base1 = (int64_t)it.j * vpb + B0;
    const int64_t base2 = sq.single ? base0 : (int64_t)it.j * vpa + A0;
    const int64_t base3 = sq.single ? base1 : (int64_t)it.i * vpb + B0;
    const int limA = vpa - A0, limB = vpb - B0;
    const int64_t ld = a.naux_pad;
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
                v = *(const v4i *)(((op & 1) ? bpb : bpa) + (base + row) * ld + q + ch * EPC);
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
                *(v4i *)&lds[buf][op][row][ch * EPC] = stg[t];
            }
        }
    };

    const int lane = tid % kWave, wid = tid / kWave;
    const Lane L{A0, B0, wid / 2, wid % 2, lane % 16, lane / 16};
    bool rv[NB], cv[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        rv[k] = A0 + L.wr * TW + k * 16 < vpa;
        cv[k] = B0 + L.wc * TW + k * 16 < vpb;
    }

    typename M::acc acc1[NB][NB], acc2[NB][NB];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            acc1[x][y] = (typename M::acc){0, 0, 0, 0};
            acc2[x][y] = (typename M::acc){0, 0, 0, 0};
        }
    // FLUSH: the FP64 sums of the finished Q panels, in the FP32 result layout.
    constexpr int NF = FLUSH ? NB : 1;
    v4d w1[NF][NF], w2[NF][NF];
#pragma unroll
    for (int x = 0; x < NF; ++x)
#pragma unroll
        for (int y = 0; y < NF; ++y) {
            w1[x][y] = (v4d){0, 0, 0, 0};
            w2[x][y] = (v4d){0, 0, 0, 0};
        }
    auto flush = [&]() {
        if constexpr (FLUSH) {
#pragma unroll
            for (int x = 0; x < NB; ++x)
#pragma unroll
                for (int y = 0; y < NB; ++y) {
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        w1[x][y][v] += (double)acc1[x][y][v];
                        w2[x][y][v] += (double)acc2[x][y][v];
                    }
                    acc1[x][y] = (typename M::acc){0, 0, 0, 0};
                    acc2[x][y] = (typename M::acc){0, 0, 0, 0};
                }
        }
    };
    const int flush_every = kMixedPanel / sq.step;  // loop iterations per 64 Q rows

    // Fragments: the lane's KPL consecutive k values, read as 16-byte chunks.
    auto ldfrag = [&](int buf, int op, int row) -> frag_t {
        frag_t f;
        const chunk_t *src = (const chunk_t *)&lds[buf][op][row][L.g * KPL];
#pragma unroll
        for (int c = 0; c < KPL / EPC; ++c) {
            const chunk_t v = src[c];
#pragma unroll
            for (int e = 0; e < EPC; ++e) f[c * EPC + e] = v[e];
        }
        return f;
    };
    // FULL: every 16x16 block of the wave lies inside nvir_pad (the common
    // case), so the MFMA sequence needs no per-block test.
    auto compute = [&](int buf, auto full, int q0) {
        frag_t f0[NB], f1[NB], f2[NB], f3[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            f0[k] = ldfrag(buf, 0, L.wr * TW + k * 16 + L.r);
            f2[k] = ldfrag(buf, 2, L.wr * TW + k * 16 + L.r);
            f1[k] = ldfrag(buf, 1, L.wc * TW + k * 16 + L.r);
            f3[k] = ldfrag(buf, 3, L.wc * TW + k * 16 + L.r);
        }
        if constexpr (SIGNED) {
            if (q0 >= a.q_neg) {
#pragma unroll
                for (int k = 0; k < NB; ++k) f0[k] = -f0[k];
            }
            if (q0 + sq.off23 >= a.q_neg) {
#pragma unroll
                for (int k = 0; k < NB; ++k) f2[k] = -f2[k];
            }
        }
#pragma unroll
        for (int s = 0; s < KPL; ++s)
#pragma unroll
            for (int x = 0; x < NB; ++x)
#pragma unroll
                for (int y = 0; y < NB; ++y)
                    if (decltype(full)::value || (rv[x] && cv[y])) {
                        acc1[x][y] = M::mma(f0[x][s], f1[y][s], acc1[x][y]);
                        acc2[x][y] = M::mma(f2[x][s], f3[y][s], acc2[x][y]);
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
            compute(cur, std::true_type{}, p * sq.step);
        else
            compute(cur, std::false_type{}, p * sq.step);
        if constexpr (FLUSH)
            if ((p + 1) % flush_every == 0 || p + 1 == sq.iters) flush();
        if (p + 1 < sq.iters) lstore(cur ^ 1);
        __syncthreads();
    }
    if constexpr (FLUSH) {
        if constexpr (RECT) {
#pragma unroll
            for (int x = 0; x < NB; ++x)
#pragma unroll
                for (int y = 0; y < NB; ++y) w1[x][y] += w2[x][y];
            rect_epilogue<TILE>(a, it, L, w1, rv, cv, RowOf<T>{});
            return;
        }
        if (sq.single) {
            static_assert(sizeof(lds) >= sizeof(double) * TW * (TILE + 1),
                          "LDS holds an FP64 C1 half-tile");
            derive_c2<TILE, double>(it, L, w1, w2, RowOf<T>{}, (double *)&lds[0][0][0][0]);
        }
        energy_epilogue<TILE>(a, it, L, w1, w2, rv, cv, RowOf<T>{}, NoCorr{});
        return;
    } else {
        if constexpr (RECT) {
#pragma unroll
            for (int x = 0; x < NB; ++x)
#pragma unroll
                for (int y = 0; y < NB; ++y) acc1[x][y] += acc2[x][y];
            rect_epilogue<TILE>(a, it, L, acc1, rv, cv, RowOf<T>{});
            return;
        }
        if (sq.single) {
            static_assert(sizeof(lds) >= sizeof(T) * TW * (TILE + 1), "LDS holds a C1 half-tile");
            derive_c2<TILE, T>(it, L, acc1, acc2, RowOf<T>{}, (T *)&lds[0][0][0][0]);
        }

        if constexpr (ADJ) {
            adjoint_epilogue<TILE>(a, it, L, acc1, acc2, rv, cv, RowOf<T>{}, NoCorr{});
        } else {
            if constexpr (AMP) amp_store<TILE>(a, it, L, acc1, acc2, rv, cv, RowOf<T>{});
            energy_epilogue<TILE>(a, it, L, acc1, acc2, rv, cv, RowOf<T>{}, NoCorr{});
        }
    }
}

// ---------------------------------------------------------------------------
// BF16 hi/lo three-pass: v_mfma_f32_16x16x32_bf16. Lane (r, g) supplies 8
// consecutive k values [8g, 8g+8) of a 32-deep panel for A rows and B columns
// alike. Each 128-byte LDS row holds the panel's 32 hi then 32 lo values as
// eight 16-byte chunks, chunk c stored at slot c ^ (row & 7) so that the
// fragment reads of 8 consecutive rows hit distinct banks.
// C += hi*hi + hi*lo + lo*hi in FP32; the omitted lo*lo term is added back
// exactly (FP64 row sums from pack.hip) on the diagonal elements K_ii[a, a],
// the only ones where it is one-signed.
// ---------------------------------------------------------------------------
template <int TILE>
using Bf16Lds = v4i[2][4][TILE][8];

template <int TILE, bool ADJ>
__device__ __forceinline__ void pair_bf16x3_body(const EnergyArgs &a, const Item &it,
                                                 Bf16Lds<TILE> &lds) {
    constexpr int KP = 32;
    constexpr int TW = TILE / 2;
    constexpr int NB = TW / 16;
    constexpr int CPR = 8;               // 4 hi + 4 lo chunks per row
    constexpr int OPCH = TILE * CPR;
    constexpr int NCH = 4 * OPCH;
    constexpr int CPT = (NCH + 255) / 256;

    const uint16_t *__restrict__ hi = (const uint16_t *)a.op0;
    const uint16_t *__restrict__ lo = (const uint16_t *)a.op1;
    const int A0 = it.ta * TILE, B0 = it.tb * TILE;
    const SplitQ sq(it, a.naux_pad, KP);
    const int64_t base0 = (int64_t)it.i * a.nvir_pad + A0;
    const int64_t base1 = (int64_t)it.j * a.nvir_pad + B0;
    const int64_t base2 = sq.single ? base0 : (int64_t)it.j * a.nvir_pad + A0;
    const int64_t base3 = sq.single ? base1 : (int64_t)it.i * a.nvir_pad + B0;
    const int limA = a.nvir_pad - A0, limB = a.nvir_pad - B0;
    const int64_t ld = a.naux_pad;
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
            if (c < NCH && row < lim && q < a.naux_pad) {
                const uint16_t *src = (ch < 4 ? hi : lo) + (base + row) * ld + q + (ch & 3) * 8;
                v = *(const v4i *)src;
            }
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
                lds[buf][op][row][ch ^ (row & 7)] = stg[t];
            }
        }
    };

    const int lane = tid % kWave, wid = tid / kWave;
    const Lane L{A0, B0, wid / 2, wid % 2, lane % 16, lane / 16};
    bool rv[NB], cv[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        rv[k] = A0 + L.wr * TW + k * 16 < a.nvir_pad;
        cv[k] = B0 + L.wc * TW + k * 16 < a.nvir_pad;
    }
    v4f acc1[NB][NB], acc2[NB][NB];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            acc1[x][y] = (v4f){0, 0, 0, 0};
            acc2[x][y] = (v4f){0, 0, 0, 0};
        }

    auto frag = [&](int buf, int op, int row, int chunk) -> v8bf {
        return __builtin_bit_cast(v8bf, lds[buf][op][row][chunk ^ (row & 7)]);
    };
    auto compute = [&](int buf) {
        v8bf h0[NB], l0[NB], h2[NB], l2[NB], h1[NB], l1[NB], h3[NB], l3[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int ra = L.wr * TW + k * 16 + L.r, rb = L.wc * TW + k * 16 + L.r;
            h0[k] = frag(buf, 0, ra, L.g);
            l0[k] = frag(buf, 0, ra, 4 + L.g);
            h2[k] = frag(buf, 2, ra, L.g);
            l2[k] = frag(buf, 2, ra, 4 + L.g);
            h1[k] = frag(buf, 1, rb, L.g);
            l1[k] = frag(buf, 1, rb, 4 + L.g);
            h3[k] = frag(buf, 3, rb, L.g);
            l3[k] = frag(buf, 3, rb, 4 + L.g);
        }
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
                if (rv[x] && cv[y]) {
                    v4f c1 = acc1[x][y], c2 = acc2[x][y];
                    c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l0[x], h1[y], c1, 0, 0, 0);
                    c2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(l2[x], h3[y], c2, 0, 0, 0);
                    c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h0[x], l1[y], c1, 0, 0, 0);
                    c2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h2[x], l3[y], c2, 0, 0, 0);
                    c1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h0[x], h1[y], c1, 0, 0, 0);
                    c2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(h2[x], h3[y], c2, 0, 0, 0);
                    acc1[x][y] = c1;
                    acc2[x][y] = c2;
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
    if (sq.single) {
        static_assert(sizeof(lds) >= sizeof(float) * TW * (TILE + 1), "LDS holds a C1 half-tile");
        derive_c2<TILE, float>(it, L, acc1, acc2, RowOf<float>{}, (float *)&lds[0][0][0][0]);
    }

    const LoSqCorr corr{a.diag_corr, a.nvir_pad};
    if constexpr (ADJ)
        adjoint_epilogue<TILE>(a, it, L, acc1, acc2, rv, cv, RowOf<float>{}, corr);
    else
        energy_epilogue<TILE>(a, it, L, acc1, acc2, rv, cv, RowOf<float>{}, corr);
}

// The same BF16 hi/lo three-pass contraction on eight waves (512 threads),
// every fragment loaded straight into registers from the operand in MFMA
// fragment order (EnergyArgs::opf, PackFmt::BF16F): per (16-row group, 32-deep
// Q panel) one 2 KB block, hi then lo, each a coalesced 1 KB wave load, so the
// main loop has no LDS traffic and no barrier (the INT8X4 kernel's scheme).
// Waves 0-3 accumulate C1 (operand rows (i, A) and (j, B)), waves 4-7 C2 ((j, A)
// and (i, B)) or, for a single-product item, C1 over the odd Q panels; each
// wave one TILE/2-square quadrant, two waves per SIMD, one panel in flight
// (double-buffered registers). Blocks past nvir_pad are neither loaded nor
// multiplied. After the loop the halves meet through the idle LDS (the
// odd-panel C1 added to C1 in the four-wave body's order, so the sums are
// bitwise those of pair_bf16x3_body), and waves 0-3 run the energy epilogue
// while waves 4-7 contribute zeros to the workgroup sum.
template <int TILE>
__device__ __forceinline__ void pair_bf16x3_w8f_body(const EnergyArgs &a, const Item &it,
                                                     Bf16Lds<TILE> &lds) {
    constexpr int KP = 32;
    constexpr int TW = TILE / 2;
    constexpr int NB = TW / 16;
    constexpr int FBLK = 2048;  // bytes of one (group, panel) block: hi and lo fragments
    static_assert(sizeof(lds) >= sizeof(float) * TILE * (TILE + 1),
                  "LDS holds a whole FP32 tile after the loop");
    const int A0 = it.ta * TILE, B0 = it.tb * TILE;
    const SplitQ sq(it, a.naux_pad, KP);
    const int np = a.naux_pad / KP, ng = a.nvir_pad / 16;
    const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
    const int prod = wid >> 2, quad = wid & 3;
    const Lane L{A0, B0, quad >> 1, quad & 1, lane % 16, lane / 16};
    const bool c2 = prod == 1 && !sq.single;
    const int occA = c2 ? it.j : it.i, occB = c2 ? it.i : it.j;
    const int pstep = sq.single ? 2 : 1, poff = (prod == 1 && sq.single) ? 1 : 0;
    const char *fbase = (const char *)a.opf;
    const char *pa[NB], *pb[NB];
    bool rv[NB], cv[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        rv[k] = A0 + L.wr * TW + k * 16 < a.nvir_pad;
        cv[k] = B0 + L.wc * TW + k * 16 < a.nvir_pad;
        pa[k] = fbase + (int64_t)(occA * ng + (A0 + L.wr * TW) / 16 + k) * np * FBLK + lane * 16;
        pb[k] = fbase + (int64_t)(occB * ng + (B0 + L.wc * TW) / 16 + k) * np * FBLK + lane * 16;
    }
    bool full = true, anyr = false, anyc = false;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        full = full && rv[k] && cv[k];
        anyr = anyr || rv[k];
        anyc = anyc || cv[k];
    }
    v4f acc[NB][NB];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) acc[x][y] = (v4f){0, 0, 0, 0};

    // [hi / lo][block]
    v4i fa[2][2][NB], fb[2][2][NB];
    auto fetch = [&](int P, v4i(&ra)[2][NB], v4i(&rb)[2][NB], auto fullc) {
        const int64_t off = (int64_t)P * FBLK;
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int k = 0; k < NB; ++k) {
                if (decltype(fullc)::value || rv[k])
                    ra[h][k] = *(const v4i *)(pa[k] + off + h * 1024);
                if (decltype(fullc)::value || cv[k])
                    rb[h][k] = *(const v4i *)(pb[k] + off + h * 1024);
            }
    };
    auto mma = [&](const v4i(&ra)[2][NB], const v4i(&rb)[2][NB], auto fullc) {
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
                if (decltype(fullc)::value || (rv[x] && cv[y])) {
                    const v8bf ha = __builtin_bit_cast(v8bf, ra[0][x]);
                    const v8bf la = __builtin_bit_cast(v8bf, ra[1][x]);
                    const v8bf hb = __builtin_bit_cast(v8bf, rb[0][y]);
                    const v8bf lb = __builtin_bit_cast(v8bf, rb[1][y]);
                    v4f c = acc[x][y];
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(la, hb, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ha, lb, c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ha, hb, c, 0, 0, 0);
                    acc[x][y] = c;
                }
    };
    // Iteration p covers panel p * pstep + poff; unrolled by two so that the
    // register buffer of each step is static.
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
    if (full)
        run(std::true_type{});
    else if (anyr && anyc)
        run(std::false_type{});
    __syncthreads();

    // Waves 4-7 hand their sums to waves 0-3 through LDS (a whole FP32 tile,
    // row-major with a one-element pad): the odd-panel C1 of a single-product
    // item is added to the even-panel C1 (acc1 += acc2, as derive_c2 does), a
    // two-product item's C2 is read back as acc2. A diagonal tile pair's C2 is
    // C1 transposed, taken from the same buffer after C1 is complete; i == j
    // gives C2 = C1.
    constexpr int LDT = TILE + 1;
    float *xb = (float *)&lds[0][0][0][0];
    auto put = [&](const v4f (&v)[NB][NB]) {
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    xb[(L.wr * TW + x * 16 + out_row<float>(L.g, e)) * LDT + L.wc * TW + y * 16 +
                       L.r] = v[x][y][e];
    };
    v4f acc2[NB][NB];
    if (prod == 1) put(acc);
    __syncthreads();
    if (prod == 0) {
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float o = xb[(L.wr * TW + x * 16 + out_row<float>(L.g, e)) * LDT +
                                       L.wc * TW + y * 16 + L.r];
                    if (sq.single)
                        acc[x][y][e] += o;
                    else
                        acc2[x][y][e] = o;
                }
    }
    const bool transpose = sq.single && it.i != it.j;
    if (transpose) {
        __syncthreads();  // every read of the odd-panel sums done
        if (prod == 0) put(acc);
        __syncthreads();
    }
    if (prod == 0 && sq.single) {
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    acc2[x][y][e] = transpose
                                        ? xb[(L.wc * TW + y * 16 + L.r) * LDT + L.wr * TW +
                                             x * 16 + out_row<float>(L.g, e)]
                                        : acc[x][y][e];
    }
    if (prod == 0) {
        const LoSqCorr corr{a.diag_corr, a.nvir_pad};
        energy_epilogue<TILE>(a, it, L, acc, acc2, rv, cv, RowOf<float>{}, corr);
    } else {
        block_store(KSum{}, KSum{}, 0.0, a.item_partials + 4 * it.id);
    }
}

}  // namespace pair
}  // namespace fh
