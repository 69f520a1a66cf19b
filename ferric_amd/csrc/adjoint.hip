// Adjoint of the RI-MP2 pair energy: Gamma = dE/dB and dE/deps (gfx950).
//
// E = c_os E_os + c_ss E_ss = sum_ijab K_ij^ab T~_ij^ab with
// T~ = (c_os + c_ss) t - c_ss t^T, t = K/D (rimp2.rs:1987-2011). Then
//   Gamma^Q_ia = dE/dB^Q_ia = 4 x[Q, ia],  x[Q, ia] = sum_jb T~_ij^ab B^Q_jb
//     (zvector.rs:651-697; x = 1/4 dE/dB, rimp2.rs:1832-1840),
//   dE/deps_k = 2 P_kk = -2 sum_jab t_kj^ab T~_kj^ab       (active occupied)
//   dE/deps_c = 2 P_cc =  2 sum_ijb t_ij^cb T~_ij^cb       (virtual)
//     (oo_rimp2.rs:1192-1300, rimp2.rs:2026-2052).
//
// One batch per occupied row i, in ascending order:
//   A  pair kernel over (i, j >= i) (energy.hip, adjoint epilogue): recomputes
//      K_ij, writes T~_ij for this row's pairs into row scratch (tbuf: a-major,
//      ubuf: b-major per j > i) and per-item t*T~ sums;
//   R  reduction of this row's energy and dE/deps partials, in a fixed order;
//   BC one grouped GEMM launch:
//      B  Gamma_i[Q, a]   += 4 sum_{j >= i, b} B_j[Q, b] T~_ij^ab   (K = (nocc-i) nvir)
//      C  Gamma_j[Q, b]   += 4 sum_a B_i[Q, a] T~_ij^ab, j > i      (K = nvir)
//      (T~_ji^ba = T~_ij^ab covers the j < i half of x).
// Every Gamma element has exactly one owner per launch and launches are
// stream-ordered, so the result is deterministic. T~ is never held for more
// than one occupied row.
//
// INT8X4 (adjoint_core_i8): groups of R occupied rows. Phase A of a group runs
// on the INT8 digit-slice kernel (energy_i8.hip) over the items of its rows and
// writes their T~ -- symmetric, T~_ji^ba = T~_ij^ab -- as BF16 hi / lo into two
// slabs over the padded (i, a) index space: S (the group's rows, every column
// from the group on) and St (the transposed blocks of the columns past the
// group), 2 x R vpg x nocc vpg x 4 bytes (5.5 GB at 100/900/3600 with R = 8,
// against the 64.8 GB amplitudes the reference holds); the per-row reductions
// follow; then two hand-written BF16x3 GEMMs (bf16x3_nt_kernel, FP32
// accumulation) add this group's part of Gamma = 4 B T~ into an FP32 gamma,
// widened to FP64 at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "fh_device.h"
#include "fh_host.h"
#include "fh_kernels.h"

namespace fh {

namespace {

// ---- Q-major padded copy: b2[Q][j*nvir_pad + b] = B[Q][j*nvir + b] ----------
template <typename Tin, typename Tout>
__global__ __launch_bounds__(256) void pack_qmajor_kernel(const Tin *__restrict__ b, int naux,
                                                          int nocc, int nvir, int nvir_pad,
                                                          int m_pad, Tout *__restrict__ out) {
    const int64_t ncol = (int64_t)nocc * nvir_pad;
    const int64_t total = (int64_t)m_pad * ncol;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < total;
         k += (int64_t)gridDim.x * 256) {
        const int64_t q = k / ncol, c = k % ncol;
        const int j = (int)(c / nvir_pad), bb = (int)(c % nvir_pad);
        Tout v = 0;
        if (q < naux && bb < nvir) v = (Tout)b[q * (int64_t)nocc * nvir + (int64_t)j * nvir + bb];
        out[k] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void fill_kernel(T *__restrict__ p, int64_t n, T v) {
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256)
        p[k] = v;
}

// ---- Grouped NT GEMM of phases B and C --------------------------------------
struct GemmTile {
    int m0, n0;        // output tile origin (Q, virtual)
    int64_t a_col;     // first k column of the A operand in b2
    const void *bm;    // B operand rows (n), k contiguous
    int64_t ldb;
    int nrows_b;       // valid B rows (nvir_pad)
    int kbeg, kend;    // contraction range
    int64_t col0;      // gamma column of n = 0 (phase C)
    int split;         // phase B: split index; phase C: -1
};

// Blocks [0, nb*splits) are phase-B (tile, split) pairs, the rest phase-C
// tiles of j = i + 1, i + 2, ...; the long phase-B blocks are dispatched first.
template <int TILE>
__device__ __forceinline__ GemmTile decode_gemm(const GemmArgs &g, int64_t bid) {
    GemmTile t;
    const int tm = g.m_pad / TILE;
    const int tn = (g.nvir_pad + TILE - 1) / TILE;
    const int64_t nb = (int64_t)tm * tn;
    const size_t esz = g.prec == Prec::F64 ? 8 : 4;
    t.a_col = (int64_t)g.i * g.nvir_pad;
    t.nrows_b = g.nvir_pad;
    const int64_t nbB = nb * g.splits;
    if (bid < nbB) {
        const int64_t tile = bid / g.splits;
        t.split = (int)(bid % g.splits);
        t.m0 = (int)(tile / tn) * TILE;
        t.n0 = (int)(tile % tn) * TILE;
        t.bm = g.tbuf;
        t.ldb = g.ld_t;
        const int k = (g.nocc - g.i) * g.nvir_pad;
        t.kbeg = min(k, t.split * g.kchunk);
        t.kend = min(k, t.kbeg + g.kchunk);
        t.col0 = (int64_t)g.i * g.nvir;
    } else {
        const int64_t c = bid - nbB;
        const int jj = 1 + (int)(c / nb);
        const int64_t tile = c % nb;
        t.split = -1;
        t.m0 = (int)(tile / tn) * TILE;
        t.n0 = (int)(tile % tn) * TILE;
        t.bm = (const char *)g.ubuf + (size_t)(jj - 1) * g.nvir_pad * g.nvir_pad * esz;
        t.ldb = g.nvir_pad;
        t.kbeg = 0;
        t.kend = g.nvir_pad;
        t.col0 = (int64_t)(g.i + jj) * g.nvir;
    }
    return t;
}

// Epilogue. Phase C: gamma[Q, col0 + n] += alpha * acc for Q < naux, n < nvir.
// Phase B: the raw tile goes to partial[split] (every element of the tile that
// lies inside (m_pad, nvir_pad)); gemm_splits_kernel adds the splits in order.
template <int TILE, typename Acc, typename RowOfT>
__device__ __forceinline__ void gemm_store(const GemmArgs &g, const GemmTile &t, int wr, int wc,
                                           int r, int gg, Acc (&acc)[TILE / 32][TILE / 32],
                                           RowOfT row_of) {
    constexpr int TW = TILE / 2, NB = TW / 16;
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            const int n = t.n0 + wc * TW + y * 16 + r;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int q = t.m0 + wr * TW + x * 16 + row_of(gg, v);
                if (t.split >= 0) {
                    if (n < g.nvir_pad)
                        g.partial[((int64_t)t.split * g.m_pad + q) * g.nvir_pad + n] =
                            (double)acc[x][y][v];
                } else if (n < g.nvir && q < g.naux) {
                    double *dst = g.gamma + (int64_t)q * g.nov + t.col0 + n;
                    *dst += g.alpha * (double)acc[x][y][v];
                }
            }
        }
}

// gamma[Q, i*nvir + a] += alpha * sum_s partial[s][Q][a], splits in order.
__global__ __launch_bounds__(256) void gemm_splits_kernel(GemmArgs g) {
    const int64_t total = (int64_t)g.naux * g.nvir;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < total;
         k += (int64_t)gridDim.x * 256) {
        const int64_t q = k / g.nvir;
        const int a = (int)(k % g.nvir);
        double s = 0.0;
        for (int sp = 0; sp < g.splits; ++sp)
            s += g.partial[((int64_t)sp * g.m_pad + q) * g.nvir_pad + a];
        g.gamma[q * g.nov + (int64_t)g.i * g.nvir + a] += g.alpha * s;
    }
}

template <typename T>
struct RowOfG {
    __device__ int operator()(int g, int v) const { return out_row<T>(g, v); }
};

// FP64: v_mfma_f64_16x16x4f64; A rows = Q of b2, B rows = virtuals of the T
// block; k contiguous in both, staged through LDS in panels of 8.
template <int TILE>
__global__ __launch_bounds__(256) void gemm_f64_kernel(GemmArgs g) {
    constexpr int KP = 8, TW = TILE / 2, NB = TW / 16, EPC = 2, CPR = KP / EPC;
    constexpr int OPCH = TILE * CPR, NCH = 2 * OPCH, CPT = (NCH + 255) / 256, LDR = KP + EPC;
    constexpr int KPL = KP / 4;
    typedef double frag_t __attribute__((ext_vector_type(KPL)));
    __shared__ __attribute__((aligned(16))) double lds[2][2][TILE][LDR];

    const GemmTile t = decode_gemm<TILE>(g, (int64_t)blockIdx.x);
    const double *__restrict__ A = (const double *)g.b2;
    const double *__restrict__ Bm = (const double *)t.bm;
    const int tid = threadIdx.x;
    const int limB = t.nrows_b - t.n0;
    v4i stg[CPT];
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int c = tid + 256 * u;
            const int op = c / OPCH, rem = c % OPCH;
            const int row = rem / CPR, ch = rem % CPR;
            v4i v = {0, 0, 0, 0};
            if (c < NCH) {
                if (op == 0)
                    v = *(const v4i *)(A + (int64_t)(t.m0 + row) * g.ld_b2 + t.a_col + k0 + ch * EPC);
                else if (row < limB)
                    v = *(const v4i *)(Bm + (int64_t)(t.n0 + row) * t.ldb + k0 + ch * EPC);
            }
            stg[u] = v;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int c = tid + 256 * u;
            if (c < NCH) {
                const int op = c / OPCH, rem = c % OPCH;
                const int row = rem / CPR, ch = rem % CPR;
                *(v4i *)&lds[buf][op][row][ch * EPC] = stg[u];
            }
        }
    };
    const int lane = tid % kWave, wid = tid / kWave;
    const int wr = wid / 2, wc = wid % 2, r = lane % 16, gg = lane / 16;
    bool cv[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) cv[k] = t.n0 + wc * TW + k * 16 < t.nrows_b;
    v4d acc[NB][NB];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) acc[x][y] = (v4d){0, 0, 0, 0};
    auto compute = [&](int buf) {
        frag_t fa[NB], fb[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            fa[k] = *(const frag_t *)&lds[buf][0][wr * TW + k * 16 + r][gg * KPL];
            fb[k] = *(const frag_t *)&lds[buf][1][wc * TW + k * 16 + r][gg * KPL];
        }
#pragma unroll
        for (int s = 0; s < KPL; ++s)
#pragma unroll
            for (int x = 0; x < NB; ++x)
#pragma unroll
                for (int y = 0; y < NB; ++y)
                    if (cv[y])
                        acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[x][s], fb[y][s],
                                                                         acc[x][y], 0, 0, 0);
    };
    const int np = (t.kend - t.kbeg) / KP;
    if (np > 0) {  // an empty split (kbeg == K) loads nothing
        gload(t.kbeg);
        lstore(0);
    }
    __syncthreads();
    for (int p = 0; p < np; ++p) {
        const int cur = p & 1;
        if (p + 1 < np) gload(t.kbeg + (p + 1) * KP);
        compute(cur);
        if (p + 1 < np) lstore(cur ^ 1);
        __syncthreads();
    }
    gemm_store<TILE>(g, t, wr, wc, r, gg, acc, RowOfG<double>{});
}

// BF16x3 from FP32 operands: each 4-float global chunk is split into BF16
// hi = RN(x), lo = RN(x - hi) while it is written to LDS (rows of 32 hi then
// 32 lo, 16-byte slots swizzled by row), then C += hi*hi + hi*lo + lo*hi with
// v_mfma_f32_16x16x32_bf16 in FP32.
__device__ __forceinline__ void split4(v4i in, uint32_t &h01, uint32_t &h23, uint32_t &l01,
                                       uint32_t &l23) {
    uint16_t h[4], l[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        // The element goes through a scalar first: with the ROCm 7.2 clang,
        // __builtin_bit_cast applied directly to an ext_vector element reads
        // element 0 for every e (every chunk then held four copies of its
        // first value).
        const int bits = in[e];
        const float x = __builtin_bit_cast(float, bits);
        const __bf16 hb = (__bf16)x;
        const __bf16 lb = (__bf16)(x - (float)hb);
        h[e] = __builtin_bit_cast(uint16_t, hb);
        l[e] = __builtin_bit_cast(uint16_t, lb);
    }
    h01 = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
    h23 = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
    l01 = (uint32_t)l[0] | ((uint32_t)l[1] << 16);
    l23 = (uint32_t)l[2] | ((uint32_t)l[3] << 16);
}

template <int TILE>
__global__ __launch_bounds__(256) void gemm_bf16x3_kernel(GemmArgs g) {
    constexpr int KP = 32, TW = TILE / 2, NB = TW / 16;
    constexpr int CPR = 8;  // 16-byte float chunks per row per panel (32 floats)
    constexpr int OPCH = TILE * CPR, NCH = 2 * OPCH, CPT = (NCH + 255) / 256;
    __shared__ __attribute__((aligned(16))) v4i lds[2][2][TILE][8];

    const GemmTile t = decode_gemm<TILE>(g, (int64_t)blockIdx.x);
    const float *__restrict__ A = (const float *)g.b2;
    const float *__restrict__ Bm = (const float *)t.bm;
    const int tid = threadIdx.x;
    const int limB = t.nrows_b - t.n0;
    v4i stg[CPT];
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int c = tid + 256 * u;
            const int op = c / OPCH, rem = c % OPCH;
            const int row = rem / CPR, ch = rem % CPR;
            v4i v = {0, 0, 0, 0};
            if (c < NCH) {
                if (op == 0)
                    v = *(const v4i *)(A + (int64_t)(t.m0 + row) * g.ld_b2 + t.a_col + k0 + ch * 4);
                else if (row < limB)
                    v = *(const v4i *)(Bm + (int64_t)(t.n0 + row) * t.ldb + k0 + ch * 4);
            }
            stg[u] = v;
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < CPT; ++u) {
            const int c = tid + 256 * u;
            if (c < NCH) {
                const int op = c / OPCH, rem = c % OPCH;
                const int row = rem / CPR, ch = rem % CPR;
                uint32_t h01, h23, l01, l23;
                split4(stg[u], h01, h23, l01, l23);
                // floats k = 4ch..4ch+3 -> hi chunk ch/2, lo chunk 4 + ch/2, half ch%2
                uint32_t *hp = (uint32_t *)&lds[buf][op][row][(ch >> 1) ^ (row & 7)];
                uint32_t *lp = (uint32_t *)&lds[buf][op][row][(4 + (ch >> 1)) ^ (row & 7)];
                const int h = (ch & 1) * 2;
                hp[h] = h01;
                hp[h + 1] = h23;
                lp[h] = l01;
                lp[h + 1] = l23;
            }
        }
    };
    const int lane = tid % kWave, wid = tid / kWave;
    const int wr = wid / 2, wc = wid % 2, r = lane % 16, gg = lane / 16;
    bool cv[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) cv[k] = t.n0 + wc * TW + k * 16 < t.nrows_b;
    v4f acc[NB][NB];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) acc[x][y] = (v4f){0, 0, 0, 0};
    auto frag = [&](int buf, int op, int row, int chunk) -> v8bf {
        return __builtin_bit_cast(v8bf, lds[buf][op][row][chunk ^ (row & 7)]);
    };
    auto compute = [&](int buf) {
        v8bf ha[NB], la[NB], hb[NB], lb[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int ra = wr * TW + k * 16 + r, rb = wc * TW + k * 16 + r;
            ha[k] = frag(buf, 0, ra, gg);
            la[k] = frag(buf, 0, ra, 4 + gg);
            hb[k] = frag(buf, 1, rb, gg);
            lb[k] = frag(buf, 1, rb, 4 + gg);
        }
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
                if (cv[y]) {
                    v4f c = acc[x][y];
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(la[x], hb[y], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ha[x], lb[y], c, 0, 0, 0);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ha[x], hb[y], c, 0, 0, 0);
                    acc[x][y] = c;
                }
    };
    const int np = (t.kend - t.kbeg) / KP;
    if (np > 0) {  // an empty split (kbeg == K) loads nothing
        gload(t.kbeg);
        lstore(0);
    }
    __syncthreads();
    for (int p = 0; p < np; ++p) {
        const int cur = p & 1;
        if (p + 1 < np) gload(t.kbeg + (p + 1) * KP);
        compute(cur);
        if (p + 1 < np) lstore(cur ^ 1);
        __syncthreads();
    }
    gemm_store<TILE>(g, t, wr, wc, r, gg, acc, RowOfG<float>{});
}

// ---- BF16x3 NT GEMM on pre-split operands (the INT8X4 adjoint's Gamma) ------
// 256 x 256 output tile per workgroup of eight waves (2 along M x 4 along N,
// each 128 x 64 = 4 x 2 blocks of v_mfma_f32_32x32x16_bf16). At this tile a
// 32-deep K panel is 64 KB of operands (A and B, hi and lo) for 12.6 MFLOP of
// MFMA work, about 190 FLOP per byte: the operand stream stays well inside the
// L2 bandwidth at the BF16 matrix rate. Per panel the workgroup stages the
// four operand slices into LDS (double-buffered, one barrier per panel); each
// LDS row holds 32 hi then 32 lo values as eight 16-byte chunks, chunk c at
// slot c ^ ((row >> 1) & 7): the 16 rows one 16-lane phase of a fragment read
// touches then cover all 16 four-bank groups (rows 128 bytes apart share a
// bank group every other row, so the slots of the even and of the odd rows
// must each be distinct). Per block and 16-deep step: C += lo*hi + hi*lo + hi*hi
// in FP32. Rows past M or N load zeros and store nothing. Tiles are walked m
// fastest after the XCD remap.
//
// 32x32x16 lane layout (checked on the device, tests/test_gpu_layout.py):
// lane l = 32 h + c supplies A[row c][k = 8 h .. 8 h + 7] and B[k = 8 h ..][col c];
// result element j of lane l is C[8 (j / 4) + 4 h + j % 4][c].
typedef float v16f __attribute__((ext_vector_type(16)));

template <int WN>
__global__ __launch_bounds__(128 * WN) void bf16x3_nt_kernel(Bf16x3Gemm g) {
    constexpr int T = 256, KP = 32;
    constexpr int NT = 128 * WN;        // threads: 2 waves along M x WN along N
    constexpr int WNE = T / WN;         // N extent of one wave
    constexpr int NY = WNE / 32;        // 32-wide blocks along N per wave
    constexpr int RP = NT / 4;          // tile rows one load instruction covers
    constexpr int LPO = 2 * (T / RP);   // loads per thread per operand (hi and lo)
    __shared__ __attribute__((aligned(16))) v4i lds[2][2][T][8];
    const int mt = (g.M + T - 1) / T, nt = (g.N + T - 1) / T;
    const int64_t tile = xcd_remap((int64_t)blockIdx.x, (int64_t)mt * nt);
    const int m0 = (int)(tile % mt) * T, n0 = (int)(tile / mt) * T;
    const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
    const int wm = wid / WN, wn = wid % WN, c = lane % 32, h = lane / 32;

    // Load u of this thread: operand u / LPO (A, then B), hi (u even) or lo
    // (u odd), row (tid >> 2) + RP ((u % LPO) >> 1) of the tile, piece
    // (u & 1) * 4 + (tid & 3) of its 64-byte LDS row (8 consecutive k each).
    // Every load instruction reads one array through a raw buffer resource
    // over the tile's rows that exist, so rows past M or N read as zeros
    // without a branch; voff is the byte offset of the row's piece.
    const int crow = tid >> 2, cpiece = tid & 3;
    const __amdgpu_buffer_rsrc_t rs[4] = {
        raw_rsrc(g.ahi + (int64_t)m0 * g.lda, (int64_t)std::min(T, g.M - m0) * g.lda * 2),
        raw_rsrc(g.alo + (int64_t)m0 * g.lda, (int64_t)std::min(T, g.M - m0) * g.lda * 2),
        raw_rsrc(g.bhi + (int64_t)n0 * g.ldb, (int64_t)std::min(T, g.N - n0) * g.ldb * 2),
        raw_rsrc(g.blo + (int64_t)n0 * g.ldb, (int64_t)std::min(T, g.N - n0) * g.ldb * 2)};
    int voff[2 * LPO];
#pragma unroll
    for (int u = 0; u < 2 * LPO; ++u) {
        const int row = crow + RP * ((u % LPO) >> 1);
        voff[u] = (int)(((int64_t)row * (u < LPO ? g.lda : g.ldb) + cpiece * 8) * 2);
    }
    v4i stg[2 * LPO];
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 2 * LPO; ++u)
            stg[u] = __builtin_bit_cast(
                v4i, __builtin_amdgcn_raw_buffer_load_b128(rs[(u / LPO) * 2 + (u & 1)],
                                                           voff[u] + 2 * k0, 0, 0));
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 2 * LPO; ++u) {
            const int row = crow + RP * ((u % LPO) >> 1), piece = (u & 1) * 4 + cpiece;
            lds[buf][u / LPO][row][piece ^ ((row >> 1) & 7)] = stg[u];
        }
    };
    v16f acc[4][NY];
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < NY; ++y)
#pragma unroll
            for (int j = 0; j < 16; ++j) acc[x][y][j] = 0.0f;
    auto frag = [&](int buf, int op, int row, int chunk) -> v8bf {
        return __builtin_bit_cast(v8bf, lds[buf][op][row][chunk ^ ((row >> 1) & 7)]);
    };
    auto compute = [&](int buf) {
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            v8bf ha[4], la[4], hb[NY], lb[NY];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                const int ra = wm * 128 + x * 32 + c;
                ha[x] = frag(buf, 0, ra, 2 * kk + h);
                la[x] = frag(buf, 0, ra, 4 + 2 * kk + h);
            }
#pragma unroll
            for (int y = 0; y < NY; ++y) {
                const int rb = wn * WNE + y * 32 + c;
                hb[y] = frag(buf, 1, rb, 2 * kk + h);
                lb[y] = frag(buf, 1, rb, 4 + 2 * kk + h);
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < NY; ++y) {
                    v16f a = acc[x][y];
                    a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(la[x], hb[y], a, 0, 0, 0);
                    a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha[x], lb[y], a, 0, 0, 0);
                    a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ha[x], hb[y], a, 0, 0, 0);
                    acc[x][y] = a;
                }
        }
    };
    const int np = g.K / KP;
    if (np > 0) {
        gload(0);
        lstore(0);
    }
    __syncthreads();
    for (int p = 0; p < np; ++p) {
        const int cur = p & 1;
        if (p + 1 < np) gload((p + 1) * KP);
        compute(cur);
        if (p + 1 < np) lstore(cur ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < NY; ++y) {
            const int n = n0 + wn * WNE + y * 32 + c;
            if (n >= g.N) continue;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int m = m0 + wm * 128 + x * 32 + 8 * (j / 4) + 4 * h + j % 4;
                if (m < g.M) g.c[(int64_t)m * g.ldc + n] += g.alpha * acc[x][y][j];
            }
        }
}

// ---- Row reduction of one batch: energy and dE/deps -------------------------
struct AdjRed {
    int i, nocc, nvir, tile, ntile, ntp, vec_stride;
    int nchunks, jc;              // pair chunks of the virtual sums (jc pairs each)
    const double *item_vec;       // (nocc - i) * ntp items
    const double *item_partials;  // 4 per item
    double *vpart;                // (nchunks, nvir, 2) virtual partial sums of this row
    double *ppart;                // (npair, 6) per-pair sums of this row
    double *rows;                 // 4 per occupied row
    double *dvir_s, *dvir_c;      // dE/deps_vir, compensated, accumulated over batches
    double *docc_s, *docc_c;      // dE/deps_occ
};

// Reduction of one row batch, in two launches with a fixed order:
//   adjoint_reduce_kernel: blocks [0, nvb*nchunks) sum, for 256 virtual
//     indices each, the t*T~ row and column sums of one chunk of jc pairs
//     (f = 1 for j = i, else 2) into vpart; blocks from nvb*nchunks on take one
//     pair each and sum its items' energy partials and t*T~ totals into ppart
//     (strided per thread, then a fixed tree);
//   adjoint_fold_kernel: adds the chunks into dE/deps_vir in chunk order, and
//     one thread folds the pairs in j order into the row's energy and the
//     occupied derivatives (dE/deps_i and dE/deps_j: -2 sum_ab t T~ per pair).
__global__ __launch_bounds__(256) void adjoint_reduce_kernel(AdjRed r) {
    const int npair = r.nocc - r.i;
    const int nvb = (r.nvir + 255) / 256;
    const int b = blockIdx.x;
    if (b < nvb * r.nchunks) {
        const int ch = b / nvb;
        const int c = (b % nvb) * 256 + threadIdx.x;
        if (c >= r.nvir) return;
        const int tc = c / r.tile, off = c % r.tile;
        auto tp_of = [&](int ta, int tb) { return encode_tile_pair(ta, tb, r.ntile); };
        KSum s;
        const int j1 = min(npair, (ch + 1) * r.jc);
        for (int jj = ch * r.jc; jj < j1; ++jj) {
            const double f = jj == 0 ? 1.0 : 2.0;
            const double *base = r.item_vec + (int64_t)jj * r.ntp * r.vec_stride;
            // tile pairs touching tile tc: (tc, u >= tc) by rows, (u <= tc, tc) by columns
            for (int u = 0; u < r.ntile; ++u) {
                if (u >= tc) s.add(f * base[(int64_t)tp_of(tc, u) * r.vec_stride + off]);
                if (u <= tc) s.add(f * base[(int64_t)tp_of(u, tc) * r.vec_stride + r.tile + off]);
            }
        }
        double *o = r.vpart + 2 * ((int64_t)ch * r.nvir + c);
        o[0] = s.s;
        o[1] = s.c;
        return;
    }
    const int jj = b - nvb * r.nchunks;
    if (jj >= npair) return;
    KSum eos, ess, sp;
    for (int t = threadIdx.x; t < r.ntp; t += 256) {
        const int64_t item = (int64_t)jj * r.ntp + t;
        const double *ip = r.item_partials + 4 * item;
        eos.merge(ip[0], ip[1]);
        ess.merge(ip[2], ip[3]);
        sp.add(r.item_vec[item * r.vec_stride + 2 * r.tile]);
    }
    wave_merge(eos);
    wave_merge(ess);
    wave_merge(sp);
    __shared__ double red[4][6];
    const int wid = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    if (lane == 0) {
        red[wid][0] = eos.s;
        red[wid][1] = eos.c;
        red[wid][2] = ess.s;
        red[wid][3] = ess.c;
        red[wid][4] = sp.s;
        red[wid][5] = sp.c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        KSum a, e, p;
        for (int w = 0; w < 4; ++w) {
            a.merge(red[w][0], red[w][1]);
            e.merge(red[w][2], red[w][3]);
            p.merge(red[w][4], red[w][5]);
        }
        double *o = r.ppart + 6 * (int64_t)jj;
        o[0] = a.s;
        o[1] = a.c;
        o[2] = e.s;
        o[3] = e.c;
        o[4] = p.s;
        o[5] = p.c;
    }
}

__global__ __launch_bounds__(256) void adjoint_fold_kernel(AdjRed r) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c < r.nvir) {
        KSum s;
        s.s = r.dvir_s[c];
        s.c = r.dvir_c[c];
        for (int ch = 0; ch < r.nchunks; ++ch) {
            const double *o = r.vpart + 2 * ((int64_t)ch * r.nvir + c);
            s.merge(o[0], o[1]);
        }
        r.dvir_s[c] = s.s;
        r.dvir_c[c] = s.c;
    }
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int npair = r.nocc - r.i;
    KSum eos, ess, di;
    di.s = r.docc_s[r.i];
    di.c = r.docc_c[r.i];
    for (int jj = 0; jj < npair; ++jj) {
        const double *o = r.ppart + 6 * (int64_t)jj;
        eos.merge(o[0], o[1]);
        ess.merge(o[2], o[3]);
        const double spair = o[4] + o[5];
        di.add(-2.0 * spair);
        if (jj > 0) {
            const int j = r.i + jj;
            KSum dj;
            dj.s = r.docc_s[j];
            dj.c = r.docc_c[j];
            dj.add(-2.0 * spair);
            r.docc_s[j] = dj.s;
            r.docc_c[j] = dj.c;
        }
    }
    r.docc_s[r.i] = di.s;
    r.docc_c[r.i] = di.c;
    r.rows[4 * r.i + 0] = eos.s;
    r.rows[4 * r.i + 1] = eos.c;
    r.rows[4 * r.i + 2] = ess.s;
    r.rows[4 * r.i + 3] = ess.c;
}

__global__ void adjoint_finish_kernel(const double *__restrict__ rows, int nocc, int nvir,
                                      const double *__restrict__ docc_s,
                                      const double *__restrict__ docc_c,
                                      const double *__restrict__ dvir_s,
                                      const double *__restrict__ dvir_c,
                                      double *__restrict__ result) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) {
        KSum os, ss;
        for (int i = 0; i < nocc; ++i) {
            os.merge(rows[4 * i], rows[4 * i + 1]);
            ss.merge(rows[4 * i + 2], rows[4 * i + 3]);
        }
        result[0] = os.value();
        result[1] = ss.value();
    }
    if (k < nocc) result[2 + k] = docc_s[k] + docc_c[k];
    if (k < nvir) result[2 + nocc + k] = dvir_s[k] + dvir_c[k];
}

// BF16 hi/lo split of FP32 values (hi = RN(x), lo = RN(x - hi), the split of
// pack.hip) over rows r < nr of cols elements each, row stride ld (in and out).
__global__ __launch_bounds__(256) void split_bf16_kernel(const float *__restrict__ x, int64_t nr,
                                                         int64_t cols, int64_t ld,
                                                         uint16_t *__restrict__ hi,
                                                         uint16_t *__restrict__ lo) {
    const int64_t n = nr * cols;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256) {
        const int64_t idx = (k / cols) * ld + k % cols;
        const float v = x[idx];
        const __bf16 h = (__bf16)v;
        const __bf16 l = (__bf16)(v - (float)h);
        hi[idx] = __builtin_bit_cast(uint16_t, h);
        lo[idx] = __builtin_bit_cast(uint16_t, l);
    }
}

// gamma[Q][i*nvir + a] = x[Q][i*nvir_pad + a] (FP32 over the padded index
// space -> FP64 over the reference layout).
__global__ __launch_bounds__(256) void widen_kernel(const float *__restrict__ x, int naux,
                                                    int nocc, int nvir, int nvir_pad,
                                                    double *__restrict__ y) {
    const int64_t nov = (int64_t)nocc * nvir, rows = (int64_t)nocc * nvir_pad;
    const int64_t n = (int64_t)naux * nov;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < n;
         k += (int64_t)gridDim.x * 256) {
        const int64_t q = k / nov, c = k % nov;
        y[k] = (double)x[q * rows + (c / nvir) * nvir_pad + c % nvir];
    }
}

void launch_split(const float *x, int64_t nr, int64_t cols, int64_t ld, uint16_t *hi,
                  uint16_t *lo, hipStream_t st) {
    const int64_t n = nr * cols;
    if (n <= 0) return;
    const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(split_bf16_kernel, dim3(blocks), dim3(256), 0, st, x, nr, cols, ld, hi, lo);
}

template <typename T>
void launch_fill(T *p, int64_t n, T v, hipStream_t st) {
    int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(fill_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
}

}  // namespace

hipError_t launch_bf16x3_nt(const Bf16x3Gemm &g, hipStream_t st) {
    if (g.M <= 0 || g.N <= 0 || g.K < 0 || g.K % 32 != 0) return hipErrorInvalidValue;
    // Byte offsets inside a 256-row tile are 32-bit (raw buffer loads).
    if ((int64_t)256 * std::max(g.lda, g.ldb) * 2 >= 0x7fffffffLL) return hipErrorInvalidValue;
    const int64_t tiles = (int64_t)((g.M + 255) / 256) * ((g.N + 255) / 256);
    if (tiles > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(bf16x3_nt_kernel<4>, dim3((unsigned)tiles), dim3(512), 0, st, g);
    return hipGetLastError();
}

hipError_t launch_adjoint_gemm(const GemmArgs &g, hipStream_t st) {
    const int tm = g.m_pad / g.tile, tn = (g.nvir_pad + g.tile - 1) / g.tile;
    const int64_t nb = (int64_t)tm * tn;
    const int64_t nblocks = nb * (int64_t)(g.splits + g.nocc - g.i - 1);
    if (nblocks <= 0 || nblocks > 0x7fffffffLL) return hipErrorInvalidValue;
    if (g.prec == Prec::F64) {
        if (g.tile == 128)
            hipLaunchKernelGGL(gemm_f64_kernel<128>, dim3((unsigned)nblocks), dim3(256), 0, st, g);
        else
            hipLaunchKernelGGL(gemm_f64_kernel<64>, dim3((unsigned)nblocks), dim3(256), 0, st, g);
    } else if (g.prec == Prec::BF16X3) {
        if (g.tile == 128)
            hipLaunchKernelGGL(gemm_bf16x3_kernel<128>, dim3((unsigned)nblocks), dim3(256), 0, st,
                               g);
        else
            hipLaunchKernelGGL(gemm_bf16x3_kernel<64>, dim3((unsigned)nblocks), dim3(256), 0, st,
                               g);
    } else {
        return hipErrorInvalidValue;
    }
    const int64_t total = (int64_t)g.naux * g.nvir;
    const unsigned rb = (unsigned)std::min<int64_t>((total + 255) / 256, 8192);
    hipLaunchKernelGGL(gemm_splits_kernel, dim3(rb), dim3(256), 0, st, g);
    return hipGetLastError();
}

// ---- Host orchestration ------------------------------------------------------

// Output tile of the phase-B/C GEMMs: 64 (several workgroups per CU) for both
// precisions -- at 100/900/3600 the energy + adjoint takes 1932 ms (FP64) and
// 692 ms (BF16x3) with 64-wide tiles against 2194 and 837 ms with 128-wide
// ones, whose FP64 kernel holds one wave per SIMD. FERRIC_HIP_GEMM_TILE (64 or
// 128) overrides.
static int gemm_tile(Prec prec) {
    (void)prec;
    if (const char *e = std::getenv("FERRIC_HIP_GEMM_TILE")) {
        const int t = std::atoi(e);
        if (t == 64 || t == 128) return t;
    }
    return 64;
}

// Occupied rows per group of the INT8X4 adjoint (FERRIC_I8_ADJ_ROWS overrides).
static int i8_adjoint_rows(size_t nocc) {
    int r = 8;
    if (const char *e = std::getenv("FERRIC_I8_ADJ_ROWS")) {
        const int v = std::atoi(e);
        if (v > 0) r = v;
    }
    return (int)std::min<size_t>(nocc, (size_t)r);
}

static fh_status adjoint_core_i8(State &s, const void *b_dev, fh_dtype b_dtype, size_t naux,
                                 size_t nocc, size_t nvir, const double *eps, size_t n_eps,
                                 size_t first_occ, size_t nocc_total, double c_os, double c_ss,
                                 hipStream_t stream, fh_spin_components *out, double *gamma_dev,
                                 double *deps) {
    const Prec prec = Prec::I8X4;
    const bool in_f32 = b_dtype == FH_DT_F32;
    EnergyPlan pl = plan_energy(nocc, nvir, naux, prec);
    if (pl.naux_pad > kI8MaxNauxPad)
        return fail(FH_RANGE, "naux_pad %d beyond the INT8 slices' int32 range (%d)", pl.naux_pad,
                    kI8MaxNauxPad);
    // GEMM index space: virtuals padded to the 32-deep K panels of the Gamma GEMMs.
    const int vpg = (int)((nvir + 31) / 32 * 32);
    const size_t ncol = nocc * (size_t)vpg;  // (i, a) columns of B2, S and gamma32
    if (ncol > (size_t)0x7fffffff || naux > (size_t)0x7fffffff)
        return fail(FH_UNSUPPORTED, "nocc * vpg = %zu beyond 32-bit GEMM sizes", ncol);
    const int R = i8_adjoint_rows(nocc);
    const size_t slab = (size_t)R * vpg * ncol;  // elements of S (and of St)
    const size_t rows8 = nocc * (size_t)pl.nvir_pad;
    const int vec_stride = 2 * pl.tile + 2;
    const size_t items = (size_t)pl.nitems;
    const int nvb = (int)((nvir + 255) / 256);
    const int nchunks_max = (int)std::min<size_t>(nocc, (size_t)std::max(1, (512 + nvb - 1) / nvb));
    std::vector<std::pair<int, size_t>> req = {
        {kSlotOp0, packed_op_bytes(pl, nocc, prec)},
        {kSlotScale, sizeof(int) * rows8},
        {kSlotStage2, sizeof(float) * naux * ncol},
        {kSlotB2Split, 2 * sizeof(uint16_t) * naux * ncol},
        {kSlotTSplit, 4 * sizeof(uint16_t) * slab},
        {kSlotG32, sizeof(float) * naux * ncol},
        {kSlotEps, sizeof(double) * (nocc + nvir)},
        {kSlotItems, sizeof(double) * 4 * items},
        {kSlotDeps, sizeof(double) * vec_stride * items},
        {kSlotRows, sizeof(double) * 4 * nocc},
        {kSlotAcc, sizeof(double) * 2 * (nocc + nvir)},
        {kSlotResult, sizeof(double) * (2 + nocc + nvir)},
        {kSlotScalar, 256},
        {kSlotRed, sizeof(double) * (2 * (size_t)nchunks_max * nvir + 6 * nocc)},
    };
    fh_status st = reserve(s, req);
    if (st != FH_OK) return st;
    // B enters the GEMMs as FP32 split into BF16: the FP32 range check applies.
    st = range_check(s, b_dev, in_f32, naux * nocc * nvir, naux, stream);
    if (st != FH_OK) return st;
    std::vector<double> e(nocc + nvir);
    for (size_t i = 0; i < nocc; ++i) e[i] = eps[first_occ + i];
    for (size_t a = 0; a < nvir; ++a) e[nocc + a] = eps[nocc_total + a];
    double *deps_dev = (double *)s.slots[kSlotEps].ptr;
    st = hip_check(hipMemcpyAsync(deps_dev, e.data(), sizeof(double) * e.size(),
                                  hipMemcpyHostToDevice, stream),
                   "H2D eps");
    if (st != FH_OK) return st;
    s.pending.h2d_bytes += sizeof(double) * e.size();

    mark(s, 0, stream);
    PackArgs pa;
    pa.b = b_dev;
    pa.in_f32 = in_f32;
    pa.naux = (int)naux;
    pa.nocc = (int)nocc;
    pa.nvir = (int)nvir;
    pa.nvir_pad = pl.nvir_pad;
    pa.naux_pad = pl.naux_pad;
    pa.fmt = PackFmt::I8X4;
    pa.out0 = s.slots[kSlotOp0].ptr;
    pa.out1 = nullptr;
    pa.scale = nullptr;
    pa.lo_sq = nullptr;
    pa.rexp = (int *)s.slots[kSlotScale].ptr;
    hipError_t he = launch_pack(pa, stream);
    if (he != hipSuccess) return fail(FH_KERNEL, "pack kernel: %s", hipGetErrorString(he));
    float *b2 = (float *)s.slots[kSlotStage2].ptr;
    {
        const int64_t total = (int64_t)naux * (int64_t)ncol;
        const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 16384);
        if (in_f32)
            hipLaunchKernelGGL((pack_qmajor_kernel<float, float>), dim3(blocks), dim3(256), 0,
                               stream, (const float *)b_dev, (int)naux, (int)nocc, (int)nvir, vpg,
                               (int)naux, b2);
        else
            hipLaunchKernelGGL((pack_qmajor_kernel<double, float>), dim3(blocks), dim3(256), 0,
                               stream, (const double *)b_dev, (int)naux, (int)nocc, (int)nvir,
                               vpg, (int)naux, b2);
    }
    uint16_t *b2hi = (uint16_t *)s.slots[kSlotB2Split].ptr, *b2lo = b2hi + naux * ncol;
    uint16_t *shi = (uint16_t *)s.slots[kSlotTSplit].ptr, *slo = shi + slab;
    uint16_t *sthi = slo + slab, *stlo = sthi + slab;
    float *x32 = (float *)s.slots[kSlotG32].ptr;
    launch_split(b2, (int64_t)naux, (int64_t)ncol, (int64_t)ncol, b2hi, b2lo, stream);
    double *acc = (double *)s.slots[kSlotAcc].ptr;
    launch_fill<double>(acc, (int64_t)(2 * (nocc + nvir)), 0.0, stream);
    launch_fill<float>(x32, (int64_t)(naux * ncol), 0.0f, stream);
    // The slabs' padding columns (nvir_pad <= y < vpg) are never written by
    // phase A and must read as zero: clear both slabs once.
    launch_fill<uint64_t>((uint64_t *)shi, (int64_t)(slab), (uint64_t)0, stream);
    he = hipGetLastError();
    if (he != hipSuccess) return fail(FH_KERNEL, "adjoint set-up kernels: %s", hipGetErrorString(he));
    mark(s, 1, stream);

    EnergyArgs ea;
    ea.prec = prec;
    ea.tile = pl.tile;
    ea.op0 = pa.out0;
    ea.op1 = nullptr;
    ea.scale = nullptr;
    ea.diag_corr = nullptr;
    ea.rexp = pa.rexp;
    ea.q_rot_ticks = i8_q_rot_ticks(s, pl.naux_pad / 64, pl.ntp);
    ea.nocc = (int)nocc;
    ea.nvir = (int)nvir;
    ea.nvir_pad = pl.nvir_pad;
    ea.naux_pad = pl.naux_pad;
    ea.nqb = 0;
    ea.eps_occ = deps_dev;
    ea.eps_vir = deps_dev + nocc;
    ea.kappa = NAN;
    ea.ntile = pl.ntile;
    ea.ntp = pl.ntp;
    ea.item_partials = (double *)s.slots[kSlotItems].ptr;
    ea.row_partials = (double *)s.slots[kSlotRows].ptr;
    ea.result = (double *)s.slots[kSlotResult].ptr;
    ea.c_os = c_os;
    ea.c_ss = c_ss;
    ea.item_vec = (double *)s.slots[kSlotDeps].ptr;
    ea.vec_stride = vec_stride;
    ea.i_fixed = -1;
    ea.s_hi = shi;
    ea.s_lo = slo;
    ea.st_hi = sthi;
    ea.st_lo = stlo;
    ea.vpg = vpg;
    ea.ld_s = (int64_t)ncol;

    AdjRed rd;
    rd.nocc = (int)nocc;
    rd.nvir = (int)nvir;
    rd.tile = pl.tile;
    rd.ntile = pl.ntile;
    rd.ntp = pl.ntp;
    rd.vec_stride = vec_stride;
    rd.rows = ea.row_partials;
    rd.docc_s = acc;
    rd.docc_c = acc + nocc;
    rd.dvir_s = acc + 2 * nocc;
    rd.dvir_c = acc + 2 * nocc + nvir;
    rd.vpart = (double *)s.slots[kSlotRed].ptr;
    rd.ppart = rd.vpart + 2 * (size_t)nchunks_max * nvir;
    auto first_item = [&](int64_t i) {
        return (i * (int64_t)nocc - i * (i - 1) / 2) * pl.ntp;
    };

    // One group of R occupied rows at a time: phase A writes its T~ slabs, the
    // per-row reductions follow, then
    //   GEMM 1: gamma[Q, (i, a)] += 4 sum_{(j, b), j >= g0} B[Q, (j, b)] S[(i, a), (j, b)]
    //   GEMM 2: gamma[Q, (j, b)] += 4 sum_{(i, a) in group} B[Q, (i, a)] St[(j, b), (i, a)],
    //           j >= g1
    // (T~ symmetric: every ordered occupied pair enters gamma once).
    for (int g0 = 0; g0 < (int)nocc; g0 += R) {
        const int g1 = std::min<int>((int)nocc, g0 + R);
        ea.g0 = g0;
        ea.g1 = g1;
        ea.ld_st = (int64_t)(g1 - g0) * vpg;
        ea.item0 = first_item(g0);
        ea.nitems = first_item(g1) - ea.item0;
        he = launch_adjoint_pairs(ea, stream);
        if (he != hipSuccess)
            return fail(FH_KERNEL, "adjoint pair kernel: %s", hipGetErrorString(he));
        for (int i = g0; i < g1; ++i) {
            const int64_t first = first_item(i);
            rd.i = i;
            rd.item_partials = ea.item_partials + 4 * first;
            rd.item_vec = ea.item_vec + first * vec_stride;
            const int npair = (int)nocc - i;
            rd.nchunks = std::min(npair, nchunks_max);
            rd.jc = (npair + rd.nchunks - 1) / rd.nchunks;
            rd.nchunks = (npair + rd.jc - 1) / rd.jc;
            hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((unsigned)(nvb * rd.nchunks + npair)),
                               dim3(256), 0, stream, rd);
            hipLaunchKernelGGL(adjoint_fold_kernel, dim3((unsigned)nvb), dim3(256), 0, stream, rd);
        }
        Bf16x3Gemm g;
        g.ahi = b2hi + (size_t)g0 * vpg;
        g.alo = b2lo + (size_t)g0 * vpg;
        g.lda = (int64_t)ncol;
        g.bhi = shi + (size_t)g0 * vpg;
        g.blo = slo + (size_t)g0 * vpg;
        g.ldb = (int64_t)ncol;
        g.c = x32 + (size_t)g0 * vpg;
        g.ldc = (int64_t)ncol;
        g.M = (int)naux;
        g.N = (g1 - g0) * vpg;
        g.K = ((int)nocc - g0) * vpg;
        g.alpha = 4.0f;
        he = launch_bf16x3_nt(g, stream);
        if (he == hipSuccess && g1 < (int)nocc) {
            g.bhi = sthi + (size_t)g1 * vpg * ea.ld_st;
            g.blo = stlo + (size_t)g1 * vpg * ea.ld_st;
            g.ldb = ea.ld_st;
            g.c = x32 + (size_t)g1 * vpg;
            g.N = ((int)nocc - g1) * vpg;
            g.K = (g1 - g0) * vpg;
            he = launch_bf16x3_nt(g, stream);
        }
        if (he != hipSuccess) return fail(FH_KERNEL, "adjoint GEMM: %s", hipGetErrorString(he));
    }
    {
        const int64_t n = (int64_t)naux * (int64_t)(nocc * nvir);
        const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 16384);
        hipLaunchKernelGGL(widen_kernel, dim3(blocks), dim3(256), 0, stream, x32, (int)naux,
                           (int)nocc, (int)nvir, vpg, gamma_dev);
    }
    mark(s, 2, stream);
    {
        const int n = (int)std::max(nocc, nvir);
        hipLaunchKernelGGL(adjoint_finish_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           stream, ea.row_partials, (int)nocc, (int)nvir, rd.docc_s, rd.docc_c,
                           rd.dvir_s, rd.dvir_c, ea.result);
    }
    he = hipGetLastError();
    if (he != hipSuccess) return fail(FH_KERNEL, "adjoint finish: %s", hipGetErrorString(he));
    mark(s, 3, stream);
    const int ngroups = (int)((nocc + R - 1) / R);
    s.pending.kernel_launches += 2 * nocc + 3 * ngroups + 8;
    std::vector<double> res(2 + nocc + nvir);
    st = hip_check(hipMemcpyAsync(res.data(), ea.result, sizeof(double) * res.size(),
                                  hipMemcpyDeviceToHost, stream),
                   "D2H adjoint result");
    if (st != FH_OK) return st;
    st = hip_check(hipStreamSynchronize(stream), "stream sync");
    if (st != FH_OK) return st;
    s.pending.d2h_bytes += sizeof(double) * res.size();
    finish_timing(s, (uint64_t)pl.nitems, pl.tile);
    out->e_os = res[0];
    out->e_ss = res[1];
    out->e_total = res[0] + res[1];
    if (deps) {
        for (size_t k = 0; k < n_eps; ++k) deps[k] = 0.0;
        for (size_t i = 0; i < nocc; ++i) deps[first_occ + i] = res[2 + i];
        for (size_t a = 0; a < nvir; ++a) deps[nocc_total + a] = res[2 + nocc + a];
    }
    return FH_OK;
}

static fh_status adjoint_core(State &s, const void *b_dev, fh_dtype b_dtype, size_t naux,
                              size_t nocc, size_t nvir, const double *eps, size_t n_eps,
                              size_t first_occ, size_t nocc_total, double c_os, double c_ss,
                              Prec prec, hipStream_t stream, fh_spin_components *out,
                              double *gamma_dev, double *deps) {
    const bool in_f32 = b_dtype == FH_DT_F32;
    const bool f64 = prec == Prec::F64;
    if (prec == Prec::I8X4)
        return adjoint_core_i8(s, b_dev, b_dtype, naux, nocc, nvir, eps, n_eps, first_occ,
                               nocc_total, c_os, c_ss, stream, out, gamma_dev, deps);
    const Prec gprec = f64 ? Prec::F64 : Prec::BF16X3;
    EnergyPlan pl = plan_energy(nocc, nvir, naux, prec);
    // The FP64 adjoint epilogue does not fit a 128-wide tile in the register
    // file (it would spill); 64-wide tiles run two workgroups per CU.
    if (f64 && pl.tile > 64) pl.tile = 64;
    // K dimensions of the phase-B/C GEMMs run over padded virtuals in panels of
    // 32 (BF16) or 8 (FP64): pad the virtual block to 32.
    pl.nvir_pad = (int)((nvir + 31) / 32 * 32);
    // Phase A launches one occupied row at a time: (nocc - i) pairs x ntp items.
    // When the average row holds fewer than 1024 items (four per CU), 32-wide
    // tiles fill the launches better (41/209/960: FP64 10.9 -> 9.0 ms, BF16x3
    // 5.0 -> 4.5 ms). An explicit FERRIC_HIP_TILE still decides.
    const char *tile_env = getenv("FERRIC_HIP_TILE");
    if (!(tile_env && *tile_env) && pl.tile > 32) {
        const size_t nt = (size_t)((pl.nvir_pad + pl.tile - 1) / pl.tile);
        if ((nocc + 1) / 2 * (nt * (nt + 1) / 2) < 1024) pl.tile = 32;
    }
    pl.ntile = (pl.nvir_pad + pl.tile - 1) / pl.tile;
    pl.ntp = pl.ntile * (pl.ntile + 1) / 2;
    const int gtile = gemm_tile(gprec);
    const int m_pad = (int)((naux + gtile - 1) / gtile * gtile);
    const size_t rows = nocc * (size_t)pl.nvir_pad;
    const size_t op_bytes = packed_op_bytes(pl, nocc, prec);
    const size_t tb = f64 ? 8 : 4;
    const size_t b2_bytes = (size_t)m_pad * rows * tb;
    const int64_t ld_t = (int64_t)rows;  // nocc * nvir_pad
    const size_t t_bytes = (size_t)pl.nvir_pad * rows * tb;
    const size_t u_bytes = (nocc > 1 ? nocc - 1 : 1) * (size_t)pl.nvir_pad * pl.nvir_pad * tb;
    const int vec_stride = 2 * pl.tile + 2;
    // Phase-B split-K: enough (tile, split) blocks to cover the CUs twice,
    // splits of at least 256 k; sized for row 0 (the longest K).
    const int kp_g = f64 ? 8 : 32;
    const int64_t nbB = (int64_t)(m_pad / gtile) * ((pl.nvir_pad + gtile - 1) / gtile);
    const int64_t k0 = (int64_t)nocc * pl.nvir_pad;
    int splits = (int)std::min<int64_t>((512 + nbB - 1) / nbB, std::max<int64_t>(1, k0 / 256));
    splits = std::max(1, splits);
    const size_t items_max = nocc * (size_t)pl.ntp;
    // Row reductions: virtual sums in chunks of pairs (enough blocks to spread
    // over the CUs), one block per pair for the energy and occupied sums.
    const int nvb = (int)((nvir + 255) / 256);
    const int nchunks_max = (int)std::min<size_t>(nocc, (size_t)std::max(1, (512 + nvb - 1) / nvb));
    std::vector<std::pair<int, size_t>> req = {
        {kSlotOp0, op_bytes},
        {kSlotStage2, b2_bytes},
        {kSlotT, t_bytes},
        {kSlotU, u_bytes},
        {kSlotEps, sizeof(double) * (nocc + nvir)},
        {kSlotItems, sizeof(double) * 4 * items_max},
        {kSlotDeps, sizeof(double) * (vec_stride * items_max)},
        {kSlotRows, sizeof(double) * 4 * nocc},
        {kSlotAcc, sizeof(double) * 2 * (nocc + nvir)},
        {kSlotResult, sizeof(double) * (2 + nocc + nvir)},
        {kSlotScalar, 256},
        {kSlotSplit, sizeof(double) * (size_t)splits * m_pad * pl.nvir_pad},
        {kSlotRed, sizeof(double) * (2 * (size_t)nchunks_max * nvir + 6 * nocc)},
    };
    if (prec == Prec::BF16X3) {
        req.push_back({kSlotOp1, op_bytes});
        req.push_back({kSlotDiag, sizeof(double) * rows});
    }
    fh_status st = reserve(s, req);
    if (st != FH_OK) return st;
    if (!f64) {
        st = range_check(s, b_dev, in_f32, naux * nocc * nvir, naux, stream);
        if (st != FH_OK) return st;
    }
    std::vector<double> e(nocc + nvir);
    for (size_t i = 0; i < nocc; ++i) e[i] = eps[first_occ + i];
    for (size_t a = 0; a < nvir; ++a) e[nocc + a] = eps[nocc_total + a];
    double *deps_dev = (double *)s.slots[kSlotEps].ptr;
    st = hip_check(hipMemcpyAsync(deps_dev, e.data(), sizeof(double) * e.size(),
                                  hipMemcpyHostToDevice, stream),
                   "H2D eps");
    if (st != FH_OK) return st;
    s.pending.h2d_bytes += sizeof(double) * e.size();

    mark(s, 0, stream);
    PackArgs pa;
    pa.b = b_dev;
    pa.in_f32 = in_f32;
    pa.naux = (int)naux;
    pa.nocc = (int)nocc;
    pa.nvir = (int)nvir;
    pa.nvir_pad = pl.nvir_pad;
    pa.naux_pad = pl.naux_pad;
    pa.fmt = f64 ? PackFmt::F64 : PackFmt::BF16X2;
    pa.out0 = s.slots[kSlotOp0].ptr;
    pa.out1 = prec == Prec::BF16X3 ? s.slots[kSlotOp1].ptr : nullptr;
    pa.scale = nullptr;
    pa.lo_sq = prec == Prec::BF16X3 ? (double *)s.slots[kSlotDiag].ptr : nullptr;
    hipError_t he = launch_pack(pa, stream);
    if (he != hipSuccess) return fail(FH_KERNEL, "pack kernel: %s", hipGetErrorString(he));
    {
        const int64_t total = (int64_t)m_pad * (int64_t)rows;
        const unsigned blocks = (unsigned)std::min<int64_t>((total + 255) / 256, 16384);
        void *b2 = s.slots[kSlotStage2].ptr;
        if (f64) {
            if (in_f32)
                hipLaunchKernelGGL((pack_qmajor_kernel<float, double>), dim3(blocks), dim3(256), 0,
                                   stream, (const float *)b_dev, (int)naux, (int)nocc, (int)nvir,
                                   pl.nvir_pad, m_pad, (double *)b2);
            else
                hipLaunchKernelGGL((pack_qmajor_kernel<double, double>), dim3(blocks), dim3(256), 0,
                                   stream, (const double *)b_dev, (int)naux, (int)nocc, (int)nvir,
                                   pl.nvir_pad, m_pad, (double *)b2);
        } else {
            if (in_f32)
                hipLaunchKernelGGL((pack_qmajor_kernel<float, float>), dim3(blocks), dim3(256), 0,
                                   stream, (const float *)b_dev, (int)naux, (int)nocc, (int)nvir,
                                   pl.nvir_pad, m_pad, (float *)b2);
            else
                hipLaunchKernelGGL((pack_qmajor_kernel<double, float>), dim3(blocks), dim3(256), 0,
                                   stream, (const double *)b_dev, (int)naux, (int)nocc, (int)nvir,
                                   pl.nvir_pad, m_pad, (float *)b2);
        }
    }
    double *acc = (double *)s.slots[kSlotAcc].ptr;
    launch_fill<double>(gamma_dev, (int64_t)naux * (int64_t)(nocc * nvir), 0.0, stream);
    launch_fill<double>(acc, (int64_t)(2 * (nocc + nvir)), 0.0, stream);
    he = hipGetLastError();
    if (he != hipSuccess) return fail(FH_KERNEL, "adjoint set-up kernels: %s", hipGetErrorString(he));
    mark(s, 1, stream);

    EnergyArgs ea;
    ea.prec = prec;
    ea.tile = pl.tile;
    ea.op0 = pa.out0;
    ea.op1 = pa.out1;
    ea.scale = nullptr;
    ea.diag_corr = pa.lo_sq;
    ea.nocc = (int)nocc;
    ea.nvir = (int)nvir;
    ea.nvir_pad = pl.nvir_pad;
    ea.naux_pad = pl.naux_pad;
    ea.nqb = 0;
    ea.eps_occ = deps_dev;
    ea.eps_vir = deps_dev + nocc;
    ea.kappa = NAN;
    ea.ntile = pl.ntile;
    ea.ntp = pl.ntp;
    ea.item_partials = (double *)s.slots[kSlotItems].ptr;
    ea.row_partials = (double *)s.slots[kSlotRows].ptr;
    ea.result = (double *)s.slots[kSlotResult].ptr;
    ea.c_os = c_os;
    ea.c_ss = c_ss;
    ea.tbuf = s.slots[kSlotT].ptr;
    ea.ubuf = s.slots[kSlotU].ptr;
    ea.ld_t = ld_t;
    ea.t_f32 = !f64;
    ea.item_vec = (double *)s.slots[kSlotDeps].ptr;
    ea.vec_stride = vec_stride;

    GemmArgs ga;
    ga.prec = gprec;
    ga.b2 = s.slots[kSlotStage2].ptr;
    ga.ld_b2 = (int64_t)rows;
    ga.tbuf = ea.tbuf;
    ga.ld_t = ld_t;
    ga.ubuf = ea.ubuf;
    ga.nocc = (int)nocc;
    ga.nvir = (int)nvir;
    ga.nvir_pad = pl.nvir_pad;
    ga.naux = (int)naux;
    ga.m_pad = m_pad;
    ga.nov = (int64_t)(nocc * nvir);
    ga.alpha = 4.0;
    ga.gamma = gamma_dev;
    ga.tile = gtile;
    ga.splits = splits;
    ga.partial = (double *)s.slots[kSlotSplit].ptr;

    AdjRed rd;
    rd.nocc = (int)nocc;
    rd.nvir = (int)nvir;
    rd.tile = pl.tile;
    rd.ntile = pl.ntile;
    rd.ntp = pl.ntp;
    rd.vec_stride = vec_stride;
    rd.item_vec = ea.item_vec;
    rd.item_partials = ea.item_partials;
    rd.rows = ea.row_partials;
    rd.docc_s = acc;
    rd.docc_c = acc + nocc;
    rd.dvir_s = acc + 2 * nocc;
    rd.dvir_c = acc + 2 * nocc + nvir;
    rd.vpart = (double *)s.slots[kSlotRed].ptr;
    rd.ppart = rd.vpart + 2 * (size_t)nchunks_max * nvir;

    int64_t total_items = 0;
    for (size_t i = 0; i < nocc; ++i) {
        ea.i_fixed = (int)i;
        ea.nitems = (int64_t)(nocc - i) * pl.ntp;
        total_items += ea.nitems;
        he = launch_adjoint_pairs(ea, stream);
        if (he != hipSuccess) return fail(FH_KERNEL, "adjoint pair kernel: %s", hipGetErrorString(he));
        rd.i = (int)i;
        {
            const int npair = (int)(nocc - i);
            rd.nchunks = std::min(npair, nchunks_max);
            rd.jc = (npair + rd.nchunks - 1) / rd.nchunks;
            rd.nchunks = (npair + rd.jc - 1) / rd.jc;
            hipLaunchKernelGGL(adjoint_reduce_kernel, dim3((unsigned)(nvb * rd.nchunks + npair)),
                               dim3(256), 0, stream, rd);
            hipLaunchKernelGGL(adjoint_fold_kernel, dim3((unsigned)nvb), dim3(256), 0, stream, rd);
        }
        ga.i = (int)i;
        {
            // Same split count for every row (fixed summation structure); the
            // chunk shrinks with the row's K.
            const int64_t k = (int64_t)(nocc - i) * pl.nvir_pad;
            const int64_t chunk = (k + splits - 1) / splits;
            ga.kchunk = (int)((chunk + kp_g - 1) / kp_g * kp_g);
        }
        he = launch_adjoint_gemm(ga, stream);
        if (he != hipSuccess) return fail(FH_KERNEL, "adjoint GEMM: %s", hipGetErrorString(he));
    }
    mark(s, 2, stream);
    {
        const int n = (int)std::max(nocc, nvir);
        hipLaunchKernelGGL(adjoint_finish_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           stream, ea.row_partials, (int)nocc, (int)nvir, rd.docc_s, rd.docc_c,
                           rd.dvir_s, rd.dvir_c, ea.result);
    }
    he = hipGetLastError();
    if (he != hipSuccess) return fail(FH_KERNEL, "adjoint reduction: %s", hipGetErrorString(he));
    mark(s, 3, stream);
    s.pending.kernel_launches += 5 * nocc + 6;
    std::vector<double> res(2 + nocc + nvir);
    st = hip_check(hipMemcpyAsync(res.data(), ea.result, sizeof(double) * res.size(),
                                  hipMemcpyDeviceToHost, stream),
                   "D2H adjoint result");
    if (st != FH_OK) return st;
    st = hip_check(hipStreamSynchronize(stream), "stream sync");
    if (st != FH_OK) return st;
    s.pending.d2h_bytes += sizeof(double) * res.size();
    finish_timing(s, (uint64_t)total_items, pl.tile);
    out->e_os = res[0];
    out->e_ss = res[1];
    out->e_total = res[0] + res[1];
    if (deps) {
        for (size_t k = 0; k < n_eps; ++k) deps[k] = 0.0;
        for (size_t i = 0; i < nocc; ++i) deps[first_occ + i] = res[2 + i];
        for (size_t a = 0; a < nvir; ++a) deps[nocc_total + a] = res[2 + nocc + a];
    }
    return FH_OK;
}

static fh_status adjoint_checks(size_t naux, size_t nocc, size_t nvir, size_t n_eps,
                                size_t first_occ, size_t nocc_total, double c_os, double c_ss,
                                int precision, Prec *prec, bool *empty) {
    if (!to_prec(precision, prec)) return fail(FH_LAYOUT, "unknown precision %d", precision);
    if (*prec != Prec::F64 && *prec != Prec::BF16X3 && *prec != Prec::I8X4)
        return fail(FH_UNSUPPORTED, "the adjoint runs in FP64, BF16x3 or INT8x4 only");
    if (!std::isfinite(c_os) || !std::isfinite(c_ss))
        return fail(FH_LAYOUT, "spin scaling must be finite, got c_os = %g, c_ss = %g", c_os, c_ss);
    return check_shapes(naux, nocc, nvir, n_eps, first_occ, nocc_total, empty);
}

}  // namespace fh

using namespace fh;

extern "C" fh_status fh_rimp2_energy_adjoint_dev(const void *b_dev, fh_dtype b_dtype, size_t naux,
                                                 size_t nocc, size_t nvir, const double *eps,
                                                 size_t n_eps, size_t first_occ,
                                                 size_t nocc_total, double c_os, double c_ss,
                                                 fh_precision precision, void *stream,
                                                 fh_spin_components *out, double *gamma_dev,
                                                 double *deps) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    if (!out) return fail(FH_LAYOUT, "null output pointer");
    if (b_dtype != FH_DT_F64 && b_dtype != FH_DT_F32)
        return fail(FH_LAYOUT, "unknown b_dtype %d", (int)b_dtype);
    Prec prec;
    bool empty = false;
    fh_status st = adjoint_checks(naux, nocc, nvir, n_eps, first_occ, nocc_total, c_os, c_ss,
                                  precision, &prec, &empty);
    if (st != FH_OK) return st;
    if (empty) {
        *out = fh_spin_components{0.0, 0.0, 0.0};
        if (deps)
            for (size_t k = 0; k < n_eps; ++k) deps[k] = 0.0;
        commit(s, prec, true, false);
        return FH_OK;
    }
    if (!b_dev || !eps || !gamma_dev) return fail(FH_LAYOUT, "null input pointer");
    st = ensure_device(s);
    if (st != FH_OK) return st;
    st = adjoint_core(s, b_dev, b_dtype, naux, nocc, nvir, eps, n_eps, first_occ, nocc_total,
                      c_os, c_ss, prec, (hipStream_t)stream, out, gamma_dev, deps);
    if (st != FH_OK) return st;
    commit(s, prec, true, false);
    return FH_OK;
}

extern "C" fh_status fh_rimp2_energy_adjoint(const double *b_ov, size_t naux, size_t nocc,
                                             size_t nvir, const double *eps, size_t n_eps,
                                             size_t first_occ, size_t nocc_total, double c_os,
                                             double c_ss, fh_precision precision,
                                             fh_spin_components *out, double *gamma,
                                             double *deps) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    if (!out) return fail(FH_LAYOUT, "null output pointer");
    Prec prec;
    bool empty = false;
    fh_status st = adjoint_checks(naux, nocc, nvir, n_eps, first_occ, nocc_total, c_os, c_ss,
                                  precision, &prec, &empty);
    if (st != FH_OK) return st;
    if (empty) {
        *out = fh_spin_components{0.0, 0.0, 0.0};
        if (deps)
            for (size_t k = 0; k < n_eps; ++k) deps[k] = 0.0;
        commit(s, prec, true, false);
        return FH_OK;
    }
    if (!b_ov || !eps || !gamma) return fail(FH_LAYOUT, "null input pointer");
    st = ensure_device(s);
    if (st != FH_OK) return st;
    const size_t nbytes = sizeof(double) * naux * nocc * nvir;
    st = reserve(s, {{kSlotStage, nbytes}, {kSlotGamma, nbytes}});
    if (st != FH_OK) return st;
    hipStream_t stream = nullptr;
    st = hip_check(hipMemcpyAsync(s.slots[kSlotStage].ptr, b_ov, nbytes, hipMemcpyHostToDevice,
                                  stream),
                   "H2D b_ov");
    if (st != FH_OK) return st;
    s.pending.h2d_bytes += nbytes;
    st = adjoint_core(s, s.slots[kSlotStage].ptr, FH_DT_F64, naux, nocc, nvir, eps, n_eps,
                      first_occ, nocc_total, c_os, c_ss, prec, stream, out,
                      (double *)s.slots[kSlotGamma].ptr, deps);
    if (st != FH_OK) return st;
    st = hip_check(hipMemcpy(gamma, s.slots[kSlotGamma].ptr, nbytes, hipMemcpyDeviceToHost),
                   "D2H gamma");
    if (st != FH_OK) return st;
    s.pending.d2h_bytes += nbytes;
    commit(s, prec, true, false);
    return FH_OK;
}
