// FP8 E4M3 variant of the fused RI-MP2 pair-energy kernel (gfx950).
//
// Same work decomposition and FP64 epilogue as energy.hip. The contraction
// uses v_mfma_scale_f32_16x16x128_f8f6f4 with OCP E4M3 operands (format 0)
// and one power-of-two (E8M0, bias 127) scale per (occupied orbital, 64-row Q
// block). Instruction layout, as measured on gfx950: bytes 0-15 of lane
// (r = l%16, g = l/16) are k = 16g..16g+15 and bytes 16-31 are
// k = 64+16g..64+16g+15, for A rows and B columns alike; the scale of
// 32-wide k block kb of row (column) r is taken from lane 16 kb + r. With
// k = Q - q0 inside a 128-deep panel, k blocks 0-1 are the panel's first
// 64-row Q block and 2-3 its second, so lane g passes the scale of Q block
// 2p + g/2 and loads panel chunks g and g + 4. Accumulation is FP32.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <type_traits>

#include "fh_device.h"
#include "fh_kernels.h"
#include "pair_body.h"

namespace fh {

namespace {

template <int TILE>
__global__ __launch_bounds__(256) void energy_fp8_kernel(EnergyArgs a) {
    constexpr int KP = 128;
    constexpr int TW = TILE / 2;
    constexpr int NB = TW / 16;
    constexpr int CPR = 8;  // 16-byte chunks per 128-byte row
    constexpr int OPCH = TILE * CPR;
    constexpr int NCH = 4 * OPCH;
    constexpr int CPT = (NCH + 255) / 256;

    __shared__ __attribute__((aligned(16))) v4i lds[2][4][TILE][CPR];

    pair::Item it;
    it.id = xcd_remap((int64_t)blockIdx.x, a.nitems);
    decode_pair(it.id / a.ntp, a.nocc, it.i, it.j);
    decode_tile_pair((int)(it.id % a.ntp), a.ntile, it.ta, it.tb);

    const uint8_t *__restrict__ bp = (const uint8_t *)a.op0;
    const int A0 = it.ta * TILE, B0 = it.tb * TILE;
    const pair::SplitQ sq(it, a.naux_pad, KP);
    // Operand rows as in pair_body.h: 0 = (i, A), 1 = (j, B), 2 = (j, A),
    // 3 = (i, B), or 0 and 1 one panel ahead for a single-product item.
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
            if (c < NCH && row < lim && q < a.naux_pad)
                v = *(const v4i *)(bp + (base + row) * ld + q + ch * 16);
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
    const pair::Lane L{A0, B0, wid / 2, wid % 2, lane % 16, lane / 16};
    const int g = L.g;
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

    const uint8_t *sc_i = a.scale + (int64_t)it.i * a.nqb;
    const uint8_t *sc_j = a.scale + (int64_t)it.j * a.nqb;

    auto frag = [&](int buf, int op, int row) -> v8i {
        v4i lo = lds[buf][op][row][g ^ (row & 7)];
        v4i hi = lds[buf][op][row][(g + 4) ^ (row & 7)];
        return (v8i){lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    };
    // Scales: operands 0/1 are panel p01; operands 2/3 are panel p01 with the
    // roles of i and j exchanged, or (single product) i and j at panel p01 + 1,
    // whose Q block index is clamped into the table where that panel lies past
    // naux_pad (its operands are zeros there).
    auto compute = [&](int buf, int p) {
        const int p01 = sq.single ? 2 * p : p;
        const int qa = 2 * p01 + g / 2;
        const int qn = min(qa + 2, a.nqb - 1);
        const int s1a = sc_i[qa], s1b = sc_j[qa];
        const int s2a = sq.single ? sc_i[qn] : s1b;
        const int s2b = sq.single ? sc_j[qn] : s1a;
        v8i f0[NB], f1[NB], f2[NB], f3[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int ra = L.wr * TW + k * 16 + L.r, rb = L.wc * TW + k * 16 + L.r;
            f0[k] = frag(buf, 0, ra);
            f2[k] = frag(buf, 2, ra);
            f1[k] = frag(buf, 1, rb);
            f3[k] = frag(buf, 3, rb);
        }
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
                if (rv[x] && cv[y]) {
                    acc1[x][y] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                        f0[x], f1[y], acc1[x][y], 0, 0, 0, s1a, 0, s1b);
                    acc2[x][y] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                        f2[x], f3[y], acc2[x][y], 0, 0, 0, s2a, 0, s2b);
                }
    };

    gload(0);
    lstore(0);
    __syncthreads();
    for (int p = 0; p < sq.iters; ++p) {
        const int cur = p & 1;
        if (p + 1 < sq.iters) gload((p + 1) * sq.step);
        compute(cur, p);
        if (p + 1 < sq.iters) lstore(cur ^ 1);
        __syncthreads();
    }
    if (sq.single) {
        static_assert(sizeof(lds) >= sizeof(float) * TW * (TILE + 1), "LDS holds a C1 half-tile");
        pair::derive_c2<TILE, float>(it, L, acc1, acc2, pair::RowOf<float>{},
                                     (float *)&lds[0][0][0][0]);
    }
    pair::energy_epilogue<TILE>(a, it, L, acc1, acc2, rv, cv, pair::RowOf<float>{},
                                pair::NoCorr{});
}

// The same contraction on eight waves (512 threads) at TILE = 128, from the
// operand in MFMA fragment order (EnergyArgs::opf, PackFmt::FP8F): per
// (16-row group, 128-deep Q panel) one 2 KB block, lane 16 g + r at byte
// 32 (16 g + r) holding the two 16-byte chunks g and g + 4 of row r's panel,
// the 32 bytes energy_fp8_kernel's LDS fragment read returns. Waves 0-3
// accumulate C1, waves 4-7 C2 or, for a single-product item, C1 over the odd
// panels; each wave one 64-square quadrant, two waves per SIMD; each fragment
// is one coalesced 2 KB wave load straight into registers (no LDS, no barrier
// in the main loop), one panel in flight. Blocks past nvir_pad are neither
// loaded nor multiplied. After the loop the halves meet through LDS in the
// four-wave kernel's order (so the sums are bitwise the same), waves 0-3 run
// the energy epilogue and waves 4-7 add zeros to the workgroup sum.
template <int TILE>
__global__ __launch_bounds__(512) void energy_fp8_w8f_kernel(EnergyArgs a) {
    constexpr int KP = 128;
    constexpr int TW = TILE / 2;
    constexpr int NB = TW / 16;
    constexpr int FBLK = 2048;
    __shared__ __attribute__((aligned(16))) float xbuf[TILE * (TILE + 1)];

    pair::Item it;
    it.id = xcd_remap((int64_t)blockIdx.x, a.nitems);
    decode_pair(it.id / a.ntp, a.nocc, it.i, it.j);
    decode_tile_pair((int)(it.id % a.ntp), a.ntile, it.ta, it.tb);

    const int A0 = it.ta * TILE, B0 = it.tb * TILE;
    const pair::SplitQ sq(it, a.naux_pad, KP);
    const int np = a.naux_pad / KP, ng = a.nvir_pad / 16;
    const int tid = threadIdx.x, lane = tid % kWave, wid = tid / kWave;
    const int prod = wid >> 2, quad = wid & 3;
    const pair::Lane L{A0, B0, quad >> 1, quad & 1, lane % 16, lane / 16};
    const int g = L.g;
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
        pa[k] = fbase + (int64_t)(occA * ng + (A0 + L.wr * TW) / 16 + k) * np * FBLK + lane * 32;
        pb[k] = fbase + (int64_t)(occB * ng + (B0 + L.wc * TW) / 16 + k) * np * FBLK + lane * 32;
    }
    bool full = true, anyr = false, anyc = false;
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        full = full && rv[k] && cv[k];
        anyr = anyr || rv[k];
        anyc = anyc || cv[k];
    }
    // Scales of panel P: the (i, 64-row Q block) bytes of this lane's rows,
    // block 2 P + g / 2, as in energy_fp8_kernel.
    const uint8_t *scA = a.scale + (int64_t)occA * a.nqb + g / 2;
    const uint8_t *scB = a.scale + (int64_t)occB * a.nqb + g / 2;
    v4f acc[NB][NB];
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) acc[x][y] = (v4f){0, 0, 0, 0};

    v8i fa[2][NB], fb[2][NB];
    int sa[2], sb[2];
    auto fetch = [&](int P, v8i(&ra)[NB], v8i(&rb)[NB], int &s0, int &s1, auto fullc) {
        const int64_t off = (int64_t)P * FBLK;
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            if (decltype(fullc)::value || rv[k]) ra[k] = *(const v8i *)(pa[k] + off);
            if (decltype(fullc)::value || cv[k]) rb[k] = *(const v8i *)(pb[k] + off);
        }
        s0 = scA[2 * P];
        s1 = scB[2 * P];
    };
    auto mma = [&](const v8i(&ra)[NB], const v8i(&rb)[NB], int s0, int s1, auto fullc) {
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
                if (decltype(fullc)::value || (rv[x] && cv[y]))
                    acc[x][y] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                        ra[x], rb[y], acc[x][y], 0, 0, 0, s0, 0, s1);
    };
    const int iters = sq.iters;
    auto valid = [&](int p) { return p < iters && p * pstep + poff < np; };
    auto run = [&](auto fullc) {
        if (valid(0)) fetch(poff, fa[0], fb[0], sa[0], sb[0], fullc);
        for (int p = 0; p < iters; p += 2) {
            if (valid(p + 1)) fetch((p + 1) * pstep + poff, fa[1], fb[1], sa[1], sb[1], fullc);
            if (valid(p)) mma(fa[0], fb[0], sa[0], sb[0], fullc);
            if (valid(p + 2)) fetch((p + 2) * pstep + poff, fa[0], fb[0], sa[0], sb[0], fullc);
            if (valid(p + 1)) mma(fa[1], fb[1], sa[1], sb[1], fullc);
        }
    };
    if (full)
        run(std::true_type{});
    else if (anyr && anyc)
        run(std::false_type{});
    __syncthreads();

    // Halves meet through LDS: odd-panel C1 added to C1 (single product), or
    // C2 read back; a diagonal tile pair's C2 is C1 transposed, i == j gives
    // C2 = C1.
    constexpr int LDT = TILE + 1;
    float *xb = xbuf;
    auto put = [&](const v4f(&v)[NB][NB]) {
#pragma unroll
        for (int x = 0; x < NB; ++x)
#pragma unroll
            for (int y = 0; y < NB; ++y)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    xb[(L.wr * TW + x * 16 + out_row<float>(g, e)) * LDT + L.wc * TW + y * 16 +
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
                    const float o = xb[(L.wr * TW + x * 16 + out_row<float>(g, e)) * LDT +
                                       L.wc * TW + y * 16 + L.r];
                    if (sq.single)
                        acc[x][y][e] += o;
                    else
                        acc2[x][y][e] = o;
                }
    }
    const bool transpose = sq.single && it.i != it.j;
    if (transpose) {
        __syncthreads();
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
                    acc2[x][y][e] = transpose ? xb[(L.wc * TW + y * 16 + L.r) * LDT + L.wr * TW +
                                                   x * 16 + out_row<float>(g, e)]
                                              : acc[x][y][e];
    }
    if (prod == 0)
        pair::energy_epilogue<TILE>(a, it, L, acc, acc2, rv, cv, pair::RowOf<float>{},
                                    pair::NoCorr{});
    else
        pair::block_store(KSum{}, KSum{}, 0.0, a.item_partials + 4 * it.id);
}

}  // namespace

hipError_t launch_energy_fp8(const EnergyArgs &a, hipStream_t st) {
    if (a.tile == 128 && a.opf)
        hipLaunchKernelGGL(energy_fp8_w8f_kernel<128>, dim3((unsigned)a.nitems), dim3(512), 0, st, a);
    else if (a.tile == 64)
        hipLaunchKernelGGL(energy_fp8_kernel<64>, dim3((unsigned)a.nitems), dim3(256), 0, st, a);
    else if (a.tile == 32)
        hipLaunchKernelGGL(energy_fp8_kernel<32>, dim3((unsigned)a.nitems), dim3(256), 0, st, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace fh
