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

#include "fh_device.h"
#include "fh_kernels.h"

namespace fh {

namespace {

struct ItemF8 {
    int i, j, ta, tb;
    int64_t id;
};

__device__ __forceinline__ void block_store_f8(KSum os, KSum ss, double w, double *out4) {
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
        for (int k = 0; k < (int)(blockDim.x / kWave); ++k) {
            o.merge(red[k][0], red[k][1]);
            s.merge(red[k][2], red[k][3]);
        }
        out4[0] = o.s;
        out4[1] = o.c;
        out4[2] = s.s;
        out4[3] = s.c;
    }
}

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

    ItemF8 it;
    it.id = xcd_remap((int64_t)blockIdx.x, a.nitems);
    decode_pair(it.id / a.ntp, a.nocc, it.i, it.j);
    decode_tile_pair((int)(it.id % a.ntp), a.ntile, it.ta, it.tb);

    const uint8_t *__restrict__ bp = (const uint8_t *)a.op0;
    const int A0 = it.ta * TILE, B0 = it.tb * TILE;
    const int64_t base0 = (int64_t)it.i * a.nvir_pad + A0;
    const int64_t base1 = (int64_t)it.j * a.nvir_pad + B0;
    const int64_t base2 = (int64_t)it.j * a.nvir_pad + A0;
    const int64_t base3 = (int64_t)it.i * a.nvir_pad + B0;
    const int limA = a.nvir_pad - A0, limB = a.nvir_pad - B0;
    const int64_t ld = a.naux_pad;
    const int tid = threadIdx.x;

    v4i stg[CPT];
    auto gload = [&](int q0) {
#pragma unroll
        for (int t = 0; t < CPT; ++t) {
            const int c = tid + 256 * t;
            const int op = c / OPCH, rem = c % OPCH;
            const int row = rem / CPR, ch = rem % CPR;
            const int64_t base = op == 0 ? base0 : op == 1 ? base1 : op == 2 ? base2 : base3;
            const int lim = (op & 1) ? limB : limA;
            v4i v = {0, 0, 0, 0};
            if (c < NCH && row < lim) v = *(const v4i *)(bp + (base + row) * ld + q0 + ch * 16);
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
    const int wr = wid / 2, wc = wid % 2;
    const int r = lane % 16, g = lane / 16;
    bool rv[NB], cv[NB];
#pragma unroll
    for (int k = 0; k < NB; ++k) {
        rv[k] = A0 + wr * TW + k * 16 < a.nvir_pad;
        cv[k] = B0 + wc * TW + k * 16 < a.nvir_pad;
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
    auto compute = [&](int buf, int p) {
        const int qb = 2 * p + g / 2;
        const int si = sc_i[qb], sj = sc_j[qb];
        v8i f0[NB], f1[NB], f2[NB], f3[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            const int ra = wr * TW + k * 16 + r, rb = wc * TW + k * 16 + r;
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
                        f0[x], f1[y], acc1[x][y], 0, 0, 0, si, 0, sj);
                    acc2[x][y] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(
                        f2[x], f3[y], acc2[x][y], 0, 0, 0, sj, 0, si);
                }
    };

    const int np = (int)(a.naux_pad / KP);
    gload(0);
    lstore(0);
    __syncthreads();
    for (int p = 0; p < np; ++p) {
        const int cur = p & 1;
        if (p + 1 < np) gload((p + 1) * KP);
        compute(cur, p);
        if (p + 1 < np) lstore(cur ^ 1);
        __syncthreads();
    }

    const double eij = a.eps_occ[it.i] + a.eps_occ[it.j];
    const bool kap = !isnan(a.kappa);
    KSum os, ss;
#pragma unroll
    for (int x = 0; x < NB; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            if (!(rv[x] && cv[y])) continue;
            const int bb = B0 + wc * TW + y * 16 + r;
            if (bb >= a.nvir) continue;
            const double eb = a.eps_vir[bb];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int aa = A0 + wr * TW + x * 16 + out_row<float>(g, v);
                if (aa < a.nvir) {
                    const double d = eij - a.eps_vir[aa] - eb;
                    const double g1 = acc1[x][y][v], g2 = acc2[x][y][v];
                    const double inv = 1.0 / d, dd = g1 - g2;
                    double t_os = (g1 * g1 + g2 * g2) * inv, t_ss = dd * dd * inv;
                    if (kap) {
                        const double d1 = 1.0 - exp(a.kappa * d);
                        t_os *= d1 * d1;
                        t_ss *= d1 * d1;
                    }
                    os.add(t_os);
                    ss.add(t_ss);
                }
            }
        }
    const double w = (it.i == it.j ? 1.0 : 2.0) * (it.ta == it.tb ? 0.5 : 1.0);
    block_store_f8(os, ss, w, a.item_partials + 4 * it.id);
}

}  // namespace

hipError_t launch_energy_fp8(const EnergyArgs &a, hipStream_t st) {
    if (a.tile == 128)
        hipLaunchKernelGGL(energy_fp8_kernel<128>, dim3((unsigned)a.nitems), dim3(256), 0, st, a);
    else if (a.tile == 64)
        hipLaunchKernelGGL(energy_fp8_kernel<64>, dim3((unsigned)a.nitems), dim3(256), 0, st, a);
    else if (a.tile == 32)
        hipLaunchKernelGGL(energy_fp8_kernel<32>, dim3((unsigned)a.nitems), dim3(256), 0, st, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace fh
