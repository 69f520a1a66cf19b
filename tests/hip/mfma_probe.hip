// Test-only probe: runs ONE wave of one MFMA instruction on caller-supplied
// per-lane operands and returns the per-lane accumulator, so the tests can
// check the operand/result lane layouts and the FP8 scale association the
// kernels of libferric_hip.so rely on. Every buffer is exactly one wave wide.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v4i __attribute__((ext_vector_type(4)));

__global__ void k_f64(const double *a, const double *b, double *c) {
    const int l = threadIdx.x;
    v4d acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[l], b[l], acc, 0, 0, 0);
    for (int v = 0; v < 4; ++v) c[4 * l + v] = acc[v];
}
__global__ void k_f32(const float *a, const float *b, float *c) {
    const int l = threadIdx.x;
    v4f acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[l], b[l], acc, 0, 0, 0);
    for (int v = 0; v < 4; ++v) c[4 * l + v] = acc[v];
}
__global__ void k_bf16(const uint16_t *a, const uint16_t *b, float *c) {
    const int l = threadIdx.x;
    v8bf x, y;
    for (int e = 0; e < 8; ++e) {
        x[e] = __builtin_bit_cast(__bf16, a[8 * l + e]);
        y[e] = __builtin_bit_cast(__bf16, b[8 * l + e]);
    }
    v4f acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, y, acc, 0, 0, 0);
    for (int v = 0; v < 4; ++v) c[4 * l + v] = acc[v];
}
__global__ void k_f8(const int *a, const int *b, const int *sa, const int *sb, float *c) {
    const int l = threadIdx.x;
    v8i x, y;
    for (int e = 0; e < 8; ++e) {
        x[e] = a[8 * l + e];
        y[e] = b[8 * l + e];
    }
    v4f acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(x, y, acc, 0, 0, 0, sa[l], 0, sb[l]);
    for (int v = 0; v < 4; ++v) c[4 * l + v] = acc[v];
}

__global__ void k_i8(const int *a, const int *b, int *c) {
    const int l = threadIdx.x;
    v4i x, y;
    for (int e = 0; e < 4; ++e) {
        x[e] = a[4 * l + e];
        y[e] = b[4 * l + e];
    }
    v4i acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(x, y, acc, 0, 0, 0);
    for (int v = 0; v < 4; ++v) c[4 * l + v] = acc[v];
}

typedef float v16f __attribute__((ext_vector_type(16)));
__global__ void k_bf16_32(const uint16_t *a, const uint16_t *b, float *c) {
    const int l = threadIdx.x;
    v8bf x, y;
    for (int e = 0; e < 8; ++e) {
        x[e] = __builtin_bit_cast(__bf16, a[8 * l + e]);
        y[e] = __builtin_bit_cast(__bf16, b[8 * l + e]);
    }
    v16f acc;
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y, acc, 0, 0, 0);
    for (int v = 0; v < 16; ++v) c[16 * l + v] = acc[v];
}

// kind 0 = f64 16x16x4, 1 = f32 16x16x4, 2 = bf16 16x16x32, 3 = scaled f8
// 16x16x128, 4 = i8 16x16x64 (16 bytes per lane, int32 result), 5 = bf16
// 32x32x16 (16 bytes per lane, 16 floats per lane).
extern "C" int probe_mfma(int kind, const void *a, const void *b, const int *sa, const int *sb,
                          void *c) {
    size_t ab = kind == 0 ? 64 * 8 : kind == 1 ? 64 * 4 : kind == 2 || kind == 4 ? 64 * 16 : 64 * 32;
    if (kind == 5) ab = 64 * 16;
    size_t cb = kind == 0 ? 64 * 4 * 8 : kind == 5 ? 64 * 16 * 4 : 64 * 4 * 4;
    void *da, *db, *dc;
    int *dsa = nullptr, *dsb = nullptr;
    if (hipMalloc(&da, ab) || hipMalloc(&db, ab) || hipMalloc(&dc, cb)) return 1;
    if (hipMemcpy(da, a, ab, hipMemcpyHostToDevice) || hipMemcpy(db, b, ab, hipMemcpyHostToDevice))
        return 2;
    if (kind == 3) {
        if (hipMalloc(&dsa, 256) || hipMalloc(&dsb, 256)) return 1;
        if (hipMemcpy(dsa, sa, 256, hipMemcpyHostToDevice) ||
            hipMemcpy(dsb, sb, 256, hipMemcpyHostToDevice))
            return 2;
    }
    switch (kind) {
    case 0: hipLaunchKernelGGL(k_f64, dim3(1), dim3(64), 0, 0, (const double *)da, (const double *)db, (double *)dc); break;
    case 1: hipLaunchKernelGGL(k_f32, dim3(1), dim3(64), 0, 0, (const float *)da, (const float *)db, (float *)dc); break;
    case 2: hipLaunchKernelGGL(k_bf16, dim3(1), dim3(64), 0, 0, (const uint16_t *)da, (const uint16_t *)db, (float *)dc); break;
    case 3: hipLaunchKernelGGL(k_f8, dim3(1), dim3(64), 0, 0, (const int *)da, (const int *)db, (const int *)dsa, (const int *)dsb, (float *)dc); break;
    case 4: hipLaunchKernelGGL(k_i8, dim3(1), dim3(64), 0, 0, (const int *)da, (const int *)db, (int *)dc); break;
    case 5: hipLaunchKernelGGL(k_bf16_32, dim3(1), dim3(64), 0, 0, (const uint16_t *)da, (const uint16_t *)db, (float *)dc); break;
    default: return 4;
    }
    if (hipGetLastError() != hipSuccess) return 5;
    if (hipMemcpy(c, dc, cb, hipMemcpyDeviceToHost)) return 6;
    hipFree(da);
    hipFree(db);
    hipFree(dc);
    if (dsa) hipFree(dsa);
    if (dsb) hipFree(dsb);
    return 0;
}

// ---- Sustained MFMA rate ----------------------------------------------------
// Every wave runs `iters` rounds of NACC independent MFMAs on register
// operands (no memory traffic inside the loop), so the launch time measures the
// rate the matrix pipes sustain on this device at its clock under that load —
// the practical ceiling for the pair kernels, next to the datasheet peak.
// Each thread writes one double so the accumulators stay live.
constexpr int kRateAcc = 16;

// NACC independent f64 MFMA accumulators per wave (fewer accumulators, more
// waves per SIMD) and, with NV > 0, NV independent vector FMA chains issued in
// the same loop (matrix and vector pipes side by side).
template <int NACC, int NV>
__global__ __launch_bounds__(256) void rate_f64(int iters, double *out) {
    const double x = 1.0 + threadIdx.x * 1e-9, y = 1.0 - blockIdx.x * 1e-9;
    v4d acc[NACC];
    double va[NV > 0 ? NV : 1];
    for (int k = 0; k < NACC; ++k) acc[k] = (v4d){0, 0, 0, 0};
    for (int k = 0; k < NV; ++k) va[k] = k * 1e-3;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int k = 0; k < NACC; ++k) {
            acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, acc[k], 0, 0, 0);
#pragma unroll
            for (int m = 0; m < NV / NACC; ++m)
                va[k * (NV / NACC) + m] = __builtin_fma(va[k * (NV / NACC) + m], x, y);
        }
    }
    double s = 0;
    for (int k = 0; k < NACC; ++k) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
    for (int k = 0; k < NV; ++k) s += va[k];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}
__global__ __launch_bounds__(256) void rate_bf16(int iters, double *out) {
    v8bf x, y;
    for (int e = 0; e < 8; ++e) {
        x[e] = (__bf16)(1.0f + threadIdx.x * 1e-3f);
        y[e] = (__bf16)(1.0f - e * 1e-3f);
    }
    v4f acc[kRateAcc];
    for (int k = 0; k < kRateAcc; ++k) acc[k] = (v4f){0, 0, 0, 0};
    for (int it = 0; it < iters; ++it)
#pragma unroll
        for (int k = 0; k < kRateAcc; ++k)
            acc[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, y, acc[k], 0, 0, 0);
    double s = 0;
    for (int k = 0; k < kRateAcc; ++k) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}
__global__ __launch_bounds__(256) void rate_f8(int iters, double *out) {
    v8i x, y;
    for (int e = 0; e < 8; ++e) {
        x[e] = 0x38383838 + (int)(threadIdx.x & 3);  // E4M3 values near 1
        y[e] = 0x38383838;
    }
    v4f acc[kRateAcc];
    for (int k = 0; k < kRateAcc; ++k) acc[k] = (v4f){0, 0, 0, 0};
    for (int it = 0; it < iters; ++it)
#pragma unroll
        for (int k = 0; k < kRateAcc; ++k)
            acc[k] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(x, y, acc[k], 0, 0, 0, 127,
                                                                        0, 127);
    double s = 0;
    for (int k = 0; k < kRateAcc; ++k) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void rate_i8(int iters, double *out) {
    v4i x, y;
    for (int e = 0; e < 4; ++e) {
        x[e] = 0x01020304 + (int)threadIdx.x;
        y[e] = 0x05060708 - e;
    }
    v4i acc[kRateAcc];
    for (int k = 0; k < kRateAcc; ++k) acc[k] = (v4i){0, 0, 0, 0};
    for (int it = 0; it < iters; ++it)
#pragma unroll
        for (int k = 0; k < kRateAcc; ++k)
            acc[k] = __builtin_amdgcn_mfma_i32_16x16x64_i8(x, y, acc[k], 0, 0, 0);
    double s = 0;
    for (int k = 0; k < kRateAcc; ++k) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

// Vector FP64 FMA: NV independent chains per lane (no MFMA), the VALU path's
// FP64 rate for comparison with the matrix path.
constexpr int kRateVec = 24;
__global__ __launch_bounds__(256) void rate_vfma_f64(int iters, double *out) {
    double acc[kRateVec];
    const double x = 1.0 + threadIdx.x * 1e-12, y = 1.0 - blockIdx.x * 1e-12;
    for (int k = 0; k < kRateVec; ++k) acc[k] = k * 1e-3;
    for (int it = 0; it < iters; ++it)
#pragma unroll
        for (int k = 0; k < kRateVec; ++k) acc[k] = __builtin_fma(acc[k], x, y);
    double s = 0;
    for (int k = 0; k < kRateVec; ++k) s += acc[k];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

// LDS-fed vector FP64 micro-kernel: a 16 x 16 thread grid, each thread an
// MR x NR register tile; per k it reads MR values of the A panel and NR of the
// B panel (k-major rows in LDS) and issues MR*NR FMAs. No global traffic in the
// loop: the rate is what LDS delivery and the FP64 VALU sustain together.
constexpr int kLdsKP = 16;
template <int MR, int NR>
__global__ __launch_bounds__(256) void rate_lds_fma(int iters, double *out) {
    constexpr int TM = 16 * MR, TN = 16 * NR;
    __shared__ __attribute__((aligned(16))) double la[kLdsKP][TM];
    __shared__ __attribute__((aligned(16))) double lb[kLdsKP][TN];
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    for (int e = tid; e < kLdsKP * TM; e += 256) (&la[0][0])[e] = 1.0 + e * 1e-9;
    for (int e = tid; e < kLdsKP * TN; e += 256) (&lb[0][0])[e] = 1.0 - e * 1e-9;
    __syncthreads();
    double acc[MR][NR];
#pragma unroll
    for (int m = 0; m < MR; ++m)
#pragma unroll
        for (int n = 0; n < NR; ++n) acc[m][n] = 0.0;
    for (int it = 0; it < iters; ++it) {
#pragma unroll 4
        for (int k = 0; k < kLdsKP; ++k) {
            double a[MR], b[NR];
#pragma unroll
            for (int m = 0; m < MR; ++m) a[m] = la[k][ty * MR + m];
#pragma unroll
            for (int n = 0; n < NR; ++n) b[n] = lb[k][tx * NR + n];
#pragma unroll
            for (int m = 0; m < MR; ++m)
#pragma unroll
                for (int n = 0; n < NR; ++n) acc[m][n] = __builtin_fma(a[m], b[n], acc[m][n]);
        }
    }
    double s = 0;
#pragma unroll
    for (int m = 0; m < MR; ++m)
#pragma unroll
        for (int n = 0; n < NR; ++n) s += acc[m][n];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

// The same with the layout of a 128 x 128 tile: thread (tx, ty) owns rows
// {2ty, 2ty+1} + 32 c and columns {2tx, 2tx+1} + 32 c (c = 0..3), read as
// 16-byte chunks that sit side by side across lanes; the LDS footprint is
// padded so one workgroup holds a CU (one wave per SIMD).
template <int PAD_KB>
__global__ __launch_bounds__(256) void rate_lds_fma_il(int iters, double *out) {
    __shared__ __attribute__((aligned(16))) double la[kLdsKP][128];
    __shared__ __attribute__((aligned(16))) double lb[kLdsKP][128];
    __shared__ double pad[PAD_KB * 128];
    const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
    for (int e = tid; e < kLdsKP * 128; e += 256) {
        (&la[0][0])[e] = 1.0 + e * 1e-9;
        (&lb[0][0])[e] = 1.0 - e * 1e-9;
    }
    if (tid == 0) pad[0] = 0.0;
    __syncthreads();
    double acc[8][8];
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int n = 0; n < 8; ++n) acc[m][n] = 0.0;
    typedef double d2 __attribute__((ext_vector_type(2)));
    for (int it = 0; it < iters; ++it) {
#pragma unroll 4
        for (int k = 0; k < kLdsKP; ++k) {
            double a[8], b[8];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const d2 va = *(const d2 *)&la[k][2 * ty + 32 * c];
                const d2 vb = *(const d2 *)&lb[k][2 * tx + 32 * c];
                a[2 * c] = va[0];
                a[2 * c + 1] = va[1];
                b[2 * c] = vb[0];
                b[2 * c + 1] = vb[1];
            }
#pragma unroll
            for (int m = 0; m < 8; ++m)
#pragma unroll
                for (int n = 0; n < 8; ++n) acc[m][n] = __builtin_fma(a[m], b[n], acc[m][n]);
        }
    }
    double s = pad[0];
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int n = 0; n < 8; ++n) s += acc[m][n];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

// kind 0 = f64 16x16x4 (16 accumulators), 2 = bf16 16x16x32, 3 = scaled f8
// 16x16x128, 4 = vector FP64 FMA (kRateVec chains per lane), 5 / 6 = f64 MFMA
// with 8 / 4 accumulators, 7 = 8 f64 MFMA + 16 vector FMA chains, 8 = 16 f64
// MFMA + 32 vector FMA chains, 9 / 10 / 11 = LDS-fed vector FMA with 8x12 /
// 8x8 / 6x8 register tiles, 12 / 13 = 8x8 interleaved layout at one / three
// workgroups per CU, 14 = i8 16x16x64 (int32 accumulation). Returns the
// FLOP rate in TFLOP/s of the second of two identical launches in *tflops.
extern "C" int probe_mfma_rate(int kind, int blocks, int iters, double *tflops) {
    if (blocks <= 0 || iters <= 0) return 4;
    double *out;
    if (hipMalloc(&out, (size_t)blocks * 256 * sizeof(double))) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) || hipEventCreate(&e1)) return 1;
    // FLOP per wave per unrolled step of the loop
    const double mf64 = 2.0 * 16 * 16 * 4, vf64 = 2.0 * 64;
    const double flop_per_step = kind == 0   ? mf64 * kRateAcc
                                 : kind == 2 ? 2.0 * 16 * 16 * 32 * kRateAcc
                                 : kind == 3 ? 2.0 * 16 * 16 * 128 * kRateAcc
                                 : kind == 4 ? vf64 * kRateVec
                                 : kind == 5 ? mf64 * 8
                                 : kind == 6 ? mf64 * 4
                                 : kind == 7 ? mf64 * 8 + vf64 * 16
                                 : kind == 8 ? mf64 * 16 + vf64 * 32
                                 : kind == 9 ? vf64 * kLdsKP * 8 * 12
                                 : kind == 10 ? vf64 * kLdsKP * 8 * 8
                                 : kind == 11 ? vf64 * kLdsKP * 6 * 8
                                 : kind == 14 ? 2.0 * 16 * 16 * 64 * kRateAcc
                                              : vf64 * kLdsKP * 8 * 8;
    float ms = 0;
    for (int rep = 0; rep < 2; ++rep) {
        hipEventRecord(e0, 0);
        switch (kind) {
        case 0: hipLaunchKernelGGL((rate_f64<16, 0>), dim3(blocks), dim3(256), 0, 0, iters, out); break;
        case 5: hipLaunchKernelGGL((rate_f64<8, 0>), dim3(blocks), dim3(256), 0, 0, iters, out); break;
        case 6: hipLaunchKernelGGL((rate_f64<4, 0>), dim3(blocks), dim3(256), 0, 0, iters, out); break;
        case 7: hipLaunchKernelGGL((rate_f64<8, 16>), dim3(blocks), dim3(256), 0, 0, iters, out); break;
        case 8: hipLaunchKernelGGL((rate_f64<16, 32>), dim3(blocks), dim3(256), 0, 0, iters, out); break;
        case 9: hipLaunchKernelGGL((rate_lds_fma<8, 12>), dim3(blocks), dim3(256), 0, 0, iters, out); break;
        case 10: hipLaunchKernelGGL((rate_lds_fma<8, 8>), dim3(blocks), dim3(256), 0, 0, iters, out); break;
        case 11: hipLaunchKernelGGL((rate_lds_fma<6, 8>), dim3(blocks), dim3(256), 0, 0, iters, out); break;
        case 12: hipLaunchKernelGGL((rate_lds_fma_il<100>), dim3(blocks), dim3(256), 0, 0, iters, out); break;
        case 13: hipLaunchKernelGGL((rate_lds_fma_il<1>), dim3(blocks), dim3(256), 0, 0, iters, out); break;
        case 2: hipLaunchKernelGGL(rate_bf16, dim3(blocks), dim3(256), 0, 0, iters, out); break;
        case 3: hipLaunchKernelGGL(rate_f8, dim3(blocks), dim3(256), 0, 0, iters, out); break;
        case 14: hipLaunchKernelGGL(rate_i8, dim3(blocks), dim3(256), 0, 0, iters, out); break;
        case 4: hipLaunchKernelGGL(rate_vfma_f64, dim3(blocks), dim3(256), 0, 0, iters, out); break;
        default: return 4;
        }
        hipEventRecord(e1, 0);
        if (hipGetLastError() != hipSuccess) return 5;
        if (hipEventSynchronize(e1)) return 6;
        hipEventElapsedTime(&ms, e0, e1);
    }
    const double flops = (double)blocks * 4 * iters * flop_per_step;
    *tflops = flops / (ms * 1e-3) / 1e12;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(out);
    return 0;
}
