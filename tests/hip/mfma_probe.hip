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

extern "C" int probe_mfma(int kind, const void *a, const void *b, const int *sa, const int *sb,
                          void *c) {
    size_t ab = kind == 0 ? 64 * 8 : kind == 1 ? 64 * 4 : kind == 2 ? 64 * 16 : 64 * 32;
    size_t cb = kind == 0 ? 64 * 4 * 8 : 64 * 4 * 4;
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
