// Device-side helpers shared by the RI-MP2 kernels (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fh {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef int v8i __attribute__((ext_vector_type(8)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned short v8u16 __attribute__((ext_vector_type(8)));

constexpr int kWave = 64;

// Raw (untyped, stride 0) buffer resource over [base, base + bytes), bytes
// clamped to [0, 2^31 - 1]: a raw buffer load at a byte offset at or past the
// range returns zeros (dword 3 of a gfx9 buffer descriptor: kRawBufferFlags).
constexpr int kRawBufferFlags = 0x00020000;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void *base, int64_t bytes) {
    const int n = bytes <= 0 ? 0 : bytes >= 0x7fffffff ? 0x7fffffff : (int)bytes;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, n,
                                             kRawBufferFlags);
}
constexpr int kXcds = 8;  // MI355X: 8 accelerator complex dies, one L2 each

// Row, inside a 16x16 result block, of accumulator element v of lane 16 g + r
// (whose column is r), per MFMA instruction, as measured on gfx950:
// v_mfma_f64_16x16x4f64 holds rows 4v + g; the FP32-accumulating 16x16
// instructions (f32 x4, bf16 x32, scaled f8 x128) hold rows 4g + v.
template <typename T>
__device__ __forceinline__ constexpr int out_row(int g, int v) {
    return 4 * g + v;
}
template <>
__device__ __forceinline__ constexpr int out_row<double>(int g, int v) {
    return 4 * v + g;
}

// Compensated (Neumaier) FP64 sum: value() = s + c.
struct KSum {
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double x) {
        double t = s + x;
        if (fabs(s) >= fabs(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
    // Merge another compensated sum: TwoSum of the heads, compensations added.
    __device__ __forceinline__ void merge(double os, double oc) {
        double t = s + os;
        double bp = t - s;
        double err = (s - (t - bp)) + (os - bp);
        s = t;
        c += oc + err;
    }
    __device__ __forceinline__ double value() const { return s + c; }
};

// Workgroups are dispatched round-robin over the XCDs (block b runs on XCD
// b % 8). Renumber so that each XCD receives a contiguous range of work
// items: consecutive items share operand panels, which then meet in the same
// L2. Bijective on [0, n).
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t n) {
    int64_t q = n / kXcds, r = n % kXcds;
    int64_t x = b % kXcds, local = b / kXcds;
    return x * q + (x < r ? x : r) + local;
}

// Unique occupied pair p (i ascending, then j from i) -> (i, j).
__device__ __forceinline__ void decode_pair(int64_t p, int nocc, int &i, int &j) {
    int ii = 0;
    int64_t rem = p;
    while (rem >= nocc - ii) {
        rem -= nocc - ii;
        ++ii;
    }
    i = ii;
    j = ii + (int)rem;
}

// Tile pair t over nt tiles with tA <= tB: the nt(nt-1)/2 off-diagonal pairs
// first (row-major strict upper triangle), then the nt diagonal ones. The
// diagonal items do half the work of the others (single-product items), so
// putting them last lets the workgroups that run one occupied pair's items
// side by side stay in step and share its operand panels in L2.
__device__ __forceinline__ void decode_tile_pair(int t, int nt, int &ta, int &tb) {
    const int noff = nt * (nt - 1) / 2;
    if (t >= noff) {
        ta = tb = t - noff;
        return;
    }
    int a = 0, rem = t;
    while (rem >= nt - 1 - a) {
        rem -= nt - 1 - a;
        ++a;
    }
    ta = a;
    tb = a + 1 + rem;
}

#ifndef FH_TILE_BAND
#define FH_TILE_BAND 4
#endif
constexpr int kTileBand = FH_TILE_BAND;  // tile rows per band of the banded tile-pair order

// Workgroup b runs on XCD b % 8. Item of workgroup b for items numbered pair
// by pair (ntp tile pairs each): with the occupied pairs in groups of eight,
// XCD x takes pair x of every group (pairs x, x + 8, ...) and walks each
// pair's tile pairs in order, so at any moment the eight XCDs work on eight
// neighbouring pairs, which mostly share their i rows: the blocks one XCD
// misses in its L2 are often in the shared MALL from another. Pairs past the
// last whole group keep the plain order. Bijective on [0, npairs * ntp).
__device__ __forceinline__ int64_t pair_interleave(int64_t b, int64_t ntp, int64_t npairs) {
    const int64_t head = npairs / kXcds * kXcds * ntp;
    if (b >= head) return b;
    const int64_t x = b % kXcds, l = b / kXcds;
    return ((l / ntp) * kXcds + x) * ntp + l % ntp;
}

// Tile pair t over nt tiles (tA <= tB), banded order: the off-diagonal pairs
// in bands of kTileBand consecutive tA rows, tB ascending within a band and tA
// ascending within each tB; then the nt diagonal pairs. Consecutive items
// share the B-side blocks of one tB and revisit the band's few A-side blocks
// every kTileBand items, so both stay in L2 between their uses (row-major
// order revisits a B-side block only one whole tA row later). Bijective on
// [0, nt (nt + 1) / 2).
__device__ __forceinline__ void decode_tile_pair_banded(int t, int nt, int &ta, int &tb) {
    const int noff = nt * (nt - 1) / 2;
    ta = tb = 0;
    if (t >= noff) {
        ta = tb = t - noff;
        return;
    }
    for (int r0 = 0; r0 < nt; r0 += kTileBand) {
        const int r1 = min(r0 + kTileBand, nt);
        for (int b = r0 + 1; b < nt; ++b) {
            const int n = min(b, r1) - r0;  // rows of the band above the diagonal in column b
            if (t < n) {
                ta = r0 + t;
                tb = b;
                return;
            }
            t -= n;
        }
    }
}

// Inverse of decode_tile_pair.
__device__ __forceinline__ int encode_tile_pair(int ta, int tb, int nt) {
    if (ta == tb) return nt * (nt - 1) / 2 + ta;
    return ta * (nt - 1) - ta * (ta - 1) / 2 + (tb - ta - 1);
}

// Wave-level butterfly reduction of a compensated pair (deterministic order).
__device__ __forceinline__ void wave_merge(KSum &k) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double os = __shfl_xor(k.s, off, kWave);
        double oc = __shfl_xor(k.c, off, kWave);
        k.merge(os, oc);
    }
}

}  // namespace fh
