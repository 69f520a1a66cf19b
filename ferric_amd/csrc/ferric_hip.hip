// Host side of the C ABI declared in include/ferric_hip.h.
//
// Mirrors the contract of the reference's device boundary
// (crates/ferric-mp2/src/rimp2_gpu.rs:105-373): shape checks first, typed
// failures instead of panics, workspace charged to a pool ceiling and
// reserved before the large upload, counters committed only on success.
// There is no CPU path here: every failure is returned to the caller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ferric_hip.h"
#include "fh_host.h"
#include "fh_kernels.h"

namespace fh {

// ---- Errors ------------------------------------------------------------------

static thread_local std::string t_last_error;

fh_status fail(fh_status st, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    t_last_error = buf;
    State &s = state();
    switch (st) {
    case FH_LAYOUT: s.stats.fail_layout++; break;
    case FH_POOL_FULL: s.stats.fail_pool_full++; break;
    case FH_HIP:
    case FH_KERNEL:
    case FH_DEVICE: s.stats.fail_hip++; break;
    case FH_RANGE: s.stats.fail_range++; break;
    case FH_UNSUPPORTED: s.stats.fail_unsupported++; break;
    default: break;
    }
    return st;
}

void clear_error() { t_last_error.clear(); }

State &state() {
    static State s;
    return s;
}

// ---- Device and pool -------------------------------------------------------

static fh_status init_locked(State &s, int device) {
    if (device < 0) {
        const char *env = getenv("FERRIC_GPU_DEVICE");
        device = 0;
        if (env && *env) {
            char *end = nullptr;
            long v = strtol(env, &end, 10);
            if (!end || *end || v < 0)
                return fail(FH_LAYOUT, "FERRIC_GPU_DEVICE=%s is not a device ordinal", env);
            device = (int)v;
        }
    }
    if (s.initialised && s.device == device) return FH_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(FH_DEVICE, "no HIP device is visible");
    if (device >= n) return fail(FH_DEVICE, "device %d requested, %d visible", device, n);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        return fail(FH_DEVICE, "hipGetDeviceProperties(%d) failed", device);
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(FH_DEVICE, "device %d is %s; this library is built for gfx950", device,
                    prop.gcnArchName);
    if (hipSetDevice(device) != hipSuccess) return fail(FH_DEVICE, "hipSetDevice(%d)", device);
    s.device = device;
    s.initialised = true;
    return FH_OK;
}

fh_status ensure_device(State &s) {
    fh_status st = init_locked(s, s.initialised ? s.device : -1);
    if (st != FH_OK) return st;
    if (hipSetDevice(s.device) != hipSuccess) return fail(FH_HIP, "hipSetDevice(%d)", s.device);
    return FH_OK;
}

static fh_status resolve_pool_limit(State &s, size_t *out) {
    if (s.pool_limit_explicit) {
        *out = s.pool_limit;
        return FH_OK;
    }
    const char *env = getenv("FERRIC_GPU_MEM_GB");
    if (env && *env) {
        char *end = nullptr;
        double gb = strtod(env, &end);
        if (!end || *end || !(gb > 0.0) || !std::isfinite(gb))
            return fail(FH_LAYOUT, "FERRIC_GPU_MEM_GB=%s is not a positive number", env);
        *out = (size_t)(gb * 1073741824.0);
        return FH_OK;
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return fail(FH_HIP, "hipMemGetInfo");
    // Default ceiling: 80% of what is free now plus what this library holds.
    *out = (size_t)(0.8 * (double)(free_b + s.held_bytes()));
    return FH_OK;
}

size_t State::held_bytes() const {
    size_t t = 0;
    for (const auto &b : slots) t += b.bytes;
    return t;
}

// Grow-only named device buffers. All sizes of a call are reserved together,
// before any transfer; a call whose total would exceed the ceiling is refused
// without allocating anything.
fh_status reserve(State &s, const std::vector<std::pair<int, size_t>> &req) {
    if (s.arena) {
        // Fixed-arena mode: 256-byte aligned slices in request order, no
        // allocation, no pool accounting.
        for (const auto &r : req) {
            const size_t off = (s.arena_used + 255) / 256 * 256;
            if (off > s.arena_bytes || r.second > s.arena_bytes - off)
                return fail(FH_POOL_FULL,
                            "the call needs more than the fixed %zu-byte scratch (%zu bytes "
                            "placed, %zu more requested)",
                            s.arena_bytes, off, r.second);
            s.slots[r.first].ptr = s.arena + off;
            s.slots[r.first].bytes = r.second;
            s.arena_used = off + r.second;
        }
        return FH_OK;
    }
    size_t limit = 0;
    fh_status st = resolve_pool_limit(s, &limit);
    if (st != FH_OK) return st;
    size_t total = 0;
    for (int k = 0; k < kNumSlots; ++k) {
        size_t want = s.slots[k].bytes;
        for (const auto &r : req)
            if (r.first == k) want = std::max(want, r.second);
        total += want;
    }
    if (total > limit)
        return fail(FH_POOL_FULL,
                    "workspace of %.3f GB exceeds the device pool ceiling of %.3f GB",
                    total / 1e9, limit / 1e9);
    for (const auto &r : req) {
        Slot &b = s.slots[r.first];
        if (b.bytes >= r.second) continue;
        // Captured graphs bake in workspace pointers: drop them on any move.
        release_graphs(s);
        if (b.ptr) {
            (void)hipFree(b.ptr);
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t n = std::max<size_t>(r.second, 256);
        if (hipMalloc(&b.ptr, n) != hipSuccess) {
            b.ptr = nullptr;
            (void)hipGetLastError();
            return fail(FH_POOL_FULL, "hipMalloc of %zu bytes failed", n);
        }
        b.bytes = n;
    }
    s.stats.pool_peak_bytes = std::max<uint64_t>(s.stats.pool_peak_bytes, s.held_bytes());
    return FH_OK;
}

void release_all(State &s) {
    for (auto &b : s.slots) {
        if (b.ptr) (void)hipFree(b.ptr);
        b.ptr = nullptr;
        b.bytes = 0;
    }
}

fh_status hip_check(hipError_t e, const char *what) {
    if (e == hipSuccess) return FH_OK;
    (void)hipGetLastError();
    return fail(FH_HIP, "%s: %s", what, hipGetErrorString(e));
}

// ---- Shape checks (rimp2_gpu.rs:211-237) -----------------------------------

fh_status check_shapes(size_t naux, size_t nocc, size_t nvir, size_t n_eps, size_t first_occ,
                       size_t nocc_total, bool *empty) {
    *empty = false;
    size_t nov = 0;
    if (__builtin_mul_overflow(nocc, nvir, &nov))
        return fail(FH_LAYOUT, "RI-MP2 dimensions nocc=%zu nvir=%zu overflow", nocc, nvir);
    size_t need_o = 0, need_v = 0;
    if (__builtin_add_overflow(first_occ, nocc, &need_o) ||
        __builtin_add_overflow(nocc_total, nvir, &need_v))
        return fail(FH_LAYOUT, "orbital index ranges overflow");
    const size_t need = std::max(need_o, need_v);
    if (n_eps < need)
        return fail(FH_LAYOUT, "eps holds %zu orbital energies but the index ranges reach %zu",
                    n_eps, need);
    if (nocc == 0 || nvir == 0) {
        *empty = true;
        return FH_OK;
    }
    if (naux == 0) return fail(FH_LAYOUT, "empty auxiliary dimension");
    size_t total = 0;
    if (__builtin_mul_overflow(naux, nov, &total))
        return fail(FH_LAYOUT, "naux*nocc*nvir overflows");
    // Kernel index arithmetic is 32-bit in the tile coordinates.
    if (nvir > (1u << 20) || nocc > (1u << 16) || naux > (1u << 24))
        return fail(FH_LAYOUT, "shape nocc=%zu nvir=%zu naux=%zu beyond the kernel's range", nocc,
                    nvir, naux);
    return FH_OK;
}

fh_status check_kappa(double kappa, bool *has) {
    *has = !std::isnan(kappa);
    if (*has && (!std::isfinite(kappa) || kappa <= 0.0))
        return fail(FH_LAYOUT, "kappa must be finite and > 0 (got %g); use NaN for plain MP2",
                    kappa);
    return FH_OK;
}

// Built-in tile side for one-problem calls (FERRIC_HIP_TILE forces one).
// Large virtual spaces: 64 for FP64 / FP32, whose 64-wide kernels run two
// workgroups (two waves per SIMD) per CU and keep the matrix pipes busier than
// the one-workgroup 128-wide kernels (FP64 at 100/900/3600: 669 vs 729 ms);
// 128 for FP8 past 80 virtuals and BF16x3 past 192, whose eight-wave 128-wide
// kernels read fragment-order operands straight into registers (at
// 100/900/3600 BF16x3 95 vs 216 ms, FP8 30 vs 58 ms against 64). Below 192
// virtuals BF16x3's eight-wave kernel at 64 is faster (40/160/960: 0.28 vs
// 0.34 ms; 41/209/960: 0.45 vs 0.42 ms).
int choose_tile(size_t nvir, Prec prec) {
    if (nvir <= 24) return 32;
    if (prec == Prec::BF16X3) return nvir <= 192 ? 64 : 128;
    if (prec == Prec::FP8 && nvir > 80) return 128;
    return 64;
}

size_t packed_elem_bytes(Prec p) {
    switch (p) {
    case Prec::F64: return 8;
    case Prec::F32: return 4;
    case Prec::MIXED: return 4;
    case Prec::BF16X3: return 2;  // per array (hi and lo)
    case Prec::FP8: return 1;
    case Prec::I8X4: return 4;  // four digit bytes
    }
    return 8;
}

// Bytes of one packed operand array. The INT8 kernel reads whole TILE-row
// blocks without a row test, so its buffer carries TILE spare rows.
size_t packed_op_bytes(const EnergyPlan &pl, size_t nocc, Prec prec) {
    const size_t rows = nocc * (size_t)pl.nvir_pad + (prec == Prec::I8X4 ? (size_t)pl.tile : 0);
    return rows * (size_t)pl.naux_pad * packed_elem_bytes(prec);
}

// Reduced-precision range check (the reference's check_f32_arithmetic_range,
// rimp2_gpu.rs:280-293): refuse when max|B| > sqrt(FLT_MAX / depth), depth
// the Q rows of one FP32 accumulation (the whole naux for the FP32-accumulating
// modes, the 64-row panel for MIXED, as the reference's b).
fh_status range_check(State &s, const void *b, bool in_f32, size_t n, size_t depth,
                      hipStream_t stream) {
    unsigned int *dbits = (unsigned int *)s.slots[kSlotScalar].ptr;
    fh_status st = hip_check(launch_absmax(b, in_f32, (int64_t)n, dbits, stream), "absmax");
    if (st != FH_OK) return st;
    unsigned int bits = 0;
    st = hip_check(hipMemcpyAsync(&bits, dbits, 4, hipMemcpyDeviceToHost, stream), "D2H absmax");
    if (st != FH_OK) return st;
    st = hip_check(hipStreamSynchronize(stream), "sync absmax");
    if (st != FH_OK) return st;
    float m;
    memcpy(&m, &bits, 4);
    const double limit = std::sqrt((double)std::numeric_limits<float>::max() / (double)depth);
    if ((double)m > limit)
        return fail(FH_RANGE,
                    "max|B_ov| = %e exceeds sqrt(FLT_MAX/%zu) = %e: an FP32 accumulation of "
                    "%zu products could overflow",
                    (double)m, depth, limit, depth);
    return FH_OK;
}

// ---- Energy core ---------------------------------------------------------------

// Wall-clock ticks per Q panel for the INT8X4 panel rotation: 1.4 us, about
// what one 64-deep panel of a 64-wide item takes when the chip is full. It pays
// when an item's Q loop is long or one occupied pair has many tile pairs (at
// least 40 panels or 32 tile pairs; measured +1 to +5 % there, -1 % below at
// 41/209/960 and 100/200/2400), so smaller problems start at panel 0.
// FERRIC_I8_QROT_TICKS overrides for every size (0 switches the rotation off).
int i8_q_rot_ticks(State &s, int npanels, int ntp) {
    if (const char *e = getenv("FERRIC_I8_QROT_TICKS")) return std::max(0, atoi(e));
    if (npanels < 40 && ntp < 32) return 0;
    if (s.wall_khz < 0) {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s.device) != hipSuccess)
            khz = 0;
        s.wall_khz = std::max(0, khz);
    }
    return (int)std::lround(s.wall_khz * 1.4e-3);
}

fh_status energy_core(State &s, const void *b_dev, fh_dtype b_dtype, size_t naux, size_t nocc,
                      size_t nvir, const double *eps, size_t first_occ, size_t nocc_total,
                      double kappa, Prec prec, hipStream_t stream, fh_spin_components *out,
                      double *row_partials, const EnergyPlan *forced, const void *neg_dev,
                      size_t naux_neg, double *amp_out) {
    const bool in_f32 = b_dtype == FH_DT_F32;
    EnergyPlan pl = forced ? *forced : plan_energy(nocc, nvir, naux, prec);
    // A signed factor stacks B- after B+ along Q, each segment padded to the
    // panel depth so that every panel lies in one of them.
    const bool signed_k = neg_dev != nullptr && naux_neg > 0;
    int q_neg = 1 << 30;
    if (signed_k) {
        if (prec != Prec::F64 && prec != Prec::F32)
            return fail(FH_UNSUPPORTED, "the signed kernel runs in FP64 or FP32");
        const size_t kp = (size_t)panel_depth(prec);
        q_neg = (int)((naux + kp - 1) / kp * kp);
        pl.naux_pad = q_neg + (int)((naux_neg + kp - 1) / kp * kp);
    }
    const size_t nov = nocc * nvir;
    const size_t rows = nocc * (size_t)pl.nvir_pad;
    const size_t eb = packed_elem_bytes(prec);
    const size_t op_bytes = packed_op_bytes(pl, nocc, prec);
    const size_t nqb = (size_t)pl.naux_pad / 64;
    std::vector<std::pair<int, size_t>> req = {
        {kSlotOp0, op_bytes},
        {kSlotEps, sizeof(double) * (nocc + nvir)},
        {kSlotItems, sizeof(double) * 4 * (size_t)pl.nitems},
        {kSlotRows, sizeof(double) * 4 * nocc},
        {kSlotResult, sizeof(double) * (2 + 2 * nocc)},
        {kSlotScalar, 256},
    };
    // BF16x3 and FP8 at tile 128: the operand (BF16: hi and lo in one buffer)
    // in MFMA fragment order, read by the eight-wave kernels straight into
    // registers.
    const bool frag = ((prec == Prec::BF16X3 && (pl.tile == 128 || pl.tile == 64)) ||
                       (prec == Prec::FP8 && pl.tile == 128)) &&
                      !signed_k && !amp_out;
    if (frag && prec == Prec::BF16X3) req[0].second = 2 * op_bytes;
    if (prec == Prec::BF16X3) {
        if (!frag) req.push_back({kSlotOp1, op_bytes});
        req.push_back({kSlotDiag, sizeof(double) * rows});
    }
    if (prec == Prec::FP8) req.push_back({kSlotScale, nocc * nqb});
    if (prec == Prec::I8X4) {
        // The int32 level sums bound naux (their own range limit, like the
        // FP32 one below): beyond it the caller retries in FP64.
        if (pl.naux_pad > kI8MaxNauxPad)
            return fail(FH_RANGE, "naux_pad %d beyond the INT8 slices' int32 range (%d)",
                        pl.naux_pad, kI8MaxNauxPad);
        req.push_back({kSlotScale, sizeof(int) * rows});
    }
    fh_status st = reserve(s, req);
    if (st != FH_OK) return st;

    // INT8 slices carry their own power-of-two scale per row: any finite
    // value is in range.
    if (prec != Prec::F64 && prec != Prec::I8X4) {
        const size_t depth = prec == Prec::MIXED ? std::min(naux, (size_t)kMixedPanel)
                                                 : naux + (signed_k ? naux_neg : 0);
        st = range_check(s, b_dev, in_f32, naux * nov, depth, stream);
        if (st != FH_OK) return st;
        if (signed_k) {
            st = range_check(s, neg_dev, in_f32, naux_neg * nov, depth, stream);
            if (st != FH_OK) return st;
        }
    }
    std::vector<double> e(nocc + nvir);
    for (size_t i = 0; i < nocc; ++i) e[i] = eps[first_occ + i];
    for (size_t a = 0; a < nvir; ++a) e[nocc + a] = eps[nocc_total + a];
    double *deps = (double *)s.slots[kSlotEps].ptr;
    st = hip_check(hipMemcpyAsync(deps, e.data(), sizeof(double) * e.size(),
                                  hipMemcpyHostToDevice, stream),
                   "H2D eps");
    if (st != FH_OK) return st;
    s.pending.h2d_bytes += sizeof(double) * e.size();

    PackArgs pa;
    pa.b = b_dev;
    pa.in_f32 = in_f32;
    pa.naux = (int)naux;
    pa.nocc = (int)nocc;
    pa.nvir = (int)nvir;
    pa.nvir_pad = pl.nvir_pad;
    pa.naux_pad = pl.naux_pad;
    pa.fmt = prec == Prec::F64      ? PackFmt::F64
             : prec == Prec::F32 || prec == Prec::MIXED ? PackFmt::F32
             : prec == Prec::BF16X3 ? (frag ? PackFmt::BF16F : PackFmt::BF16X2)
             : prec == Prec::I8X4   ? PackFmt::I8X4
                                    : (frag ? PackFmt::FP8F : PackFmt::FP8);
    pa.out0 = s.slots[kSlotOp0].ptr;
    pa.out1 = prec == Prec::BF16X3 && !frag ? s.slots[kSlotOp1].ptr : nullptr;
    pa.scale = prec == Prec::FP8 ? (uint8_t *)s.slots[kSlotScale].ptr : nullptr;
    pa.lo_sq = prec == Prec::BF16X3 ? (double *)s.slots[kSlotDiag].ptr : nullptr;
    pa.rexp = prec == Prec::I8X4 ? (int *)s.slots[kSlotScale].ptr : nullptr;
    mark(s, 0, stream);
    hipError_t he;
    if (signed_k) {
        PackArgs pn = pa;
        pa.ld = pl.naux_pad;
        pa.naux_pad = q_neg;
        pn.ld = pl.naux_pad;
        pn.b = neg_dev;
        pn.naux = (int)naux_neg;
        pn.naux_pad = pl.naux_pad - q_neg;
        pn.out0 = (char *)pa.out0 + (size_t)q_neg * eb;
        he = launch_pack(pa, stream);
        if (he == hipSuccess) he = launch_pack(pn, stream);
    } else {
        he = launch_pack(pa, stream);
    }
    if (he != hipSuccess) {
        (void)hipGetLastError();
        return fail(FH_KERNEL, "pack kernel: %s", hipGetErrorString(he));
    }

    EnergyArgs ea;
    ea.prec = prec;
    ea.tile = pl.tile;
    ea.op0 = pa.out0;
    ea.op1 = pa.out1;
    ea.scale = pa.scale;
    ea.diag_corr = pa.lo_sq;
    ea.rexp = pa.rexp;
    ea.opf = frag ? pa.out0 : nullptr;
    ea.q_rot_ticks = prec == Prec::I8X4 ? i8_q_rot_ticks(s, pl.naux_pad / 64, pl.ntp) : 0;
    ea.nocc = (int)nocc;
    ea.nvir = (int)nvir;
    ea.nvir_pad = pl.nvir_pad;
    ea.naux_pad = pl.naux_pad;
    ea.nqb = (int)nqb;
    ea.eps_occ = deps;
    ea.eps_vir = deps + nocc;
    ea.kappa = kappa;
    ea.ntile = pl.ntile;
    ea.ntp = pl.ntp;
    ea.nitems = pl.nitems;
    ea.item_partials = (double *)s.slots[kSlotItems].ptr;
    ea.row_partials = (double *)s.slots[kSlotRows].ptr;
    ea.result = (double *)s.slots[kSlotResult].ptr;
    ea.q_neg = q_neg;
    ea.tamp = amp_out;
    mark(s, 1, stream);
    he = launch_energy(ea, stream);
    if (he != hipSuccess) {
        (void)hipGetLastError();
        return fail(FH_KERNEL, "energy kernel: %s", hipGetErrorString(he));
    }
    mark(s, 2, stream);
    he = launch_energy_reduce(ea, stream);
    if (he != hipSuccess) {
        (void)hipGetLastError();
        return fail(FH_KERNEL, "reduction kernels: %s", hipGetErrorString(he));
    }
    mark(s, 3, stream);
    s.pending.kernel_launches += 4;
    std::vector<double> res(2 + 2 * nocc);
    st = hip_check(hipMemcpyAsync(res.data(), ea.result, sizeof(double) * res.size(),
                                  hipMemcpyDeviceToHost, stream),
                   "D2H result");
    if (st != FH_OK) return st;
    st = hip_check(hipStreamSynchronize(stream), "stream sync");
    if (st != FH_OK) return st;
    s.pending.d2h_bytes += sizeof(double) * res.size();
    finish_timing(s, (uint64_t)pl.nitems, pl.tile);
    out->e_os = res[0];
    out->e_ss = res[1];
    out->e_total = res[0] + res[1];
    if (row_partials)
        for (size_t i = 0; i < 2 * nocc; ++i) row_partials[i] = res[2 + i];
    return FH_OK;
}

EnergyPlan plan_energy(size_t nocc, size_t nvir, size_t naux, Prec prec) {
    EnergyPlan p;
    // The tile table is measured on the batched kernel and applies to batched
    // domains; one-problem calls take the built-in choice unless
    // FERRIC_HIP_TILE forces a tile.
    (void)naux;
    p.tile = choose_tile(nvir, prec);
    const int forced = forced_tile();
    if (forced > 0) p.tile = forced;
    if (prec == Prec::I8X4 && p.tile > 64) p.tile = 64;  // 64 and 32 are built
    p.nvir_pad = (int)((nvir + 15) / 16 * 16);
    const int kp = std::max(panel_depth(prec), prec == Prec::FP8 ? 128 : 0);
    p.naux_pad = (int)((naux + kp - 1) / kp * kp);
    p.ntile = (p.nvir_pad + p.tile - 1) / p.tile;
    p.ntp = p.ntile * (p.ntile + 1) / 2;
    const int64_t npairs = (int64_t)nocc * (int64_t)(nocc + 1) / 2;
    p.nitems = npairs * p.ntp;
    return p;
}

void commit(State &s, Prec prec, bool adjoint, bool batched) {
    s.stats.offloaded++;
    s.stats.h2d_bytes += s.pending.h2d_bytes;
    s.stats.d2h_bytes += s.pending.d2h_bytes;
    s.stats.kernel_launches += s.pending.kernel_launches;
    switch (prec) {
    case Prec::F64: s.stats.fp64_calls++; break;
    case Prec::F32: s.stats.fp32_calls++; break;
    case Prec::MIXED: s.stats.mixed_calls++; break;
    case Prec::BF16X3: s.stats.bf16x3_calls++; break;
    case Prec::FP8: s.stats.fp8_calls++; break;
    case Prec::I8X4: s.stats.int8x4_calls++; break;
    }
    if (adjoint) s.stats.adjoint_calls++;
    if (batched) s.stats.batched_calls++;
}

bool timing_on(State &s) {
    if (s.timing < 0) {
        const char *env = getenv("FERRIC_TIMING");
        s.timing = (env && *env && strcmp(env, "0") != 0) ? 1 : 0;
    }
    return s.timing == 1;
}

void mark(State &s, int k, hipStream_t st) {
    if (!timing_on(s)) return;
    if (!s.ev[k] && hipEventCreate(&s.ev[k]) != hipSuccess) {
        s.ev[k] = nullptr;
        return;
    }
    (void)hipEventRecord(s.ev[k], st);
}

void finish_timing(State &s, uint64_t items, int tile) {
    s.last_timing = fh_timing{};
    if (!timing_on(s)) return;
    float t[3] = {0, 0, 0}, tot = 0;
    for (int k = 0; k < 3; ++k)
        if (s.ev[k] && s.ev[k + 1]) (void)hipEventElapsedTime(&t[k], s.ev[k], s.ev[k + 1]);
    if (s.ev[0] && s.ev[3]) (void)hipEventElapsedTime(&tot, s.ev[0], s.ev[3]);
    s.last_timing.pack_ms = t[0];
    s.last_timing.contract_ms = t[1];
    s.last_timing.reduce_ms = t[2];
    s.last_timing.total_ms = tot;
    s.last_timing.items = items;
    s.last_timing.tile = (uint32_t)tile;
}

bool to_prec(int p, Prec *out) {
    switch (p) {
    case FH_FP64: *out = Prec::F64; return true;
    case FH_FP32: *out = Prec::F32; return true;
    case FH_BF16X3: *out = Prec::BF16X3; return true;
    case FH_FP8_E4M3: *out = Prec::FP8; return true;
    case FH_INT8X4: *out = Prec::I8X4; return true;
    case FH_MIXED: *out = Prec::MIXED; return true;
    default: return false;
    }
}

}  // namespace fh

using namespace fh;

// ---- C ABI -----------------------------------------------------------------------

extern "C" {

int fh_abi_version(void) { return FERRIC_HIP_ABI_VERSION; }

const char *fh_last_error(void) { return t_last_error.c_str(); }

fh_status fh_init(int device) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    fh_status st = init_locked(s, device);
    return st;
}

fh_status fh_set_pool_limit(size_t bytes) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    s.pool_limit = bytes;
    s.pool_limit_explicit = bytes != 0;
    return FH_OK;
}

fh_status fh_get_pool_limit(size_t *bytes) {
    if (!bytes) return fail(FH_LAYOUT, "null output pointer");
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    fh_status st = ensure_device(s);
    if (st != FH_OK) return st;
    return resolve_pool_limit(s, bytes);
}

fh_status fh_release_workspace(void) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.initialised) {
        (void)hipSetDevice(s.device);
        release_all(s);
    }
    release_graphs(s);
    return FH_OK;
}

fh_status fh_get_stats(fh_stats *out) {
    if (!out) return fail(FH_LAYOUT, "null output pointer");
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    *out = s.stats;
    return FH_OK;
}

fh_status fh_reset_stats(void) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    memset(&s.stats, 0, sizeof s.stats);
    return FH_OK;
}

fh_status fh_set_timing(int on) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    s.timing = on ? 1 : 0;
    return FH_OK;
}

fh_status fh_last_timing(fh_timing *out) {
    if (!out) return fail(FH_LAYOUT, "null output pointer");
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    *out = s.last_timing;
    return FH_OK;
}

fh_status fh_rimp2_energy_dev(const void *b_dev, fh_dtype b_dtype, size_t naux, size_t nocc,
                              size_t nvir, const double *eps, size_t n_eps, size_t first_occ,
                              size_t nocc_total, double kappa, fh_precision precision,
                              void *stream, fh_spin_components *out, double *row_partials) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    if (!out) return fail(FH_LAYOUT, "null output pointer");
    Prec prec;
    if (!to_prec(precision, &prec)) return fail(FH_LAYOUT, "unknown precision %d", (int)precision);
    if (b_dtype != FH_DT_F64 && b_dtype != FH_DT_F32)
        return fail(FH_LAYOUT, "unknown b_dtype %d", (int)b_dtype);
    bool empty = false, has_k = false;
    fh_status st = check_shapes(naux, nocc, nvir, n_eps, first_occ, nocc_total, &empty);
    if (st != FH_OK) return st;
    st = check_kappa(kappa, &has_k);
    if (st != FH_OK) return st;
    if (empty) {
        *out = fh_spin_components{0.0, 0.0, 0.0};
        if (row_partials)
            for (size_t i = 0; i < 2 * nocc; ++i) row_partials[i] = 0.0;
        commit(s, prec, false, false);
        return FH_OK;
    }
    if (!b_dev || !eps) return fail(FH_LAYOUT, "null input pointer");
    st = ensure_device(s);
    if (st != FH_OK) return st;
    st = energy_core(s, b_dev, b_dtype, naux, nocc, nvir, eps, first_occ, nocc_total, kappa, prec,
                     (hipStream_t)stream, out, row_partials, nullptr);
    if (st != FH_OK) return st;
    commit(s, prec, false, false);
    return FH_OK;
}

fh_status fh_rimp2_energy(const double *b_ov, size_t naux, size_t nocc, size_t nvir,
                          const double *eps, size_t n_eps, size_t first_occ, size_t nocc_total,
                          double kappa, fh_precision precision, fh_spin_components *out) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    if (!out) return fail(FH_LAYOUT, "null output pointer");
    Prec prec;
    if (!to_prec(precision, &prec)) return fail(FH_LAYOUT, "unknown precision %d", (int)precision);
    bool empty = false, has_k = false;
    fh_status st = check_shapes(naux, nocc, nvir, n_eps, first_occ, nocc_total, &empty);
    if (st != FH_OK) return st;
    st = check_kappa(kappa, &has_k);
    if (st != FH_OK) return st;
    if (empty) {
        *out = fh_spin_components{0.0, 0.0, 0.0};
        commit(s, prec, false, false);
        return FH_OK;
    }
    if (!b_ov || !eps) return fail(FH_LAYOUT, "null input pointer");
    st = ensure_device(s);
    if (st != FH_OK) return st;
    // Reserve the staging copy of B together with the call's scratch so that
    // a refusal happens before the upload (rimp2_gpu.rs:326-335).
    const size_t bbytes = sizeof(double) * naux * nocc * nvir;
    EnergyPlan pl = plan_energy(nocc, nvir, naux, prec);
    const size_t eb = packed_elem_bytes(prec);
    const size_t op_bytes = packed_op_bytes(pl, nocc, prec);
    std::vector<std::pair<int, size_t>> req = {{kSlotStage, bbytes}, {kSlotOp0, op_bytes}};
    if (prec == Prec::BF16X3) req.push_back({kSlotOp1, op_bytes});
    st = reserve(s, req);
    if (st != FH_OK) return st;
    hipStream_t stream = nullptr;
    st = hip_check(hipMemcpyAsync(s.slots[kSlotStage].ptr, b_ov, bbytes, hipMemcpyHostToDevice,
                                  stream),
                   "H2D b_ov");
    if (st != FH_OK) return st;
    s.pending.h2d_bytes += bbytes;
    st = energy_core(s, s.slots[kSlotStage].ptr, FH_DT_F64, naux, nocc, nvir, eps, first_occ,
                     nocc_total, kappa, prec, stream, out, nullptr, &pl);
    if (st != FH_OK) return st;
    commit(s, prec, false, false);
    return FH_OK;
}

}  // extern "C"
