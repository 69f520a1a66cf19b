// fh_rimp2_energy_dev_noalloc: the device-resident energy entry with its
// workspace in a fixed device array instead of the library's pool, for callers
// that may not allocate device memory (code under graph capture, serving
// loops, harnesses that forbid hipMalloc).
//
// The workspace is `scratch`, a file-scope __device__ array of
// FH_NOALLOC_SCRATCH_BYTES. Its host-side name is not a device address: the
// address is looked up once per device with hipGetSymbolAddress, with that
// device current, and only that pointer is handed to kernels and copies. Every
// call on one device shares the array, and a call leaves its contents behind;
// no workspace byte left by an earlier call reaches a result (the packer
// writes every padded operand element the MFMAs use, each item writes its own
// partial, the row and result slots are written before they are read, the
// absmax word is cleared by a kernel; the INT8 operand's spare rows feed only
// blocks the kernel switches off), so nothing needs clearing. Calls that
// overlap on one device race on it: the library serialises its own calls,
// and a caller running the energy on several streams or processes of one
// device at once must not use this entry.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "../../include/ferric_hip.h"
#include "fh_host.h"
#include "fh_kernels.h"

namespace fh {
namespace {

__device__ __align__(256) unsigned char scratch[FH_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on the
// first call there). Called with the library's lock held.
fh_status scratch_address(int device, void **out) {
    if (device < 0 || device >= kMaxDevices)
        return fail(FH_DEVICE, "device ordinal %d beyond %d", device, kMaxDevices);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(FH_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                        hipGetErrorString(e));
        }
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return FH_OK;
}

// Points the state's slots into the arena for one call and puts the pool's
// slots back afterwards, whatever the call returned.
struct ArenaScope {
    State &s;
    Slot saved[kNumSlots];
    ArenaScope(State &st, void *base, size_t bytes) : s(st) {
        for (int k = 0; k < kNumSlots; ++k) {
            saved[k] = s.slots[k];
            s.slots[k] = Slot{};
        }
        s.arena = (char *)base;
        s.arena_bytes = bytes;
        s.arena_used = 0;
    }
    ~ArenaScope() {
        for (int k = 0; k < kNumSlots; ++k) s.slots[k] = saved[k];
        s.arena = nullptr;
        s.arena_bytes = 0;
        s.arena_used = 0;
    }
};

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" fh_status fh_rimp2_energy_dev_noalloc(const void *b_dev, fh_dtype b_dtype, size_t naux,
                                                 size_t nocc, size_t nvir, const double *eps,
                                                 size_t n_eps, size_t first_occ,
                                                 size_t nocc_total, double kappa,
                                                 fh_precision precision, void *stream,
                                                 fh_spin_components *out,
                                                 double *row_partials) {
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
    void *base = nullptr;
    st = scratch_address(s.device, &base);
    if (st != FH_OK) return st;
    {
        ArenaScope scope(s, base, (size_t)FH_NOALLOC_SCRATCH_BYTES);
        st = energy_core(s, b_dev, b_dtype, naux, nocc, nvir, eps, first_occ, nocc_total, kappa,
                         prec, (hipStream_t)stream, out, row_partials, nullptr);
    }
    if (st != FH_OK) return st;
    commit(s, prec, false, false);
    return FH_OK;
}
