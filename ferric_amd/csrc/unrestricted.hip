// Unrestricted (open-shell) RI-MP2 blocks and the signed (terf) factor: host
// orchestration and C ABI.
//
// Same spin (u_rimp2.rs:693-729): E_ss' = 1/4 sum_{i<=j} f_ij sum_ab
// (g_ab - g_ba)^2 / D of one spin channel. The closed-shell pair kernel already
// accumulates E_ss = sum_{i<=j} f_ij * 1/2 sum_ab (g_ab - g_ba)^2 / D (its
// symmetric tile pairs see both orientations), so the same-spin energy is
// E_ss / 2 of that kernel run on the channel's B and eps.
//
// Opposite spin (u_rimp2.rs:731-760): sum over every alpha i, beta J and
// virtuals a, B of (ia|JB)^2 / D — a rectangle without the pair symmetry. Its
// items (i, J, tA, tB) run the K4 body in RECT mode (pair_body.h): one
// product per item, alpha operand rows from one packed array and beta rows
// from another, both packed over the same Q padding.
//
// Signed factor (rimp2.rs:1559-1607): K = B+^T B+ - B-^T B-, the two factors
// stacked along Q in one packed operand (energy_core with neg_dev).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "fh_host.h"
#include "fh_kernels.h"

namespace fh {

namespace {

fh_status check_channel(const fh_spin_channel *ch, size_t naux, bool *empty) {
    if (!ch) return fail(FH_LAYOUT, "null spin channel");
    fh_status st = check_shapes(naux, ch->nocc, ch->nvir, ch->n_eps, ch->first_occ,
                                ch->nocc_total, empty);
    if (st != FH_OK) return st;
    if (!*empty && (!ch->b || !ch->eps)) return fail(FH_LAYOUT, "null spin-channel pointer");
    return FH_OK;
}

fh_status opposite_core(State &s, const fh_spin_channel &ca, const fh_spin_channel &cb,
                        const void *ba, const void *bb, fh_dtype dt, size_t naux, Prec prec,
                        hipStream_t stream, double *energy) {
    const bool in_f32 = dt == FH_DT_F32;
    const EnergyPlan pa_plan = plan_energy(ca.nocc, std::max(ca.nvir, cb.nvir), naux, prec);
    const int tile = pa_plan.tile;
    const int vpa = (int)((ca.nvir + 15) / 16 * 16), vpb = (int)((cb.nvir + 15) / 16 * 16);
    const int ntile_a = (vpa + tile - 1) / tile, ntile_b = (vpb + tile - 1) / tile;
    const int naux_pad = pa_plan.naux_pad;
    const int64_t nitems = (int64_t)ca.nocc * (int64_t)cb.nocc * ntile_a * ntile_b;
    if (nitems > 0x7fffffffLL) return fail(FH_UNSUPPORTED, "%lld work items", (long long)nitems);
    const size_t eb = packed_elem_bytes(prec);
    // INT8X4 operands carry TILE spare rows after their last (the fetch of a
    // narrow last tile reads whole 16-row blocks)
    const bool i8 = prec == Prec::I8X4;
    const size_t rows_a = ca.nocc * (size_t)vpa + (i8 ? (size_t)tile : 0);
    const size_t rows_b = cb.nocc * (size_t)vpb + (i8 ? (size_t)tile : 0);
    const size_t bytes_a = rows_a * naux_pad * eb;
    const size_t bytes_b = rows_b * naux_pad * eb;
    const size_t neps = ca.nocc + ca.nvir + cb.nocc + cb.nvir;
    std::vector<std::pair<int, size_t>> req = {{kSlotOp0, bytes_a},
                                               {kSlotOp1, bytes_b},
                                               {kSlotEps, sizeof(double) * neps},
                                               {kSlotItems, sizeof(double) * 4 * (size_t)nitems},
                                               {kSlotRows, sizeof(double) * 4 * ca.nocc},
                                               {kSlotResult, sizeof(double) * (2 + 2 * ca.nocc)},
                                               {kSlotScalar, 256}};
    if (i8) {
        if (naux_pad > kI8MaxNauxPad)
            return fail(FH_RANGE, "naux_pad %d beyond the INT8 slices' int32 range (%d)",
                        naux_pad, kI8MaxNauxPad);
        req.push_back({kSlotScale, sizeof(int) * rows_a});
        req.push_back({kSlotRexpB, sizeof(int) * rows_b});
    }
    fh_status st = reserve(s, req);
    if (st != FH_OK) return st;
    if (prec != Prec::F64 && !i8) {
        const size_t depth = prec == Prec::MIXED ? std::min(naux, (size_t)kMixedPanel) : naux;
        st = range_check(s, ba, in_f32, naux * ca.nocc * ca.nvir, depth, stream);
        if (st == FH_OK) st = range_check(s, bb, in_f32, naux * cb.nocc * cb.nvir, depth, stream);
        if (st != FH_OK) return st;
    }
    std::vector<double> e;
    e.reserve(neps);
    for (size_t i = 0; i < ca.nocc; ++i) e.push_back(ca.eps[ca.first_occ + i]);
    for (size_t a = 0; a < ca.nvir; ++a) e.push_back(ca.eps[ca.nocc_total + a]);
    for (size_t i = 0; i < cb.nocc; ++i) e.push_back(cb.eps[cb.first_occ + i]);
    for (size_t a = 0; a < cb.nvir; ++a) e.push_back(cb.eps[cb.nocc_total + a]);
    double *deps = (double *)s.slots[kSlotEps].ptr;
    st = hip_check(hipMemcpyAsync(deps, e.data(), sizeof(double) * e.size(),
                                  hipMemcpyHostToDevice, stream),
                   "H2D eps");
    if (st != FH_OK) return st;
    s.pending.h2d_bytes += sizeof(double) * e.size();

    PackArgs pk;
    pk.in_f32 = in_f32;
    pk.naux = (int)naux;
    pk.naux_pad = naux_pad;
    pk.fmt = prec == Prec::F64 ? PackFmt::F64 : i8 ? PackFmt::I8X4 : PackFmt::F32;
    pk.out1 = nullptr;
    pk.scale = nullptr;
    pk.lo_sq = nullptr;
    int *rexp_a = i8 ? (int *)s.slots[kSlotScale].ptr : nullptr;
    int *rexp_b = i8 ? (int *)s.slots[kSlotRexpB].ptr : nullptr;
    mark(s, 0, stream);
    pk.b = ba;
    pk.nocc = (int)ca.nocc;
    pk.nvir = (int)ca.nvir;
    pk.nvir_pad = vpa;
    pk.out0 = s.slots[kSlotOp0].ptr;
    pk.rexp = rexp_a;
    hipError_t he = launch_pack(pk, stream);
    if (he == hipSuccess) {
        pk.b = bb;
        pk.nocc = (int)cb.nocc;
        pk.nvir = (int)cb.nvir;
        pk.nvir_pad = vpb;
        pk.out0 = s.slots[kSlotOp1].ptr;
        pk.rexp = rexp_b;
        he = launch_pack(pk, stream);
    }
    if (he != hipSuccess) {
        (void)hipGetLastError();
        return fail(FH_KERNEL, "pack kernel: %s", hipGetErrorString(he));
    }

    EnergyArgs ea;
    ea.prec = prec;
    ea.tile = tile;
    ea.op0 = s.slots[kSlotOp0].ptr;
    ea.op1 = s.slots[kSlotOp1].ptr;
    ea.scale = nullptr;
    ea.diag_corr = nullptr;
    ea.nocc = (int)ca.nocc;
    ea.nvir = (int)ca.nvir;
    ea.nvir_pad = vpa;
    ea.naux_pad = naux_pad;
    ea.nqb = 0;
    ea.eps_occ = deps;
    ea.eps_vir = deps + ca.nocc;
    ea.kappa = NAN;
    ea.ntile = ntile_a;
    ea.ntp = ntile_a * ntile_b;
    ea.nitems = nitems;
    ea.item_partials = (double *)s.slots[kSlotItems].ptr;
    ea.row_partials = (double *)s.slots[kSlotRows].ptr;
    ea.result = (double *)s.slots[kSlotResult].ptr;
    ea.nocc_b = (int)cb.nocc;
    ea.nvir_b = (int)cb.nvir;
    ea.nvir_pad_b = vpb;
    ea.ntile_b = ntile_b;
    ea.eps_occ_b = deps + ca.nocc + ca.nvir;
    ea.eps_vir_b = deps + ca.nocc + ca.nvir + cb.nocc;
    ea.rexp = rexp_a;
    ea.rexp_b = rexp_b;
    ea.q_rot_ticks = i8 ? i8_q_rot_ticks(s, ea.naux_pad / 64, ea.ntile * ea.ntile_b) : 0;
    mark(s, 1, stream);
    he = launch_energy_rect(ea, stream);
    if (he != hipSuccess) {
        (void)hipGetLastError();
        return fail(FH_KERNEL, "opposite-spin kernel: %s", hipGetErrorString(he));
    }
    mark(s, 2, stream);
    he = launch_rect_reduce(ea, stream);
    if (he != hipSuccess) {
        (void)hipGetLastError();
        return fail(FH_KERNEL, "reduction kernels: %s", hipGetErrorString(he));
    }
    mark(s, 3, stream);
    s.pending.kernel_launches += 6;
    double res[2] = {0.0, 0.0};
    st = hip_check(hipMemcpyAsync(res, ea.result, sizeof res, hipMemcpyDeviceToHost, stream),
                   "D2H result");
    if (st != FH_OK) return st;
    st = hip_check(hipStreamSynchronize(stream), "stream sync");
    if (st != FH_OK) return st;
    s.pending.d2h_bytes += sizeof res;
    finish_timing(s, (uint64_t)nitems, tile);
    *energy = res[0];
    return FH_OK;
}

// Shared front end of the same-spin entries.
fh_status same_spin_call(State &s, const fh_spin_channel *ch, const void *b_dev, fh_dtype dt,
                         size_t naux, Prec prec, hipStream_t stream, double *energy) {
    fh_spin_components sc{};
    fh_status st = energy_core(s, b_dev, dt, naux, ch->nocc, ch->nvir, ch->eps, ch->first_occ,
                               ch->nocc_total, NAN, prec, stream, &sc, nullptr, nullptr);
    if (st != FH_OK) return st;
    *energy = 0.5 * sc.e_ss;
    return FH_OK;
}

// Host entries stage their inputs through these slots, reserved with the
// call's scratch so that a refusal happens before any upload.
fh_status stage(State &s, int slot, const double *host, size_t n, hipStream_t stream) {
    fh_status st = hip_check(hipMemcpyAsync(s.slots[slot].ptr, host, sizeof(double) * n,
                                            hipMemcpyHostToDevice, stream),
                             "H2D b");
    if (st == FH_OK) s.pending.h2d_bytes += sizeof(double) * n;
    return st;
}

}  // namespace
}  // namespace fh

using namespace fh;

extern "C" {

fh_status fh_u_same_spin_dev(const fh_spin_channel *ch, fh_dtype b_dtype, size_t naux,
                             fh_precision precision, void *stream, double *energy) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    if (!energy) return fail(FH_LAYOUT, "null output pointer");
    Prec prec;
    if (!to_prec(precision, &prec)) return fail(FH_LAYOUT, "unknown precision %d", (int)precision);
    if (b_dtype != FH_DT_F64 && b_dtype != FH_DT_F32)
        return fail(FH_LAYOUT, "unknown b_dtype %d", (int)b_dtype);
    bool empty = false;
    fh_status st = check_channel(ch, naux, &empty);
    if (st != FH_OK) return st;
    if (empty) {
        *energy = 0.0;
        commit(s, prec, false, false);
        return FH_OK;
    }
    st = ensure_device(s);
    if (st != FH_OK) return st;
    st = same_spin_call(s, ch, ch->b, b_dtype, naux, prec, (hipStream_t)stream, energy);
    if (st != FH_OK) return st;
    commit(s, prec, false, false);
    return FH_OK;
}

fh_status fh_u_same_spin(const fh_spin_channel *ch, size_t naux, fh_precision precision,
                         double *energy) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    if (!energy) return fail(FH_LAYOUT, "null output pointer");
    Prec prec;
    if (!to_prec(precision, &prec)) return fail(FH_LAYOUT, "unknown precision %d", (int)precision);
    bool empty = false;
    fh_status st = check_channel(ch, naux, &empty);
    if (st != FH_OK) return st;
    if (empty) {
        *energy = 0.0;
        commit(s, prec, false, false);
        return FH_OK;
    }
    st = ensure_device(s);
    if (st != FH_OK) return st;
    const size_t n = naux * ch->nocc * ch->nvir;
    st = reserve(s, {{kSlotStage, sizeof(double) * n}});
    if (st != FH_OK) return st;
    st = stage(s, kSlotStage, (const double *)ch->b, n, nullptr);
    if (st != FH_OK) return st;
    st = same_spin_call(s, ch, s.slots[kSlotStage].ptr, FH_DT_F64, naux, prec, nullptr, energy);
    if (st != FH_OK) return st;
    commit(s, prec, false, false);
    return FH_OK;
}

static fh_status opposite_entry(const fh_spin_channel *alpha, const fh_spin_channel *beta,
                                fh_dtype b_dtype, size_t naux, fh_precision precision,
                                void *stream, double *energy, bool host) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    if (!energy) return fail(FH_LAYOUT, "null output pointer");
    Prec prec;
    if (!to_prec(precision, &prec)) return fail(FH_LAYOUT, "unknown precision %d", (int)precision);
    if (prec != Prec::F64 && prec != Prec::F32 && prec != Prec::MIXED && prec != Prec::I8X4)
        return fail(FH_UNSUPPORTED,
                    "the opposite-spin kernel runs in FP64, FP32, MIXED or INT8X4");
    if (b_dtype != FH_DT_F64 && b_dtype != FH_DT_F32)
        return fail(FH_LAYOUT, "unknown b_dtype %d", (int)b_dtype);
    bool ea = false, eb = false;
    fh_status st = check_channel(alpha, naux, &ea);
    if (st != FH_OK) return st;
    st = check_channel(beta, naux, &eb);
    if (st != FH_OK) return st;
    if (ea || eb) {
        *energy = 0.0;
        commit(s, prec, false, false);
        return FH_OK;
    }
    st = ensure_device(s);
    if (st != FH_OK) return st;
    const void *ba = alpha->b, *bb = beta->b;
    hipStream_t strm = (hipStream_t)stream;
    if (host) {
        const size_t na = naux * alpha->nocc * alpha->nvir, nb = naux * beta->nocc * beta->nvir;
        st = reserve(s, {{kSlotStage, sizeof(double) * na}, {kSlotStage2, sizeof(double) * nb}});
        if (st != FH_OK) return st;
        st = stage(s, kSlotStage, (const double *)ba, na, strm);
        if (st == FH_OK) st = stage(s, kSlotStage2, (const double *)bb, nb, strm);
        if (st != FH_OK) return st;
        ba = s.slots[kSlotStage].ptr;
        bb = s.slots[kSlotStage2].ptr;
    }
    st = opposite_core(s, *alpha, *beta, ba, bb, b_dtype, naux, prec, strm, energy);
    if (st != FH_OK) return st;
    commit(s, prec, false, false);
    return FH_OK;
}

fh_status fh_u_opposite_spin_dev(const fh_spin_channel *alpha, const fh_spin_channel *beta,
                                 fh_dtype b_dtype, size_t naux, fh_precision precision,
                                 void *stream, double *energy) {
    return opposite_entry(alpha, beta, b_dtype, naux, precision, stream, energy, false);
}

fh_status fh_u_opposite_spin(const fh_spin_channel *alpha, const fh_spin_channel *beta,
                             size_t naux, fh_precision precision, double *energy) {
    return opposite_entry(alpha, beta, FH_DT_F64, naux, precision, nullptr, energy, true);
}

static fh_status signed_entry(const void *pos, size_t naux_pos, const void *neg, size_t naux_neg,
                              fh_dtype b_dtype, size_t nocc, size_t nvir, const double *eps,
                              size_t n_eps, size_t first_occ, size_t nocc_total,
                              fh_precision precision, void *stream, fh_spin_components *out,
                              bool host) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    if (!out) return fail(FH_LAYOUT, "null output pointer");
    Prec prec;
    if (!to_prec(precision, &prec)) return fail(FH_LAYOUT, "unknown precision %d", (int)precision);
    if (prec != Prec::F64 && prec != Prec::F32)
        return fail(FH_UNSUPPORTED, "the signed kernel runs in FP64 or FP32");
    if (b_dtype != FH_DT_F64 && b_dtype != FH_DT_F32)
        return fail(FH_LAYOUT, "unknown b_dtype %d", (int)b_dtype);
    bool empty = false;
    fh_status st = check_shapes(naux_pos, nocc, nvir, n_eps, first_occ, nocc_total, &empty);
    if (st != FH_OK) return st;
    if (empty) {
        *out = fh_spin_components{0.0, 0.0, 0.0};
        commit(s, prec, false, false);
        return FH_OK;
    }
    if (!pos || !eps || (naux_neg > 0 && !neg)) return fail(FH_LAYOUT, "null input pointer");
    st = ensure_device(s);
    if (st != FH_OK) return st;
    hipStream_t strm = (hipStream_t)stream;
    if (host) {
        const size_t np = naux_pos * nocc * nvir, nn = naux_neg * nocc * nvir;
        std::vector<std::pair<int, size_t>> req = {{kSlotStage, sizeof(double) * np}};
        if (nn) req.push_back({kSlotStage2, sizeof(double) * nn});
        st = reserve(s, req);
        if (st != FH_OK) return st;
        st = stage(s, kSlotStage, (const double *)pos, np, strm);
        if (st == FH_OK && nn) st = stage(s, kSlotStage2, (const double *)neg, nn, strm);
        if (st != FH_OK) return st;
        pos = s.slots[kSlotStage].ptr;
        neg = nn ? s.slots[kSlotStage2].ptr : nullptr;
    }
    st = energy_core(s, pos, b_dtype, naux_pos, nocc, nvir, eps, first_occ, nocc_total, NAN, prec,
                     strm, out, nullptr, nullptr, naux_neg ? neg : nullptr, naux_neg);
    if (st != FH_OK) return st;
    commit(s, prec, false, false);
    return FH_OK;
}

fh_status fh_rimp2_energy_signed_dev(const void *pos, size_t naux_pos, const void *neg,
                                     size_t naux_neg, fh_dtype b_dtype, size_t nocc, size_t nvir,
                                     const double *eps, size_t n_eps, size_t first_occ,
                                     size_t nocc_total, fh_precision precision, void *stream,
                                     fh_spin_components *out) {
    return signed_entry(pos, naux_pos, neg, naux_neg, b_dtype, nocc, nvir, eps, n_eps,
                        first_occ, nocc_total, precision, stream, out, false);
}

fh_status fh_rimp2_energy_signed(const double *pos, size_t naux_pos, const double *neg,
                                 size_t naux_neg, size_t nocc, size_t nvir, const double *eps,
                                 size_t n_eps, size_t first_occ, size_t nocc_total,
                                 fh_precision precision, fh_spin_components *out) {
    return signed_entry(pos, naux_pos, neg, naux_neg, FH_DT_F64, nocc, nvir, eps, n_eps,
                        first_occ, nocc_total, precision, nullptr, out, true);
}

}  // extern "C"
