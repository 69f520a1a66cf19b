// Host-side state and helpers shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <utility>
#include <vector>

#include "../../include/ferric_hip.h"
#include "fh_kernels.h"

namespace fh {

enum SlotId {
    kSlotStage = 0,  // FP64 staging copy of a host b_ov
    kSlotOp0,        // packed operand (or BF16 hi)
    kSlotOp1,        // BF16 lo
    kSlotScale,      // FP8 scales
    kSlotEps,        // eps_occ | eps_vir
    kSlotItems,      // per-item partial sums
    kSlotRows,       // per-row partial sums
    kSlotResult,     // energies + row values
    kSlotScalar,     // small scalars (absmax)
    kSlotGamma,      // adjoint: FP64 gamma staging for host calls
    kSlotT,          // adjoint: per-row T~ block
    kSlotDeps,       // adjoint: dE/deps partials
    kSlotBatch,      // batched: descriptors and work list
    kSlotBatchOut,   // batched: per-domain partials
    kSlotDiag,       // BF16: lo^2 row sums
    kSlotStage2,     // adjoint: Q-major padded copy of B
    kSlotU,          // adjoint: b-major T~ blocks of one row
    kSlotAcc,        // adjoint: dE/deps accumulators
    kSlotSplit,      // adjoint: split-K partial tiles of phase B
    kSlotRed,        // adjoint: per-row reduction partials
    kSlotDressX,     // dressing: (Q|nu i) half-transformed rows
    kSlotDressY,     // dressing: undressed MO rows (Q|ia)
    kSlotB2Split,    // INT8X4 adjoint: BF16 hi and lo of the Q-major copy of B
    kSlotTSplit,     // INT8X4 adjoint: BF16 hi and lo of a row group's T~ slabs S and St
    kSlotG32,        // INT8X4 adjoint: FP32 gamma over the padded index space
    kSlotPairs,      // pair energies: (e_os, e_ss) per unique pair
    kSlotPnoWork,    // PNO: five nvir x nvir matrices and three vectors per batched pair
    kSlotPnoMeta,    // PNO: pair list, per-pair results, masks, GEMM pointer arrays
    kSlotFitWork,    // domain fits: padded V_DD and [A_i | A_j] per batched pair
    kSlotFitMeta,    // domain fits: domain index lists, sizes, pairs, Cholesky info
    kSlotGPairs,     // energies from explicit (ia|jb) blocks: the pair list
    kSlotRexpB,      // INT8X4 opposite spin: row exponents of the beta rows
    kSlotDensAmp,    // MP2 densities: amplitude blocks of every ordered occupied pair
    kSlotDensPart,   // MP2 densities: split-K partial tiles
    kSlotDensOut,    // MP2 densities: P_vv (padded) then P_oo
    kNumSlots
};

struct Slot {
    void *ptr = nullptr;
    size_t bytes = 0;
};

struct Pending {
    uint64_t h2d_bytes = 0, d2h_bytes = 0, kernel_launches = 0;
};

struct GraphCache;  // batched.hip

struct State {
    std::mutex mu;
    bool initialised = false;
    int device = 0;
    size_t pool_limit = 0;
    bool pool_limit_explicit = false;
    Slot slots[kNumSlots];
    fh_stats stats{};
    Pending pending;
    GraphCache *graphs = nullptr;
    void *blas = nullptr;  // rocBLAS handle of the device, made on first use (blas_handle)
    // Phase timing (FERRIC_TIMING): events on the call's stream.
    int timing = -1;  // -1: not yet read from the environment
    int wall_khz = -1;  // device wall-clock rate (kHz); -1: not yet queried
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    fh_timing last_timing{};
    // Fixed arena (fh_rimp2_energy_dev_noalloc): while set, reserve() carves
    // the requested slots out of it in request order and never allocates.
    char *arena = nullptr;
    size_t arena_bytes = 0, arena_used = 0;
    size_t held_bytes() const;
};

// Phase markers: record(k) for k = 0 (start), 1 (after pack), 2 (after the
// contraction), 3 (after the reductions); finish() after the stream sync.
bool timing_on(State &s);
// Wall-clock ticks per Q panel for the INT8X4 kernels' panel rotation (0: off)
// for a problem of npanels 64-deep Q panels and ntp tile pairs per occupied pair.
int i8_q_rot_ticks(State &s, int npanels, int ntp);
void mark(State &s, int k, hipStream_t st);
void finish_timing(State &s, uint64_t items, int tile);

struct EnergyPlan {
    int tile = 128;
    int nvir_pad = 16;
    int naux_pad = 128;
    int ntile = 1;
    int ntp = 1;
    int64_t nitems = 0;
};

State &state();
fh_status fail(fh_status st, const char *fmt, ...);
void clear_error();
fh_status ensure_device(State &s);
fh_status reserve(State &s, const std::vector<std::pair<int, size_t>> &req);
void release_all(State &s);
fh_status hip_check(hipError_t e, const char *what);
fh_status check_shapes(size_t naux, size_t nocc, size_t nvir, size_t n_eps, size_t first_occ,
                       size_t nocc_total, bool *empty);
fh_status check_kappa(double kappa, bool *has);
fh_status range_check(State &s, const void *b, bool in_f32, size_t n, size_t depth,
                      hipStream_t stream);
EnergyPlan plan_energy(size_t nocc, size_t nvir, size_t naux, Prec prec);
// FERRIC_HIP_TILE's tile side, or 0 when it is unset.
int forced_tile();
// Tile side of n batched domains: FERRIC_HIP_TILE's if set, else the first
// tile-table entry whose bounds hold the domain, else 0 (one environment
// read, one lock).
void tuned_tiles(Prec prec, size_t n, const size_t *nvir, const size_t *naux, int *out);
size_t packed_elem_bytes(Prec p);
// The device's rocBLAS handle (created on first use, atomics off so that its
// GEMMs are deterministic), bound to stream.
fh_status blas_handle(State &s, hipStream_t stream, void **handle);
// Bytes of one packed operand array of plan pl (INT8X4 adds TILE spare rows).
size_t packed_op_bytes(const EnergyPlan &pl, size_t nocc, Prec prec);
// neg_dev (naux_neg rows, same layout and dtype as b_dev) makes the kernel
// signed: K = B^T B - N^T N (FP64 / FP32 only).
fh_status energy_core(State &s, const void *b_dev, fh_dtype b_dtype, size_t naux, size_t nocc,
                      size_t nvir, const double *eps, size_t first_occ, size_t nocc_total,
                      double kappa, Prec prec, hipStream_t stream, fh_spin_components *out,
                      double *row_partials, const EnergyPlan *forced,
                      const void *neg_dev = nullptr, size_t naux_neg = 0,
                      double *amp_out = nullptr);
void commit(State &s, Prec prec, bool adjoint, bool batched);
bool to_prec(int p, Prec *out);
void release_graphs(State &s);

}  // namespace fh
