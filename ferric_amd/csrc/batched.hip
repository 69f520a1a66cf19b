// Batched local-MP2 pair domains: many independent RI-MP2 pair-energy
// problems of uneven shape in one persistent launch per tile class (gfx950).
//
// The reference evaluates its pair domains one after another on the host
// (crates/ferric-mp2/src/dlpno_mp2.rs:718-745, one nvir x nvir GEMM and pair
// loop per pair). Here every (domain, occupied pair i <= j, virtual tile pair)
// is one work item of a device-wide queue; workgroups of a persistent kernel
// pull items, most expensive domains first, and run the same fused
// contraction + FP64 epilogue as the one-problem kernels (pair_body.h). Each
// item writes its own compensated partial, and one workgroup per domain folds
// its items in item order, so the energies do not depend on which workgroup
// ran which item. The tile side of each domain comes from the tile table
// (tuning.hip, filled by ferric_amd/autotune.py); domains are grouped by tile
// side, one persistent launch per group. The whole sequence (queue reset,
// packing, persistent launches, reductions) can be captured once in a hipGraph
// and replayed for the same descriptor set.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "fh_device.h"
#include "fh_host.h"
#include "fh_kernels.h"
#include "pair_body.h"

namespace fh {

namespace {

struct DomDev {
    int64_t b_off;      // caller B (row-major (naux, nocc*nvir)) offset, elements
    int64_t eps_off;    // [occupied..., virtual...] offset
    int64_t bp_off;     // packed rows offset (elements), same in the lo arena
    int64_t row_off;    // first packed row (for the BF16 diagonal correction)
    int64_t item0;      // first work item of the domain
    int nocc, nvir, naux, nvir_pad, naux_pad, ntile, ntp, tile;
};

struct PackTile {
    int d, r0, q0, pad;
};

struct BatchArgs {
    Prec prec;
    const double *b;          // caller B arena (FP64)
    const double *eps;        // caller eps arena
    const DomDev *dom;
    const int *item_dom;      // domain of each work item
    const PackTile *ptiles;
    int64_t nptiles;
    void *op0;                // packed FP64 rows, or BF16 hi
    void *op1;                // BF16 lo
    double *corr;             // BF16 sum_Q lo^2 per packed row
    int64_t nrows;            // packed rows over all domains
    const int64_t *row_dom_start;  // first packed row of each domain (ndom + 1)
    int ndom;
    double *item_partials;
    int *counters;            // one queue counter per tile class
    double *out;              // (e_os, e_ss) per domain
};

// ---- packing -------------------------------------------------------------------
// One 64 x 64 tile of one domain: rows r of the domain's packed block, Q in
// [q0, q0 + 64). Same layout and zero padding as pack.hip.
__global__ __launch_bounds__(256) void batched_pack_kernel(BatchArgs a) {
    __shared__ double tile[64][65];
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    for (int64_t t = blockIdx.x; t < a.nptiles; t += gridDim.x) {
        const PackTile pt = a.ptiles[t];
        const DomDev dm = a.dom[pt.d];
        const int64_t nrows = (int64_t)dm.nocc * dm.nvir_pad;
        const int64_t nov = (int64_t)dm.nocc * dm.nvir;
        const int64_t r = pt.r0 + tx;
        int64_t col = -1;
        if (r < nrows) {
            const int64_t i = r / dm.nvir_pad;
            const int aa = (int)(r % dm.nvir_pad);
            if (aa < dm.nvir) col = i * dm.nvir + aa;
        }
        __syncthreads();
        for (int q = ty; q < 64; q += 4) {
            const int qq = pt.q0 + q;
            double v = 0.0;
            if (col >= 0 && qq < dm.naux) v = a.b[dm.b_off + (int64_t)qq * nov + col];
            tile[q][tx] = v;
        }
        __syncthreads();
        for (int rr = ty; rr < 64; rr += 4) {
            const int64_t orow = pt.r0 + rr;
            if (orow < nrows && pt.q0 + tx < dm.naux_pad) {
                const int64_t o = dm.bp_off + orow * dm.naux_pad + pt.q0 + tx;
                const double x = tile[tx][rr];
                if (a.prec == Prec::F64) {
                    ((double *)a.op0)[o] = x;
                } else {
                    const float xf = (float)x;
                    const __bf16 h = (__bf16)xf;
                    const __bf16 l = (__bf16)(xf - (float)h);
                    ((uint16_t *)a.op0)[o] = __builtin_bit_cast(uint16_t, h);
                    ((uint16_t *)a.op1)[o] = __builtin_bit_cast(uint16_t, l);
                }
            }
        }
    }
}

// BF16: sum over Q of lo^2 per packed row (one wave per row).
__global__ __launch_bounds__(256) void batched_lo_sq_kernel(BatchArgs a) {
    const int lane = threadIdx.x % kWave;
    for (int64_t row = (int64_t)blockIdx.x * 4 + threadIdx.x / kWave; row < a.nrows;
         row += (int64_t)gridDim.x * 4) {
        // domain of this packed row: last d with row_dom_start[d] <= row
        int lo = 0, hi = a.ndom - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (a.row_dom_start[mid] <= row) lo = mid; else hi = mid - 1;
        }
        const DomDev dm = a.dom[lo];
        const int64_t local = row - dm.row_off;
        const uint16_t *p = (const uint16_t *)a.op1 + dm.bp_off + local * dm.naux_pad;
        double s = 0.0;
        for (int q = lane; q < dm.naux_pad; q += kWave) {
            const float v = (float)__builtin_bit_cast(__bf16, p[q]);
            s += (double)(v * v);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, kWave);
        if (lane == 0) a.corr[row] = s;
    }
}

__global__ void reset_counters_kernel(int *c, int n) {
    if (threadIdx.x < n) c[threadIdx.x] = 0;
}

// ---- persistent pair kernel ------------------------------------------------------
// Items [item_begin, item_end) of one tile class, taken from a device-wide
// counter. Every workgroup exits once the counter passes item_end.
template <int TILE, bool BF16>
__global__ __launch_bounds__(256) void batched_pair_kernel(BatchArgs a, int64_t item_begin,
                                                           int64_t item_end, int cls) {
    __shared__ __attribute__((aligned(16))) pair::K4Lds<double, TILE, 16> lds64;
    __shared__ __attribute__((aligned(16))) pair::Bf16Lds<TILE> ldsbf;
    __shared__ int64_t s_item;
    EnergyArgs e;
    e.prec = BF16 ? Prec::BF16X3 : Prec::F64;
    e.tile = TILE;
    e.scale = nullptr;
    e.nqb = 0;
    e.kappa = NAN;
    e.item_partials = a.item_partials;
    e.row_partials = nullptr;
    e.result = nullptr;
    for (;;) {
        if (threadIdx.x == 0) s_item = item_begin + atomicAdd(&a.counters[cls], 1);
        __syncthreads();
        const int64_t item = s_item;
        __syncthreads();
        if (item >= item_end) break;
        const DomDev dm = a.dom[a.item_dom[item]];
        const int64_t local = item - dm.item0;
        pair::Item it;
        it.id = item;
        decode_pair(local / dm.ntp, dm.nocc, it.i, it.j);
        decode_tile_pair((int)(local % dm.ntp), dm.ntile, it.ta, it.tb);
        e.nocc = dm.nocc;
        e.nvir = dm.nvir;
        e.nvir_pad = dm.nvir_pad;
        e.naux_pad = dm.naux_pad;
        e.ntile = dm.ntile;
        e.ntp = dm.ntp;
        e.eps_occ = a.eps + dm.eps_off;
        e.eps_vir = a.eps + dm.eps_off + dm.nocc;
        if constexpr (BF16) {
            e.op0 = (const uint16_t *)a.op0 + dm.bp_off;
            e.op1 = (const uint16_t *)a.op1 + dm.bp_off;
            e.diag_corr = a.corr + dm.row_off;
            pair::pair_bf16x3_body<TILE, false>(e, it, ldsbf);
        } else {
            e.op0 = (const double *)a.op0 + dm.bp_off;
            e.op1 = nullptr;
            e.diag_corr = nullptr;
            pair::pair_k4_body<double, TILE, 16, false>(e, it, lds64);
        }
        __syncthreads();
    }
}

// One workgroup per domain: its items in order.
__global__ __launch_bounds__(256) void batched_reduce_kernel(BatchArgs a) {
    const int d = blockIdx.x;
    const DomDev dm = a.dom[d];
    const int64_t n = (int64_t)dm.nocc * (dm.nocc + 1) / 2 * dm.ntp;
    KSum os, ss;
    for (int64_t w = dm.item0 + threadIdx.x; w < dm.item0 + n; w += 256) {
        os.merge(a.item_partials[4 * w], a.item_partials[4 * w + 1]);
        ss.merge(a.item_partials[4 * w + 2], a.item_partials[4 * w + 3]);
    }
    wave_merge(os);
    wave_merge(ss);
    __shared__ double red[4][4];
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
        for (int k = 0; k < 4; ++k) {
            o.merge(red[k][0], red[k][1]);
            s.merge(red[k][2], red[k][3]);
        }
        a.out[2 * d] = o.value();
        a.out[2 * d + 1] = s.value();
    }
}

}  // namespace

// ---- graph cache ------------------------------------------------------------------
struct GraphCache {
    std::map<uint64_t, hipGraphExec_t> execs;
    hipStream_t capture = nullptr;
    ~GraphCache() {
        for (auto &kv : execs) (void)hipGraphExecDestroy(kv.second);
        if (capture) (void)hipStreamDestroy(capture);
    }
};

void release_graphs(State &s) {
    delete s.graphs;
    s.graphs = nullptr;
}

static uint64_t fnv1a(const void *p, size_t n, uint64_t h = 1469598103934665603ull) {
    const unsigned char *c = (const unsigned char *)p;
    for (size_t k = 0; k < n; ++k) {
        h ^= c[k];
        h *= 1099511628211ull;
    }
    return h;
}

struct Plan {
    std::vector<DomDev> dom;
    std::vector<int> item_dom;
    std::vector<PackTile> ptiles;
    std::vector<int64_t> row_start;
    std::vector<std::pair<int, std::pair<int64_t, int64_t>>> classes;  // tile, [begin, end)
    int64_t nitems = 0, packed = 0, nrows = 0;
};

static fh_status make_plan(const fh_domain *d, size_t ndom, size_t b_len, size_t eps_len,
                           Prec prec, Plan &pl) {
    const int kp = panel_depth(prec);
    pl.dom.resize(ndom);
    std::vector<int> order(ndom);
    for (size_t k = 0; k < ndom; ++k) {
        const fh_domain &x = d[k];
        if (x.nocc == 0 || x.nvir == 0 || x.naux == 0)
            return fail(FH_LAYOUT, "domain %zu has an empty dimension (nocc %u nvir %u naux %u)",
                        k, x.nocc, x.nvir, x.naux);
        if (x.nocc > 4096 || x.nvir > 65536 || x.naux > (1u << 20))
            return fail(FH_LAYOUT, "domain %zu is beyond the kernel's range", k);
        const uint64_t nb = (uint64_t)x.naux * x.nocc * x.nvir;
        if (x.b_offset > b_len || nb > b_len - x.b_offset)
            return fail(FH_LAYOUT, "domain %zu: B [%llu, +%llu) outside the %zu-element buffer", k,
                        (unsigned long long)x.b_offset, (unsigned long long)nb, b_len);
        if (x.eps_offset > eps_len || (uint64_t)x.nocc + x.nvir > eps_len - x.eps_offset)
            return fail(FH_LAYOUT, "domain %zu: eps outside the %zu-element buffer", k, eps_len);
        DomDev &m = pl.dom[k];
        m.b_off = (int64_t)x.b_offset;
        m.eps_off = (int64_t)x.eps_offset;
        m.nocc = (int)x.nocc;
        m.nvir = (int)x.nvir;
        m.naux = (int)x.naux;
        m.nvir_pad = (m.nvir + 15) / 16 * 16;
        m.naux_pad = (m.naux + kp - 1) / kp * kp;
        int tile = tuned_tile(prec, x.nvir, x.naux);
        if (tile <= 0) tile = x.nvir <= 40 ? 32 : 64;
        m.tile = tile;
        m.ntile = (m.nvir_pad + tile - 1) / tile;
        m.ntp = m.ntile * (m.ntile + 1) / 2;
        order[k] = (int)k;
    }
    // Items grouped by tile class; inside a class, most expensive domains first
    // (cost of an item ~ naux_pad * tile^2, times the domain's item count).
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const DomDev &a = pl.dom[x], &b = pl.dom[y];
        if (a.tile != b.tile) return a.tile > b.tile;
        return (double)a.naux_pad > (double)b.naux_pad;
    });
    int64_t item = 0, packed = 0, rows = 0;
    int cur_tile = -1;
    for (int k : order) {
        DomDev &m = pl.dom[k];
        if (m.tile != cur_tile) {
            if (cur_tile > 0) pl.classes.back().second.second = item;
            pl.classes.push_back({m.tile, {item, item}});
            cur_tile = m.tile;
        }
        m.item0 = item;
        const int64_t n = (int64_t)m.nocc * (m.nocc + 1) / 2 * m.ntp;
        pl.item_dom.insert(pl.item_dom.end(), (size_t)n, k);
        item += n;
    }
    if (!pl.classes.empty()) pl.classes.back().second.second = item;
    pl.nitems = item;
    // Packed offsets in descriptor order (row_start must be monotone in d).
    pl.row_start.resize(ndom + 1);
    for (size_t k = 0; k < ndom; ++k) {
        DomDev &m = pl.dom[k];
        m.bp_off = packed;
        m.row_off = rows;
        pl.row_start[k] = rows;
        const int64_t nr = (int64_t)m.nocc * m.nvir_pad;
        packed += nr * m.naux_pad;
        rows += nr;
        for (int64_t r0 = 0; r0 < nr; r0 += 64)
            for (int q0 = 0; q0 < m.naux_pad; q0 += 64)
                pl.ptiles.push_back(PackTile{(int)k, (int)r0, q0, 0});
    }
    pl.row_start[ndom] = rows;
    pl.packed = packed;
    pl.nrows = rows;
    if (pl.nitems > 0x7fffffffLL) return fail(FH_LAYOUT, "too many work items (%lld)",
                                              (long long)pl.nitems);
    return FH_OK;
}

static void enqueue(const BatchArgs &ba, const Plan &pl, Prec prec, hipStream_t st) {
    const int ncls = (int)pl.classes.size();
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, st, ba.counters, ncls);
    const unsigned pblocks = (unsigned)std::min<int64_t>(std::max<int64_t>(pl.ptiles.size(), 1),
                                                         4096);
    hipLaunchKernelGGL(batched_pack_kernel, dim3(pblocks), dim3(256), 0, st, ba);
    if (prec == Prec::BF16X3) {
        const unsigned rb = (unsigned)std::min<int64_t>((pl.nrows + 3) / 4, 8192);
        hipLaunchKernelGGL(batched_lo_sq_kernel, dim3(std::max(rb, 1u)), dim3(256), 0, st, ba);
    }
    for (int c = 0; c < ncls; ++c) {
        const int tile = pl.classes[c].first;
        const int64_t b0 = pl.classes[c].second.first, b1 = pl.classes[c].second.second;
        // Persistent grid: a few workgroups per CU, never more than the items.
        const int64_t grid = std::min<int64_t>(b1 - b0, tile == 128 ? 256 : tile == 64 ? 768 : 1536);
        if (grid <= 0) continue;
        const bool bf = prec == Prec::BF16X3;
#define FH_BLAUNCH(T, B)                                                                     \
    hipLaunchKernelGGL((batched_pair_kernel<T, B>), dim3((unsigned)grid), dim3(256), 0, st, ba, \
                       b0, b1, c)
        if (tile == 32) {
            if (bf) FH_BLAUNCH(32, true); else FH_BLAUNCH(32, false);
        } else if (tile == 64) {
            if (bf) FH_BLAUNCH(64, true); else FH_BLAUNCH(64, false);
        } else {
            if (bf) FH_BLAUNCH(128, true); else FH_BLAUNCH(128, false);
        }
#undef FH_BLAUNCH
    }
    hipLaunchKernelGGL(batched_reduce_kernel, dim3((unsigned)pl.dom.size()), dim3(256), 0, st, ba);
}

}  // namespace fh

using namespace fh;

extern "C" fh_status fh_rimp2_batched_dev(const fh_domain *domains, size_t ndom,
                                          const double *b_dev, size_t b_len,
                                          const double *eps_dev, size_t eps_len,
                                          fh_precision precision, int use_graph, void *stream,
                                          double *e_os, double *e_ss) {
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    clear_error();
    s.stats.calls++;
    s.pending = Pending{};
    Prec prec;
    if (!to_prec(precision, &prec)) return fail(FH_LAYOUT, "unknown precision %d", (int)precision);
    if (prec != Prec::F64 && prec != Prec::BF16X3)
        return fail(FH_UNSUPPORTED, "batched domains run in FP64 or BF16x3");
    if (ndom == 0) {
        commit(s, prec, false, true);
        return FH_OK;
    }
    if (!domains || !b_dev || !eps_dev || !e_os || !e_ss)
        return fail(FH_LAYOUT, "null pointer argument");
    if (ndom > (1u << 20)) return fail(FH_LAYOUT, "too many domains (%zu)", ndom);
    Plan pl;
    fh_status st = make_plan(domains, ndom, b_len, eps_len, prec, pl);
    if (st != FH_OK) return st;
    st = ensure_device(s);
    if (st != FH_OK) return st;
    const bool bf = prec == Prec::BF16X3;
    const size_t eb = bf ? 2 : 8;
    const size_t meta = sizeof(DomDev) * ndom + sizeof(int) * pl.item_dom.size() +
                        sizeof(PackTile) * pl.ptiles.size() + sizeof(int64_t) * (ndom + 1) + 256;
    std::vector<std::pair<int, size_t>> req = {
        {kSlotBatch, meta},
        {kSlotOp0, pl.packed * eb},
        {kSlotItems, sizeof(double) * 4 * pl.nitems},
        {kSlotBatchOut, sizeof(double) * 2 * ndom + 64 * sizeof(int)},
    };
    if (bf) {
        req.push_back({kSlotOp1, pl.packed * eb});
        req.push_back({kSlotDiag, sizeof(double) * pl.nrows});
    }
    st = reserve(s, req);
    if (st != FH_OK) return st;
    // Metadata: [DomDev...][item_dom...][PackTile...][row_start...]
    char *mbase = (char *)s.slots[kSlotBatch].ptr;
    std::vector<char> host(meta, 0);
    size_t off = 0;
    auto put = [&](const void *p, size_t n, size_t align) {
        off = (off + align - 1) / align * align;
        memcpy(host.data() + off, p, n);
        const size_t at = off;
        off += n;
        return at;
    };
    const size_t o_dom = put(pl.dom.data(), sizeof(DomDev) * ndom, 16);
    const size_t o_item = put(pl.item_dom.data(), sizeof(int) * pl.item_dom.size(), 16);
    const size_t o_pt = put(pl.ptiles.data(), sizeof(PackTile) * pl.ptiles.size(), 16);
    const size_t o_rs = put(pl.row_start.data(), sizeof(int64_t) * (ndom + 1), 16);
    hipStream_t ust = (hipStream_t)stream;
    st = hip_check(hipMemcpyAsync(mbase, host.data(), off, hipMemcpyHostToDevice, ust),
                   "H2D batch metadata");
    if (st != FH_OK) return st;
    s.pending.h2d_bytes += off;

    BatchArgs ba;
    ba.prec = prec;
    ba.b = b_dev;
    ba.eps = eps_dev;
    ba.dom = (const DomDev *)(mbase + o_dom);
    ba.item_dom = (const int *)(mbase + o_item);
    ba.ptiles = (const PackTile *)(mbase + o_pt);
    ba.nptiles = (int64_t)pl.ptiles.size();
    ba.op0 = s.slots[kSlotOp0].ptr;
    ba.op1 = bf ? s.slots[kSlotOp1].ptr : nullptr;
    ba.corr = bf ? (double *)s.slots[kSlotDiag].ptr : nullptr;
    ba.nrows = pl.nrows;
    ba.row_dom_start = (const int64_t *)(mbase + o_rs);
    ba.ndom = (int)ndom;
    ba.item_partials = (double *)s.slots[kSlotItems].ptr;
    ba.out = (double *)s.slots[kSlotBatchOut].ptr;
    ba.counters = (int *)(ba.out + 2 * ndom);

    mark(s, 0, ust);
    if (use_graph) {
        if (!s.graphs) s.graphs = new GraphCache();
        GraphCache &gc = *s.graphs;
        // Key: everything the captured launches bake in -- the resolved plan
        // (shapes, offsets, tile choices) and every device pointer.
        uint64_t key = fnv1a(pl.dom.data(), sizeof(DomDev) * ndom);
        const void *ptrs[8] = {b_dev, eps_dev, ba.op0, ba.op1, ba.item_partials, mbase, ba.corr,
                               ba.out};
        key = fnv1a(ptrs, sizeof ptrs, key);
        key = fnv1a(&prec, sizeof prec, key);
        key = fnv1a(&pl.nitems, sizeof pl.nitems, key);
        auto itg = gc.execs.find(key);
        if (itg == gc.execs.end()) {
            if (!gc.capture) {
                st = hip_check(hipStreamCreateWithFlags(&gc.capture, hipStreamNonBlocking),
                               "create capture stream");
                if (st != FH_OK) return st;
            }
            // The metadata upload above must land before the graph first runs.
            st = hip_check(hipStreamSynchronize(ust), "sync before capture");
            if (st != FH_OK) return st;
            st = hip_check(hipStreamBeginCapture(gc.capture, hipStreamCaptureModeThreadLocal),
                           "begin capture");
            if (st != FH_OK) return st;
            enqueue(ba, pl, prec, gc.capture);
            hipGraph_t g = nullptr;
            hipError_t he = hipStreamEndCapture(gc.capture, &g);
            if (he != hipSuccess || !g) return fail(FH_HIP, "end capture: %s", hipGetErrorString(he));
            hipGraphExec_t ex = nullptr;
            he = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (he != hipSuccess) return fail(FH_HIP, "graph instantiate: %s", hipGetErrorString(he));
            itg = gc.execs.emplace(key, ex).first;
        }
        st = hip_check(hipGraphLaunch(itg->second, ust), "graph launch");
        if (st != FH_OK) return st;
    } else {
        enqueue(ba, pl, prec, ust);
        hipError_t he = hipGetLastError();
        if (he != hipSuccess) return fail(FH_KERNEL, "batched kernels: %s", hipGetErrorString(he));
    }
    mark(s, 1, ust);
    mark(s, 2, ust);
    std::vector<double> res(2 * ndom);
    st = hip_check(hipMemcpyAsync(res.data(), ba.out, sizeof(double) * res.size(),
                                  hipMemcpyDeviceToHost, ust),
                   "D2H batched energies");
    if (st != FH_OK) return st;
    mark(s, 3, ust);
    st = hip_check(hipStreamSynchronize(ust), "stream sync");
    if (st != FH_OK) return st;
    s.pending.d2h_bytes += sizeof(double) * res.size();
    s.pending.kernel_launches += 3 + pl.classes.size();
    finish_timing(s, (uint64_t)pl.nitems, pl.classes.size() == 1 ? pl.classes[0].first : 0);
    for (size_t k = 0; k < ndom; ++k) {
        e_os[k] = res[2 * k];
        e_ss[k] = res[2 * k + 1];
    }
    commit(s, prec, false, true);
    return FH_OK;
}
