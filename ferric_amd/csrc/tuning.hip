// Tile selection for the pair-energy kernels.
//
// A tile table (written by ferric_amd/autotune.py, which times the batched
// domain kernel, and installed through fh_tune_set) maps (precision, nvir
// bucket, naux bucket) to the measured best tile side of a batched domain;
// one-problem calls keep their built-in choice. FERRIC_HIP_TILE forces one
// tile for every call of either kind.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>
#include <vector>

#include "fh_host.h"

namespace fh {

namespace {
struct TuneEntry {
    int prec;
    size_t nvir_max, naux_max;
    int tile;
};
std::mutex g_tune_mu;
std::vector<TuneEntry> g_table;
}  // namespace

int forced_tile() {
    const char *env = getenv("FERRIC_HIP_TILE");
    if (env && *env) {
        const int t = atoi(env);
        if (t == 32 || t == 64 || t == 128) return t;
    }
    return 0;
}

static int table_tile(Prec prec, size_t nvir, size_t naux) {
    // First entry (in installation order) whose bucket holds the shape.
    for (const auto &e : g_table)
        if (e.prec == (int)prec && nvir <= e.nvir_max && naux <= e.naux_max) return e.tile;
    return 0;
}

void tuned_tiles(Prec prec, size_t n, const size_t *nvir, const size_t *naux, int *out) {
    if (const int t = forced_tile()) {
        for (size_t k = 0; k < n; ++k) out[k] = t;
        return;
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (size_t k = 0; k < n; ++k) out[k] = table_tile(prec, nvir[k], naux[k]);
}

void tune_set(int prec, size_t nvir_max, size_t naux_max, int tile) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_table.push_back(TuneEntry{prec, nvir_max, naux_max, tile});
}

void tune_clear() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_table.clear();
}

}  // namespace fh

using namespace fh;

extern "C" fh_status fh_tune_set(fh_precision precision, size_t nvir_max, size_t naux_max,
                                 int tile) {
    Prec p;
    if (!to_prec(precision, &p)) return fail(FH_LAYOUT, "unknown precision %d", (int)precision);
    if (tile != 32 && tile != 64 && tile != 128)
        return fail(FH_LAYOUT, "tile side %d is not 32, 64 or 128", tile);
    tune_set((int)p, nvir_max, naux_max, tile);
    // Captured batched graphs baked in the previous tile choice.
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    release_graphs(s);
    return FH_OK;
}

extern "C" fh_status fh_tune_clear(void) {
    tune_clear();
    State &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    release_graphs(s);
    return FH_OK;
}
