"""Tile autotuner for the pair kernels.

For each precision and each (virtual size, auxiliary size) bucket it times
every candidate tile side on a synthetic set of local-MP2 pair domains drawn
from that bucket (nocc_d uniform on {2..12}; nvir_d uniform and naux_d
log-uniform inside the bucket) through the batched path, and keeps the
fastest. The table lives in ferric_amd/tuning/tiles.json and is installed into
the library (fh_tune_set) whenever the library loads; for each batched domain
the library takes the first entry whose bounds hold it (nvir <= nvir_max and
naux <= naux_max, entries in table order), and domains outside every bucket
keep the built-in choice. The last auxiliary bucket of each virtual bucket is
open upwards. One-problem calls do not consult the table: their kernels (and
the 128-wide eight-wave ones among them) are chosen by the library's measured
built-in rule.

    python -m ferric_amd.autotune [--out PATH]   # measure on the current GPU, write the table
"""
from __future__ import annotations

import json
import math
import os
import time

import numpy as np

from ._lib import Precision

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning", "tiles.json")
BUCKETS = [(20, 48), (48, 96), (96, 160), (160, 256)]
NAUX_BUCKETS = [(80, 240), (240, 721)]
CANDIDATES = (32, 64)
NAUX_MAX = 1 << 20


def load_table(path: str = TABLE) -> list:
    if not os.path.exists(path):
        return []
    with open(path) as f:
        return json.load(f).get("entries", [])


def install(entries) -> None:
    from ._lib import check, lib

    L = lib()
    check(L.fh_tune_clear(), "fh_tune_clear")
    for e in entries:
        check(L.fh_tune_set(int(Precision[e["precision"]]), int(e["nvir_max"]),
                            int(e.get("naux_max", NAUX_MAX)), int(e["tile"])), "fh_tune_set")


def _bucket_set(lo: int, hi: int, ndom: int, seed: int, naux_lo: int = 80, naux_hi: int = 721):
    from .synthetic import DomainSet, _b, _loguniform

    rng = np.random.default_rng(seed)
    nocc = rng.integers(2, 13, ndom)
    nvir = rng.integers(lo, hi, ndom)
    naux = np.clip(np.floor(_loguniform(rng, naux_lo, naux_hi, ndom)), naux_lo,
                   naux_hi - 1).astype(np.int64)
    sizes = naux * nocc * nvir
    b_off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    e_off = np.concatenate([[0], np.cumsum(nocc + nvir)[:-1]]).astype(np.int64)
    b = np.empty(int(sizes.sum()))
    eps = np.empty(int((nocc + nvir).sum()))
    for d in range(ndom):
        b[b_off[d]:b_off[d] + sizes[d]] = _b(rng, int(naux[d]), int(nocc[d]), int(nvir[d])).ravel()
        eps[e_off[d]:e_off[d] + nocc[d] + nvir[d]] = np.concatenate(
            [np.sort(rng.uniform(-1.2, -0.35, nocc[d])), np.sort(_loguniform(rng, 0.12, 5.0, nvir[d]))])
    return DomainSet(seed, nocc, nvir, naux, b_off, e_off, b, eps)


def measure(precisions=(Precision.FP64, Precision.BF16X3), ndom: int = 128, reps: int = 5,
            log=print) -> list:
    import torch

    from ._lib import check, lib
    from .batched import domain_set_energies

    check(lib().fh_tune_clear(), "fh_tune_clear")
    entries = []
    for prec in precisions:
        for k, (lo, hi) in enumerate(BUCKETS):
            for m, (qlo, qhi) in enumerate(NAUX_BUCKETS):
                ds = _bucket_set(lo, hi, ndom, seed=4000 + 10 * k + m, naux_lo=qlo, naux_hi=qhi)
                b = torch.from_numpy(ds.b_flat).cuda()
                e = torch.from_numpy(ds.eps_flat).cuda()
                best = None
                for tile in CANDIDATES:
                    os.environ["FERRIC_HIP_TILE"] = str(tile)
                    try:
                        domain_set_energies(ds, prec, use_graph=False, b_dev=b, eps_dev=e)
                        torch.cuda.synchronize()
                        ts = []
                        for _ in range(reps):
                            t0 = time.perf_counter()
                            domain_set_energies(ds, prec, use_graph=False, b_dev=b, eps_dev=e)
                            ts.append(time.perf_counter() - t0)
                    finally:
                        del os.environ["FERRIC_HIP_TILE"]
                    t = min(ts)
                    log(f"tune {prec.name} nvir [{lo},{hi}) naux [{qlo},{qhi}) tile {tile}: "
                        f"{t * 1e3:.3f} ms")
                    if best is None or t < best[1]:
                        best = (tile, t)
                last = m == len(NAUX_BUCKETS) - 1
                entries.append({"precision": prec.name, "nvir_max": hi - 1,
                                "naux_max": NAUX_MAX if last else qhi - 1,
                                "tile": best[0], "ms": round(best[1] * 1e3, 3)})
    return entries


def main(argv=None):
    import argparse

    ap = argparse.ArgumentParser(description="measure the tile table on the current GPU")
    ap.add_argument("--out", default=TABLE, help="where to write the table (default: the "
                    "table the library installs)")
    args = ap.parse_args(argv)
    entries = measure()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump({"device": "MI355X (gfx950)", "entries": entries}, f, indent=1)
    print(json.dumps(entries))
    install(entries)


if __name__ == "__main__":
    main()
