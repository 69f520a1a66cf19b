"""Parity on the mixed-precision workload's own distribution (run with -m gpu).

* FP64 at 100/900/3600 (B stored FP32, 1 % of the Q rows x20): the device's
  per-occupied-row partials of rows 0, 50 and 99 against the oracle's partials
  of those rows (spin_components_from_blocks' per-i body, rimp2.rs:1520-1530,
  on BLAS tail blocks), each within the energy bound of that row alone
  (tests/common/rimp2_error_bound.rs restated, 2 gamma_naux(u64) S_ab per
  element, the reference's f64 device-vs-CPU bound of
  crates/ferric-mp2/tests/gpu_rimp2_resident.rs). At |E| ~ 1.7e7 Eh an
  absolute 1e-10 Eh would be below one FP64 ulp of the total; the element
  bound is the bar the reference itself uses.
* Every reduced precision on a scaled-down copy of that distribution
  (synthetic.mixed_large at 12/120/480, the same x20 hot rows that stress the
  FP8 per-(64-row Q block, i) scale): the device energy against the FP64
  oracle energy of exactly the values the format contracts (tests/quant.py),
  within the FP32-accumulation bound, and INT8X4 to 1e-13 relative.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from ferric_amd import Precision, rimp2, synthetic  # noqa: E402
from tests import quant  # noqa: E402

U32 = 2.0 ** -24


def test_fp64_rows_against_oracle_at_100_900_3600(oracle_mod):
    p = synthetic.mixed_large()
    print("fp64 rows 100/900/3600", flush=True)
    t = torch.from_numpy(p.b_ov).cuda()
    sc, rows = rimp2.spin_components(t, p.eps, p.nocc, p.nvir, precision=Precision.FP64,
                                     row_partials=True)
    del t
    torch.cuda.empty_cache()
    # the partials fold to the energy (the device folds them compensated, in
    # ascending i, as the reference's fold order)
    assert math.fsum(rows[:, 0]) == pytest.approx(sc.e_os, rel=1e-15)
    assert math.fsum(rows[:, 1]) == pytest.approx(sc.e_ss, rel=1e-15)
    b64 = p.b_ov.astype(np.float64)
    want = oracle_mod.row_partials_blas(b64, p.eps, p.nocc, p.nvir, (0, 50, 99))
    for i, (eos, ess, bos, bss) in want.items():
        d_os, d_ss = rows[i, 0] - eos, rows[i, 1] - ess
        print(f"  row {i}: E_os {rows[i, 0]:.6f} (dE {d_os:+.2e}, bound {bos:.2e})  "
              f"E_ss {rows[i, 1]:.6f} (dE {d_ss:+.2e}, bound {bss:.2e})", flush=True)
        assert abs(d_os) <= bos and abs(d_ss) <= bss, (i, d_os, bos, d_ss, bss)


@pytest.fixture(scope="module")
def hot_small():
    return synthetic.mixed_large(seed=1013, nocc=12, nvir=120, naux=480)


@pytest.mark.parametrize("prec", [Precision.FP32, Precision.BF16X3, Precision.FP8_E4M3])
def test_reduced_precision_exact_formats_on_hot_rows(hot_small, oracle_mod, prec):
    p = hot_small
    b, eps, nocc, nvir = p.b_ov.astype(np.float64), p.eps, p.nocc, p.nvir
    print("hot rows", prec.name, flush=True)
    sc = rimp2.spin_components(torch.from_numpy(p.b_ov).cuda(), eps, nocc, nvir,
                               precision=prec)
    eps_g = (p.naux + 16) * U32
    if prec == Precision.FP32:
        want = oracle_mod.spin_components(b, eps, nocc, nvir)
        bnd = oracle_mod.energy_bound(b, eps, nocc, nvir, eps_g=eps_g)
    elif prec == Precision.FP8_E4M3:
        bq = quant.fp8_dequantized(b, nocc, nvir)
        want = oracle_mod.spin_components(bq, eps, nocc, nvir)
        bnd = oracle_mod.energy_bound(bq, eps, nocc, nvir, eps_g=eps_g)
    else:
        hi, lo = quant.bf16_split(b)
        g = hi.T @ hi + hi.T @ lo + lo.T @ hi
        g += np.diag(np.sum(lo * lo, axis=0))
        want = oracle_mod.spin_components_from_g(g, eps, nocc, nvir)
        bnd = oracle_mod.energy_bound(hi + lo, eps, nocc, nvir, eps_g=eps_g)
    f64 = oracle_mod.spin_components(b, eps, nocc, nvir)
    print(f"  dE vs exact format {sc.e_os - want[0]:+.2e} / {sc.e_ss - want[1]:+.2e} "
          f"(bound {bnd[0]:.2e} / {bnd[1]:.2e}); dE vs FP64 {sc.e_total - f64[2]:+.3e}",
          flush=True)
    assert abs(sc.e_os - want[0]) <= bnd[0]
    assert abs(sc.e_ss - want[1]) <= bnd[1]


def test_int8x4_exact_format_on_hot_rows(hot_small, oracle_mod):
    p = hot_small
    b, eps, nocc, nvir = p.b_ov.astype(np.float64), p.eps, p.nocc, p.nvir
    print("hot rows INT8X4", flush=True)
    sc = rimp2.spin_components(torch.from_numpy(p.b_ov).cuda(), eps, nocc, nvir,
                               precision=Precision.INT8X4)
    want = oracle_mod.spin_components_from_g(quant.int8x4_g(b), eps, nocc, nvir)
    tol = 1e-13 * (abs(want[0]) + abs(want[1]))
    assert abs(sc.e_os - want[0]) <= tol and abs(sc.e_ss - want[1]) <= tol, (sc, want)
    f64 = oracle_mod.spin_components(b, eps, nocc, nvir)
    bos, bss = quant.int8x4_energy_bound(b, eps, nocc, nvir)
    print(f"  dE vs FP64 {sc.e_os - f64[0]:+.2e} / {sc.e_ss - f64[1]:+.2e} (bound {bos:.2e} / "
          f"{bss:.2e})", flush=True)
    assert abs(sc.e_os - f64[0]) <= bos and abs(sc.e_ss - f64[1]) <= bss


def _eight_wave_check(oracle_mod, p, prec, tile=128):
    """Energy of one BF16x3 / FP8 call at the given tile side against the FP64
    oracle energy of exactly the values the format contracts, within the
    FP32-accumulation bound."""
    from ferric_amd._lib import last_timing, set_timing

    b, eps, nocc, nvir = p.b_ov.astype(np.float64), p.eps, p.nocc, p.nvir
    set_timing(True)
    try:
        sc = rimp2.spin_components(torch.from_numpy(p.b_ov).cuda(), eps, nocc, nvir,
                                   precision=prec)
        assert last_timing()["tile"] == tile
    finally:
        set_timing(False)
    eps_g = (p.naux + 32) * U32
    if prec == Precision.FP8_E4M3:
        bq = quant.fp8_dequantized(b, nocc, nvir)
        want = oracle_mod.spin_components(bq, eps, nocc, nvir)
        bnd = oracle_mod.energy_bound(bq, eps, nocc, nvir, eps_g=eps_g)
    else:
        hi, lo = quant.bf16_split(b)
        g = hi.T @ hi + hi.T @ lo + lo.T @ hi
        g += np.diag(np.sum(lo * lo, axis=0))
        want = oracle_mod.spin_components_from_g(g, eps, nocc, nvir)
        bnd = oracle_mod.energy_bound(hi + lo, eps, nocc, nvir, eps_g=eps_g)
    print(f"  dE vs exact format {sc.e_os - want[0]:+.2e} / {sc.e_ss - want[1]:+.2e} "
          f"(bound {bnd[0]:.2e} / {bnd[1]:.2e})", flush=True)
    assert abs(sc.e_os - want[0]) <= bnd[0]
    assert abs(sc.e_ss - want[1]) <= bnd[1]


@pytest.mark.parametrize("prec", [Precision.BF16X3, Precision.FP8_E4M3])
def test_eight_wave_kernels_odd_panels_thin_tile(oracle_mod, prec, monkeypatch):
    """The eight-wave BF16x3 / FP8 kernels at the 128-wide tile: 150 virtuals
    leave a 32-row last tile, and 330 auxiliary functions give an odd panel
    count (11 BF16 panels of 32, 3 FP8 panels of 128), so the waves of a
    single-product item that take the odd panels meet one past naux_pad."""
    monkeypatch.setenv("FERRIC_HIP_TILE", "128")
    print("eight-wave", prec.name, flush=True)
    _eight_wave_check(oracle_mod, synthetic.mixed_large(seed=1017, nocc=6, nvir=150, naux=330),
                      prec)


@pytest.mark.parametrize("prec", [Precision.BF16X3, Precision.FP8_E4M3])
@pytest.mark.parametrize("shape", [(1, 97, 33), (2, 128, 32), (3, 257, 129), (5, 81, 200)])
@pytest.mark.parametrize("tile", [128, 64])
def test_eight_wave_fragment_order_shapes(oracle_mod, prec, shape, tile, monkeypatch):
    """The fragment-order operands of the eight-wave kernels at awkward shapes:
    one occupied orbital (every item single-product), a virtual count of
    exactly one tile, one past two tiles (a 16-row last block group) and one
    past 80, a single BF16 panel, auxiliary counts just past a panel; at the
    128-wide and the 64-wide tile (the BF16x3 eight-wave kernel at both, FP8's
    eight-wave kernel at 128 and its four-wave kernel at 64)."""
    monkeypatch.setenv("FERRIC_HIP_TILE", str(tile))
    nocc, nvir, naux = shape
    print("eight-wave fragment order", prec.name, shape, tile, flush=True)
    _eight_wave_check(oracle_mod,
                      synthetic.mixed_large(seed=1100 + nocc, nocc=nocc, nvir=nvir, naux=naux),
                      prec, tile)
