"""Per-pair MP2 energies and the local-MP2 pair layer on the GPU, run with -m gpu.

* fh_rimp2_pair_energies[_dev] (estimate_pair_energies_from_b_ov,
  dlpno_mp2.rs:527-594) against the oracle's per-pair restatement: FP64 to
  1e-12 relative per pair on the golden cases (frozen core included), over
  several tiles at decane-scale shapes, from FP32 storage, and on real water,
  whose pairs sum to the ORCA RI-MP2 correlation energy.
* The reduced precisions: each one's pairs sum to that precision's energy,
  and the INT8X4 pairs stay within the format's derived bound of the FP64
  pairs.
* The pair layer end to end (dlpno_mp2.rs:670-760 at t_cut_pno = 0): the
  exact configuration reproduces the dense energy; on a chain of localized
  orbitals the default 1e-5 Eh screen drops the far pairs and the energy
  moves by exactly the discarded pairs' energy (pair_energy_screen.rs:90-105).
* Refusals: a non-negative denominator (pair_energy_screen.rs:494-502) and
  PNO truncation, which is not on the device.
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from ferric_amd import LayoutError, Precision, UnsupportedError, rimp2, synthetic  # noqa: E402
from ferric_amd.dlpno import (DlpnoConfig, build_pair_domains_by_energy,  # noqa: E402
                              dlpno_mp2_from_b_ov, estimate_pair_energies_from_b_ov)
from tests import quant  # noqa: E402

FIX = os.path.join(os.path.dirname(__file__), "golden", "water_ccpvdz_real.npz")


def _cuda(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _close(got, want, scale, rel=1e-12):
    """max |got - want| <= rel * max |scale| (the opposite-spin pairs set the scale)."""
    err = np.max(np.abs(got - want)) / np.max(np.abs(scale))
    assert err <= rel, err


def _upper_sum(m):
    n = m.shape[0]
    return math.fsum(m[i, j] for i in range(n) for j in range(i, n))


def test_fp64_pairs_against_oracle_golden(golden, oracle_mod):
    for name, c in golden.items():
        print("pair energies fp64", name, flush=True)
        nocc, nvir, fc, nt = c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"]
        w_os, w_ss = oracle_mod.pair_energies(c["b"], c["eps"], nocc, nvir, fc, nt)
        dev = estimate_pair_energies_from_b_ov(_cuda(c["b"]), c["eps"], nocc, nvir, fc, nt)
        _close(dev.e_os, w_os, w_os)
        _close(dev.e_ss, w_ss, w_os)
        host = estimate_pair_energies_from_b_ov(c["b"], c["eps"], nocc, nvir, fc, nt)
        assert np.array_equal(host.e_os, dev.e_os) and np.array_equal(host.e_ss, dev.e_ss)
        sc = rimp2.spin_components(_cuda(c["b"]), c["eps"], nocc, nvir, fc, nt)
        assert abs(dev.total() - sc.e_total) <= 1e-12 * abs(sc.e_total)


def test_fp64_pairs_decane_scale_and_fp32_storage(oracle_mod):
    p = synthetic.decane_scale()
    print("pair energies fp64 decane-scale", flush=True)
    w_os, w_ss = oracle_mod.pair_energies(p.b_ov, p.eps, p.nocc, p.nvir)
    dev = estimate_pair_energies_from_b_ov(_cuda(p.b_ov), p.eps, p.nocc, p.nvir)
    _close(dev.e_os, w_os, w_os)
    _close(dev.e_ss, w_ss, w_os)
    b32 = p.b_ov.astype(np.float32)
    print("pair energies fp32 storage", flush=True)
    s32 = estimate_pair_energies_from_b_ov(_cuda(b32), p.eps, p.nocc, p.nvir)
    s64 = estimate_pair_energies_from_b_ov(_cuda(b32.astype(np.float64)), p.eps, p.nocc, p.nvir)
    assert np.array_equal(s32.e, s64.e)


def test_real_water_pairs_sum_to_orca():
    z = np.load(FIX)
    nocc, nvir = int(z["nocc"]), int(z["nvir"])
    print("pair energies real water", flush=True)
    pe = estimate_pair_energies_from_b_ov(_cuda(z["dist_b_ov"]), z["dist_eps"], nocc, nvir)
    assert abs(pe.total() - float(z["dist_e_corr_ref"])) < 1e-8
    assert abs(_upper_sum(pe.e_os) - float(z["dist_e_os_ref"])) < 1e-8
    assert abs(_upper_sum(pe.e_ss) - float(z["dist_e_ss_ref"])) < 1e-8


@pytest.mark.parametrize("prec", [Precision.FP32, Precision.BF16X3, Precision.INT8X4,
                                  Precision.FP8_E4M3])
def test_reduced_precision_pairs_sum_to_their_energy(prec):
    p = synthetic.decane_scale()
    print("pair energies", prec.name, flush=True)
    t = _cuda(p.b_ov)
    pe = estimate_pair_energies_from_b_ov(t, p.eps, p.nocc, p.nvir, precision=prec)
    sc = rimp2.spin_components(t, p.eps, p.nocc, p.nvir, precision=prec)
    assert abs(_upper_sum(pe.e_os) - sc.e_os) <= 1e-12 * abs(sc.e_total)
    assert abs(_upper_sum(pe.e_ss) - sc.e_ss) <= 1e-12 * abs(sc.e_total)
    assert np.array_equal(pe.e, pe.e.T)


def test_int8x4_pairs_within_the_format_bound(oracle_mod):
    p = synthetic.decane_scale()
    print("pair energies int8x4 vs oracle", flush=True)
    w_os, w_ss = oracle_mod.pair_energies(p.b_ov, p.eps, p.nocc, p.nvir)
    pe = estimate_pair_energies_from_b_ov(_cuda(p.b_ov), p.eps, p.nocc, p.nvir,
                                          precision=Precision.INT8X4)
    bos, bss = quant.int8x4_energy_bound(p.b_ov, p.eps, p.nocc, p.nvir)
    # every pair's error is one of the non-negative terms the bound sums
    assert np.max(np.abs(pe.e_os - w_os)) <= bos
    assert np.max(np.abs(pe.e_ss - w_ss)) <= bss


def test_dlpno_exact_config_reproduces_the_dense_energy():
    p = synthetic.decane_scale()
    print("dlpno exact config", flush=True)
    t = _cuda(p.b_ov)
    pe = estimate_pair_energies_from_b_ov(t, p.eps, p.nocc, p.nvir)
    centers = np.random.default_rng(5).uniform(-8, 8, (p.nocc, 3))
    d = build_pair_domains_by_energy(centers, pe, 0.0)
    assert d.is_complete()
    sc, diag = dlpno_mp2_from_b_ov(t, p.eps, p.nocc, p.nvir, 0, p.nocc, d, DlpnoConfig.exact())
    dense = rimp2.spin_components(t, p.eps, p.nocc, p.nvir)
    assert abs(sc.e_total - dense.e_total) <= 1e-12 * abs(dense.e_total)
    assert abs(sc.e_os - dense.e_os) <= 1e-12 * abs(dense.e_total)
    assert diag.pair_retention == 1.0 and diag.n_pairs == p.nocc * (p.nocc + 1) // 2


def test_dlpno_default_screen_on_a_local_chain(oracle_mod):
    p, centers = synthetic.local_chain()
    print("dlpno default screen, local chain", flush=True)
    t = _cuda(p.b_ov)
    pe = estimate_pair_energies_from_b_ov(t, p.eps, p.nocc, p.nvir)
    w_os, w_ss = oracle_mod.pair_energies(p.b_ov, p.eps, p.nocc, p.nvir)
    _close(pe.e_os, w_os, w_os)
    _close(pe.e_ss, w_ss, w_os)
    cfg = DlpnoConfig()
    d = build_pair_domains_by_energy(centers, pe, cfg.t_cut_pairs)
    # the oracle's pairs select the same pair list
    keep = [(i, j) for i in range(p.nocc) for j in range(i, p.nocc)
            if i == j or abs(w_os[i, j] + w_ss[i, j]) >= cfg.t_cut_pairs]
    assert d.pairs == keep and 0.0 < d.pair_retention() < 1.0
    sc, diag = dlpno_mp2_from_b_ov(t, p.eps, p.nocc, p.nvir, 0, p.nocc, d, cfg, pair_energies=pe)
    dense = rimp2.spin_components(t, p.eps, p.nocc, p.nvir)
    drop = pe.discarded_energy(cfg.t_cut_pairs)
    print(f"  kept {len(d.pairs)} of {p.nocc * (p.nocc + 1) // 2} pairs, discarded {drop:.3e} Eh",
          flush=True)
    assert abs(sc.e_total - (dense.e_total - drop)) <= 1e-12 * abs(dense.e_total)
    assert 0.0 > drop > -1e-5 * (p.nocc * (p.nocc + 1) // 2 - len(d.pairs))
    assert diag.n_pairs == len(d.pairs) and diag.virtual_retention == 1.0
    # computing the pair energies inside the driver gives the same energy
    sc2, _ = dlpno_mp2_from_b_ov(p.b_ov, p.eps, p.nocc, p.nvir, 0, p.nocc, d, cfg)
    assert sc2.e_total == sc.e_total


def test_refusals():
    rng = np.random.default_rng(7)
    b = rng.standard_normal((6, 12))
    eps = np.array([-1.0, -0.8, -0.6, -0.6, 0.7, 1.1, 1.6])
    print("pair energies refusals", flush=True)
    with pytest.raises(LayoutError):
        estimate_pair_energies_from_b_ov(b, eps, 3, 4)
    with pytest.raises(LayoutError):
        estimate_pair_energies_from_b_ov(_cuda(b), eps, 3, 4)
    with pytest.raises(LayoutError):
        estimate_pair_energies_from_b_ov(_cuda(b), eps, 3, 5)
    p, centers = synthetic.local_chain(nocc=4, nvir=8, rows_per_site=4)
    pe = estimate_pair_energies_from_b_ov(p.b_ov, p.eps, p.nocc, p.nvir)
    d = build_pair_domains_by_energy(centers, pe, 0.0)
    with pytest.raises(UnsupportedError):
        dlpno_mp2_from_b_ov(p.b_ov, p.eps, p.nocc, p.nvir, 0, p.nocc, d,
                            DlpnoConfig(t_cut_pno=1e-7))
