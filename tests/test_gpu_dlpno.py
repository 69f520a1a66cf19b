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
* PNO truncation (fh_dlpno_pno_pair_energies): at t_cut_pno = 0 equal to the
  canonical pairs (the exactness contract, dlpno_mp2.rs:797-830); above it,
  per pair against the oracle's restatement at 1e-10 of the largest pair, with
  the same kept PNO counts and discarded weights, on a local chain over four
  thresholds, at 300 virtuals where the semicanonical step runs in the 128-
  and 256-wide classes, the reference's own toy system and frozen-core cases;
  a huge threshold keeps one PNO per pair; batching and the host entry change
  nothing; the driver with the default screen plus truncation.
* Domain-local fits (fh_domain_fit_pair_blocks): every pair's block against
  the oracle's restatement at 1e-12 at three radii (uneven domains in one
  batch), host and device entries equal, batching; an unbounded radius gives
  the canonical pair energies.
* Refusals: a non-negative denominator (pair_energy_screen.rs:494-502),
  FP32 input to the PNO path, reduced precision with truncation, an empty
  aux domain, a metric that is not positive definite.
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from ferric_amd import (KernelError, LayoutError, Precision, UnsupportedError, rimp2,  # noqa: E402
                        synthetic)
from ferric_amd.dlpno import (DlpnoConfig, build_pair_domains_by_energy,  # noqa: E402
                              dlpno_mp2_from_b_ov, domain_fit_pair_blocks,
                              estimate_pair_energies_from_b_ov, pno_pair_energies)
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
    with pytest.raises(LayoutError):  # (ia|jb) with a zero gap in the listed pair
        pno_pair_energies(b, eps, 3, 4, 0, 3, [(2, 2)], 1e-3)
    with pytest.raises(LayoutError):
        pno_pair_energies(_cuda(p.b_ov).float(), p.eps, p.nocc, p.nvir, 0, p.nocc, [(0, 0)], 0.0)
    pe = estimate_pair_energies_from_b_ov(p.b_ov, p.eps, p.nocc, p.nvir)
    d = build_pair_domains_by_energy(centers, pe, 0.0)
    with pytest.raises(UnsupportedError):
        dlpno_mp2_from_b_ov(p.b_ov, p.eps, p.nocc, p.nvir, 0, p.nocc, d,
                            DlpnoConfig(t_cut_pno=1e-7), precision=Precision.FP32)


# -- PNO truncation (dlpno_mp2.rs:249-352, :700-744; local_pno.rs:132-200) ------

def _pairs(nocc):
    return [(i, j) for i in range(nocc) for j in range(i, nocc)]


def _pno_check(oracle_mod, b, eps, nocc, nvir, fo, nt, pairs, t_cut, rel=1e-10):
    w_os, w_ss, w_kept, w_disc, w_kocc = oracle_mod.pno_pair_energies(b, eps, nocc, nvir, fo,
                                                                      nt, pairs, t_cut)
    g_os, g_ss, kept, worst = pno_pair_energies(_cuda(b), eps, nocc, nvir, fo, nt, pairs, t_cut)
    # each discarded occupation carries the eigensolvers' absolute error, a few
    # units of rounding of the largest occupation (at most tr D = kept + discarded)
    occ_tol = 64 * nvir * 2.2e-16 * np.max(w_kocc + w_disc)
    scale = np.max(np.abs(w_os))
    print(f"  t_cut {t_cut:g}: kept {kept} of {len(pairs) * nvir}, max |dE_pair| "
          f"{np.max(np.abs(g_os - w_os)):.2e} {np.max(np.abs(g_ss - w_ss)):.2e}", flush=True)
    assert kept == int(w_kept.sum())
    assert abs(worst - np.max(w_disc)) <= 1e-10 * np.max(w_disc) + occ_tol
    assert np.max(np.abs(g_os - w_os)) <= rel * scale
    assert np.max(np.abs(g_ss - w_ss)) <= rel * scale
    return g_os, g_ss, kept, worst


def test_pno_zero_truncation_equals_the_canonical_pairs():
    p = synthetic.decane_scale()
    print("pno t_cut 0 vs canonical pairs, decane-scale", flush=True)
    t = _cuda(p.b_ov)
    pe = estimate_pair_energies_from_b_ov(t, p.eps, p.nocc, p.nvir)
    pairs = [(i, j) for i in range(0, p.nocc, 4) for j in range(i, p.nocc, 5)]
    g_os, g_ss, kept, worst = pno_pair_energies(t, p.eps, p.nocc, p.nvir, 0, p.nocc, pairs, 0.0)
    want_os = np.array([pe.e_os[i, j] for i, j in pairs])
    want_ss = np.array([pe.e_ss[i, j] for i, j in pairs])
    scale = np.max(np.abs(want_os))
    assert kept == len(pairs) * p.nvir and worst == 0.0
    assert np.max(np.abs(g_os - want_os)) <= 1e-11 * scale
    assert np.max(np.abs(g_ss - want_ss)) <= 1e-11 * scale


@pytest.mark.parametrize("t_cut", [1e-8, 1e-6, 1e-4, 1e-2])
def test_pno_truncation_against_oracle_local_chain(oracle_mod, t_cut):
    p, _ = synthetic.local_chain(nocc=10, nvir=40, rows_per_site=12)
    print("pno local chain 10/40", flush=True)
    _pno_check(oracle_mod, p.b_ov, p.eps, p.nocc, p.nvir, 0, p.nocc, _pairs(p.nocc), t_cut)


@pytest.mark.parametrize("nocc,nvir,rows,seed,t_cut", [(6, 300, 12, 1010, 1e-8),
                                                       (6, 300, 12, 1010, 1e-6),
                                                       (5, 130, 40, 1011, 1e-8)])
def test_pno_width_classes(oracle_mod, monkeypatch, nocc, nvir, rows, seed, t_cut):
    """nvir 300: kept counts 1..192 put the pairs of one batch into the 128- and
    256-wide classes of the semicanonical step; nvir 130: kept counts 82..130
    straddle the 128-wide class and the full width. Batches of 4 pairs (classes
    formed per batch) give the same pair energies."""
    p, _ = synthetic.local_chain(nocc=nocc, nvir=nvir, rows_per_site=rows, seed=seed)
    print(f"pno width classes {nocc}/{nvir}", t_cut, flush=True)
    one = _pno_check(oracle_mod, p.b_ov, p.eps, p.nocc, p.nvir, 0, p.nocc, _pairs(p.nocc), t_cut)
    monkeypatch.setenv("FERRIC_PNO_BATCH", "4")
    few = pno_pair_energies(_cuda(p.b_ov), p.eps, p.nocc, p.nvir, 0, p.nocc, _pairs(p.nocc),
                            t_cut)
    assert few[2] == one[2]
    scale = np.max(np.abs(one[0]))
    assert np.max(np.abs(few[0] - one[0])) <= 1e-12 * scale
    assert np.max(np.abs(few[1] - one[1])) <= 1e-12 * scale


def test_pno_reference_toy_and_frozen_core(oracle_mod, golden):
    from tests.test_oracle_pno import toy_b_ov

    b, eps = toy_b_ov(14, 4, 6)
    print("pno reference toy 14/4/6", flush=True)
    _pno_check(oracle_mod, b, eps, 4, 6, 0, 4, _pairs(4), 1e-2)
    for name, c in golden.items():
        if c["first_occ"] == 0:
            continue
        print("pno frozen core", name, flush=True)
        nocc, nvir = c["nocc"], c["nvir"]
        _pno_check(oracle_mod, c["b"], c["eps"], nocc, nvir, c["first_occ"], c["nocc_total"],
                   _pairs(nocc), 1e-5)


def test_pno_never_empties_a_pair(oracle_mod):
    p, _ = synthetic.local_chain(nocc=5, nvir=24, rows_per_site=8)
    print("pno huge threshold", flush=True)
    g_os, g_ss, kept, _ = _pno_check(oracle_mod, p.b_ov, p.eps, p.nocc, p.nvir, 0, p.nocc,
                                     _pairs(p.nocc), 1e300)
    assert kept == len(_pairs(p.nocc))
    assert np.all(np.abs(g_ss) <= 1e-12 * np.abs(g_os))


def test_pno_batches_host_entry_and_driver(oracle_mod, monkeypatch):
    p, centers = synthetic.local_chain(nocc=8, nvir=32, rows_per_site=10)
    print("pno batches / host entry / driver", flush=True)
    pairs = _pairs(p.nocc)
    one = pno_pair_energies(_cuda(p.b_ov), p.eps, p.nocc, p.nvir, 0, p.nocc, pairs, 1e-5)
    host = pno_pair_energies(p.b_ov, p.eps, p.nocc, p.nvir, 0, p.nocc, pairs, 1e-5)
    assert np.array_equal(one[0], host[0]) and np.array_equal(one[1], host[1])
    assert one[2:] == host[2:]
    monkeypatch.setenv("FERRIC_PNO_BATCH", "5")
    few = pno_pair_energies(_cuda(p.b_ov), p.eps, p.nocc, p.nvir, 0, p.nocc, pairs, 1e-5)
    assert few[2] == one[2]
    scale = np.max(np.abs(one[0]))
    assert np.max(np.abs(few[0] - one[0])) <= 1e-12 * scale
    assert np.max(np.abs(few[1] - one[1])) <= 1e-12 * scale
    monkeypatch.delenv("FERRIC_PNO_BATCH")
    # the driver: default pair screen plus PNO truncation, against the oracle's sum
    pe = estimate_pair_energies_from_b_ov(p.b_ov, p.eps, p.nocc, p.nvir)
    cfg = DlpnoConfig(t_cut_pno=1e-6)
    d = build_pair_domains_by_energy(centers, pe, cfg.t_cut_pairs)
    sc, diag = dlpno_mp2_from_b_ov(_cuda(p.b_ov), p.eps, p.nocc, p.nvir, 0, p.nocc, d, cfg)
    w_os, w_ss, w_kept, w_disc, _ = oracle_mod.pno_pair_energies(p.b_ov, p.eps, p.nocc, p.nvir,
                                                                 0, p.nocc, d.pairs, 1e-6)
    assert abs(sc.e_os - np.sum(w_os)) <= 1e-10 * abs(sc.e_total)
    assert abs(sc.e_ss - np.sum(w_ss)) <= 1e-10 * abs(sc.e_total)
    assert diag.virtual_retention == w_kept.sum() / (len(d.pairs) * p.nvir) < 1.0
    assert diag.n_pairs == len(d.pairs) and diag.max_discarded_weight > 0.0
    dense = rimp2.spin_components(_cuda(p.b_ov), p.eps, p.nocc, p.nvir)
    assert sc.e_total > dense.e_total  # truncation loses correlation


# -- domain-local fits (lmp2_amplitude.rs:1707-1758) --------------------------------

def test_domain_fit_blocks_against_oracle(oracle_mod, monkeypatch):
    a, v, xyz, cen, eps = synthetic.local_fit_chain(nocc=8, nvir=16, aux_per_site=6)
    nocc, nvir = cen.shape[0], 16
    pairs = _pairs(nocc)
    for radius in (3.0, 6.0, math.inf):
        print(f"domain fit radius {radius}", flush=True)
        jd, dmax = domain_fit_pair_blocks(_cuda(a), _cuda(v), xyz, cen, nocc, nvir, radius, pairs)
        jd = jd.cpu().numpy()
        sizes = []
        for k, (i, j) in enumerate(pairs):
            want, d = oracle_mod.domain_fit_pair(a, v, xyz, cen, i, j, nvir, radius)
            sizes.append(d)
            assert np.max(np.abs(jd[k] - want)) <= 1e-12 * np.max(np.abs(want)), (radius, i, j)
        assert dmax == max(sizes)
        if math.isinf(radius):
            assert min(sizes) == a.shape[0]
        else:
            assert min(sizes) < dmax  # uneven domains in one batch
    jh, _ = domain_fit_pair_blocks(a, v, xyz, cen, nocc, nvir, 3.0, pairs)
    jd, _ = domain_fit_pair_blocks(_cuda(a), _cuda(v), xyz, cen, nocc, nvir, 3.0, pairs)
    assert np.array_equal(jh, jd.cpu().numpy())
    monkeypatch.setenv("FERRIC_FIT_BATCH", "7")
    jb, _ = domain_fit_pair_blocks(_cuda(a), _cuda(v), xyz, cen, nocc, nvir, 3.0, pairs)
    assert np.max(np.abs(jb.cpu().numpy() - jh)) <= 1e-13 * np.max(np.abs(jh))


def test_domain_fit_unbounded_gives_the_canonical_pairs(oracle_mod):
    a, v, xyz, cen, eps = synthetic.local_fit_chain(nocc=6, nvir=12, aux_per_site=5)
    nocc, nvir = 6, 12
    print("domain fit unbounded vs canonical pair energies", flush=True)
    jd, _ = domain_fit_pair_blocks(_cuda(a), _cuda(v), xyz, cen, nocc, nvir, math.inf)
    b = np.linalg.solve(np.linalg.cholesky(v), a)
    pe = estimate_pair_energies_from_b_ov(_cuda(b), eps, nocc, nvir)
    ev = eps[nocc:]
    for k, (i, j) in enumerate(_pairs(nocc)):
        g = jd[k].cpu().numpy()
        d = eps[i] + eps[j] - ev[:, None] - ev[None, :]
        fac = 1.0 if i == j else 2.0
        e = fac * np.sum((2 * g - g.T) * g / d)
        assert abs(e - pe.e[i, j]) <= 1e-11 * np.max(np.abs(pe.e))


def test_domain_fit_refusals():
    a, v, xyz, cen, _ = synthetic.local_fit_chain(nocc=3, nvir=4, aux_per_site=4)
    print("domain fit refusals", flush=True)
    far = cen + np.array([0.0, 100.0, 0.0])
    with pytest.raises(LayoutError):
        domain_fit_pair_blocks(_cuda(a), _cuda(v), xyz, far, 3, 4, 1.0)
    bad = v.copy()
    bad[0, 0] = -1.0  # not positive definite in any domain holding aux 0
    with pytest.raises(KernelError):
        domain_fit_pair_blocks(_cuda(a), _cuda(bad), xyz, cen, 3, 4, math.inf, [(0, 1)])
    with pytest.raises(LayoutError):
        domain_fit_pair_blocks(_cuda(a).float(), _cuda(v), xyz, cen, 3, 4, 1.0)


# -- the dense-g entries (pair_energy_screen.rs:125-190, dlpno_mp2.rs:354-440) ------

def test_dense_g_pair_energies_and_pno(oracle_mod):
    from ferric_amd.dlpno import (dlpno_mp2_spin_components, estimate_pair_energies,
                                  pno_pair_energies_from_g)

    p, centers = synthetic.local_chain(nocc=8, nvir=40, rows_per_site=10)
    nocc, nvir = p.nocc, p.nvir
    rng = np.random.default_rng(77)
    c = rng.standard_normal(p.b_ov.shape) * 0.02
    for name, g in (("gram", p.b_ov.T @ p.b_ov), ("non-gram", p.b_ov.T @ (p.b_ov + c))):
        print("dense g", name, flush=True)
        gd = _cuda(g)
        pe = estimate_pair_energies(gd, p.eps, nocc, nvir)
        w_os, w_ss = oracle_mod.pair_energies_from_g(g, p.eps, nocc, nvir)
        _close(pe.e_os, w_os, w_os)
        _close(pe.e_ss, w_ss, w_os)
        pairs = _pairs(nocc)
        for t in (0.0, 1e-6):
            g_os, g_ss, kept, worst = pno_pair_energies_from_g(gd, p.eps, nocc, nvir, 0, nocc,
                                                               pairs, t)
            v_os, v_ss, v_kept, v_disc, _ = oracle_mod.pno_pair_energies(
                None, p.eps, nocc, nvir, 0, nocc, pairs, t, g=g)
            assert kept == int(v_kept.sum())
            assert np.max(np.abs(g_os - v_os)) <= 1e-10 * np.max(np.abs(v_os))
            assert np.max(np.abs(g_ss - v_ss)) <= 1e-10 * np.max(np.abs(v_os))
        d = build_pair_domains_by_energy(centers, pe, 1e-5)
        sc, diag = dlpno_mp2_spin_components(gd, p.eps, nocc, nvir, 0, nocc, d,
                                             DlpnoConfig(t_cut_pno=1e-6))
        v_os, v_ss, v_kept, _, _ = oracle_mod.pno_pair_energies(None, p.eps, nocc, nvir, 0, nocc,
                                                                d.pairs, 1e-6, g=g)
        assert abs(sc.e_total - (np.sum(v_os) + np.sum(v_ss))) <= 1e-10 * abs(sc.e_total)
        assert diag.virtual_retention == v_kept.sum() / (len(d.pairs) * nvir)
        sc0, _ = dlpno_mp2_spin_components(gd, p.eps, nocc, nvir, 0, nocc, d)
        assert abs(sc0.e_total - sum(pe.e_ij(i, j) for i, j in d.pairs)) <= 1e-12 * abs(sc0.e_total)
    # the Gram case against the B entry
    pb = estimate_pair_energies_from_b_ov(_cuda(p.b_ov), p.eps, nocc, nvir)
    pg = estimate_pair_energies(_cuda(p.b_ov.T @ p.b_ov), p.eps, nocc, nvir)
    _close(pg.e, pb.e, pb.e_os)
