"""Pins the oracle's per-pair energies and tests the pair screen (host logic).

The pair energies of estimate_pair_energies_from_b_ov (dlpno_mp2.rs:527-594)
are pinned by what fixes them independently of their formula: they sum to the
pinned closed-shell energy over the golden cases (frozen core included) and to
the ORCA RI-MP2 correlation energy of real water; B blocks on disjoint aux rows
make every off-diagonal pair vanish and each diagonal pair the one-orbital
energy; scaling one occupied's columns by s scales its pairs by s^2 and its
diagonal by s^4, which fixes which entry belongs to which pair. The screen
follows the reference's own tests in pair_energy_screen.rs:320-505.
"""
import math
import os

import numpy as np
import pytest

from ferric_amd import LayoutError, Precision, UnsupportedError
from ferric_amd.dlpno import (DlpnoConfig, PairDomains, PairEnergies,
                              build_pair_domains_by_energy, dlpno_mp2_from_b_ov)

FIX = os.path.join(os.path.dirname(__file__), "golden", "water_ccpvdz_real.npz")


def _pe(oracle_mod, b, eps, nocc, nvir, fo=0, nt=None):
    os_, ss = oracle_mod.pair_energies(b, eps, nocc, nvir, fo, nt)
    return PairEnergies(nocc, os_ + ss, os_, ss)


def _upper_sum(m):
    n = m.shape[0]
    return math.fsum(m[i, j] for i in range(n) for j in range(i, n))


def test_pair_energies_sum_to_golden_energy(oracle_mod, golden):
    """pair_energy_screen.rs:345-378: the pairs sum to the MP2 components."""
    for name, c in golden.items():
        os_, ss = oracle_mod.pair_energies(c["b"], c["eps"], c["nocc"], c["nvir"],
                                           c["first_occ"], c["nocc_total"])
        w_os, w_ss = c["energy"]
        assert abs(_upper_sum(os_) - w_os) <= 1e-12 * abs(w_os), name
        assert abs(_upper_sum(ss) - w_ss) <= 1e-12 * abs(w_os), name
        assert np.array_equal(os_, os_.T) and np.array_equal(ss, ss.T)


def test_pair_energies_sum_to_orca_water(oracle_mod):
    z = np.load(FIX)
    nocc, nvir = int(z["nocc"]), int(z["nvir"])
    os_, ss = oracle_mod.pair_energies(z["dist_b_ov"], z["dist_eps"], nocc, nvir)
    assert abs(_upper_sum(os_ + ss) - float(z["dist_e_corr_ref"])) < 1e-8


def _ladder(nocc, nvir):
    return np.concatenate([-1.0 - 0.15 * np.arange(nocc), 0.4 + 0.2 * np.arange(nvir)])


def test_disjoint_aux_support_isolates_the_diagonal(oracle_mod):
    rng = np.random.default_rng(31)
    nocc, nvir, rows = 4, 5, 7
    b = np.zeros((nocc * rows, nocc * nvir))
    eps = _ladder(nocc, nvir)
    for i in range(nocc):
        b[i * rows:(i + 1) * rows, i * nvir:(i + 1) * nvir] = rng.standard_normal((rows, nvir))
    os_, ss = oracle_mod.pair_energies(b, eps, nocc, nvir)
    e = os_ + ss
    assert np.all(e[~np.eye(nocc, dtype=bool)] == 0.0)
    for i in range(nocc):
        one = oracle_mod.spin_components(np.ascontiguousarray(b[:, i * nvir:(i + 1) * nvir]),
                                         np.concatenate([eps[i:i + 1], eps[nocc:]]), 1, nvir)
        assert abs(e[i, i] - one[2]) <= 1e-13 * abs(one[2])
        # i == j: g is symmetric, so the same-spin part is exactly zero.
        assert abs(ss[i, i]) <= 1e-15 * abs(os_[i, i])


def test_scaling_one_orbital_fixes_the_pair_index(oracle_mod):
    rng = np.random.default_rng(32)
    nocc, nvir, naux, k, s = 5, 4, 19, 2, 1.7
    b = rng.standard_normal((naux, nocc * nvir))
    eps = _ladder(nocc, nvir)
    e0 = sum(oracle_mod.pair_energies(b, eps, nocc, nvir))
    b2 = b.copy()
    b2[:, k * nvir:(k + 1) * nvir] *= s
    e1 = sum(oracle_mod.pair_energies(b2, eps, nocc, nvir))
    want = e0.copy()
    want[k, :] *= s * s
    want[:, k] *= s * s
    np.testing.assert_allclose(e1, want, rtol=1e-13, atol=0)


def test_non_negative_denominator_is_refused(oracle_mod):
    """pair_energy_screen.rs:494-502: HOMO and LUMO degenerate."""
    rng = np.random.default_rng(33)
    b = rng.standard_normal((6, 12))
    eps = np.array([-1.0, -0.8, -0.6, -0.6, 0.7, 1.1, 1.6])
    with pytest.raises(ValueError):
        oracle_mod.pair_energies(b, eps, 3, 4)


# -- the screen (pair_energy_screen.rs:206-282) -----------------------------------

def _synthetic(oracle_mod):
    rng = np.random.default_rng(0xC0FFEE)
    nocc, nvir = 3, 4
    eps = np.array([-1.0, -0.8, -0.6, 0.4, 0.7, 1.1, 1.6])
    return _pe(oracle_mod, rng.standard_normal((9, nocc * nvir)), eps, nocc, nvir)


CENTERS3 = np.array([[0.0, 0.0, 0.0], [1.5, 0.0, 0.0], [20.0, 0.0, 0.0]])


def test_zero_threshold_retains_every_pair(oracle_mod):
    pe = _synthetic(oracle_mod)
    d = build_pair_domains_by_energy(CENTERS3, pe, 0.0)
    assert len(d.pairs) == 6 and d.is_complete()
    assert d.pair_retention() == 1.0 and d.coupling_retention() == 1.0
    assert math.isinf(d.pair_cutoff)


def test_retention_falls_monotonically_and_keeps_the_diagonal(oracle_mod):
    pe = _synthetic(oracle_mod)
    last = 10 ** 9
    for t in (0.0, 1e-6, 1e-4, 1e-2, 1.0, 1e6, np.finfo(float).max):
        d = build_pair_domains_by_energy(CENTERS3, pe, t)
        assert d.nocc <= len(d.pairs) <= last
        last = len(d.pairs)
    assert last == 3 and all((i, i) in d.pairs for i in range(3))
    assert all(len(row) >= 1 for row in d.coupled)


def test_discarded_energy_matches_the_dropped_pairs(oracle_mod):
    pe = _synthetic(oracle_mod)
    for t in (0.0, 1e-4, 1e-2, 1.0):
        d = build_pair_domains_by_energy(CENTERS3, pe, t)
        kept = sum(pe.e_ij(i, j) for i, j in d.pairs)
        assert abs(kept + pe.discarded_energy(t) - pe.total()) < 1e-12


def test_a_mid_threshold_drops_some_but_not_all(oracle_mod):
    pe = _synthetic(oracle_mod)
    offs = sorted(abs(pe.e_ij(i, j)) for i in range(3) for j in range(i + 1, 3))
    t = 0.5 * (offs[0] + offs[-1])
    d = build_pair_domains_by_energy(CENTERS3, pe, t)
    assert not d.is_complete() and len(d.pairs) > 3


def test_coupling_screen_on_pair_centroids(oracle_mod):
    pe = _synthetic(oracle_mod)
    d = build_pair_domains_by_energy(CENTERS3, pe, 0.0, 2.0)
    # Centroids: (0,0)=0 (0,1)=0.75 (0,2)=10 (1,1)=1.5 (1,2)=10.75 (2,2)=20.
    x = {p: 0.5 * (CENTERS3[p[0], 0] + CENTERS3[p[1], 0]) for p in d.pairs}
    for p, row in enumerate(d.coupled):
        want = [q for q in range(len(d.pairs))
                if q == p or abs(x[d.pairs[q]] - x[d.pairs[p]]) <= 2.0]
        assert row == want
    assert d.coupled[0] == [0, 1, 3] and d.coupled[5] == [5]
    assert 0.0 < d.coupling_retention() < 1.0


def test_invalid_inputs_are_errors(oracle_mod):
    pe = _synthetic(oracle_mod)
    for t, c in ((-1.0, math.inf), (math.nan, math.inf), (0.0, -1.0)):
        with pytest.raises(LayoutError):
            build_pair_domains_by_energy(CENTERS3, pe, t, c)
    with pytest.raises(LayoutError):
        build_pair_domains_by_energy(np.zeros((3, 2)), pe, 0.0)
    with pytest.raises(LayoutError):
        build_pair_domains_by_energy(np.zeros((2, 3)), pe, 0.0)


def test_dlpno_config_checks_before_any_device_call(oracle_mod):
    pe = _synthetic(oracle_mod)
    d = build_pair_domains_by_energy(CENTERS3, pe, 0.0)
    b = np.zeros((9, 12))
    eps = np.array([-1.0, -0.8, -0.6, 0.4, 0.7, 1.1, 1.6])
    with pytest.raises(LayoutError):
        dlpno_mp2_from_b_ov(b, eps, 3, 4, 0, 3, d, DlpnoConfig(t_cut_pno=-1.0), pair_energies=pe)
    with pytest.raises(UnsupportedError):  # PNO truncation runs in FP64
        dlpno_mp2_from_b_ov(b, eps, 3, 4, 0, 3, d, DlpnoConfig(t_cut_pno=1e-8),
                            precision=Precision.BF16X3, pair_energies=pe)
    with pytest.raises(LayoutError):
        dlpno_mp2_from_b_ov(b, eps, 4, 3, 0, 4, d, pair_energies=pe)
    assert DlpnoConfig.exact().is_exact() and not DlpnoConfig().is_exact()
    assert DlpnoConfig().t_cut_pairs == 1e-5


def test_dlpno_sum_over_retained_pairs(oracle_mod):
    """With the pair energies supplied, the t_cut_pno = 0 energy is host logic:
    the exact config reproduces the total, a screen drops the discarded energy."""
    pe = _synthetic(oracle_mod)
    b = np.zeros((9, 12))
    eps = np.array([-1.0, -0.8, -0.6, 0.4, 0.7, 1.1, 1.6])
    d = build_pair_domains_by_energy(CENTERS3, pe, 0.0)
    sc, diag = dlpno_mp2_from_b_ov(b, eps, 3, 4, 0, 3, d, DlpnoConfig.exact(), pair_energies=pe)
    assert abs(sc.e_total - pe.total()) < 1e-14 and diag.pair_retention == 1.0
    assert abs(sc.e_os + sc.e_ss - sc.e_total) == 0.0
    offs = sorted(abs(pe.e_ij(i, j)) for i in range(3) for j in range(i + 1, 3))
    t = 0.5 * (offs[0] + offs[-1])
    d = build_pair_domains_by_energy(CENTERS3, pe, t)
    sc, diag = dlpno_mp2_from_b_ov(b, eps, 3, 4, 0, 3, d, pair_energies=pe)
    assert abs(sc.e_total - (pe.total() - pe.discarded_energy(t))) < 1e-14
    assert diag.n_pairs == len(d.pairs) and diag.virtual_retention == 1.0
