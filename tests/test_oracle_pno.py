"""Pins the oracle's PNO pair energies (dlpno_mp2.rs:249-340, :700-744;
local_pno.rs:132-200) by what fixes them independently of their formula.

* Exactness (dlpno_mp2.rs:797-830): at t_cut_pno = 0 every pair energy equals
  the pinned canonical pair energy.
* Trace identity: kept plus discarded occupations equal tr D = 2 ||T||_F^2 for
  T = g / D, which fails for a density built from g, or from one of its two
  terms.
* An independent route to the truncated energy: the kept subspace from
  scipy's subset eigensolver and the semicanonical energy as a Sylvester
  solve, (F - e_ij/2) Y + Y (F - e_ij/2) = -X with X = Q^T g Q, F = Q^T e Q,
  e_os = f <X, Y>, e_ss = f <X - X^T, Y> -- no second eigensolve, basis-free.
* The reference's own properties (local_pno.rs:281-420, dlpno_mp2.rs:972-1003,
  :1415-1460): zero threshold keeps everything, a huge one keeps one PNO per
  pair (whose same-spin energy is then zero), retention falls with the
  threshold, truncation loses correlation, a global amplitude sign changes
  nothing.
"""
import numpy as np
import pytest
import scipy.linalg

from ferric_amd import synthetic


def toy_b_ov(naux, nocc, nvir):
    """dlpno_mp2.rs:1026-1043, re-implemented bit for bit (64-bit LCG)."""
    s = 0x9E3779B97F4A7C15
    m = (1 << 64) - 1
    b = np.empty(naux * nocc * nvir)
    for k in range(b.size):
        s = (s * 6364136223846793005 + 1442695040888963407) & m
        b[k] = (s >> 33) / float(1 << 31) - 1.0
    eps = [-1.0 - 0.1 * i for i in range(nocc)] + [0.5 + 0.1 * a for a in range(nvir)]
    return b.reshape(naux, nocc * nvir), np.array(eps)


def _all_pairs(nocc):
    return [(i, j) for i in range(nocc) for j in range(i, nocc)]


def _chain():
    p, _ = synthetic.local_chain(nocc=6, nvir=20, rows_per_site=10)
    return p


def test_zero_truncation_is_exact(oracle_mod):
    for b, eps, nocc, nvir in (toy_b_ov(14, 4, 6) + (4, 6), (lambda p: (p.b_ov, p.eps, p.nocc,
                                                                          p.nvir))(_chain())):
        pairs = _all_pairs(nocc)
        os0, ss0 = oracle_mod.pair_energies(b, eps, nocc, nvir)
        os_, ss, kept, disc, _ = oracle_mod.pno_pair_energies(b, eps, nocc, nvir, 0, nocc, pairs,
                                                              0.0)
        scale = np.max(np.abs(os0))
        for k, (i, j) in enumerate(pairs):
            assert abs(os_[k] - os0[i, j]) <= 1e-12 * scale
            assert abs(ss[k] - ss0[i, j]) <= 1e-12 * scale
        assert np.all(kept == nvir) and np.all(disc == 0.0)


@pytest.mark.parametrize("t_cut", [1e-6, 1e-4, 1e-2])
def test_trace_identity_and_sylvester_route(oracle_mod, t_cut):
    p = _chain()
    b, eps, nocc, nvir = p.b_ov, p.eps, p.nocc, p.nvir
    pairs = _all_pairs(nocc)
    os_, ss, kept, disc, kocc = oracle_mod.pno_pair_energies(b, eps, nocc, nvir, 0, nocc, pairs,
                                                             t_cut)
    ev = eps[nocc:]
    for k, (i, j) in enumerate(pairs):
        g = b[:, i * nvir:(i + 1) * nvir].T @ b[:, j * nvir:(j + 1) * nvir]
        eij = eps[i] + eps[j]
        t = g / (eij - ev[:, None] - ev[None, :])
        tr = 2.0 * np.sum(t * t)
        assert abs(kocc[k] + disc[k] - tr) <= 1e-12 * tr
        d = t @ t.T + t.T @ t
        w, q = scipy.linalg.eigh(0.5 * (d + d.T), subset_by_value=(t_cut, np.inf), driver="evr")
        if w.size == 0:  # never empty: the largest occupation
            w, q = scipy.linalg.eigh(0.5 * (d + d.T), subset_by_index=(nvir - 1, nvir - 1))
        assert w.size == kept[k]
        x = q.T @ g @ q
        f = q.T @ (ev[:, None] * q)
        a = f - 0.5 * eij * np.eye(w.size)
        y = scipy.linalg.solve_sylvester(a, a, -x)
        fac = 1.0 if i == j else 2.0
        scale = abs(os_[k]) + 1e-300
        assert abs(fac * np.sum(x * y) - os_[k]) <= 1e-11 * scale
        assert abs(fac * np.sum((x - x.T) * y) - ss[k]) <= 1e-11 * scale


def test_reference_properties(oracle_mod):
    b, eps = toy_b_ov(14, 4, 6)
    nocc, nvir = 4, 6
    pairs = _all_pairs(nocc)
    last = None
    for t in (0.0, 1e-4, 1e-2, 1.0, 1e300):
        os_, ss, kept, disc, _ = oracle_mod.pno_pair_energies(b, eps, nocc, nvir, 0, nocc, pairs, t)
        if last is not None:
            assert np.all(kept <= last)
        last = kept
    assert np.all(kept == 1) and np.all(np.abs(ss) <= 1e-12 * np.abs(os_))
    os0, ss0 = oracle_mod.pair_energies(b, eps, nocc, nvir)
    full = np.sum((os0 + ss0)[np.triu_indices(nocc)])
    os_, ss, kept, disc, _ = oracle_mod.pno_pair_energies(b, eps, nocc, nvir, 0, nocc, pairs, 1e-2)
    assert kept.sum() < len(pairs) * nvir and disc.max() > 0.0
    assert np.sum(os_ + ss) > full  # truncation loses (negative) correlation energy
    b2 = b.copy()
    b2[:, 2 * nvir:3 * nvir] *= -1.0  # T_2j -> -T_2j for every j != 2
    o2, s2, k2, d2, _ = oracle_mod.pno_pair_energies(b2, eps, nocc, nvir, 0, nocc, pairs, 1e-2)
    assert np.array_equal(k2, kept)
    np.testing.assert_allclose(o2, os_, rtol=1e-12)
    np.testing.assert_allclose(s2, ss, rtol=1e-11, atol=1e-14 * np.max(np.abs(os_)))


def test_dense_g_source_matches_the_b_source(oracle_mod):
    """The dense-g restatements against the pinned B ones: with g = B^T B the
    block source changes nothing; with a non-Gram g the pairs sum to the pinned
    spin_components_from_g."""
    b, eps = toy_b_ov(14, 4, 6)
    nocc, nvir = 4, 6
    pairs = _all_pairs(nocc)
    g = b.T @ b
    os0, ss0 = oracle_mod.pair_energies(b, eps, nocc, nvir)
    os1, ss1 = oracle_mod.pair_energies_from_g(g, eps, nocc, nvir)
    np.testing.assert_allclose(os1, os0, rtol=1e-12)
    np.testing.assert_allclose(ss1, ss0, rtol=1e-12, atol=1e-13 * np.max(np.abs(os0)))
    for t in (0.0, 1e-2):
        w = oracle_mod.pno_pair_energies(b, eps, nocc, nvir, 0, nocc, pairs, t)
        v = oracle_mod.pno_pair_energies(None, eps, nocc, nvir, 0, nocc, pairs, t, g=g)
        assert np.array_equal(w[2], v[2])
        np.testing.assert_allclose(v[0], w[0], rtol=1e-10)
    rng = np.random.default_rng(9)
    c = rng.uniform(-1, 1, b.shape)
    gn = b.T @ c + 0.3 * c.T @ b
    os2, ss2 = oracle_mod.pair_energies_from_g(gn, eps, nocc, nvir)
    tot = oracle_mod.spin_components_from_g(gn, eps, nocc, nvir)
    iu = np.triu_indices(nocc)
    assert abs(np.sum(os2[iu]) - tot[0]) <= 1e-12 * abs(tot[0])
    assert abs(np.sum(ss2[iu]) - tot[1]) <= 1e-12 * abs(tot[0])
