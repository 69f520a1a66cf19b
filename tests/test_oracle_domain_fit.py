"""Pins the oracle's domain-local fit (domain_fit_pair, lmp2_amplitude.rs:1707-1758).

* An unbounded radius puts every aux function in every domain, and the fit is
  the whole-basis one: J_ij = B_i^T B_j with B = L^-1 A, V = L L^T -- whose
  pair energies are the pinned canonical ones.
* Locality: when A_i and A_j live on aux functions near their centres and the
  metric does not couple them to the rest, the domain fit equals the
  whole-basis fit exactly (to rounding), at a radius that leaves most aux
  functions out.
* Permuting the aux functions (positions, metric, rows of A) changes nothing.
* An empty domain is an error (lmp2_amplitude.rs:1736-1741).
"""
import math

import numpy as np
import pytest
import scipy.linalg

from ferric_amd import synthetic


def _whole(a, v):
    l = np.linalg.cholesky(v)
    return scipy.linalg.solve_triangular(l, a, lower=True)


def test_unbounded_radius_is_the_whole_basis_fit(oracle_mod):
    a, v, xyz, cen, eps = synthetic.local_fit_chain()
    nocc, nvir = cen.shape[0], a.shape[1] // cen.shape[0]
    b = _whole(a, v)
    for i in range(nocc):
        for j in range(i, nocc):
            jb, d = oracle_mod.domain_fit_pair(a, v, xyz, cen, i, j, nvir, math.inf)
            want = b[:, i * nvir:(i + 1) * nvir].T @ b[:, j * nvir:(j + 1) * nvir]
            assert d == a.shape[0]
            assert np.max(np.abs(jb - want)) <= 1e-12 * np.max(np.abs(want))
    # and the pair energies of those blocks are the pinned canonical ones
    os_, ss = oracle_mod.pair_energies(b, eps, nocc, nvir)
    jb, _ = oracle_mod.domain_fit_pair(a, v, xyz, cen, 1, 3, nvir, math.inf)
    ev = eps[nocc:]
    dd = eps[1] + eps[3] - ev[:, None] - ev[None, :]
    assert abs(2 * np.sum(jb * jb / dd) - os_[1, 3]) <= 1e-12 * abs(os_[1, 3])


def test_local_support_and_decoupled_metric_give_the_whole_fit(oracle_mod):
    rng = np.random.default_rng(41)
    nocc, nvir, per = 4, 5, 6
    naux = nocc * per
    xyz = np.zeros((naux, 3))
    xyz[:, 0] = np.repeat(10.0 * np.arange(nocc), per) + rng.uniform(-1, 1, naux)
    cen = np.zeros((nocc, 3))
    cen[:, 0] = 10.0 * np.arange(nocc)
    v = np.zeros((naux, naux))
    a = np.zeros((naux, nocc * nvir))
    for s in range(nocc):
        blk = slice(s * per, (s + 1) * per)
        m = rng.standard_normal((per, per))
        v[blk, blk] = m @ m.T + per * np.eye(per)
        a[blk, s * nvir:(s + 1) * nvir] = rng.standard_normal((per, nvir))
    b = _whole(a, v)
    for i, j in ((0, 0), (0, 1), (1, 3), (2, 2)):
        jb, d = oracle_mod.domain_fit_pair(a, v, xyz, cen, i, j, nvir, 2.0)
        assert d == (per if i == j else 2 * per)
        want = b[:, i * nvir:(i + 1) * nvir].T @ b[:, j * nvir:(j + 1) * nvir]
        assert np.max(np.abs(jb - want)) <= 1e-13 * (np.max(np.abs(want)) + 1.0)


def test_aux_permutation_changes_nothing(oracle_mod):
    a, v, xyz, cen, _ = synthetic.local_fit_chain(nocc=5, nvir=6)
    nvir = 6
    perm = np.random.default_rng(3).permutation(a.shape[0])
    for i, j in ((0, 2), (1, 1), (3, 4)):
        j0, d0 = oracle_mod.domain_fit_pair(a, v, xyz, cen, i, j, nvir, 4.0)
        j1, d1 = oracle_mod.domain_fit_pair(a[perm], v[np.ix_(perm, perm)], xyz[perm], cen, i, j,
                                            nvir, 4.0)
        assert d0 == d1 < a.shape[0]
        assert np.max(np.abs(j1 - j0)) <= 1e-12 * np.max(np.abs(j0))


def test_empty_domain_is_an_error(oracle_mod):
    a, v, xyz, cen, _ = synthetic.local_fit_chain(nocc=3, nvir=4)
    cen = cen + np.array([0.0, 100.0, 0.0])
    with pytest.raises(ValueError):
        oracle_mod.domain_fit_pair(a, v, xyz, cen, 0, 1, 4, 1.0)
