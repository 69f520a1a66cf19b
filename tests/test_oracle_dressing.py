"""Pins of the oracle's dressing / MO-transform stage (rimp2.rs:648-756), CPU only.

* against an independent numpy einsum of B = M C_left^T (Q|mu nu) C_right;
* output bands stitch into the full tensor;
* end to end on real water (tests/golden/water_ccpvdz_real.npz, distorted
  geometry): B dressed with V^-1/2 from this project's (Q|mu nu), MO
  coefficients and metric gives the ORCA/PySCF RI-MP2 energy to 1e-10 Eh
  (a different factor from the fixture's Cholesky one: the energy is invariant).
"""
import os

import numpy as np
import pytest

import oracle

FIX = os.path.join(os.path.dirname(__file__), "golden", "water_ccpvdz_real.npz")


def _inv_sqrt(v):
    w, u = np.linalg.eigh(v)
    return (u / np.sqrt(w)) @ u.T


def test_against_einsum_and_bands():
    rng = np.random.default_rng(3)
    naux, nao, nl, nr = 300, 11, 3, 7
    a = rng.standard_normal((naux, nao, nao))
    a = a + a.transpose(0, 2, 1)
    m = rng.standard_normal((naux, naux))
    cl, cr = rng.standard_normal((nao, nl)), rng.standard_normal((nao, nr))
    want = m @ np.einsum("Qmn,mi,na->Qia", a, cl, cr).reshape(naux, nl * nr)
    got = oracle.dress_b_ov(a, m, cl, cr)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    band = oracle.dress_b_ov(a, m, cl, cr, output_band=(100, 260))
    assert np.max(np.abs(band - want[100:260])) <= 1e-12 * np.max(np.abs(want))
    # a transposed C_left fails (the pin has teeth)
    wrong = m @ np.einsum("Qmn,mi,na->Qia", a, cl[::-1], cr).reshape(naux, nl * nr)
    assert np.max(np.abs(wrong - got)) > 1e-3 * np.max(np.abs(got))


def test_real_water_energy_through_the_dressing():
    w = np.load(FIX)
    j3, c, j2 = w["dist_j3"], w["dist_c"], w["dist_j2"]
    b = oracle.dress_b_ov(j3, _inv_sqrt(j2), c[:, :5], c[:, 5:])
    e = oracle.spin_components(b, w["dist_eps"], 5, 19)
    assert abs(e[0] - float(w["dist_e_os_ref"])) <= 1e-10
    assert abs(e[1] - float(w["dist_e_ss_ref"])) <= 1e-10
    assert abs(e[2] - float(w["dist_e_corr_orca"])) <= 1e-10
