"""Real water (cc-pVDZ / cc-pVDZ-RI) through the oracle, against published energies.

The fixture tests/golden/water_ccpvdz_real.npz holds B and the orbital energies
built by tests/golden/make_water_real.py (this project's own integrals and RHF).
The distorted geometry's references were computed by two independent codes
(ORCA 6.1.1 and PySCF 2.13.1, agreeing to 5e-11 Eh; reference file
testdata/reference/validation/{rimp2_gradient,scs_mp2_gradient}/
h2o_distorted_cc-pvdz.json), so the oracle is pinned there at 1e-10 Eh,
including the OS/SS split and one frozen core orbital. The equilibrium geometry's
known answer (testdata/reference/h2o_cc-pvdz_rimp2.json) is met within 1e-7 Eh;
the reference's own test of that number allows 1e-6 (rimp2.rs:3047-3052).
"""
import os

import numpy as np
import pytest

import oracle

FIX = os.path.join(os.path.dirname(__file__), "golden", "water_ccpvdz_real.npz")


@pytest.fixture(scope="module")
def water():
    return np.load(FIX)


def test_fixture_rhf_energies(water):
    assert abs(float(water["e_rhf"]) - float(water["e_rhf_reference"])) <= 1e-9
    assert abs(float(water["dist_e_rhf"]) - float(water["dist_e_rhf_ref"])) <= 1e-9


def test_distorted_water_all_electron(water):
    b, eps = water["dist_b_ov"], water["dist_eps"]
    e_os, e_ss, e = oracle.spin_components(b, eps, 5, 19)
    assert abs(e_os - float(water["dist_e_os_ref"])) <= 1e-10
    assert abs(e_ss - float(water["dist_e_ss_ref"])) <= 1e-10
    assert abs(e - float(water["dist_e_corr_ref"])) <= 1e-10
    assert abs(e - float(water["dist_e_corr_orca"])) <= 1e-10


def test_distorted_water_frozen_core(water):
    b, eps = water["dist_b_ov"], water["dist_eps"]
    b_fc = np.ascontiguousarray(b[:, 19:])          # active occupied 1..4
    e_os, e_ss, e = oracle.spin_components(b_fc, eps, 4, 19, first_occ=1, nocc_total=5)
    assert abs(e_os - float(water["dist_e_os_fc_ref"])) <= 1e-10
    assert abs(e_ss - float(water["dist_e_ss_fc_ref"])) <= 1e-10
    assert abs(e - float(water["dist_e_corr_fc_orca"])) <= 1e-10


def test_equilibrium_water_known_answer(water):
    e = oracle.spin_components(water["b_ov"], water["eps"], 5, 19)[2]
    assert abs(e - float(water["e_corr_reference"])) <= 1e-7
    assert e == pytest.approx(float(water["e_corr_numpy"]), abs=1e-13)


def test_h2_known_answer(water):
    """H2 / cc-pVDZ at 0.74 Angstrom, the reference's test_rimp2_h2_ccpvdz
    (rimp2.rs:3187-3201): RI-MP2 within 1e-4 of PySCF's canonical MP2."""
    nocc, nvir = int(water["h2_nocc"]), int(water["h2_nvir"])
    e = oracle.spin_components(water["h2_b_ov"], water["h2_eps"], nocc, nvir)[2]
    assert abs(e - float(water["h2_e_corr_canonical"])) < 1e-4
    assert e == pytest.approx(float(water["h2_e_corr_numpy"]), abs=1e-14)
    # one occupied orbital: only the (0, 0) pair, whose same-spin part is zero
    e_os, e_ss, _ = oracle.spin_components(water["h2_b_ov"], water["h2_eps"], nocc, nvir)
    assert nocc == 1 and abs(e_ss) <= 1e-15 * abs(e_os)


def test_rijk_water_known_answer(water):
    """Water with a density-fitted SCF (def2-universal-jkfit) and cc-pVDZ-RI for
    MP2: the reference's RI-JK known answer (testdata/reference/
    h2o_cc-pvdz_rijk-def2-universal-jkfit_dfmp2.json, PySCF 2.13.1), which the
    fixture's own integrals and SCF reproduce to 1e-12 Eh."""
    assert abs(float(water["rijk_e_rhf"]) - float(water["rijk_e_rhf_reference"])) <= 1e-9
    e = oracle.spin_components(water["rijk_b_ov"], water["rijk_eps"], 5, 19)[2]
    assert abs(e - float(water["rijk_e_corr_reference"])) <= 1e-10
    assert e == pytest.approx(float(water["rijk_e_corr_numpy"]), abs=1e-13)


def test_conventional_water_known_answer(water):
    """Conventional MP2 of the equilibrium water (testdata/reference/
    h2o_cc-pvdz_mp2.json, PySCF 2.13.0): B is an exact factor of the fixture's
    four-centre (ia|jb) (95 x 95, full rank), so the RI-MP2 path returns the
    unfitted energy. The reference's SCF energy is the RI file's to 3e-12 Eh
    (same geometry); the energy is met to 2e-10 Eh."""
    assert abs(float(water["conv_e_rhf_reference"]) - float(water["e_rhf"])) <= 1e-9
    b = water["conv_b_ov"]
    assert b.shape == (95, 95)
    e = oracle.spin_components(b, water["eps"], 5, 19)[2]
    assert abs(e - float(water["conv_e_corr_reference"])) <= 1e-9
    assert e == pytest.approx(float(water["conv_e_corr_numpy"]), abs=1e-13)
    # the fit error: RI-MP2 recovers all but about 1.5e-5 Eh of it here
    ri = oracle.spin_components(water["b_ov"], water["eps"], 5, 19)[2]
    assert -1e-4 < e - ri < 0
