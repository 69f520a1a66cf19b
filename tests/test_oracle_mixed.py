"""The restatement of the reference's mixed arithmetic and error model
(oracle/mixed.py), pinned on the CPU.

* The error factors against the reference's own published numbers: the panel
  bound eps_G(64, naux) = 3.93e-6 for naux 84..3600, and the energy bound of
  H2O / cc-pVDZ (5/19/84), published as 1.6e-6 (OS) / 2.8e-6 (SS) Eh
  (site/src/reference/validation.md:527), from this project's RI-JK water B.
* The host twin (mixed_host.rs:54-86): one panel is a plain f32 product; the
  reference's cancelling construction (gpu_rimp2_mixed.rs:411-447) keeps 2^-9
  under the f64 flush and loses it under the f32-across-panels mutant; every
  element lies within eps_accumulation * S_ab of the f64 product of the rounded
  operands, on the hot-row distribution too; the RMS error grows with the
  panel width.
"""
import os

import numpy as np
import pytest

from oracle import mixed as mx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("naux", [84, 364, 700, 1036, 3600])
def test_panel_bound_matches_the_published_value(naux):
    assert mx.eps_device(naux, 64) == pytest.approx(3.93e-6, abs=0.005e-6)
    # the f32 part dominates: independent of naux to within the u64 terms
    assert mx.eps_accumulation(naux, 64) < mx.eps_device(naux, 64)
    # the host twin's first-order factor plus the f64 reference's own term
    assert mx.mixed_error_factor(naux, 64) + mx.gamma(naux, mx.U64) == pytest.approx(
        mx.eps_device(naux, 64), rel=1e-3)


def test_water_energy_bound_matches_the_published_bound(oracle_mod):
    w = np.load(os.path.join(ROOT, "tests", "golden", "water_ccpvdz_real.npz"))
    b, eps = w["rijk_b_ov"], w["rijk_eps"]
    naux = b.shape[0]
    bos, bss = oracle_mod.energy_bound(b, eps, 5, 19, eps_g=mx.eps_device(naux, 64),
                                       eta=mx.eta_abs(naux, float(np.abs(b).max())))
    assert float(f"{bos:.1e}") == 1.6e-6 and float(f"{bss:.1e}") == 2.8e-6, (bos, bss)


def test_one_panel_is_a_plain_f32_product():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((7, 50))
    b = rng.standard_normal((50, 9))
    want = (a.astype(np.float32) @ b.astype(np.float32)).astype(np.float64)
    assert np.array_equal(mx.gemm_f32_f64acc_host(a, b, 50), want)
    assert np.array_equal(mx.gemm_f32_f64acc_host(a, b, 500), want)


def test_flush_construction_keeps_what_an_f32_sum_loses():
    x = mx.flush_construction(64)
    left = x[:, 0:1].T
    assert mx.gemm_f32_f64acc_host(left, x, 64)[0, 1] == 2.0 ** -9
    assert mx.f32_across_panels_host(left, x, 64)[0, 1] == 0.0


@pytest.mark.parametrize("hot", [False, True])
def test_every_element_within_the_accumulation_bound(hot):
    rng = np.random.default_rng(2)
    naux, n = 480, 60
    b = rng.standard_normal((naux, n)) * np.sqrt(0.09 / naux)
    if hot:
        b[rng.choice(naux, 5, replace=False)] *= 20.0
    br = b.astype(np.float32).astype(np.float64)
    got = mx.gemm_f32_f64acc_host(br.T, br, 64)
    ref = br.T @ br
    s = np.abs(br).T @ np.abs(br)
    bound = mx.eps_accumulation(naux, 64) * s + mx.eta_abs(naux, float(np.abs(b).max()))
    ratio = np.abs(got - ref) / bound
    assert ratio.max() <= 1.0
    assert ratio.max() >= 1e-3     # a bound loose by orders would fail


def test_rms_error_grows_with_the_panel_width():
    rng = np.random.default_rng(3)
    naux, n = 2048, 48
    b = rng.standard_normal((naux, n)).astype(np.float32).astype(np.float64)
    ref = b.T @ b
    s = np.abs(b).T @ np.abs(b)
    rms = []
    for kb in (8, 64, 512, naux):
        err = (mx.gemm_f32_f64acc_host(b.T, b, kb) - ref) / s
        rms.append(float(np.sqrt(np.mean(err ** 2))))
    assert all(y > x for x, y in zip(rms, rms[1:])), rms
