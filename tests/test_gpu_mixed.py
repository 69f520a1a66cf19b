"""Precision.MIXED, the reference's mixed arithmetic, on the GPU (run with -m gpu).

MIXED rounds B to FP32, accumulates each (ia|jb) over Q panels of at most 64
rows in FP32 MFMA registers and adds every panel's sums into FP64 registers
(crates/ferric-core/src/gpu/mixed.rs:70,122-180). Checks:

* the reference's cancelling construction (gpu_rimp2_mixed.rs:411-447): three
  panels summing to +2^20, 2^-9 and -2^20 keep the 2^-9 through the FP64
  flush (read here through the pair energy it produces), where the FP32 mode,
  which accumulates the whole Q range in FP32, loses it;
* energies within the reference's error model (tests/common/rimp2_eps.rs,
  restated in oracle/mixed.py): against the FP64 oracle of the FP32-rounded B
  within eps_accumulation(naux, 64) S_ab per element, against the FP64 oracle
  of B itself within eps_device(naux, 64) S_ab, propagated through the energy
  bound; on the golden fixtures, real water (the published H2O / cc-pVDZ
  comparison: -7.0e-9 Eh inside 1.6e-6 / 2.8e-6, validation.md:527), the
  decane-scale workload and the hot-row distribution;
* the range rule of check_f32_arithmetic_range with b = 64, and the
  dispatcher's mixed_ok rule (rimp2.rs:1277, 1293-1300).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import ferric_amd  # noqa: E402
from ferric_amd import Precision, dlpno, rimp2, synthetic  # noqa: E402
from oracle import mixed as mx  # noqa: E402

MIXED = Precision.MIXED
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _gpu(b, eps, nocc, nvir, fc=0, nt=None, kappa=None, prec=MIXED):
    t = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    return rimp2.spin_components(t, eps, nocc, nvir, fc, nt, kappa, prec)


def _bounds(oracle_mod, b, eps, nocc, nvir, fc=0, nt=None, kappa=None):
    naux = b.shape[0]
    eta = mx.eta_abs(naux, float(np.abs(b).max()))
    br = b.astype(np.float32).astype(np.float64)
    acc = oracle_mod.energy_bound(br, eps, nocc, nvir, fc, nt, kappa=kappa,
                                  eps_g=mx.eps_accumulation(naux, 64), eta=eta)
    dev = oracle_mod.energy_bound(b, eps, nocc, nvir, fc, nt, kappa=kappa,
                                  eps_g=mx.eps_device(naux, 64), eta=eta)
    return br, acc, dev


def _check(oracle_mod, sc, b, eps, nocc, nvir, fc=0, nt=None, kappa=None, label=""):
    br, acc, dev = _bounds(oracle_mod, b, eps, nocc, nvir, fc, nt, kappa)
    want_r = oracle_mod.spin_components(br, eps, nocc, nvir, fc, nt, kappa)
    want = oracle_mod.spin_components(b, eps, nocc, nvir, fc, nt, kappa)
    print(f"  {label}: dE vs FP64 {sc.e_os - want[0]:+.2e} / {sc.e_ss - want[1]:+.2e} "
          f"(total {sc.e_total - want[2]:+.2e}; bound {dev[0]:.1e} / {dev[1]:.1e}); "
          f"accumulation part {sc.e_os - want_r[0]:+.2e} / {sc.e_ss - want_r[1]:+.2e}",
          flush=True)
    assert abs(sc.e_os - want_r[0]) <= acc[0] and abs(sc.e_ss - want_r[1]) <= acc[1], label
    assert abs(sc.e_os - want[0]) <= dev[0] and abs(sc.e_ss - want[1]) <= dev[1], label
    return sc.e_os - want[0], sc.e_ss - want[1]


def test_flush_construction_keeps_the_middle_panel():
    x = mx.flush_construction(64)            # (192, 2): ones, then +2^20 / 2^-9 / -2^20
    eps = np.array([-1.0, -1.0, 1.0])        # two occupied, one virtual: D = -4
    print("mixed flush construction", flush=True)
    pe = dlpno.estimate_pair_energies_from_b_ov(x, eps, 2, 1, precision=MIXED)
    g = 2.0 ** -9
    assert pe.e_os[0, 1] == 2.0 * g * g / -4.0, pe.e_os
    f32 = dlpno.estimate_pair_energies_from_b_ov(x, eps, 2, 1, precision=Precision.FP32)
    assert f32.e_os[0, 1] == 0.0, f32.e_os    # whole-range FP32 accumulation loses it


def test_golden_cases_within_the_mixed_error_model(golden, oracle_mod):
    for name, c in golden.items():
        print("mixed golden", name, flush=True)
        args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
        _check(oracle_mod, _gpu(*args), *args, label=name)


def test_real_water_against_the_published_mixed_comparison(oracle_mod):
    """RI-JK water (5/19/84): the reference measured dE_total = -7.0e-9 Eh with
    its cuBLAS panels (bound 1.6e-6 / 2.8e-6). The rounding of another kernel
    gives another small number; it must sit inside the same bound and at the
    published scale (|dE| <= 1e-7, a sixteenth of the OS bound)."""
    w = np.load(os.path.join(ROOT, "tests", "golden", "water_ccpvdz_real.npz"))
    for tag in ("rijk_", "", "dist_"):
        b, eps = w[f"{tag}b_ov"], w[f"{tag}eps"]
        d_os, d_ss = _check(oracle_mod, _gpu(b, eps, 5, 19), b, eps, 5, 19, label=f"water {tag}")
        assert abs(d_os + d_ss) <= 1e-7
        # one frozen core orbital
        bf = np.ascontiguousarray(b[:, 19:])
        _check(oracle_mod, _gpu(bf, eps, 4, 19, 1, 5), bf, eps, 4, 19, 1, 5,
               label=f"water {tag}fc1")


def test_decane_scale_and_hot_rows(oracle_mod):
    p = synthetic.decane_scale()
    print("mixed decane-scale", flush=True)
    sc = _gpu(p.b_ov, p.eps, p.nocc, p.nvir)
    f64 = oracle_mod.spin_components_blas(p.b_ov, p.eps, p.nocc, p.nvir, nthreads=16)
    rows = oracle_mod.row_partials_blas(p.b_ov, p.eps, p.nocc, p.nvir, range(p.nocc),
                                        eps_g=mx.eps_device(p.naux, 64),
                                        eta=mx.eta_abs(p.naux, float(np.abs(p.b_ov).max())))
    bos = sum(r[2] for r in rows.values())
    bss = sum(r[3] for r in rows.values())
    print(f"  dE {sc.e_os - f64[0]:+.2e} / {sc.e_ss - f64[1]:+.2e} (bound {bos:.1e} / "
          f"{bss:.1e})", flush=True)
    assert abs(sc.e_os - f64[0]) <= bos and abs(sc.e_ss - f64[1]) <= bss
    h = synthetic.mixed_large(seed=1013, nocc=12, nvir=120, naux=480)
    print("mixed hot rows 12/120/480", flush=True)
    b = h.b_ov.astype(np.float64)
    _check(oracle_mod, _gpu(h.b_ov, h.eps, h.nocc, h.nvir), b, h.eps, h.nocc, h.nvir,
           label="hot rows")


def test_kappa_and_tiles(golden, oracle_mod, monkeypatch):
    c = golden["res_150_9_31_2_5"]
    args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    for tile in ("32", "64", "128"):
        monkeypatch.setenv("FERRIC_HIP_TILE", tile)
        print("mixed tile", tile, flush=True)
        _check(oracle_mod, _gpu(*args), *args, label=f"tile {tile}")
        _check(oracle_mod, _gpu(*args, kappa=1.3), *args, kappa=1.3, label=f"kappa tile {tile}")


def test_range_rule_uses_the_64_row_panel():
    flt_max = float(np.finfo(np.float32).max)
    naux, nocc, nvir = 1024, 2, 3
    eps = np.array([-1.0, -0.9, 0.5, 0.7, 0.9])
    b = np.full((naux, nocc * nvir), 1e-3)
    b[3, 2] = np.sqrt(flt_max / 512)          # above sqrt(FLT_MAX / naux), below / 64
    print("mixed range rule", flush=True)
    with pytest.raises(ferric_amd.RangeError):
        _gpu(b, eps, nocc, nvir, prec=Precision.FP32)
    assert np.isfinite(_gpu(b, eps, nocc, nvir).e_total)
    b[3, 2] = 2.0 * np.sqrt(flt_max / 64)
    with pytest.raises(ferric_amd.RangeError):
        _gpu(b, eps, nocc, nvir)
    os.environ["FERRIC_GPU_PRECISION"] = "mixed"
    try:
        got = rimp2.spin_components_from_b_ov_kappa(b, eps, nocc, nvir, 0, nocc)
    finally:
        del os.environ["FERRIC_GPU_PRECISION"]
    want = _gpu(b, eps, nocc, nvir, prec=Precision.FP64)
    assert (got.e_os, got.e_ss) == (want.e_os, want.e_ss)


def test_mixed_ok_rule_through_the_dispatcher(golden, monkeypatch):
    c = golden["res_120_7_23_1_3"]
    args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    monkeypatch.setenv("FERRIC_GPU_PRECISION", "mixed")
    mixed = _gpu(*args)
    f64 = _gpu(*args, prec=Precision.FP64)
    f64k = _gpu(*args, kappa=1.3, prec=Precision.FP64)
    assert (mixed.e_os, mixed.e_ss) != (f64.e_os, f64.e_ss)
    got = rimp2.spin_components_fitted(*args)
    assert (got.e_os, got.e_ss) == (mixed.e_os, mixed.e_ss)
    for op in ("erfc_coulomb", "terfc"):
        got = rimp2.spin_components_fitted(*args, op_kind=op)
        assert (got.e_os, got.e_ss) == (mixed.e_os, mixed.e_ss), op
    for kw in ({"op_kind": "erf_coulomb"}, {"op_kind": "yukawa"}, {"composite": True},
               {"fit_metric": "erfc"}):
        got = rimp2.spin_components_fitted(*args, **kw)
        assert (got.e_os, got.e_ss) == (f64.e_os, f64.e_ss), kw
    got = rimp2.spin_components_fitted(*args, kappa=1.3)
    assert (got.e_os, got.e_ss) == (f64k.e_os, f64k.e_ss)
    with pytest.raises(ferric_amd.LayoutError):
        rimp2.spin_components_fitted(*args, kappa=-1.0)


def test_opposite_and_same_spin_blocks(oracle_mod):
    from ferric_amd import unrestricted as u

    rng = np.random.default_rng(2031)

    def chan(naux, nocc, nvir):
        b = rng.standard_normal((naux, nocc * nvir)) * np.sqrt(0.09 / naux)
        occ = np.sort(rng.uniform(-1.2, -0.35, nocc))
        vir = np.sort(np.exp(rng.uniform(np.log(0.12), np.log(5.0), nvir)))
        return u.SpinChannel(b, np.concatenate([occ, vir]), nocc, nvir)

    a, b = chan(300, 9, 70), chan(300, 8, 75)
    print("mixed uhf blocks", flush=True)
    ss = u.u_same_spin(a, precision=MIXED)
    os_ = u.u_opposite_spin(a, b, precision=MIXED)
    want_ss = oracle_mod.u_same_spin(a.b, a.eps, a.nocc, a.nvir)
    want_os = oracle_mod.u_opposite_spin((a.b, a.eps, a.nocc, a.nvir, 0, a.nocc),
                                         (b.b, b.eps, b.nocc, b.nvir, 0, b.nocc))
    print(f"  same spin dE {ss - want_ss:+.2e}, opposite spin dE {os_ - want_os:+.2e}",
          flush=True)
    # FP32-class accuracy: within eps_device(naux, 64) of every element's S,
    # i.e. a few 1e-6 relative at most
    assert abs(ss - want_ss) <= 1e-5 * abs(want_ss)
    assert abs(os_ - want_os) <= 1e-5 * abs(want_os)
