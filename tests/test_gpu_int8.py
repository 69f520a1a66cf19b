"""The INT8 digit-slice kernel (Precision.INT8X4) against the oracle, run with -m gpu.

* Exact format: on the golden fixtures (tile-32 and tile-64 shapes, frozen
  core, single-product and two-product items) the GPU energy equals the oracle
  energy (rimp2.rs:984-1018 restated) of quant.int8x4_g(b), the g the format
  forms bit for bit, up to FP64 summation order. This pins the digit packer,
  the lane layout of the int32 MFMA result and the level combination.
* Against FP64: the golden cases and the decane-scale workload within the
  format's derived bound (quant.int8x4_energy_bound: half a unit of rounding
  per element and the dropped levels, propagated through the pair sums), and
  the 100/900/3600 workload within 1 uEh of the FP64 kernel.
* Errors: naux beyond the int32 level-sum range is a RangeError (after which the
  dispatcher's FP64 retry returns the FP64 energy); a NaN in B gives NaN.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import ferric_amd  # noqa: E402
from ferric_amd import Precision, rimp2, synthetic  # noqa: E402
from tests import quant  # noqa: E402

I8 = Precision.INT8X4


def _gpu(b, eps, nocc, nvir, fc=0, nt=None, kappa=None, prec=I8):
    t = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    return rimp2.spin_components(t, eps, nocc, nvir, fc, nt, kappa, prec)


def test_exact_format_golden(golden, oracle_mod):
    for name, c in golden.items():
        print("int8x4 exact format", name, flush=True)
        nocc, nvir, fc, nt = c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"]
        b, eps = c["b"], c["eps"]
        sc = _gpu(b, eps, nocc, nvir, fc, nt)
        want = oracle_mod.spin_components_from_g(quant.int8x4_g(b), eps, nocc, nvir, fc, nt)
        tol = 1e-13 * (abs(want[0]) + abs(want[1])) + 1e-15
        assert abs(sc.e_os - want[0]) <= tol, (name, sc, want)
        assert abs(sc.e_ss - want[1]) <= tol, (name, sc, want)
        # and against the FP64 energy of the fixture (its generator's brute force)
        d_os, d_ss = sc.e_os - c["energy"][0], sc.e_ss - c["energy"][1]
        bos, bss = quant.int8x4_energy_bound(b, eps, nocc, nvir, fc, nt)
        print(f"  dE vs fp64 {d_os:.2e} {d_ss:.2e}  bound {bos:.2e} {bss:.2e}", flush=True)
        assert abs(d_os) <= bos + 1e-14 and abs(d_ss) <= bss + 1e-14, name
        sk = _gpu(b, eps, nocc, nvir, fc, nt, kappa=1.3)
        kos, kss = c["energy_kappa"]
        assert abs(sk.e_os - kos) <= bos + 1e-14 and abs(sk.e_ss - kss) <= bss + 1e-14, name


def test_decane_scale_within_the_format_bound(oracle_mod):
    p = synthetic.decane_scale()
    print("int8x4 decane-scale", flush=True)
    ref = oracle_mod.spin_components_blas(p.b_ov, p.eps, p.nocc, p.nvir, nthreads=16)
    sc = _gpu(p.b_ov, p.eps, p.nocc, p.nvir)
    bos, bss = quant.int8x4_energy_bound(p.b_ov, p.eps, p.nocc, p.nvir)
    print(f"  E_os {sc.e_os:.13f} ({sc.e_os - ref[0]:.2e}, bound {bos:.2e})  E_ss "
          f"{sc.e_ss:.13f} ({sc.e_ss - ref[1]:.2e}, bound {bss:.2e})", flush=True)
    assert abs(sc.e_os - ref[0]) <= bos and abs(sc.e_ss - ref[1]) <= bss
    # the error is FP64-class: far inside 1 uEh
    assert abs(sc.e_total - ref[2]) <= 1e-8
    # FP32 storage of the same values goes through the same digits
    s32 = _gpu(p.b_ov.astype(np.float32), p.eps, p.nocc, p.nvir)
    s64 = _gpu(p.b_ov.astype(np.float32).astype(np.float64), p.eps, p.nocc, p.nvir)
    assert (s32.e_os, s32.e_ss) == (s64.e_os, s64.e_ss)


def test_panel_rotation_changes_no_bit(monkeypatch):
    """Each item may start its Q panels anywhere (a wall-clock phase): the int32
    level sums are exact, so energies and the adjoint's Gamma / dE/deps are the
    same bit for bit with the rotation off (0), changing every tick (1), at an
    odd rate (7) and at the built-in rate. Shapes: an odd panel count with
    single- and two-product items (41/209/960, 15 panels), the 32-wide tile, and
    the hot-row distribution."""
    cases = [synthetic.decane_scale(), synthetic.mixed_large(seed=11, nocc=12, nvir=20, naux=200),
             synthetic.mixed_large(seed=12, nocc=9, nvir=130, naux=520)]
    for p in cases:
        t = torch.from_numpy(p.b_ov).cuda()
        got = {}
        for ticks in ("0", "1", "7", None):
            if ticks is None:
                monkeypatch.delenv("FERRIC_I8_QROT_TICKS", raising=False)
            else:
                monkeypatch.setenv("FERRIC_I8_QROT_TICKS", ticks)
            sc = rimp2.spin_components(t, p.eps, p.nocc, p.nvir, precision=I8)
            got[ticks] = (sc.e_os, sc.e_ss)
        print("int8x4 rotation", p.nocc, p.nvir, p.b_ov.shape[0], got["0"], flush=True)
        assert len(set(got.values())) == 1, got
    # the opposite-spin rectangle (alpha and beta halves split the panels by parity)
    from ferric_amd import unrestricted as un
    pa = synthetic.mixed_large(seed=13, nocc=9, nvir=70, naux=520)
    pb = synthetic.mixed_large(seed=14, nocc=8, nvir=75, naux=520)
    ca = un.SpinChannel(torch.from_numpy(pa.b_ov).cuda(), pa.eps, pa.nocc, pa.nvir)
    cb = un.SpinChannel(torch.from_numpy(pb.b_ov).cuda(), pb.eps, pb.nocc, pb.nvir)
    got = set()
    for ticks in ("0", "1", "7"):
        monkeypatch.setenv("FERRIC_I8_QROT_TICKS", ticks)
        got.add(un.u_opposite_spin(ca, cb, I8))
    assert len(got) == 1, got
    p = synthetic.decane_scale()
    t = torch.from_numpy(p.b_ov).cuda()
    outs = []
    for ticks in ("0", "1"):
        monkeypatch.setenv("FERRIC_I8_QROT_TICKS", ticks)
        sc, e, gam, deps = rimp2.energy_adjoint(t, p.eps, p.nocc, p.nvir, 0, p.nocc, 1.2, 1 / 3,
                                                I8)
        outs.append((e, gam.cpu().numpy(), np.asarray(deps)))
    assert outs[0][0] == outs[1][0]
    assert np.array_equal(outs[0][1], outs[1][1]) and np.array_equal(outs[0][2], outs[1][2])


def test_mixed_workload_within_one_micro_hartree():
    p = synthetic.mixed_large()
    print("int8x4 mixed 100/900/3600", flush=True)
    t = torch.from_numpy(p.b_ov).cuda()
    f64 = rimp2.spin_components(t, p.eps, p.nocc, p.nvir, precision=Precision.FP64)
    i8 = rimp2.spin_components(t, p.eps, p.nocc, p.nvir, precision=I8)
    d = i8.e_total - f64.e_total
    print(f"  E fp64 {f64.e_total:.9f} int8x4 {i8.e_total:.9f} dE {d:.3e}", flush=True)
    assert abs(d) <= 1e-6
    del t
    torch.cuda.empty_cache()


def test_range_and_non_finite():
    rng = np.random.default_rng(7)
    nocc, nvir = 2, 6
    eps = np.concatenate([[-1.0, -0.8], np.linspace(0.3, 1.5, nvir)])
    big = rng.standard_normal((32900, nocc * nvir)) * 1e-2
    print("int8x4 range", flush=True)
    with pytest.raises(ferric_amd.RangeError):
        _gpu(big, eps, nocc, nvir)
    ref = _gpu(big, eps, nocc, nvir, prec=Precision.FP64)
    tr = rimp2.try_spin_components_on_device(big, eps, nocc, nvir, 0, nocc, mixed_ok=True)
    assert tr is not None
    os.environ["FERRIC_GPU_PRECISION"] = "int8x4"
    try:
        tr = rimp2.try_spin_components_on_device(big, eps, nocc, nvir, 0, nocc, mixed_ok=True)
    finally:
        del os.environ["FERRIC_GPU_PRECISION"]
    assert (tr.e_os, tr.e_ss) == (ref.e_os, ref.e_ss)
    b = rng.standard_normal((100, nocc * nvir)) * 0.03
    b[17, 4] = np.nan
    print("int8x4 nan", flush=True)
    sc = _gpu(b, eps, nocc, nvir)
    assert np.isnan(sc.e_total)


def test_non_finite_row_meets_finite_rows():
    """A NaN in one packed row (i, a) reaches every energy that reads it, also
    where it meets a finite row of another occupied (pair (i, j), i != j) or of
    the other spin channel: the row exponent marking it must survive the sum of
    two exponents the epilogue forms, whatever the finite row's exponent."""
    from ferric_amd import dlpno
    from ferric_amd import unrestricted as u

    rng = np.random.default_rng(11)
    nocc, nvir, naux = 3, 20, 96
    eps = np.concatenate([[-1.2, -1.0, -0.8], np.linspace(0.3, 1.5, nvir)])
    b = rng.standard_normal((naux, nocc * nvir)) * 0.03   # row maxima ~0.1: exponents ~ -3
    b[5, 1 * nvir + 7] = np.nan                           # occupied 1, virtual 7
    print("int8x4 nan pairs", flush=True)
    pe = dlpno.estimate_pair_energies_from_b_ov(b, eps, nocc, nvir, precision=I8)
    for i in range(nocc):
        for j in range(i, nocc):
            touches = 1 in (i, j)
            assert np.isnan(pe.e_os[i, j]) == touches, (i, j, pe.e_os)
            assert np.isnan(pe.e_ss[i, j]) == touches, (i, j, pe.e_ss)
    print("int8x4 nan opposite spin", flush=True)
    beta = rng.standard_normal((naux, 2 * 24)) * 0.03
    eps_b = np.concatenate([[-1.1, -0.9], np.linspace(0.25, 1.4, 24)])
    ca = u.SpinChannel(b, eps, nocc, nvir)
    cb = u.SpinChannel(beta, eps_b, 2, 24)
    assert np.isnan(u.u_opposite_spin(ca, cb, precision=I8))
    clean = b.copy()
    clean[5, 1 * nvir + 7] = 0.0
    assert np.isfinite(u.u_opposite_spin(u.SpinChannel(clean, eps, nocc, nvir), cb, precision=I8))
