"""Unrestricted blocks and the signed factor on the GPU, against the oracle (-m gpu).

FP64 results match the oracle's restatement of u_rimp2.rs:693-760 and
rimp2.rs:1559-1607 to 1e-10 Eh; FP32 results match the oracle on the
FP32-rounded inputs within the FP32-accumulation scale (naux * 2^-24
relative per element, stated per test). Shapes are uneven on purpose (alpha and
beta of different sizes, frozen cores, aux counts off the panel depth).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import ferric_amd  # noqa: E402
from ferric_amd import Precision, rimp2  # noqa: E402
from ferric_amd import unrestricted as u  # noqa: E402


def _channel(rng, naux, nocc, nvir, fc=0):
    b = rng.standard_normal((naux, nocc * nvir)) * np.sqrt(0.09 / naux)
    occ = np.sort(np.concatenate([rng.uniform(-11.3, -11.1, fc), rng.uniform(-1.2, -0.35, nocc)]))
    vir = np.sort(np.exp(rng.uniform(np.log(0.12), np.log(5.0), nvir)))
    return u.SpinChannel(b, np.concatenate([occ, vir]), nocc, nvir, fc, fc + nocc)


def _oracle_ch(ch):
    return (ch.b, ch.eps, ch.nocc, ch.nvir, ch.first_occ, ch.nocc_total)


@pytest.fixture(scope="module")
def pair_small():
    rng = np.random.default_rng(2024)
    return _channel(rng, 93, 5, 37, fc=1), _channel(rng, 93, 4, 41, fc=1)


@pytest.fixture(scope="module")
def pair_mid():
    rng = np.random.default_rng(2025)
    return _channel(rng, 500, 12, 120, fc=2), _channel(rng, 500, 11, 125, fc=2)


@pytest.mark.parametrize("tile", ["32", "64", "128"])
def test_same_and_opposite_spin_fp64_every_tile(pair_small, oracle_mod, tile, monkeypatch):
    monkeypatch.setenv("FERRIC_HIP_TILE", tile)
    a, b = pair_small
    print("uhf small tile", tile, flush=True)
    for ch in (a, b):
        got = u.u_same_spin(ch)
        want = oracle_mod.u_same_spin(*_oracle_ch(ch))
        assert abs(got - want) <= 1e-10, (got, want)
    got = u.u_opposite_spin(a, b)
    want = oracle_mod.u_opposite_spin(_oracle_ch(a), _oracle_ch(b))
    assert abs(got - want) <= 1e-10, (got, want)


def test_uhf_mid_size_and_device_entries(pair_mid, oracle_mod):
    a, b = pair_mid
    print("uhf mid", flush=True)
    e_ab = u.u_opposite_spin(a, b)
    want = oracle_mod.u_opposite_spin(_oracle_ch(a), _oracle_ch(b))
    assert abs(e_ab - want) <= 1e-10
    assert u.u_opposite_spin(b, a) == pytest.approx(e_ab, rel=1e-12)
    e_aa = u.u_same_spin(a)
    assert abs(e_aa - oracle_mod.u_same_spin(*_oracle_ch(a))) <= 1e-10
    # device tensors on the current stream give the host entries' results
    da = u.SpinChannel(torch.from_numpy(a.b).cuda(), a.eps, a.nocc, a.nvir, a.first_occ)
    db = u.SpinChannel(torch.from_numpy(b.b).cuda(), b.eps, b.nocc, b.nvir, b.first_occ)
    assert u.u_opposite_spin(da, db) == e_ab
    assert u.u_same_spin(da) == e_aa


def test_closed_shell_reductions_on_device(pair_small):
    a, _ = pair_small
    sc = rimp2.spin_components(a.b, a.eps, a.nocc, a.nvir, a.first_occ, a.nocc_total)
    assert u.u_same_spin(a) == 0.5 * sc.e_ss
    assert u.u_opposite_spin(a, a) == pytest.approx(sc.e_os, rel=1e-12)


def test_fp32_opposite_spin_exact_format(pair_mid, oracle_mod):
    a, b = pair_mid
    fa = u.SpinChannel(a.b.astype(np.float32).astype(np.float64), a.eps, a.nocc, a.nvir,
                       a.first_occ)
    fb = u.SpinChannel(b.b.astype(np.float32).astype(np.float64), b.eps, b.nocc, b.nvir,
                       b.first_occ)
    want = oracle_mod.u_opposite_spin(_oracle_ch(fa), _oracle_ch(fb))
    got = u.u_opposite_spin(a, b, Precision.FP32)
    # FP32 accumulation over naux = 500: each (ia|JB) within ~naux * 2^-24 of its
    # scale, the energy (sum of squares) within twice that relative
    assert abs(got - want) <= 2 * 500 * 2.0 ** -24 * abs(want)
    print(f"fp32 opposite-spin rel dE {abs(got - want) / abs(want):.2e}", flush=True)


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.FP32])
def test_signed_factor(oracle_mod, prec):
    rng = np.random.default_rng(31)
    nocc, nvir, fc = 6, 29, 1
    ch = _channel(rng, 37, nocc, nvir, fc)
    neg = rng.standard_normal((23, nocc * nvir)) * np.sqrt(0.02 / 23)
    print("signed", prec.name, flush=True)
    got = u.spin_components_from_signed_b_ov(ch.b, neg, ch.eps, nocc, nvir, fc, fc + nocc, prec)
    if prec == Precision.FP64:
        want = oracle_mod.signed_spin_components(ch.b, neg, ch.eps, nocc, nvir, fc, fc + nocc)
        assert abs(got.e_os - want[0]) <= 1e-10 and abs(got.e_ss - want[1]) <= 1e-10
    else:
        want = oracle_mod.signed_spin_components(ch.b.astype(np.float32).astype(np.float64),
                                                 neg.astype(np.float32).astype(np.float64),
                                                 ch.eps, nocc, nvir, fc, fc + nocc)
        assert abs(got.e_total - want[2]) <= 2 * 60 * 2.0 ** -24 * abs(want[2])
    # no negative rows: the plain energy
    plain = rimp2.spin_components(ch.b, ch.eps, nocc, nvir, fc, fc + nocc, precision=prec)
    none = u.spin_components_from_signed_b_ov(ch.b, np.zeros((0, nocc * nvir)), ch.eps, nocc,
                                              nvir, fc, fc + nocc, prec)
    assert none.e_total == plain.e_total
    # rows in both factors cancel
    y = rng.standard_normal((16, nocc * nvir)) * 0.01
    both = u.spin_components_from_signed_b_ov(np.vstack([ch.b, y]), y, ch.eps, nocc, nvir, fc,
                                              fc + nocc, Precision.FP64)
    ref = rimp2.spin_components(ch.b, ch.eps, nocc, nvir, fc, fc + nocc)
    assert both.e_total == pytest.approx(ref.e_total, rel=1e-12)


def test_signed_device_entry_and_errors(oracle_mod):
    rng = np.random.default_rng(5)
    nocc, nvir = 4, 33
    ch = _channel(rng, 50, nocc, nvir)
    neg = rng.standard_normal((18, nocc * nvir)) * 0.01
    host = u.spin_components_from_signed_b_ov(ch.b, neg, ch.eps, nocc, nvir)
    dev = u.spin_components_from_signed_b_ov(torch.from_numpy(ch.b).cuda(),
                                             torch.from_numpy(neg).cuda(), ch.eps, nocc, nvir)
    assert dev.e_total == host.e_total
    with pytest.raises(ferric_amd.UnsupportedError):
        u.spin_components_from_signed_b_ov(ch.b, neg, ch.eps, nocc, nvir,
                                           precision=Precision.BF16X3)
    with pytest.raises(ferric_amd.UnsupportedError):
        u.u_opposite_spin(ch, ch, Precision.FP8_E4M3)
    assert u.try_u_opposite_spin_on_device(ch, ch) == pytest.approx(
        oracle_mod.u_opposite_spin(_oracle_ch(ch), _oracle_ch(ch)), abs=1e-10)


def test_same_spin_int8x4_within_the_format_bound(pair_mid, oracle_mod):
    """The same-spin block through the INT8X4 kernel: half the closed-shell
    same-spin sum, so within half of the format's derived same-spin bound of
    the oracle (u_rimp2.rs:693-729), and FP64-class (1e-8 Eh)."""
    from tests import quant

    for ch in pair_mid:
        print("uhf same-spin int8x4", ch.nocc, ch.nvir, flush=True)
        got = u.u_same_spin(ch, Precision.INT8X4)
        want = oracle_mod.u_same_spin(*_oracle_ch(ch))
        _, bss = quant.int8x4_energy_bound(ch.b, ch.eps, ch.nocc, ch.nvir, ch.first_occ,
                                           ch.nocc_total)
        assert abs(got - want) <= 0.5 * bss + 1e-14, (got, want, bss)
        assert abs(got - want) <= 1e-8


def _rect_energy(g, ch_a, ch_b):
    """sum (ia|JB)^2 / D over every alpha i, a and beta J, B of g (nov_a, nov_b)."""
    eo_a = ch_a.eps[ch_a.first_occ:ch_a.first_occ + ch_a.nocc]
    ev_a = ch_a.eps[ch_a.nocc_total:ch_a.nocc_total + ch_a.nvir]
    eo_b = ch_b.eps[ch_b.first_occ:ch_b.first_occ + ch_b.nocc]
    ev_b = ch_b.eps[ch_b.nocc_total:ch_b.nocc_total + ch_b.nvir]
    d = (eo_a[:, None, None, None] + eo_b[None, None, :, None] - ev_a[None, :, None, None]
         - ev_b[None, None, None, :])
    g4 = g.reshape(ch_a.nocc, ch_a.nvir, ch_b.nocc, ch_b.nvir)
    return g4, d


@pytest.mark.parametrize("tile", ["32", "64"])
def test_opposite_spin_int8x4(pair_small, pair_mid, oracle_mod, tile, monkeypatch):
    """The opposite-spin rectangle through INT8 digit slices: equal to the energy
    of exactly the g the format forms (quant.int8x4_g_rect) to 1e-13, and within
    the format's derived bound (and 1e-8 Eh) of the oracle (u_rimp2.rs:731-760)."""
    from tests import quant

    monkeypatch.setenv("FERRIC_HIP_TILE", tile)
    for a, b in (pair_small, pair_mid):
        print("uhf opposite-spin int8x4", tile, a.nocc, a.nvir, b.nocc, b.nvir, flush=True)
        got = u.u_opposite_spin(a, b, Precision.INT8X4)
        g4, d = _rect_energy(quant.int8x4_g_rect(a.b, b.b), a, b)
        fmt = float(np.sum(g4 * g4 / d))
        assert abs(got - fmt) <= 1e-13 * abs(fmt), (got, fmt)
        ref = oracle_mod.u_opposite_spin(_oracle_ch(a), _oracle_ch(b))
        G4, _ = _rect_energy(a.b.T @ b.b, a, b)
        dl4, _ = _rect_energy(quant.int8x4_g_bound_rect(a.b, b.b), a, b)
        bound = float(np.sum((2 * np.abs(G4) * dl4 + dl4 * dl4) / np.abs(d)))
        print(f"  dE {got - ref:.2e} bound {bound:.2e}", flush=True)
        assert abs(got - ref) <= bound + 1e-14 and abs(got - ref) <= 1e-8
    da = u.SpinChannel(torch.from_numpy(a.b).cuda(), a.eps, a.nocc, a.nvir, a.first_occ)
    db = u.SpinChannel(torch.from_numpy(b.b).cuda(), b.eps, b.nocc, b.nvir, b.first_occ)
    assert u.u_opposite_spin(da, db, Precision.INT8X4) == got


def test_opposite_spin_int8x4_odd_shapes(oracle_mod):
    """One Q panel (waves 4-7 idle), one beta occupied, virtual counts off the
    16-row groups and across tiles: still the exact format's energy."""
    from tests import quant

    rng = np.random.default_rng(77)
    for naux, (na, va), (nb, vb) in ((40, (3, 5), (1, 70)), (130, (2, 67), (3, 9))):
        a = _channel(rng, naux, na, va)
        b = _channel(rng, naux, nb, vb)
        print("uhf opposite-spin int8x4 odd", naux, na, va, nb, vb, flush=True)
        got = u.u_opposite_spin(a, b, Precision.INT8X4)
        g4, d = _rect_energy(quant.int8x4_g_rect(a.b, b.b), a, b)
        fmt = float(np.sum(g4 * g4 / d))
        assert abs(got - fmt) <= 1e-13 * abs(fmt), (got, fmt)
        ref = oracle_mod.u_opposite_spin(_oracle_ch(a), _oracle_ch(b))
        assert abs(got - ref) <= 1e-8
