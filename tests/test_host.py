"""Host-side logic that needs no GPU: option parsing, SCS, generators, FP8 format."""
import math

import numpy as np
import pytest

from ferric_amd import rimp2, synthetic
from ferric_amd._lib import LayoutError, Precision


def test_active_occ():
    assert rimp2.active_occ(5, 0) == 5
    assert rimp2.active_occ(5, 1) == 4
    assert rimp2.active_occ(0, 0) == 0
    with pytest.raises(LayoutError):
        rimp2.active_occ(3, 3)
    with pytest.raises(LayoutError):
        rimp2.active_occ(3, 7)


def test_settings_precision(monkeypatch):
    monkeypatch.delenv("FERRIC_GPU_PRECISION", raising=False)
    assert rimp2.settings_precision() == Precision.FP64
    for v, p in (("mixed", Precision.MIXED), ("fp32", Precision.FP32), ("BF16X3", Precision.BF16X3),
                 ("fp8", Precision.FP8_E4M3), ("f64", Precision.FP64)):
        monkeypatch.setenv("FERRIC_GPU_PRECISION", v)
        assert rimp2.settings_precision() == p
    monkeypatch.setenv("FERRIC_GPU_PRECISION", "half")
    with pytest.raises(LayoutError):
        rimp2.settings_precision()


def test_mixed_ok_rule():
    """rimp2.rs:1277 and :1293-1300: reduced precision only for a primitive
    Coulomb, erfc or terfc operator, the Coulomb fit metric and no kappa."""
    assert rimp2.mixed_ok()
    for op in ("coulomb", "erfc_coulomb", "terfc"):
        assert rimp2.mixed_energy_operator(op)
        assert not rimp2.mixed_energy_operator(op, composite=True)
    for op in ("erf_coulomb", "terf", "yukawa", "slater_geminal", "cgtg"):
        assert not rimp2.mixed_energy_operator(op)
        assert not rimp2.mixed_ok(op_kind=op)
    assert not rimp2.mixed_ok(kappa=1.3)
    assert not rimp2.mixed_ok(fit_metric="erfc")
    with pytest.raises(LayoutError):
        rimp2.mixed_ok(op_kind="coloumb")
    with pytest.raises(LayoutError):
        rimp2.mixed_ok(fit_metric="gaussian")


def test_scs():
    sc = rimp2.SpinComponents(-1.0, -0.5, -1.5)
    assert rimp2.scs_mp2(sc) == pytest.approx(6 / 5 * -1.0 + 1 / 3 * -0.5)
    assert rimp2.scs_mp2(sc, rimp2.ScsMp2Config(1.0, 1.0)) == -1.5
    with pytest.raises(LayoutError):
        rimp2.scs_mp2(sc, rimp2.ScsMp2Config(math.nan, 1.0))


def test_water_shapes_distribution():
    p = synthetic.water_shapes()
    assert p.b_ov.shape == (84, 5 * 19) and p.b_ov.dtype == np.float64
    eo, ev = p.eps_occ, p.eps_vir
    assert -20.6 <= eo[0] <= -20.4 and np.all((eo[1:] >= -1.4) & (eo[1:] <= -0.45))
    assert np.all(np.diff(eo) >= 0) and np.all(np.diff(ev) >= 0)
    assert np.all((ev >= 0.15) & (ev <= 4.0))
    assert abs(p.b_ov.std() - math.sqrt(0.09 / 84)) < 0.1 * math.sqrt(0.09 / 84)
    assert np.array_equal(p.b_ov, synthetic.water_shapes().b_ov)


def test_decane_scale_distribution():
    p = synthetic.decane_scale()
    assert p.b_ov.shape == (960, 41 * 209)
    eo = p.eps_occ
    assert np.sum(eo < -11) == 10 and np.all(eo[:10] >= -11.3) and np.all(eo[10:] <= -0.35)
    assert np.all((p.eps_vir >= 0.12) & (p.eps_vir <= 5.0))
    assert p.flops == pytest.approx(41 * 42 * 209 ** 2 * 960)


def test_mixed_large_hot_rows():
    p = synthetic.mixed_large(nocc=6, nvir=40, naux=400)
    assert p.b_ov.dtype == np.float32
    rs = np.sqrt(np.mean(p.b_ov.astype(np.float64) ** 2, axis=1))
    base = math.sqrt(0.09 / 400)
    hot = rs > 5 * base
    assert hot.sum() == 4  # 1% of 400 rows
    assert np.allclose(rs[hot] / base, 20, rtol=0.15)


def test_batched_domains_shapes():
    d = synthetic.batched_domains(64)
    assert len(d) == 64
    assert d.nocc.min() >= 2 and d.nocc.max() <= 12
    assert d.nvir.min() >= 20 and d.nvir.max() <= 240
    assert d.naux.min() >= 80 and d.naux.max() <= 720
    p = d.domain(5)
    assert p.b_ov.shape == (d.naux[5], d.nocc[5] * d.nvir[5])
    assert d.flops == pytest.approx(sum(d.domain(k).flops for k in range(64)))


def test_e4m3_encoder_is_round_to_nearest_even():
    """The numpy E4M3 restatement (tests/quant.py) the FP8 parity test relies on:
    every finite code round-trips, random values land on the nearest code."""
    from tests.quant import e4m3_decode, e4m3_encode, fp8_scale_exponents

    codes = np.arange(256, dtype=np.uint8)
    codes = codes[(codes & 0x7F) != 0x7F]
    vals = e4m3_decode(codes)
    assert np.array_equal(e4m3_decode(e4m3_encode(vals.astype(np.float32))), vals)
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(100_000) * rng.choice([1e-3, 1.0, 100.0], 100_000)).astype(np.float32)
    x = np.clip(x, -447.0, 447.0)
    pos = np.unique(np.abs(vals))
    a = np.abs(x.astype(np.float64))
    idx = np.searchsorted(pos, a)
    lo = pos[np.clip(idx - 1, 0, len(pos) - 1)]
    hi = pos[np.clip(idx, 0, len(pos) - 1)]
    dist = np.minimum(np.abs(a - lo), np.abs(hi - a))
    got = np.abs(e4m3_decode(e4m3_encode(x)))
    assert np.all(np.abs(got - a) <= dist + 1e-12)
    assert e4m3_decode(e4m3_encode(np.float32(1000.0))) == 448.0   # saturates
    e = fp8_scale_exponents(np.array([448.0, 449.0, 1.0, 0.0]))
    assert list(e) == [0, 1, -8, 0]


def test_bf16_split_residual():
    from tests.quant import bf16_split

    rng = np.random.default_rng(1)
    x = rng.standard_normal(10_000)
    hi, lo = bf16_split(x)
    x32 = x.astype(np.float32).astype(np.float64)
    assert np.all(np.abs(x32 - hi) <= 2.0 ** -8 * np.abs(x32))
    assert np.all(np.abs(x32 - hi - lo) <= 2.0 ** -16 * np.abs(x32))


def test_autotune_table_roundtrip(tmp_path):
    from ferric_amd import autotune

    path = tmp_path / "tiles.json"
    path.write_text('{"entries": [{"precision": "FP64", "nvir_max": 47, "tile": 32}]}')
    entries = autotune.load_table(str(path))
    assert entries == [{"precision": "FP64", "nvir_max": 47, "tile": 32}]
    assert autotune.load_table(str(tmp_path / "missing.json")) == []
    autotune.install(entries)            # accepted by the library (no device needed)
    autotune.install(autotune.load_table())
