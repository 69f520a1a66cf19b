"""Parity of the HIP pair-energy kernels with the oracle (run with -m gpu).

FP64: within 1e-10 Eh and within the derived element bound 2 gamma_naux(u64) S_ab
(the reference's f64 device-vs-CPU bound, crates/ferric-mp2/tests/gpu_rimp2_resident.rs)
on the golden fixtures, the decane-scale workload and per occupied row.
Reduced precision: the oracle run on exactly the values each format contracts
(FP32 rounding, BF16 hi/lo split, scaled E4M3, restated in tests/quant.py) must
match within the FP32-accumulation bound, which pins the quantisers and the
operand layouts bit-exactly; |dE| against FP64 is reported.
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import ferric_amd  # noqa: E402
from ferric_amd import Precision, rimp2, synthetic  # noqa: E402
from tests import quant  # noqa: E402

U32 = 2.0 ** -24


def _gpu(b, eps, nocc, nvir, fc=0, nt=None, kappa=None, prec=Precision.FP64, rows=False):
    t = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    return rimp2.spin_components(t, eps, nocc, nvir, fc, nt, kappa, prec, row_partials=rows)


def _bound(oracle_mod, b, eps, nocc, nvir, fc, nt, kappa=None, eps_g=None):
    return oracle_mod.energy_bound(b, eps, nocc, nvir, fc, nt, kappa=kappa, eps_g=eps_g)


def test_fp64_golden_cases(golden, oracle_mod):
    for name, c in golden.items():
        print("fp64 golden", name, flush=True)
        args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
        sc = _gpu(*args)
        bos, bss = _bound(oracle_mod, *args)
        want_os, want_ss = c["energy"]
        assert abs(sc.e_os - want_os) <= min(1e-10, bos + 1e-15), (name, sc, c["energy"])
        assert abs(sc.e_ss - want_ss) <= min(1e-10, bss + 1e-15), (name, sc, c["energy"])
        assert sc.e_total == sc.e_os + sc.e_ss
        sk = _gpu(*args, kappa=1.3)
        kos, kss = c["energy_kappa"]
        assert abs(sk.e_os - kos) <= 1e-10 and abs(sk.e_ss - kss) <= 1e-10, name


def test_fp64_host_entry_matches_device_entry(golden):
    c = golden["lcg_120_8_40_2"]
    args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    host = rimp2.spin_components(*args)
    dev = _gpu(*args)
    assert (host.e_os, host.e_ss) == (dev.e_os, dev.e_ss)


@pytest.mark.parametrize("tile", ["32", "64", "128"])
def test_fp64_every_tile_shape(golden, tile, monkeypatch):
    monkeypatch.setenv("FERRIC_HIP_TILE", tile)
    for name in ("res_150_9_31_2_5", "lcg_120_8_40_2", "water_shapes", "lcg_40_3_1_0"):
        print("tile", tile, name, flush=True)
        c = golden[name]
        sc = _gpu(c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
        assert abs(sc.e_os - c["energy"][0]) <= 1e-10, (tile, name)
        assert abs(sc.e_ss - c["energy"][1]) <= 1e-10, (tile, name)


@pytest.fixture(scope="module")
def decane():
    return synthetic.decane_scale()


@pytest.fixture(scope="module")
def decane_oracle(decane, oracle_mod):
    p = decane
    return oracle_mod.spin_components_blas(p.b_ov, p.eps, p.nocc, p.nvir, nthreads=16)


def test_fp64_decane_scale(decane, decane_oracle, oracle_mod):
    p = decane
    print("fp64 decane-scale", flush=True)
    sc, rows = _gpu(p.b_ov, p.eps, p.nocc, p.nvir, rows=True)
    print(f"E_os {sc.e_os:.15f} E_ss {sc.e_ss:.15f} oracle {decane_oracle}", flush=True)
    assert abs(sc.e_os - decane_oracle[0]) <= 1e-10
    assert abs(sc.e_ss - decane_oracle[1]) <= 1e-10
    # per-row partials against the oracle's per-i partials
    for i in (0, 9, 10, 25, 40):
        g = p.b_ov[:, i * p.nvir:(i + 1) * p.nvir].T @ p.b_ov[:, i * p.nvir:]
        os_i, ss_i = oracle_mod.i_partial(g, i, p.nocc, p.nvir, p.eps, 0, p.nocc)
        assert abs(rows[i, 0] - os_i) <= 1e-11 and abs(rows[i, 1] - ss_i) <= 1e-11, i
    # a checksum of checksums: the rows fold to the energy
    assert math.fsum(rows[:, 0]) == pytest.approx(sc.e_os, abs=1e-13)


def test_fp64_deterministic(decane):
    p = decane
    t = torch.from_numpy(p.b_ov).cuda()
    r1 = rimp2.spin_components(t, p.eps, p.nocc, p.nvir)
    r2 = rimp2.spin_components(t, p.eps, p.nocc, p.nvir)
    assert (r1.e_os, r1.e_ss) == (r2.e_os, r2.e_ss)


def test_fp32_input_in_fp64_mode_is_exact_promotion(decane):
    p = decane
    b32 = p.b_ov.astype(np.float32)
    a = _gpu(b32, p.eps, p.nocc, p.nvir)
    b = _gpu(b32.astype(np.float64), p.eps, p.nocc, p.nvir)
    assert (a.e_os, a.e_ss) == (b.e_os, b.e_ss)


@pytest.mark.parametrize("prec", [Precision.FP32, Precision.BF16X3, Precision.FP8_E4M3])
def test_reduced_precision_exact_formats(golden, oracle_mod, prec):
    """The GPU result equals the FP64 energy of the values the format
    contracts, up to FP32 accumulation."""
    for name in ("res_150_9_31_2_5", "res_120_7_23_1_3", "water_shapes", "lcg_120_8_40_2"):
        print("format", prec.name, name, flush=True)
        c = golden[name]
        nocc, nvir, fc, nt = c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"]
        b, eps = c["b"], c["eps"]
        naux = b.shape[0]
        sc = _gpu(b, eps, nocc, nvir, fc, nt, prec=prec)
        eps_g = (naux + 16) * U32
        if prec == Precision.FP32:
            bq = b.astype(np.float32).astype(np.float64)
            want = oracle_mod.spin_components(bq, eps, nocc, nvir, fc, nt)
            bnd = _bound(oracle_mod, bq, eps, nocc, nvir, fc, nt, eps_g=eps_g)
        elif prec == Precision.FP8_E4M3:
            bq = quant.fp8_dequantized(b, nocc, nvir)
            want = oracle_mod.spin_components(bq, eps, nocc, nvir, fc, nt)
            bnd = _bound(oracle_mod, bq, eps, nocc, nvir, fc, nt, eps_g=eps_g)
        else:
            hi, lo = quant.bf16_split(b)
            g = hi.T @ hi + hi.T @ lo + lo.T @ hi
            g += np.diag(np.sum(lo * lo, axis=0))   # the diagonal lo*lo correction
            want = oracle_mod.spin_components_from_g(g, eps, nocc, nvir, fc, nt)
            bnd = _bound(oracle_mod, hi + lo, eps, nocc, nvir, fc, nt, eps_g=eps_g)
        assert abs(sc.e_os - want[0]) <= bnd[0], (prec, name, sc, want, bnd)
        assert abs(sc.e_ss - want[1]) <= bnd[1], (prec, name, sc, want, bnd)
        print(f"  dE vs fp64 {sc.e_total - sum(c['energy']):.3e}", flush=True)


def test_reduced_precision_decane_scale(decane, decane_oracle, oracle_mod):
    p = decane
    ref = sum(decane_oracle[:2])
    for prec in (Precision.FP32, Precision.BF16X3, Precision.FP8_E4M3):
        print("decane", prec.name, flush=True)
        sc = _gpu(p.b_ov, p.eps, p.nocc, p.nvir, prec=prec)
        print(f"  {prec.name}: E {sc.e_total:.12f} dE {sc.e_total - ref:.3e}", flush=True)
        if prec == Precision.FP8_E4M3:
            bq = quant.fp8_dequantized(p.b_ov, p.nocc, p.nvir)
            want = oracle_mod.spin_components_blas(bq, p.eps, p.nocc, p.nvir, nthreads=16)
            bnd = _bound(oracle_mod, bq, p.eps, p.nocc, p.nvir, 0, p.nocc,
                         eps_g=(p.naux + 16) * U32)
            assert abs(sc.e_os - want[0]) <= bnd[0]
            assert abs(sc.e_ss - want[1]) <= bnd[1]
        else:
            assert abs(sc.e_total - ref) <= 1e-6 * abs(ref)


def test_errors_are_typed(decane):
    p = decane
    t = torch.from_numpy(p.b_ov).cuda()
    lib = ferric_amd.lib()
    assert lib.fh_set_pool_limit(1 << 20) == 0
    try:
        with pytest.raises(ferric_amd.PoolFullError):
            rimp2.spin_components(t, p.eps, p.nocc, p.nvir)
    finally:
        lib.fh_set_pool_limit(0)
    big = torch.full((64, 12), 1e30, dtype=torch.float64, device="cuda")
    with pytest.raises(ferric_amd.RangeError):
        rimp2.spin_components(big, np.arange(7.0) - 3.5, 3, 4, precision=Precision.FP32)
    # the f64 path takes the same operand, and the dispatcher falls back to it
    rimp2.spin_components(big, np.arange(7.0) - 3.5, 3, 4)
    os.environ["FERRIC_GPU_PRECISION"] = "bf16x3"
    try:
        sc = rimp2.spin_components_from_b_ov_kappa(big, np.arange(7.0) - 3.5, 3, 4, 0, 3)
        assert math.isfinite(sc.e_total)
    finally:
        del os.environ["FERRIC_GPU_PRECISION"]
    s = ferric_amd.stats()
    assert s["fail_pool_full"] >= 1 and s["fail_range"] >= 1


def test_real_water_against_published_energies(oracle_mod):
    """cc-pVDZ / cc-pVDZ-RI water (tests/golden/water_ccpvdz_real.npz): the
    distorted geometry against ORCA/PySCF (1e-10 Eh, all-electron and one frozen
    core), the equilibrium geometry against the reference's known answer (1e-7),
    H2 against the reference's H2 known answer (1e-4) in FP64 and INT8X4, the
    RI-JK water against its PySCF known answer (1e-10 Eh), and conventional MP2
    of the equilibrium water from an exact factor of (ia|jb) (1e-9 Eh)."""
    import os

    w = np.load(os.path.join(os.path.dirname(__file__), "golden", "water_ccpvdz_real.npz"))
    print("real water", flush=True)
    b = torch.from_numpy(w["dist_b_ov"]).cuda()
    sc = rimp2.spin_components(b, w["dist_eps"], 5, 19)
    assert abs(sc.e_os - float(w["dist_e_os_ref"])) <= 1e-10
    assert abs(sc.e_ss - float(w["dist_e_ss_ref"])) <= 1e-10
    assert abs(sc.e_total - float(w["dist_e_corr_orca"])) <= 1e-10
    fc = rimp2.spin_components(torch.from_numpy(np.ascontiguousarray(w["dist_b_ov"][:, 19:])).cuda(),
                               w["dist_eps"], 4, 19, first_occ=1, nocc_total=5)
    assert abs(fc.e_os - float(w["dist_e_os_fc_ref"])) <= 1e-10
    assert abs(fc.e_ss - float(w["dist_e_ss_fc_ref"])) <= 1e-10
    eq = rimp2.spin_components(torch.from_numpy(w["b_ov"]).cuda(), w["eps"], 5, 19)
    assert abs(eq.e_total - float(w["e_corr_reference"])) <= 1e-7
    want = oracle_mod.spin_components(w["b_ov"], w["eps"], 5, 19)
    assert abs(eq.e_total - want[2]) <= 1e-12
    # RI-JK water (PySCF DF-RHF + DF-MP2): 1e-10 Eh in FP64, INT8X4 within 1e-8
    for prec, tol in ((Precision.FP64, 1e-10), (Precision.INT8X4, 1e-8)):
        rj = rimp2.spin_components(torch.from_numpy(w["rijk_b_ov"]).cuda(), w["rijk_eps"], 5, 19,
                                   precision=prec)
        assert abs(rj.e_total - float(w["rijk_e_corr_reference"])) <= tol
    # conventional MP2 (PySCF): B^T B is the four-centre (ia|jb), 95 rows
    for prec, tol in ((Precision.FP64, 1e-9), (Precision.INT8X4, 1e-8)):
        cv = rimp2.spin_components(torch.from_numpy(w["conv_b_ov"]).cuda(), w["eps"], 5, 19,
                                   precision=prec)
        assert abs(cv.e_total - float(w["conv_e_corr_reference"])) <= tol
    # H2 / cc-pVDZ (rimp2.rs:3187-3201): within 1e-4 of canonical MP2, and the oracle
    nh, vh = int(w["h2_nocc"]), int(w["h2_nvir"])
    for prec in (Precision.FP64, Precision.INT8X4):
        h2 = rimp2.spin_components(torch.from_numpy(w["h2_b_ov"]).cuda(), w["h2_eps"], nh, vh,
                                   precision=prec)
        assert abs(h2.e_total - float(w["h2_e_corr_canonical"])) < 1e-4
        tol = 1e-12 if prec == Precision.FP64 else 1e-10  # INT8X4: 2^-32 per element
        assert abs(h2.e_total - float(w["h2_e_corr_numpy"])) <= tol
