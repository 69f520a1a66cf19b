"""Parity of the adjoint kernel (Gamma = dE/dB, dE/deps) on the GPU (run with -m gpu).

FP64 against the closed-form golden fixtures (Gamma = 4 x_ov, dE/deps = 2 diag P,
crates/ferric-mp2/src/rimp2.rs:1832-1840, zvector.rs:651-697, oo_rimp2.rs:1192-1300),
against the oracle's restatement of the reference's intermediates, against
central finite differences of the device energy (h = 1e-5, 128 seeded entries)
at 41/209/960, and through Euler homogeneity at that size. BF16x3 against the
FP64 adjoint. The torch.autograd.Function against both.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import ferric_amd  # noqa: E402
from ferric_amd import Precision, rimp2, synthetic  # noqa: E402
from ferric_amd.autograd import ri_mp2_energy  # noqa: E402


def _adj(b, eps, nocc, nvir, fc=0, nt=None, c_os=1.0, c_ss=1.0, prec=Precision.FP64,
         device=True):
    bb = torch.from_numpy(np.ascontiguousarray(b)).cuda() if device else b
    sc, e, gam, deps = rimp2.energy_adjoint(bb, eps, nocc, nvir, fc, nt, c_os, c_ss, prec)
    if device:
        gam = gam.cpu().numpy()
    return sc, e, gam, deps


@pytest.mark.parametrize("tag,scaling", [("mp2", (1.0, 1.0)), ("scs", (1.2, 1.0 / 3.0))])
def test_fp64_adjoint_matches_golden_closed_form(golden, tag, scaling):
    for name, c in golden.items():
        if f"gamma_{tag}" not in c:
            continue
        print("adjoint golden", tag, name, flush=True)
        sc, e, gam, deps = _adj(c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"],
                                c["nocc_total"], *scaling)
        want_g, want_d = c[f"gamma_{tag}"], c[f"deps_{tag}"]
        assert np.max(np.abs(gam - want_g)) <= 1e-12 * np.max(np.abs(want_g)), name
        assert np.max(np.abs(deps - want_d)) <= 1e-12 * np.max(np.abs(want_d)), name
        assert abs(sc.e_os - c["energy"][0]) <= 1e-12 and abs(sc.e_ss - c["energy"][1]) <= 1e-12
        assert e == pytest.approx(scaling[0] * sc.e_os + scaling[1] * sc.e_ss, rel=1e-15)


@pytest.mark.parametrize("tile", ["32", "64", "128"])
def test_fp64_adjoint_matches_oracle(golden, oracle_mod, tile, monkeypatch):
    monkeypatch.setenv("FERRIC_HIP_TILE", tile)
    for name in ("lcg_120_8_40_2", "res_150_9_31_2_5", "res_300_5_17_0_9", "water_shapes"):
        print("adjoint oracle", tile, name, flush=True)
        c = golden[name]
        args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
        _, _, e_o, g_o, d_o = oracle_mod.adjoint(*args, 1.2, 0.4)
        sc, e, gam, deps = _adj(*args, 1.2, 0.4)
        assert abs(e - e_o) <= 1e-10
        assert np.max(np.abs(gam - g_o)) <= 1e-11 * np.max(np.abs(g_o)), name
        assert np.max(np.abs(deps - d_o)) <= 1e-11 * np.max(np.abs(d_o)), name
        assert np.all(deps[:c["first_occ"]] == 0.0)


def test_host_entry_matches_device_entry(golden):
    c = golden["res_120_7_23_1_3"]
    args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    d = _adj(*args, 1.0, 1.0, device=True)
    h = _adj(*args, 1.0, 1.0, device=False)
    assert np.array_equal(d[2], h[2]) and np.array_equal(d[3], h[3])


@pytest.fixture(scope="module")
def decane():
    return synthetic.decane_scale()


@pytest.fixture(scope="module")
def decane_adj(decane):
    p = decane
    return _adj(p.b_ov, p.eps, p.nocc, p.nvir, 0, p.nocc, 1.2, 1.0 / 3.0)


def test_fp64_adjoint_central_differences_decane(decane, decane_adj):
    """128 seeded entries of B and 8 orbital energies against central
    differences of the FP64 device energy, h = 1e-5."""
    p = decane
    sc, e, gam, deps = decane_adj
    t = torch.from_numpy(p.b_ov).cuda()

    def energy(bt, eps):
        s = rimp2.spin_components(bt, eps, p.nocc, p.nvir)
        return 1.2 * s.e_os + (1.0 / 3.0) * s.e_ss

    rng = np.random.default_rng(128)
    h = 1e-5
    worst = 0.0
    for k in range(128):
        q, col = int(rng.integers(p.naux)), int(rng.integers(p.nocc * p.nvir))
        old = t[q, col].item()
        t[q, col] = old + h
        ep = energy(t, p.eps)
        t[q, col] = old - h
        em = energy(t, p.eps)
        t[q, col] = old
        fd = (ep - em) / (2 * h)
        err = abs(fd - gam[q, col])
        worst = max(worst, err / max(1e-3, abs(gam[q, col])))
        assert err <= 1e-6 * max(1.0, abs(gam[q, col])), (k, q, col, fd, gam[q, col])
    print(f"worst relative FD mismatch {worst:.2e}", flush=True)
    for idx in (0, 9, 10, 40, 41, 100, 249):
        ep, em = p.eps.copy(), p.eps.copy()
        ep[idx] += h
        em[idx] -= h
        fd = (energy(t, ep) - energy(t, em)) / (2 * h)
        assert abs(fd - deps[idx]) <= 1e-6 * max(1.0, abs(deps[idx])), (idx, fd, deps[idx])


def test_adjoint_euler_identities_decane(decane, decane_adj):
    """E is homogeneous of degree 4 in B and -1 in eps; a common eps shift
    leaves it unchanged: sum B Gamma = 4E, sum eps dE/deps = -E, sum dE/deps = 0."""
    p = decane
    sc, e, gam, deps = decane_adj
    assert math.fsum((p.b_ov * gam).ravel()) == pytest.approx(4 * e, rel=1e-11)
    assert math.fsum(p.eps * deps) == pytest.approx(-e, rel=1e-11)
    assert abs(math.fsum(deps)) <= 1e-11 * math.fsum(np.abs(deps))
    energy = rimp2.spin_components(torch.from_numpy(p.b_ov).cuda(), p.eps, p.nocc, p.nvir)
    assert e == pytest.approx(1.2 * energy.e_os + energy.e_ss / 3.0, rel=1e-13)


def test_bf16x3_adjoint_against_fp64(decane, decane_adj):
    p = decane
    _, e64, g64, d64 = decane_adj
    sc, e, gam, deps = _adj(p.b_ov, p.eps, p.nocc, p.nvir, 0, p.nocc, 1.2, 1.0 / 3.0,
                            Precision.BF16X3)
    rel_g = np.max(np.abs(gam - g64)) / np.max(np.abs(g64))
    rel_d = np.max(np.abs(deps - d64)) / np.max(np.abs(d64))
    print(f"bf16x3 adjoint: dE {e - e64:.3e}, max rel dGamma {rel_g:.2e}, dEps {rel_d:.2e}",
          flush=True)
    assert abs(e - e64) <= 1e-6 * abs(e64)
    assert rel_g <= 1e-4 and rel_d <= 1e-4
    # Euler identities hold to the format's accuracy
    assert math.fsum((p.b_ov * gam).ravel()) == pytest.approx(4 * e, rel=1e-5)


def test_int8x4_adjoint_against_fp64(decane, decane_adj):
    """INT8X4 adjoint: phase A on the INT8 digit-slice kernel, the gamma GEMMs in
    BF16x3. Its energy is the INT8X4 energy (the same g, another summation
    order) and within 1e-8 Eh of FP64; dE/deps comes from the FP64 epilogue on
    that g; gamma has the BF16x3 GEMMs' accuracy."""
    p = decane
    _, e64, g64, d64 = decane_adj
    sc, e, gam, deps = _adj(p.b_ov, p.eps, p.nocc, p.nvir, 0, p.nocc, 1.2, 1.0 / 3.0,
                            Precision.INT8X4)
    en = rimp2.spin_components(torch.from_numpy(p.b_ov).cuda(), p.eps, p.nocc, p.nvir,
                               precision=Precision.INT8X4)
    rel_g = np.max(np.abs(gam - g64)) / np.max(np.abs(g64))
    rel_d = np.max(np.abs(deps - d64)) / np.max(np.abs(d64))
    print(f"int8x4 adjoint: dE {e - e64:.3e}, max rel dGamma {rel_g:.2e}, dEps {rel_d:.2e}",
          flush=True)
    assert e == pytest.approx(1.2 * en.e_os + en.e_ss / 3.0, rel=1e-12)
    assert abs(e - e64) <= 1e-8
    assert rel_d <= 1e-7
    assert rel_g <= 1e-4
    assert math.fsum((p.b_ov * gam).ravel()) == pytest.approx(4 * e, rel=1e-5)


def test_int8x4_adjoint_golden_frozen_core(golden):
    for name in ("res_150_9_31_2_5", "water_shapes"):
        c = golden[name]
        args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
        print("int8x4 adjoint", name, flush=True)
        _, e64, g64, d64 = _adj(*args, 1.0, 1.0)
        _, e8, g8, d8 = _adj(*args, 1.0, 1.0, Precision.INT8X4)
        assert abs(e8 - e64) <= 1e-10 * max(1.0, abs(e64)), name
        assert np.max(np.abs(d8 - d64)) <= 1e-7 * np.max(np.abs(d64)), name
        assert np.max(np.abs(g8 - g64)) <= 1e-4 * np.max(np.abs(g64)), name


def test_adjoint_deterministic(golden):
    c = golden["res_150_9_31_2_5"]
    args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    a = _adj(*args, 1.0, 1.0)
    b = _adj(*args, 1.0, 1.0)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])


def test_autograd_function(golden):
    c = golden["water_shapes"]
    b = torch.from_numpy(c["b"]).cuda().requires_grad_(True)
    eps = torch.from_numpy(c["eps"]).requires_grad_(True)
    e_os, e_ss = ri_mp2_energy(b, eps, c["nocc"], c["nvir"])
    cfg = rimp2.ScsMp2Config()
    e = cfg.c_os * e_os + cfg.c_ss * e_ss
    e.backward()
    want_g, want_d = c["gamma_scs"], c["deps_scs"]   # (1.2, 1/3): the SCS weights
    assert np.max(np.abs(b.grad.cpu().numpy() - want_g)) <= 1e-12 * np.max(np.abs(want_g))
    assert np.max(np.abs(eps.grad.numpy() - want_d)) <= 1e-12 * np.max(np.abs(want_d))
    assert float(e.detach()) == pytest.approx(1.2 * c["energy"][0] + c["energy"][1] / 3.0, rel=1e-12)


def test_adjoint_errors():
    b = torch.zeros((8, 6), dtype=torch.float64, device="cuda")
    with pytest.raises(ferric_amd.UnsupportedError):
        rimp2.energy_adjoint(b, np.arange(5.0) - 2.5, 2, 3, precision=Precision.FP8_E4M3)
    with pytest.raises(ferric_amd.LayoutError):
        rimp2.energy_adjoint(b, np.arange(5.0) - 2.5, 2, 3, c_os=math.inf)
    with pytest.raises(ferric_amd.LayoutError):
        rimp2.energy_adjoint(b, np.arange(3.0), 2, 3)


def _water_cases():
    import os

    w = np.load(os.path.join(os.path.dirname(__file__), "golden", "water_ccpvdz_real.npz"))
    nocc, nvir = int(w["nocc"]), int(w["nvir"])
    out = {}
    for tag, bkey, ekey in (("equilibrium", "b_ov", "eps"), ("distorted", "dist_b_ov", "dist_eps")):
        b, eps = w[bkey], w[ekey]
        out[f"{tag}"] = (b, eps, nocc, nvir, 0, nocc)
        # one frozen core orbital: drop occupied 0's columns of B
        out[f"{tag}_fc1"] = (np.ascontiguousarray(b[:, nvir:]), eps, nocc - 1, nvir, 1, nocc)
    return out


@pytest.mark.parametrize("prec", [Precision.FP64, Precision.BF16X3, Precision.INT8X4])
@pytest.mark.parametrize("case", ["equilibrium", "equilibrium_fc1", "distorted",
                                  "distorted_fc1"])
def test_adjoint_real_water(oracle_mod, prec, case):
    """The three-index density Gamma = 4 x_ov and dE/deps = 2 diag P on real
    cc-pVDZ / cc-pVDZ-RI water B (this project's integrals and RHF), all-electron
    and with one frozen core orbital, MP2 and SCS weights, against the oracle's
    restatement of the reference's intermediates (zvector.rs:651-697,
    oo_rimp2.rs:1192-1300). FP64 to 1e-11 relative; BF16x3 and INT8X4 to
    their formats' accuracy (Gamma 1e-4 relative, as the decane-scale tests)."""
    args = _water_cases()[case]
    for c_os, c_ss in ((1.0, 1.0), (1.2, 1.0 / 3.0)):
        print("water adjoint", case, prec.name, c_os, flush=True)
        _, _, e_o, g_o, d_o = oracle_mod.adjoint(*args, c_os, c_ss)
        sc, e, gam, deps = _adj(*args, c_os, c_ss, prec)
        rel_g = np.max(np.abs(gam - g_o)) / np.max(np.abs(g_o))
        rel_d = np.max(np.abs(deps - d_o)) / np.max(np.abs(d_o))
        print(f"  dE {e - e_o:+.2e}, max rel dGamma {rel_g:.2e}, dEps {rel_d:.2e}", flush=True)
        assert np.all(deps[:args[4]] == 0.0)
        if prec == Precision.FP64:
            assert abs(e - e_o) <= 1e-12
            assert rel_g <= 1e-11 and rel_d <= 1e-11
        elif prec == Precision.INT8X4:
            assert abs(e - e_o) <= 1e-10
            assert rel_g <= 1e-4 and rel_d <= 1e-7
        else:  # BF16x3: FP32 accumulation of hi/lo products, within the 1 uEh bar
            assert abs(e - e_o) <= 1e-6
            assert rel_g <= 1e-4 and rel_d <= 1e-4


@pytest.mark.parametrize("prec,tol_g", [(Precision.BF16X3, 1e-4), (Precision.INT8X4, 1e-4)])
def test_autograd_function_reduced_precision(golden, prec, tol_g):
    """ri_mp2_energy(..., precision) backward in BF16x3 and INT8X4 against the
    FP64 closed-form SCS gradients of the fixture."""
    c = golden["water_shapes"]
    b = torch.from_numpy(c["b"]).cuda().requires_grad_(True)
    eps = torch.from_numpy(c["eps"]).requires_grad_(True)
    e_os, e_ss = ri_mp2_energy(b, eps, c["nocc"], c["nvir"], precision=prec)
    e = 1.2 * e_os + e_ss / 3.0
    e.backward()
    want_g, want_d = c["gamma_scs"], c["deps_scs"]
    rel_g = np.max(np.abs(b.grad.cpu().numpy() - want_g)) / np.max(np.abs(want_g))
    rel_d = np.max(np.abs(eps.grad.numpy() - want_d)) / np.max(np.abs(want_d))
    print(f"autograd {prec.name}: rel dGamma {rel_g:.2e}, rel dEps {rel_d:.2e}", flush=True)
    assert rel_g <= tol_g and rel_d <= 1e-4
    assert float(e.detach()) == pytest.approx(1.2 * c["energy"][0] + c["energy"][1] / 3.0,
                                              rel=1e-6)


def test_autograd_function_host_tensor(golden):
    """A CPU tensor B takes the host entry (uploaded for the call); its
    gradient comes back as a CPU tensor equal to the device-tensor one."""
    c = golden["water_shapes"]
    bh = torch.from_numpy(c["b"].copy()).requires_grad_(True)
    bd = torch.from_numpy(c["b"]).cuda().requires_grad_(True)
    for bt in (bh, bd):
        e_os, e_ss = ri_mp2_energy(bt, torch.from_numpy(c["eps"]), c["nocc"], c["nvir"])
        (e_os + e_ss).backward()
    assert bh.grad.device.type == "cpu"
    assert torch.equal(bh.grad, bd.grad.cpu())
