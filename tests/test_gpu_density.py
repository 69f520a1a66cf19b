"""MP2 one-particle densities P_oo, P_vv on the GPU (run with -m gpu).

fh_rimp2_mp2_density[_dev] against the oracle's build_mp2_density
(crates/ferric-mp2/src/oo_rimp2.rs:1192-1300) and its spin-scaled form
(apply_spin_scaling / opposite_spin_density, rimp2.rs:1917-2052): on every
golden fixture with the C oracle's dense-amplitude loops, at decane scale
(41/209/960) and on real water with the BLAS restatement (itself pinned to the
loops on the CPU), to 1e-11 of the largest element; symmetry, the trace
identity tr P_oo + tr P_vv = 0, diag(P) equal to the adjoint's dE/deps / 2,
host and device entries identical, FP32 storage equal to the FP64 values it
holds.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import ferric_amd  # noqa: E402
from ferric_amd import rimp2, synthetic  # noqa: E402

SCALINGS = [(1.0, 1.0), (1.2, 1.0 / 3.0)]


def _rel(got, want):
    return float(np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-300))


def _dens(b, eps, nocc, nvir, fc=0, nt=None, c_os=1.0, c_ss=1.0, device=True):
    bb = torch.from_numpy(np.ascontiguousarray(b)).cuda() if device else b
    return rimp2.mp2_density(bb, eps, nocc, nvir, fc, nt, c_os, c_ss)


@pytest.mark.parametrize("scaling", SCALINGS)
def test_golden_cases_against_the_loop_oracle(golden, oracle_mod, scaling):
    for name, c in golden.items():
        print("density golden", name, scaling, flush=True)
        args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
        sc, p_oo, p_vv = _dens(*args, *scaling)
        t2, _ = oracle_mod.t2_and_integrals(*args)
        w_oo, w_vv = oracle_mod.mp2_density(t2, c["nocc"], c["nvir"], *scaling)
        assert _rel(p_oo, w_oo) <= 1e-11, (name, _rel(p_oo, w_oo))
        assert _rel(p_vv, w_vv) <= 1e-11, (name, _rel(p_vv, w_vv))
        assert abs(sc.e_os - c["energy"][0]) <= 1e-12 and abs(sc.e_ss - c["energy"][1]) <= 1e-12


@pytest.fixture(scope="module")
def decane():
    return synthetic.decane_scale()


@pytest.mark.parametrize("scaling", SCALINGS)
def test_decane_scale(decane, oracle_mod, scaling):
    p = decane
    print("density decane-scale", scaling, flush=True)
    sc, p_oo, p_vv = _dens(p.b_ov, p.eps, p.nocc, p.nvir, 0, p.nocc, *scaling)
    w_oo, w_vv = oracle_mod.mp2_density_blas(p.b_ov, p.eps, p.nocc, p.nvir, 0, p.nocc, *scaling)
    print(f"  rel dP_oo {_rel(p_oo, w_oo):.2e}, rel dP_vv {_rel(p_vv, w_vv):.2e}", flush=True)
    assert _rel(p_oo, w_oo) <= 1e-11 and _rel(p_vv, w_vv) <= 1e-11
    assert _rel(p_oo, p_oo.T) <= 1e-12 and _rel(p_vv, p_vv.T) <= 1e-12
    tr = np.trace(p_oo) + np.trace(p_vv)
    assert abs(tr) <= 1e-12 * np.sum(np.abs(np.diag(p_vv)))
    # diag(P) is the adjoint's orbital-energy gradient / 2
    _, _, _, deps = rimp2.energy_adjoint(torch.from_numpy(p.b_ov).cuda(), p.eps, p.nocc, p.nvir,
                                         0, p.nocc, *scaling)
    assert _rel(2 * np.diag(p_oo), deps[:p.nocc]) <= 1e-11
    assert _rel(2 * np.diag(p_vv), deps[p.nocc:]) <= 1e-11


def test_real_water_frozen_core(oracle_mod):
    w = np.load(os.path.join(os.path.dirname(__file__), "golden", "water_ccpvdz_real.npz"))
    for tag in ("", "dist_"):
        b, eps = w[f"{tag}b_ov"], w[f"{tag}eps"]
        for args in ((b, eps, 5, 19, 0, 5), (np.ascontiguousarray(b[:, 19:]), eps, 4, 19, 1, 5)):
            print("density water", tag, args[4], flush=True)
            for scaling in SCALINGS:
                _, p_oo, p_vv = _dens(*args, *scaling)
                w_oo, w_vv = oracle_mod.mp2_density_blas(*args, *scaling)
                assert _rel(p_oo, w_oo) <= 1e-11 and _rel(p_vv, w_vv) <= 1e-11


def test_host_entry_fp32_storage_and_errors(golden):
    c = golden["res_120_7_23_1_3"]
    args = (c["b"], c["eps"], c["nocc"], c["nvir"], c["first_occ"], c["nocc_total"])
    d = _dens(*args, 1.2, 0.4)
    h = _dens(*args, 1.2, 0.4, device=False)
    assert np.array_equal(d[1], h[1]) and np.array_equal(d[2], h[2])
    b32 = c["b"].astype(np.float32)
    a32 = _dens(b32, *args[1:])
    a64 = _dens(b32.astype(np.float64), *args[1:])
    assert np.array_equal(a32[1], a64[1]) and np.array_equal(a32[2], a64[2])
    with pytest.raises(ferric_amd.LayoutError):
        _dens(*args, float("nan"), 1.0)
    with pytest.raises(ferric_amd.LayoutError):
        rimp2.mp2_density(torch.zeros((8, 6), dtype=torch.float64, device="cuda"),
                          np.arange(3.0), 2, 3)
